// mlp.hip -- BF16 MFMA multilayer perceptron for the far-wake surrogate.
//
// Every product is one NT GEMM, C[M][N] = A[M][K] * B[N][K]^T with BF16
// operands (K contiguous in both) and FP32 accumulation in gfx950's
// v_mfma_f32_32x32x16_bf16: lane l feeds A row l%32 and B row l%32 with
// eight consecutive k of each 16-k step (the same k for A and B, which is
// all the sum over k needs) and receives C[8*(v/4) + 4*(l/32) + v%4][l%32] in
// accumulator v.  A workgroup computes a 64 x 64 block of C from 64-wide K
// slices staged in LDS (double buffered, one barrier per slice), each wave a
// 32 x 32 block; the epilogue applies the layer's bias and SiLU (forward) or
// SiLU's derivative (backward) and writes the BF16 activations or deltas in
// both [batch][feature] and [feature][batch] layouts, so that the forward
// pass (X W^T), the input gradient (dZ W) and the weight gradient (dZ^T X)
// all use this one kernel.  Where the hidden width is at most 256 the
// row-local passes run in one kernel per 32 rows instead, with bitwise the
// same results: the forward (mlp_forward_fused, for inference and for
// training) and the backward deltas (mlp_backward_fused); the weight
// gradients stay GEMMs, every layer's in one launch (gemm_bf16_group).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "bs_common.hpp"
#include "winds_mlp.h"

namespace winds {
namespace nn {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t bf16;

// round to nearest even (gfx950's v_cvt_pk_bf16_f32)
__device__ __forceinline__ bf16 to_bf16(float x) { return __builtin_bit_cast(bf16, (__bf16)x); }
__device__ __forceinline__ float from_bf16(bf16 h) { return __uint_as_float((uint32_t)h << 16); }

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));

// logistic function with the hardware exponential and reciprocal (~1 ulp
// each): the epilogues evaluate it for every activation, and the accurate
// library forms cost ~4x the instructions there
__device__ __forceinline__ float sigm(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

// What the GEMM epilogue does with each element x of the product.
enum Epi : int {
    kEpiF32 = 0,       // C = x
    kEpiBias = 1,      // C = x + bias[col]
    kEpiBiasSilu = 2,  // z = x + bias[col]; C = z (if C); a = SiLU(z) -> Ab, AT
    kEpiSiluBack = 3,  // d = x SiLU'(Zin); C = d; d -> Ab, AT
};

struct EpiArgs {
    float *C;
    int ldc;
    const float *bias;
    const float *Zin;
    int ldz;
    bf16 *Ab;   // [M][N] BF16 copy (row stride lda_b)
    int lda_b;
    bf16 *AT;   // [N][M] BF16 copy (row stride ldt), may be null
    int ldt;
};

constexpr int kTile = 64;            // C block of a workgroup (rows and columns)
constexpr int kSlice = 128;          // K per LDS slice
constexpr int kLdsRow = kSlice + 8;  // padded LDS row: 272 bytes, staggers the banks
constexpr int kMaxSplit = 16;        // K splits of the weight-gradient product

// C[M][N] = A[M][K] * B[N][K]^T with the epilogue EPI.  K a multiple of 32;
// lda, ldb multiples of 8; rows of A beyond M and of B beyond N are read as
// zeros and their results dropped.  256 threads.  Split K (kEpiF32 only):
// workgroup z sums k in [z kspan, (z+1) kspan) into C + z slab.
template <int EPI>
__device__ __forceinline__ void gemm_block(const bf16 *__restrict__ A, int lda,
                                           const bf16 *__restrict__ B, int ldb, int M, int N,
                                           int K, int kspan, int64_t slab, EpiArgs e, int bx,
                                           int by, int bz) {
    __shared__ __attribute__((aligned(16))) bf16 sA[2][kTile][kLdsRow];
    __shared__ __attribute__((aligned(16))) bf16 sB[2][kTile][kLdsRow];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int row0 = by * kTile, col0 = bx * kTile;
    const int kb = bz * kspan, ke = min(K, kb + kspan);
    // loader: a 64 x 128 slice of each operand is 1024 pieces of 8 BF16, four
    // per thread (piece idx: row idx / 16, k 8 (idx % 16)); K and the split
    // bounds are multiples of 32, so a piece is wholly inside or outside
    bf16x8 ra[4], rb[4];
    const bf16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int idx = t + 256 * i, r = idx >> 4, k = k0 + (idx & 15) * 8;
            ra[i] = (k < ke && row0 + r < M) ? *(const bf16x8 *)(A + (int64_t)(row0 + r) * lda + k)
                                             : zero;
            rb[i] = (k < ke && col0 + r < N) ? *(const bf16x8 *)(B + (int64_t)(col0 + r) * ldb + k)
                                             : zero;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int idx = t + 256 * i, r = idx >> 4, c = (idx & 15) * 8;
            *(bf16x8 *)&sA[buf][r][c] = ra[i];
            *(bf16x8 *)&sB[buf][r][c] = rb[i];
        }
    };
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    f32x16 acc = {0};
    load(kb);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += kSlice) {
        const bool more = k0 + kSlice < ke;
        if (more) load(k0 + kSlice);
#pragma unroll
        for (int s = 0; s < kSlice / 16; ++s) {
            const bf16x8 a = *(const bf16x8 *)&sA[buf][wr + (l & 31)][16 * s + 8 * (l >> 5)];
            const bf16x8 b = *(const bf16x8 *)&sB[buf][wc + (l & 31)][16 * s + 8 * (l >> 5)];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8v, a),
                                                          __builtin_bit_cast(bf16x8v, b), acc, 0,
                                                          0, 0);
        }
        // the other buffer was last read before the previous barrier
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    if constexpr (EPI == kEpiF32) e.C += bz * slab;
    const int col = col0 + wc + (l & 31);
    if constexpr (EPI == kEpiF32 || EPI == kEpiBias) {
        if (col >= N) return;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int row = row0 + wr + 8 * (v >> 2) + 4 * (l >> 5) + (v & 3);
            if (row >= M) continue;
            const float x = EPI == kEpiBias ? acc[v] + e.bias[col] : acc[v];
            e.C[(int64_t)row * e.ldc + col] = x;
        }
    } else {
        // FP32 results straight out; the BF16 ones staged as the 64 x 64
        // tile in LDS (the slices are no longer read) and written to both
        // copies with 16-byte stores
        bf16(*tile)[kTile + 8] = reinterpret_cast<bf16(*)[kTile + 8]>(&sA[0][0][0]);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int rl = wr + 8 * (v >> 2) + 4 * (l >> 5) + (v & 3), row = row0 + rl;
            bf16 h = 0;
            if (row < M && col < N) {
                const int64_t o = (int64_t)row * e.ldc + col;
                if constexpr (EPI == kEpiBiasSilu) {
                    const float z = acc[v] + e.bias[col];
                    if (e.C) e.C[o] = z;
                    h = to_bf16(z * sigm(z));
                } else {
                    const float z = e.Zin[(int64_t)row * e.ldz + col];
                    const float sg = sigm(z);
                    const float d = acc[v] * (sg * (1.0f + z * (1.0f - sg)));
                    e.C[o] = d;
                    h = to_bf16(d);
                }
            }
            tile[rl][wc + (l & 31)] = h;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int idx = t + 256 * i;
            // [M][N] copy: row idx / 8, columns 8 (idx % 8) .. +7
            const int r = idx >> 3, c8 = (idx & 7) * 8;
            if (row0 + r < M && col0 + c8 < N)
                *(bf16x8 *)(e.Ab + (int64_t)(row0 + r) * e.lda_b + col0 + c8) =
                    *(const bf16x8 *)&tile[r][c8];
            // [N][M] copy: column idx / 8, rows 8 (idx % 8) .. +7
            const int c = idx >> 3, r8 = (idx & 7) * 8;
            if (e.AT && col0 + c < N && row0 + r8 < M) {
                bf16x8 q;
#pragma unroll
                for (int j = 0; j < 8; ++j) q[j] = (short)tile[r8 + j][c];
                *(bf16x8 *)(e.AT + (int64_t)(col0 + c) * e.ldt + row0 + r8) = q;
            }
        }
    }
}

template <int EPI>
__global__ __launch_bounds__(256) void gemm_bf16(const bf16 *__restrict__ A, int lda,
                                                 const bf16 *__restrict__ B, int ldb, int M,
                                                 int N, int K, int kspan, int64_t slab,
                                                 EpiArgs e) {
    gemm_block<EPI>(A, lda, B, ldb, M, N, K, kspan, slab, e, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Several FP32-epilogue products in one launch (the layers' weight
// gradients): workgroup g belongs to the last problem whose first <= g and
// takes that problem's block (g - first) in x-fastest order.
struct GemmProblem {
    const bf16 *A, *B;
    float *C;
    int lda, ldb, ldc, M, N, K, kspan, gx, gy;
    int64_t slab;
    int first;
};
constexpr int kMaxProblems = 17;
struct GemmGroup {
    GemmProblem P[kMaxProblems];
    int n;
};
__global__ __launch_bounds__(256) void gemm_bf16_group(GemmGroup g) {
    int pi = 0;
    while (pi + 1 < g.n && g.P[pi + 1].first <= (int)blockIdx.x) ++pi;
    const GemmProblem &P = g.P[pi];
    const int q = (int)blockIdx.x - P.first;
    const int bx = q % P.gx, by = (q / P.gx) % P.gy, bz = q / (P.gx * P.gy);
    EpiArgs e{P.C, P.ldc, nullptr, nullptr, 0, nullptr, 0, nullptr, 0};
    gemm_block<kEpiF32>(P.A, P.lda, P.B, P.ldb, P.M, P.N, P.K, P.kspan, P.slab, e, bx, by, bz);
}

// The whole forward pass for 32 rows in one workgroup (inference): the
// rows' activations stay in LDS between layers (BF16, [32][K], K up to
// 256), each wave computes its 32-column block of the next layer from them
// with the weights' fragments read straight from device memory (the waves
// use different weight rows, so staging them would share nothing), and the
// epilogue writes bias + SiLU in BF16 back to LDS; the output layer writes
// FP32 rows of n_out.  Eight waves, one 32 x 32 block each: a small tile
// spreads a batch over many CUs, which at these sizes bounds the time
// (each layer's MFMAs and SiLU epilogue run on the CU that holds the rows).  The MFMA sequence of every block -- 16-k steps in
// increasing k over K rounded up to the 128-k slices of gemm_bf16, zero
// operands past the padded width -- and the epilogue arithmetic are those
// of gemm_bf16 with pack_input and take_cols, so the result is bitwise the
// per-layer forward's.  512 threads; hidden widths up to 256.
constexpr int kFRows = 32;
constexpr int kFMaxK = 256;
constexpr int kFLdsRow = kFMaxK + 8;

struct FusedLayer {
    const bf16 *Wb;      // [Np][Kp]
    const float *bias;   // [Np]
    int Kp, Np;
};
constexpr int kFMaxLayers = 17;
struct FusedNet {
    FusedLayer L[kFMaxLayers];
    int nl, n_in, n_out;
};
// What the training forward also leaves for the backward pass, as the
// per-layer GEMMs do: each layer's FP32 pre-activations Z [Bp][Np], the
// hidden layers' BF16 activations transposed AT [Np][Bp] and the input
// transposed XT [Kp][Bp].
struct FusedTrain {
    float *Z[kFMaxLayers];
    bf16 *AT[kFMaxLayers];
    bf16 *XT;
    int Bp;
};

// Wave w (of 8) computes column block w of layer li when it exists (the
// output layer has one).
__device__ __forceinline__ bool fused_active(const FusedNet &net, int li, int w) {
    return w < net.L[li].Np / 32;
}

constexpr int kFSteps = kFMaxK / 16;  // 16-k steps of the widest layer

constexpr int kFThreads = 512;

// rows [32] x columns [n] of an LDS tile -> out [n][ld] at column row0,
// 16-byte stores of eight rows
__device__ __forceinline__ void store_transposed(const bf16 (*tile)[kFLdsRow], int n, bf16 *out,
                                                 int ld, int row0) {
    for (int i = threadIdx.x; i < n * (kFRows / 8); i += kFThreads) {
        const int c = i / (kFRows / 8), r8 = (i % (kFRows / 8)) * 8;
        bf16x8 q;
#pragma unroll
        for (int j = 0; j < 8; ++j) q[j] = (short)tile[r8 + j][c];
        *(bf16x8 *)(out + (int64_t)c * ld + row0 + r8) = q;
    }
}

template <bool TRAIN>
__global__ __launch_bounds__(kFThreads) void mlp_forward_fused(FusedNet net, const float *__restrict__ X,
                                                         int B, float *__restrict__ Y,
                                                         FusedTrain tr) {
    __shared__ __attribute__((aligned(16))) bf16 act[2][kFRows][kFLdsRow];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int row0 = blockIdx.x * kFRows;
    const bf16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
    // a layer's weight fragments for this wave, all requested at once (one
    // device-memory latency per layer, overlapped with the previous layer's
    // epilogue)
    bf16x8 bw[kFSteps];
    auto load_w = [&](int li) {
        const FusedLayer L = net.L[li];
        const bool on = fused_active(net, li, w);
        const int K = (L.Kp + kSlice - 1) / kSlice * kSlice;
        const bf16 *row = L.Wb + (int64_t)(w * 32 + (l & 31)) * L.Kp;
#pragma unroll
        for (int s = 0; s < kFSteps; ++s) {
            const int k = 16 * s + 8 * (l >> 5);
            bw[s] = (16 * s < K && on && k < L.Kp) ? *(const bf16x8 *)(row + k) : zero;
        }
    };
    load_w(0);
    // the input rows in BF16 (they are contiguous in X: eight loads per
    // thread in flight at a time), zero past n_in up to the first layer's
    // slices and in the rows past B
    const int K0 = (net.L[0].Kp + kSlice - 1) / kSlice * kSlice;
    const int nrows = B - row0 < kFRows ? B - row0 : kFRows;
    const int nflat = nrows * net.n_in;
    const float *Xr = X + (int64_t)row0 * net.n_in;
    for (int base = 0; base < nflat; base += kFThreads * 8) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int f = base + t + kFThreads * j;
            v[j] = f < nflat ? Xr[f] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int f = base + t + kFThreads * j;
            if (f < nflat) {
                const int r = f / net.n_in;
                act[0][r][f - r * net.n_in] = to_bf16(v[j]);
            }
        }
    }
    for (int i = t; i < kFRows * K0; i += kFThreads) {
        const int r = i / K0, k = i - r * K0;
        if (r >= nrows || k >= net.n_in) act[0][r][k] = 0;
    }
    __syncthreads();
    if constexpr (TRAIN) store_transposed(act[0], net.L[0].Kp, tr.XT, tr.Bp, row0);
    for (int li = 0; li < net.nl; ++li) {
        const FusedLayer L = net.L[li];
        const bool last = li + 1 == net.nl;
        const int K = (L.Kp + kSlice - 1) / kSlice * kSlice;
        const bf16(*in)[kFLdsRow] = act[li & 1];
        bf16(*out)[kFLdsRow] = act[(li + 1) & 1];
        const bool on = fused_active(net, li, w);
        f32x16 acc = {0};
        if (on) {
#pragma unroll
            for (int s = 0; s < kFSteps; ++s) {
                if (16 * s >= K) break;
                const int k = 16 * s + 8 * (l >> 5);
                const bf16x8 a = *(const bf16x8 *)&in[l & 31][k];
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    __builtin_bit_cast(bf16x8v, a), __builtin_bit_cast(bf16x8v, bw[s]), acc, 0, 0,
                    0);
            }
        }
        if (!last) load_w(li + 1);
        if (on) {
            const int col = w * 32 + (l & 31);
            const float bias = L.bias[col];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int rl = 8 * (v >> 2) + 4 * (l >> 5) + (v & 3);
                const float z = acc[v] + bias;
                if constexpr (TRAIN) tr.Z[li][(int64_t)(row0 + rl) * L.Np + col] = z;
                if (!last) {
                    out[rl][col] = to_bf16(z * sigm(z));
                } else if (!TRAIN && col < net.n_out && row0 + rl < B) {
                    Y[(int64_t)(row0 + rl) * net.n_out + col] = z;
                }
            }
        }
        // the next layer's slices past this layer's width are zeros
        if (!last) {
            const int Kn = (L.Np + kSlice - 1) / kSlice * kSlice;
            for (int i = t; i < kFRows * (Kn - L.Np); i += kFThreads) {
                const int r = i / (Kn - L.Np), k = L.Np + i % (Kn - L.Np);
                out[r][k] = 0;
            }
        }
        __syncthreads();
        if constexpr (TRAIN)
            if (!last) store_transposed(out, L.Np, tr.AT[li], tr.Bp, row0);
    }
}

// The backward pass's deltas for 32 rows in one workgroup: the output
// layer's delta from the squared error (mse_grad's arithmetic), then layer
// by layer dZ_prev = (D W) SiLU'(Z_prev) with the deltas D held in LDS as
// BF16 and W^T's fragments read from device memory, as gemm_bf16 with
// kEpiSiluBack computes them; each layer's FP32 delta (for the bias sums)
// and its transposed BF16 copy (for the weight gradients) go to device
// memory.  The loss adds one partial per workgroup atomically.
struct FusedBack {
    const float *Z[kFMaxLayers];  // pre-activations [Bp][Np]
    float *dZ[kFMaxLayers];       // FP32 deltas [Bp][Np]
    bf16 *DT[kFMaxLayers];        // BF16 deltas transposed [Np][Bp]
    const bf16 *WT[kFMaxLayers];  // W^T [Kp][Np]
    int Np[kFMaxLayers], Kp[kFMaxLayers];
    const float *T;
    float *loss;
    int B, Bp, nl, n_out;
};

__global__ __launch_bounds__(kFThreads) void mlp_backward_fused(FusedBack fb) {
    __shared__ __attribute__((aligned(16))) bf16 dl[2][kFRows][kFLdsRow];
    __shared__ float part[kFThreads / 64];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int row0 = blockIdx.x * kFRows;
    const bf16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
    const int lo = fb.nl - 1;
    // W^T fragments of layer i for this wave's column block, all at once
    bf16x8 bw[kFSteps];
    auto load_w = [&](int i) {
        const int K = (fb.Np[i] + kSlice - 1) / kSlice * kSlice;
        const bool on = w < fb.Kp[i] / 32;
        const bf16 *row = fb.WT[i] + (int64_t)(w * 32 + (l & 31)) * fb.Np[i];
#pragma unroll
        for (int s = 0; s < kFSteps; ++s) {
            const int k = 16 * s + 8 * (l >> 5);
            bw[s] = (16 * s < K && on && k < fb.Np[i]) ? *(const bf16x8 *)(row + k) : zero;
        }
    };
    if (lo >= 1) load_w(lo);
    // the output layer's delta: 2 (y - t) / (B n_out) on real rows and
    // outputs, zero elsewhere
    {
        const int Npo = fb.Np[lo];
        const int K = (Npo + kSlice - 1) / kSlice * kSlice;
        const float sc = 1.0f / ((float)fb.B * (float)fb.n_out);
        float contrib = 0.0f;
        for (int i = t; i < kFRows * K; i += kFThreads) {
            const int r = i / K, c = i - r * K, b = row0 + r;
            if (c >= Npo) {
                dl[0][r][c] = 0;
                continue;
            }
            float g = 0.0f;
            if (b < fb.B && c < fb.n_out) {
                const float rr = fb.Z[lo][(int64_t)b * Npo + c] - fb.T[(int64_t)b * fb.n_out + c];
                g = 2.0f * rr * sc;
                contrib += rr * rr * sc;
            }
            fb.dZ[lo][(int64_t)b * Npo + c] = g;
            dl[0][r][c] = to_bf16(g);
        }
        if (fb.loss) {
            for (int off = 32; off > 0; off >>= 1) contrib += __shfl_down(contrib, off);
            if (l == 0) part[w] = contrib;
        }
        __syncthreads();
        if (fb.loss && t == 0) {
            float s = 0.0f;
            for (int q = 0; q < kFThreads / 64; ++q) s += part[q];
            if (s != 0.0f) atomicAdd(fb.loss, s);
        }
        store_transposed(dl[0], Npo, fb.DT[lo], fb.Bp, row0);
    }
    for (int i = lo; i >= 1; --i) {
        const bf16(*in)[kFLdsRow] = dl[(lo - i) & 1];
        bf16(*out)[kFLdsRow] = dl[(lo - i + 1) & 1];
        const int K = (fb.Np[i] + kSlice - 1) / kSlice * kSlice;
        const int Nout = fb.Kp[i];  // = the previous layer's Np
        const bool on = w < Nout / 32;
        f32x16 acc = {0};
        if (on) {
#pragma unroll
            for (int s = 0; s < kFSteps; ++s) {
                if (16 * s >= K) break;
                const int k = 16 * s + 8 * (l >> 5);
                const bf16x8 a = *(const bf16x8 *)&in[l & 31][k];
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    __builtin_bit_cast(bf16x8v, a), __builtin_bit_cast(bf16x8v, bw[s]), acc, 0, 0,
                    0);
            }
        }
        if (i > 1) load_w(i - 1);
        if (on) {
            const int col = w * 32 + (l & 31);
            const float *Zp = fb.Z[i - 1];
            float *dZp = fb.dZ[i - 1];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int rl = 8 * (v >> 2) + 4 * (l >> 5) + (v & 3);
                const int64_t o = (int64_t)(row0 + rl) * Nout + col;
                const float z = Zp[o];
                const float sg = sigm(z);
                const float d = acc[v] * (sg * (1.0f + z * (1.0f - sg)));
                dZp[o] = d;
                out[rl][col] = to_bf16(d);
            }
        }
        // the next product's slices past this width are zeros
        const int Kn = (Nout + kSlice - 1) / kSlice * kSlice;
        for (int q = t; q < kFRows * (Kn - Nout); q += kFThreads) {
            const int r = q / (Kn - Nout), k = Nout + q % (Kn - Nout);
            out[r][k] = 0;
        }
        __syncthreads();
        store_transposed(out, Nout, fb.DT[i - 1], fb.Bp, row0);
    }
}

// Row and column of flat index i in rows of N: 32-bit division whenever i
// fits (every batch these tables hold; the 64-bit form costs ~4x the
// instructions in these memory-bound elementwise kernels).
__device__ __forceinline__ void row_col(int64_t i, int N, int &r, int &c) {
    if (i < (int64_t)0xffffffff) {
        const uint32_t u = (uint32_t)i;
        r = (int)(u / (uint32_t)N);
        c = (int)(u - (uint32_t)r * (uint32_t)N);
    } else {
        r = (int)(i / N);
        c = (int)(i % N);
    }
}

// FP32 rows [B][n] -> BF16 [Bp][Kp] and [Kp][Bp], zero padded
__global__ void pack_input(const float *__restrict__ X, int B, int n, int Bp, int Kp,
                           bf16 *__restrict__ Xb, bf16 *__restrict__ XT) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)Bp * Kp) return;
    int b, k;
    row_col(i, Kp, b, k);
    const float x = (b < B && k < n) ? X[(int64_t)b * n + k] : 0.0f;
    const bf16 h = to_bf16(x);
    Xb[i] = h;
    if (XT) XT[(int64_t)k * Bp + b] = h;
}

// dZ = 2 (Y - T) / (B n_out) on the real rows/columns, 0 elsewhere; loss sum
__global__ void mse_grad(const float *__restrict__ Y, const float *__restrict__ T, int B, int nout,
                         int Bp, int Np, float *__restrict__ dZ, float *__restrict__ loss) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float contrib = 0.0f;
    if (i < (int64_t)Bp * Np) {
        int b, c;
        row_col(i, Np, b, c);
        float g = 0.0f;
        if (b < B && c < nout) {
            const float r = Y[i] - T[(int64_t)b * nout + c];
            const float s = 1.0f / ((float)B * (float)nout);
            g = 2.0f * r * s;
            contrib = r * r * s;
        }
        dZ[i] = g;
    }
    if (!loss) return;
    // workgroup sum, one atomic per workgroup (one per wave contended on the
    // single address: ~28 us per step at batch 4,096)
    __shared__ float part[16];
    for (int off = 32; off > 0; off >>= 1) contrib += __shfl_down(contrib, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = contrib;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
        if (t != 0.0f) atomicAdd(loss, t);
    }
}

// BF16 copies in both layouts (training), over 64 x 64 tiles through LDS so
// that both copies are written with consecutive lanes on consecutive
// addresses (per-element transposed stores touch one 2-byte word per lane
// per line).  MODE 0: input rows X [B][n] (zero padded) -> [Bp][N] and
// [N][Bp]; MODE 2: the output layer's delta dZ -> [Bp][N] and [N][Bp].  Bp
// is a multiple of 64.
template <int MODE>
__global__ __launch_bounds__(256) void tile_pack(float *__restrict__ Z, const float *__restrict__ X,
                                                 const float *__restrict__ bias, int B, int n,
                                                 int Bp, int N, bf16 *__restrict__ Ab,
                                                 bf16 *__restrict__ AT) {
    __shared__ bf16 tile[64][66];
    const int t = threadIdx.x;
    const int b0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int c = t & 63;
    const int col = n0 + c;
#pragma unroll 4
    for (int k = 0; k < 16; ++k) {
        const int r = (t >> 6) + 4 * k;
        const int row = b0 + r;
        if (col < N) {
            const int64_t i = (int64_t)row * N + col;
            float v;
            if constexpr (MODE == 0) {
                v = (row < B && col < n) ? X[(int64_t)row * n + col] : 0.0f;
            } else {
                v = Z[i];
            }
            const bf16 h = to_bf16(v);
            Ab[i] = h;
            tile[r][c] = h;
        }
    }
    __syncthreads();
    const int rr = t & 63;
#pragma unroll 4
    for (int k = 0; k < 16; ++k) {
        const int cc = (t >> 6) + 4 * k;
        if (n0 + cc < N) AT[(int64_t)(n0 + cc) * Bp + b0 + rr] = tile[rr][cc];
    }
}

// Partial column sums of dZ [Bp][N] over blocks of 64 rows: workgroup
// (x, z) sums rows 64z..64z+63 of columns 64x..64x+63 into db [z][N]; each
// thread adds 16 rows in order, the four quarters are then added in order
__device__ __forceinline__ void col_sum_block(const float *__restrict__ dZ, int N,
                                              float *__restrict__ db, int bx, int by) {
    const int c = bx * 64 + (threadIdx.x & 63), q = threadIdx.x >> 6;
    const int r0 = by * 64 + 16 * q;
    float s = 0.0f;
    if (c < N) {
        float v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = dZ[(int64_t)(r0 + j) * N + c];
#pragma unroll
        for (int j = 0; j < 16; ++j) s += v[j];
    }
    __shared__ float red[4][64];
    red[q][threadIdx.x & 63] = s;
    __syncthreads();
    if (q == 0 && c < N) {
        const int j = threadIdx.x & 63;
        db[(int64_t)by * N + c] = ((red[0][j] + red[1][j]) + red[2][j]) + red[3][j];
    }
}

// Every layer's bias-gradient partials in one launch: workgroup g belongs
// to the last layer whose first <= g.
struct ColSumLayer {
    const float *dZ;
    float *db;
    int N, gx, first;
};
struct ColSumGroup {
    ColSumLayer L[kMaxProblems];
    int n;
};
__global__ __launch_bounds__(256) void col_sum_group(ColSumGroup g) {
    int li = 0;
    while (li + 1 < g.n && g.L[li + 1].first <= (int)blockIdx.x) ++li;
    const ColSumLayer &L = g.L[li];
    const int q = (int)blockIdx.x - L.first;
    col_sum_block(L.dZ, L.N, L.db, q % L.gx, q / L.gx);
}

// Adam's per-step scalars {learning rate, 1 - b1^t, 1 - b2^t} into device
// memory, so that a replayed graph of the step reads the current ones
__global__ void set_hyper(float *__restrict__ h, float lr, float c1, float c2) {
    if (threadIdx.x == 0) {
        h[0] = lr;
        h[1] = c1;
        h[2] = c2;
    }
}

__global__ void zero_scalar(float *__restrict__ p) {
    if (threadIdx.x == 0) *p = 0.0f;
}

// Adam on one layer: weights i < N*K (then their BF16 copies W [N][K] and
// W^T [K][N]), biases after them; gradients as partial slabs (gW [S][N][K],
// gb [Sb][N]); hyp = {lr, c1, c2}
// Adam on element i of one layer's weights (i < N K) or biases, with the
// gradient summed from its partials in order (S split slabs of the weights,
// Sb row-block partials of the biases).
struct AdamLayer {
    float *W, *mW, *vW, *b, *mb, *vb;
    const float *gW, *gb;
    bf16 *Wb, *WT;
    int N, K, S, Sb;
    int64_t off;  // first element of this layer in the flat launch
};
__device__ __forceinline__ void adam_elem(const AdamLayer &L, int64_t i, float lr, float c1,
                                          float c2, float b1, float b2, float eps) {
    const int64_t nw = (int64_t)L.N * L.K;
    float *p = i < nw ? L.W : L.b;
    const float *g = i < nw ? L.gW : L.gb;
    float *m = i < nw ? L.mW : L.mb;
    float *v = i < nw ? L.vW : L.vb;
    const int64_t j = i < nw ? i : i - nw;
    const int64_t stride = i < nw ? nw : L.N;
    const int np = i < nw ? L.S : L.Sb;
    float gi = 0.0f;
    for (int z0 = 0; z0 < np; z0 += 16) {
        // sixteen loads in flight, then their sum in order
        float u[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) u[q] = z0 + q < np ? g[(int64_t)(z0 + q) * stride + j] : 0.0f;
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if (z0 + q < np) gi = z0 + q == 0 ? u[q] : gi + u[q];
    }
    const float mi = b1 * m[j] + (1.0f - b1) * gi;
    const float vi = b2 * v[j] + (1.0f - b2) * gi * gi;
    m[j] = mi;
    v[j] = vi;
    const float x = p[j] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[j] = x;
    if (i < nw) {
        const int n = (int)(i / L.K), k = (int)(i - (int64_t)n * L.K);
        const bf16 h = to_bf16(x);
        L.Wb[i] = h;
        L.WT[(int64_t)k * L.N + n] = h;
    }
}

// One Adam step over every layer in one launch: element e of the flat
// range belongs to the last layer whose off <= e.
constexpr int kMaxAdamLayers = 17;
struct AdamNet {
    AdamLayer L[kMaxAdamLayers];
    int nl;
    int64_t total;
};
__global__ void adam_all(AdamNet net, const float *__restrict__ hyp, float b1, float b2,
                         float eps) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= net.total) return;
    int li = 0;
    while (li + 1 < net.nl && net.L[li + 1].off <= e) ++li;
    adam_elem(net.L[li], e - net.L[li].off, hyp[0], hyp[1], hyp[2], b1, b2, eps);
}

// FP32 master W [N][K] -> BF16 W [N][K] and W^T [K][N]
__global__ void refresh_w(const float *__restrict__ W, int N, int K, bf16 *__restrict__ Wb,
                          bf16 *__restrict__ WT) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)N * K) return;
    const int n = (int)(i / K), k = (int)(i % K);
    const bf16 h = to_bf16(W[i]);
    Wb[i] = h;
    WT[(int64_t)k * N + n] = h;
}

__global__ void take_cols(const float *__restrict__ Y, int Np, int B, int nout,
                          float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * nout) return;
    int b, c;
    row_col(i, nout, b, c);
    out[i] = Y[(int64_t)b * Np + c];
}

}  // namespace nn
}  // namespace winds

using namespace winds::nn;

namespace {

int rup(int x, int m) { return (x + m - 1) / m * m; }
unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

struct Buf {
    void *p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t b) {
        if (p) (void)hipFree(p);
        p = nullptr;
        return winds::dev_alloc((void **)&p, b < 256 ? 256 : b);
    }
    template <class T>
    T *as() const { return (T *)p; }
};

struct Layer {
    int K = 0, N = 0;    // real sizes
    int Kp = 0, Np = 0;  // padded sizes
    Buf W, b, Wb, WT, gW, gb, mW, vW, mb, vb;
    Buf Z, Ab, AT, dZ, Db, DT;
};

}  // namespace

struct winds_mlp {
    int n_in = 0, n_out = 0, hidden = 0, nl = 0, Bmax = 0;
    std::vector<Layer> L;
    Buf Xb, XT, loss, hyp;
    int64_t step = 0;
    hipStream_t st = nullptr;
    // the training step as a graph, replayed while the batch pointers, the
    // batch size and the stream repeat
    const float *gX = nullptr, *gT = nullptr;
    int gB = 0;
    hipStream_t gst = nullptr;
    int repeats = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // inference through mlp_forward_fused (hidden widths up to 256)
    bool fused = false;
};

namespace {

using winds::set_error;

#define MLP_HIP(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_error(WINDS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <int EPI>
void gemm(hipStream_t st, const bf16 *A, int lda, const bf16 *B, int ldb, int M, int N, int K,
          const EpiArgs &e, int splits = 1, int64_t slab = 0) {
    const int kspan = splits > 1 ? rup((K + splits - 1) / splits, kSlice) : K;
    dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile, splits);
    gemm_bf16<EPI><<<grid, 256, 0, st>>>(A, lda, B, ldb, M, N, K, kspan, slab, e);
}

// K splits of the weight-gradient product over a batch of Bp rows: enough
// workgroups for the chip at 256 x 256 weights, each split at least 256 rows
int splits_of(int Bp) { return std::max(1, std::min(kMaxSplit, Bp / 256)); }

EpiArgs epi_f32(float *C, int ldc) { return EpiArgs{C, ldc, nullptr, nullptr, 0, nullptr, 0, nullptr, 0}; }

// forward with the caches training needs (keep_t: also the [feature][batch]
// layouts and the pre-activations)
int forward(winds_mlp *m, const float *X, int B, int Bp, bool keep_t, hipStream_t st) {
    Layer &l0 = m->L[0];
    if (keep_t && m->fused) {
        // the whole forward per 32 rows, leaving Z, AT and XT as below
        FusedNet net{};
        FusedTrain tr{};
        net.nl = m->nl;
        net.n_in = m->n_in;
        net.n_out = m->n_out;
        for (int i = 0; i < m->nl; ++i) {
            Layer &l = m->L[i];
            net.L[i] = FusedLayer{l.Wb.as<bf16>(), l.b.as<float>(), l.Kp, l.Np};
            tr.Z[i] = l.Z.as<float>();
            tr.AT[i] = l.AT.as<bf16>();
        }
        tr.XT = m->XT.as<bf16>();
        tr.Bp = Bp;
        mlp_forward_fused<true><<<(unsigned)(Bp / kFRows), kFThreads, 0, st>>>(net, X, B, nullptr,
                                                                              tr);
        MLP_HIP(hipGetLastError());
        return WINDS_OK;
    }
    if (keep_t)
        tile_pack<0><<<dim3((l0.Kp + 63) / 64, Bp / 64), 256, 0, st>>>(
            nullptr, X, nullptr, B, m->n_in, Bp, l0.Kp, m->Xb.as<bf16>(), m->XT.as<bf16>());
    else
        pack_input<<<nblk((int64_t)Bp * l0.Kp), 256, 0, st>>>(X, B, m->n_in, Bp, l0.Kp,
                                                              m->Xb.as<bf16>(), nullptr);
    MLP_HIP(hipGetLastError());
    const bf16 *in = m->Xb.as<bf16>();
    for (int i = 0; i < m->nl; ++i) {
        Layer &l = m->L[i];
        if (i + 1 < m->nl) {
            const EpiArgs e{keep_t ? l.Z.as<float>() : nullptr, l.Np, l.b.as<float>(), nullptr, 0,
                            l.Ab.as<bf16>(), l.Np, keep_t ? l.AT.as<bf16>() : nullptr, Bp};
            gemm<kEpiBiasSilu>(st, in, l.Kp, l.Wb.as<bf16>(), l.Kp, Bp, l.Np, l.Kp, e);
        } else {
            EpiArgs e = epi_f32(l.Z.as<float>(), l.Np);
            e.bias = l.b.as<float>();
            gemm<kEpiBias>(st, in, l.Kp, l.Wb.as<bf16>(), l.Kp, Bp, l.Np, l.Kp, e);
        }
        MLP_HIP(hipGetLastError());
        in = l.Ab.as<bf16>();
    }
    return WINDS_OK;
}

int weight_gradients(winds_mlp *m, int Bp, hipStream_t st);

int backward(winds_mlp *m, const float *T, int B, int Bp, hipStream_t st, float *loss) {
    Layer &lo = m->L[m->nl - 1];
    if (m->fused) {
        FusedBack fb{};
        for (int i = 0; i < m->nl; ++i) {
            Layer &l = m->L[i];
            fb.Z[i] = l.Z.as<float>();
            fb.dZ[i] = l.dZ.as<float>();
            fb.DT[i] = l.DT.as<bf16>();
            fb.WT[i] = l.WT.as<bf16>();
            fb.Np[i] = l.Np;
            fb.Kp[i] = l.Kp;
        }
        fb.T = T;
        fb.loss = loss;
        fb.B = B;
        fb.Bp = Bp;
        fb.nl = m->nl;
        fb.n_out = m->n_out;
        mlp_backward_fused<<<(unsigned)(Bp / kFRows), kFThreads, 0, st>>>(fb);
        MLP_HIP(hipGetLastError());
        return weight_gradients(m, Bp, st);
    }
    mse_grad<<<nblk((int64_t)Bp * lo.Np), 256, 0, st>>>(lo.Z.as<float>(), T, B, m->n_out, Bp, lo.Np,
                                                       lo.dZ.as<float>(), loss);
    MLP_HIP(hipGetLastError());
    // the output layer's delta in BF16; the hidden layers' come from the
    // input-gradient GEMM's epilogue
    tile_pack<2><<<dim3((lo.Np + 63) / 64, Bp / 64), 256, 0, st>>>(
        lo.dZ.as<float>(), nullptr, nullptr, Bp, lo.Np, Bp, lo.Np, lo.Db.as<bf16>(),
        lo.DT.as<bf16>());
    MLP_HIP(hipGetLastError());
    // the deltas, layer by layer: dZ_prev [Bp][Kp] = (dZ [Bp][Np] * W [Np][Kp])
    // SiLU'(Z_prev) (B operand W^T [Kp][Np])
    for (int i = m->nl - 1; i > 0; --i) {
        Layer &l = m->L[i];
        Layer &p = m->L[i - 1];
        const EpiArgs e{p.dZ.as<float>(), p.Np, nullptr, p.Z.as<float>(), p.Np,
                        p.Db.as<bf16>(), p.Np, p.DT.as<bf16>(), Bp};
        gemm<kEpiSiluBack>(st, l.Db.as<bf16>(), l.Np, l.WT.as<bf16>(), l.Np, Bp, l.Kp, l.Np, e);
        MLP_HIP(hipGetLastError());
    }
    return weight_gradients(m, Bp, st);
}

// Every layer's weight gradient in one launch and the bias partials in
// another, from the deltas and activations the passes left.
int weight_gradients(winds_mlp *m, int Bp, hipStream_t st) {
    // every layer's weight gradient in one launch, dW [Np][Kp] = dZ^T
    // [Np][Bp] * X^T [Kp][Bp]^T as S partial slabs over the batch
    // (S = splits_of(Bp)), and the bias gradients as one partial per 64 rows
    const int S = splits_of(Bp);
    GemmGroup gg{};
    ColSumGroup cg{};
    gg.n = cg.n = m->nl;
    int gfirst = 0, cfirst = 0;
    for (int i = 0; i < m->nl; ++i) {
        Layer &l = m->L[i];
        const bf16 *xt = i == 0 ? m->XT.as<bf16>() : m->L[i - 1].AT.as<bf16>();
        const int kspan = S > 1 ? rup((Bp + S - 1) / S, kSlice) : Bp;
        const int gx = (l.Kp + kTile - 1) / kTile, gy = (l.Np + kTile - 1) / kTile;
        gg.P[i] = GemmProblem{l.DT.as<bf16>(), xt, l.gW.as<float>(), Bp, Bp, l.Kp, l.Np, l.Kp,
                              Bp, kspan, gx, gy, (int64_t)l.Np * l.Kp, gfirst};
        gfirst += gx * gy * S;
        const int cx = (l.Np + 63) / 64;
        cg.L[i] = ColSumLayer{l.dZ.as<float>(), l.gb.as<float>(), l.Np, cx, cfirst};
        cfirst += cx * (Bp / 64);
    }
    gemm_bf16_group<<<gfirst, 256, 0, st>>>(gg);
    MLP_HIP(hipGetLastError());
    col_sum_group<<<cfirst, 256, 0, st>>>(cg);
    MLP_HIP(hipGetLastError());
    return WINDS_OK;
}

// One training step on stream st: the loss into the network's own scalar,
// forward, backward, Adam with the scalars set_hyper stored.  Launches only,
// so that it can be recorded into a graph.
int train_body(winds_mlp *m, const float *X, const float *T, int B, hipStream_t st) {
    const int Bp = rup(B, 64);
    zero_scalar<<<1, 64, 0, st>>>(m->loss.as<float>());
    int rc = forward(m, X, B, Bp, true, st);
    if (!rc) rc = backward(m, T, B, Bp, st, m->loss.as<float>());
    if (rc) return rc;
    AdamNet an{};
    an.nl = m->nl;
    int64_t off = 0;
    for (int i = 0; i < m->nl; ++i) {
        Layer &l = m->L[i];
        an.L[i] = AdamLayer{l.W.as<float>(), l.mW.as<float>(), l.vW.as<float>(), l.b.as<float>(),
                            l.mb.as<float>(), l.vb.as<float>(), l.gW.as<float>(), l.gb.as<float>(),
                            l.Wb.as<bf16>(), l.WT.as<bf16>(), l.Np, l.Kp, splits_of(Bp), Bp / 64,
                            off};
        off += (int64_t)l.Np * l.Kp + l.Np;
    }
    an.total = off;
    adam_all<<<nblk(off), 256, 0, st>>>(an, m->hyp.as<const float>(), 0.9f, 0.999f, 1e-8f);
    MLP_HIP(hipGetLastError());
    return WINDS_OK;
}

int refresh(winds_mlp *m, hipStream_t st) {
    for (Layer &l : m->L) {
        refresh_w<<<nblk((int64_t)l.Np * l.Kp), 256, 0, st>>>(l.W.as<float>(), l.Np, l.Kp,
                                                             l.Wb.as<bf16>(), l.WT.as<bf16>());
        MLP_HIP(hipGetLastError());
    }
    return WINDS_OK;
}

uint64_t sm64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

void winds_mlp_destroy(winds_mlp *m) {
    if (!m) return;
    (void)hipDeviceSynchronize();
    if (m->exec) (void)hipGraphExecDestroy(m->exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

int winds_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_hidden_layers, int32_t n_out,
                     int32_t max_batch, uint64_t seed, winds_mlp **out) {
    winds::clear_error();
    if (!out) return set_error(WINDS_E_ARG, "null output");
    *out = nullptr;
    if (n_in < 1 || n_in > 256 || n_out < 1 || n_out > 32 || n_hidden < 32 || n_hidden > 512 ||
        n_hidden % 32 || n_hidden_layers < 1 || n_hidden_layers > 16 || max_batch < 1)
        return set_error(WINDS_E_DIMS, "bad network sizes");
    winds_mlp *m = new winds_mlp();
    m->n_in = n_in;
    m->n_out = n_out;
    m->hidden = n_hidden;
    m->nl = n_hidden_layers + 1;
    m->Bmax = rup(max_batch, 64);
    m->L.resize(m->nl);
    auto fail = [&](int code) {
        winds_mlp_destroy(m);
        return code;
    };
#define MC(expr)                                                                              \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(set_error(WINDS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
    } while (0)
    MC(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
    const int Bp = m->Bmax;
    uint64_t s = seed;
    for (int i = 0; i < m->nl; ++i) {
        Layer &l = m->L[i];
        l.K = i == 0 ? n_in : n_hidden;
        l.N = i + 1 < m->nl ? n_hidden : n_out;
        l.Kp = rup(l.K, 32);
        l.Np = rup(l.N, 32);
        const size_t w = (size_t)l.Np * l.Kp * 4, b = (size_t)l.Np * 4;
        MC(l.W.alloc(w));
        MC(l.b.alloc(b));
        MC(l.Wb.alloc(w / 2));
        MC(l.WT.alloc(w / 2));
        MC(l.gW.alloc(w * splits_of(Bp)));  // split partial slabs
        MC(l.gb.alloc(b * (Bp / 64)));
        MC(l.mW.alloc(w));
        MC(l.vW.alloc(w));
        MC(l.mb.alloc(b));
        MC(l.vb.alloc(b));
        const size_t a = (size_t)Bp * l.Np;
        MC(l.Z.alloc(a * 4));
        MC(l.Ab.alloc(a * 2));
        MC(l.AT.alloc(a * 2));
        MC(l.dZ.alloc(a * 4));
        MC(l.Db.alloc(a * 2));
        MC(l.DT.alloc(a * 2));
        // He normal (Box-Muller over SplitMix64), zero bias, padding zero
        std::vector<float> hw((size_t)l.Np * l.Kp, 0.0f);
        const double sd = std::sqrt(2.0 / l.K) * (i + 1 == m->nl ? 0.1 : 1.0);
        for (int n = 0; n < l.N; ++n)
            for (int k = 0; k < l.K; ++k) {
                const double u1 = ((sm64(s) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
                const double u2 = (sm64(s) >> 11) * (1.0 / 9007199254740992.0);
                hw[(size_t)n * l.Kp + k] =
                    (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
            }
        MC(hipMemcpy(l.W.p, hw.data(), w, hipMemcpyHostToDevice));
        MC(hipMemset(l.b.p, 0, b));
        MC(hipMemset(l.mW.p, 0, w));
        MC(hipMemset(l.vW.p, 0, w));
        MC(hipMemset(l.mb.p, 0, b));
        MC(hipMemset(l.vb.p, 0, b));
    }
    MC(m->Xb.alloc((size_t)Bp * m->L[0].Kp * 2));
    MC(m->XT.alloc((size_t)Bp * m->L[0].Kp * 2));
    MC(m->loss.alloc(256));
    MC(m->hyp.alloc(256));
#undef MC
    int rc = refresh(m, m->st);
    if (!rc && hipStreamSynchronize(m->st) != hipSuccess) rc = set_error(WINDS_E_HIP, "sync");
    if (rc) return fail(rc);
    m->fused = n_hidden <= kFMaxK && m->nl <= kFMaxLayers;
    *out = m;
    return WINDS_OK;
}

int64_t winds_mlp_num_params(const winds_mlp *m) {
    if (!m) return 0;
    int64_t n = 0;
    for (const Layer &l : m->L) n += (int64_t)l.N * l.K + l.N;
    return n;
}

int winds_mlp_info(const winds_mlp *m, int64_t out[6]) {
    if (!m || !out) return set_error(WINDS_E_ARG, "null argument");
    out[0] = m->n_in;
    out[1] = m->hidden;
    out[2] = m->nl - 1;
    out[3] = m->n_out;
    out[4] = m->Bmax;
    out[5] = winds_mlp_num_params(m);
    return WINDS_OK;
}

int winds_mlp_get_params(winds_mlp *m, float *host) {
    winds::clear_error();
    if (!m || !host) return set_error(WINDS_E_ARG, "null argument");
    MLP_HIP(hipStreamSynchronize(m->st));
    for (const Layer &l : m->L) {
        std::vector<float> w((size_t)l.Np * l.Kp), b(l.Np);
        MLP_HIP(hipMemcpy(w.data(), l.W.p, w.size() * 4, hipMemcpyDeviceToHost));
        MLP_HIP(hipMemcpy(b.data(), l.b.p, b.size() * 4, hipMemcpyDeviceToHost));
        for (int n = 0; n < l.N; ++n)
            for (int k = 0; k < l.K; ++k) *host++ = w[(size_t)n * l.Kp + k];
        for (int n = 0; n < l.N; ++n) *host++ = b[n];
    }
    return WINDS_OK;
}

int winds_mlp_set_params(winds_mlp *m, const float *host) {
    winds::clear_error();
    if (!m || !host) return set_error(WINDS_E_ARG, "null argument");
    MLP_HIP(hipStreamSynchronize(m->st));
    for (Layer &l : m->L) {
        std::vector<float> w((size_t)l.Np * l.Kp, 0.0f), b(l.Np, 0.0f);
        for (int n = 0; n < l.N; ++n)
            for (int k = 0; k < l.K; ++k) w[(size_t)n * l.Kp + k] = *host++;
        for (int n = 0; n < l.N; ++n) b[n] = *host++;
        MLP_HIP(hipMemcpy(l.W.p, w.data(), w.size() * 4, hipMemcpyHostToDevice));
        MLP_HIP(hipMemcpy(l.b.p, b.data(), b.size() * 4, hipMemcpyHostToDevice));
    }
    int rc = refresh(m, m->st);
    if (rc) return rc;
    MLP_HIP(hipStreamSynchronize(m->st));
    return WINDS_OK;
}

int winds_mlp_forward(winds_mlp *m, const float *X, int32_t B, float *Y, void *stream) {
    winds::clear_error();
    if (!m) return set_error(WINDS_E_ARG, "null network");
    if (B < 0 || B > m->Bmax) return set_error(WINDS_E_DIMS, "batch larger than max_batch");
    if (B == 0) return WINDS_OK;
    if (!X || !Y) return set_error(WINDS_E_ARG, "null input/output");
    hipStream_t st = (hipStream_t)stream;  // NULL: the default stream
    if (m->fused) {
        FusedNet net{};
        net.nl = m->nl;
        net.n_in = m->n_in;
        net.n_out = m->n_out;
        for (int i = 0; i < m->nl; ++i)
            net.L[i] = FusedLayer{m->L[i].Wb.as<bf16>(), m->L[i].b.as<float>(), m->L[i].Kp,
                                  m->L[i].Np};
        mlp_forward_fused<false><<<(unsigned)((B + kFRows - 1) / kFRows), kFThreads, 0, st>>>(
            net, X, B, Y, FusedTrain{});
        MLP_HIP(hipGetLastError());
        return WINDS_OK;
    }
    const int Bp = rup(B, 64);
    int rc = forward(m, X, B, Bp, false, st);
    if (rc) return rc;
    const Layer &lo = m->L[m->nl - 1];
    take_cols<<<nblk((int64_t)B * m->n_out), 256, 0, st>>>(lo.Z.as<float>(), lo.Np, B, m->n_out, Y);
    MLP_HIP(hipGetLastError());
    return WINDS_OK;
}

int winds_mlp_set_fused(winds_mlp *m, int32_t on) {
    winds::clear_error();
    if (!m) return set_error(WINDS_E_ARG, "null network");
    if (on && !(m->hidden <= kFMaxK && m->nl <= kFMaxLayers && m->L[0].Kp <= kFMaxK))
        return set_error(WINDS_E_DIMS, "fused forward needs hidden width and inputs up to 256");
    m->fused = on != 0;
    return WINDS_OK;
}

int winds_mlp_train_step(winds_mlp *m, const float *X, const float *T, int32_t B, float lr,
                         float *loss, void *stream) {
    winds::clear_error();
    if (!m || !X || !T) return set_error(WINDS_E_ARG, "null argument");
    if (B < 1 || B > m->Bmax) return set_error(WINDS_E_DIMS, "batch out of range");
    hipStream_t st = (hipStream_t)stream;  // NULL: the default stream
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
    set_hyper<<<1, 64, 0, st>>>(m->hyp.as<float>(), lr, c1, c2);
    MLP_HIP(hipGetLastError());
    const bool same = X == m->gX && T == m->gT && B == m->gB && st == m->gst;
    if (!same) {
        if (m->exec) MLP_HIP(hipGraphExecDestroy(m->exec));
        if (m->graph) MLP_HIP(hipGraphDestroy(m->graph));
        m->exec = nullptr;
        m->graph = nullptr;
        m->gX = X;
        m->gT = T;
        m->gB = B;
        m->gst = st;
        m->repeats = 0;
    } else {
        m->repeats++;
    }
    if (!m->exec && m->repeats >= 1) {
        // the second call with the same batch buffers: record the step (on
        // the network's own stream, which capture allows; nothing runs)
        MLP_HIP(hipStreamBeginCapture(m->st, hipStreamCaptureModeThreadLocal));
        int rc = train_body(m, X, T, B, m->st);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(m->st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess)
            return set_error(WINDS_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        m->graph = g;
        MLP_HIP(hipGraphInstantiateWithFlags(&m->exec, m->graph, 0));
    }
    if (m->exec) {
        MLP_HIP(hipGraphLaunch(m->exec, st));
    } else {
        const int rc = train_body(m, X, T, B, st);
        if (rc) return rc;
    }
    if (loss)
        MLP_HIP(hipMemcpyAsync(loss, m->loss.p, sizeof(float), hipMemcpyDeviceToDevice, st));
    return WINDS_OK;
}

int winds_mlp_gradients(winds_mlp *m, const float *X, const float *T, int32_t B, float *grads,
                        float *loss) {
    winds::clear_error();
    if (!m || !X || !T || !grads) return set_error(WINDS_E_ARG, "null argument");
    if (B < 1 || B > m->Bmax) return set_error(WINDS_E_DIMS, "batch out of range");
    const int Bp = rup(B, 64);
    MLP_HIP(hipMemsetAsync(m->loss.p, 0, sizeof(float), m->st));
    int rc = forward(m, X, B, Bp, true, m->st);
    if (!rc) rc = backward(m, T, B, Bp, m->st, m->loss.as<float>());
    if (rc) return rc;
    MLP_HIP(hipStreamSynchronize(m->st));
    const int S = splits_of(Bp);
    for (const Layer &l : m->L) {
        // the split partials added in split order, as the update does
        const size_t nw = (size_t)l.Np * l.Kp;
        const int Sb = Bp / 64;
        std::vector<float> wp(nw * S), bp((size_t)l.Np * Sb), w(nw), b(l.Np);
        MLP_HIP(hipMemcpy(wp.data(), l.gW.p, wp.size() * 4, hipMemcpyDeviceToHost));
        MLP_HIP(hipMemcpy(bp.data(), l.gb.p, bp.size() * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nw; ++i) {
            float g = wp[i];
            for (int z = 1; z < S; ++z) g += wp[z * nw + i];
            w[i] = g;
        }
        for (int n = 0; n < l.Np; ++n) {
            float g = bp[n];
            for (int z = 1; z < Sb; ++z) g += bp[(size_t)z * l.Np + n];
            b[n] = g;
        }
        for (int n = 0; n < l.N; ++n)
            for (int k = 0; k < l.K; ++k) *grads++ = w[(size_t)n * l.Kp + k];
        for (int n = 0; n < l.N; ++n) *grads++ = b[n];
    }
    if (loss) MLP_HIP(hipMemcpy(loss, m->loss.p, sizeof(float), hipMemcpyDeviceToHost));
    return WINDS_OK;
}

}  // extern "C"
