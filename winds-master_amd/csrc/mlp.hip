// mlp.hip -- BF16 MFMA multilayer perceptron for the far-wake surrogate.
//
// Every product is one NT GEMM, C[M][N] = A[M][K] * B[N][K]^T with BF16
// operands (K contiguous in both) and FP32 accumulation in
// v_mfma_f32_32x32x8bf16_1k: lane l feeds A row l%32 and B row l%32 at
// k = 4*(l/32)..+3 and receives C[8*(v/4) + 4*(l/32) + v%4][l%32] in
// accumulator v.  Keeping activations and deltas in both [batch][feature]
// and [feature][batch] layouts lets the forward pass (X W^T), the input
// gradient (dZ W) and the weight gradient (dZ^T X) all use that one kernel.

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

// round to nearest even
__device__ __forceinline__ bf16 to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16 h) { return __uint_as_float((uint32_t)h << 16); }

typedef short bf16x8 __attribute__((ext_vector_type(8)));

// C[M][N] = A[M][K] * B[N][K]^T; M, N multiples of 32, K a multiple of 32,
// lda/ldb multiples of 32.  Workgroup: 4 waves, a 64x64 block of C, one
// 32x32 block per wave.  Each 32-wide K block is fed as four MFMAs from two
// 16-byte loads per lane and operand: lane l holds k in [16*(l/32), +16) of
// the block, and MFMA q takes its elements 4q..4q+3 (A and B use the same
// assignment of k to MFMA steps, which the sum over k does not see).  The
// next block's loads are issued before the current block's MFMAs.
__global__ __launch_bounds__(256) void gemm_nt(const bf16 *__restrict__ A, int lda,
                                               const bf16 *__restrict__ B, int ldb,
                                               float *__restrict__ C, int ldc, int M, int N,
                                               int K) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int tr = blockIdx.y * 64 + (w >> 1) * 32;
    const int tc = blockIdx.x * 64 + (w & 1) * 32;
    if (tr >= M || tc >= N) return;
    const bf16 *ap = A + (int64_t)(tr + (l & 31)) * lda + 16 * (l >> 5);
    const bf16 *bp = B + (int64_t)(tc + (l & 31)) * ldb + 16 * (l >> 5);
    f32x16 acc = {0};
    bf16x8 a0 = *(const bf16x8 *)ap, a1 = *(const bf16x8 *)(ap + 8);
    bf16x8 b0 = *(const bf16x8 *)bp, b1 = *(const bf16x8 *)(bp + 8);
    for (int k = 0; k < K; k += 32) {
        bf16x8 na0 = a0, na1 = a1, nb0 = b0, nb1 = b1;
        if (k + 32 < K) {
            na0 = *(const bf16x8 *)(ap + k + 32);
            na1 = *(const bf16x8 *)(ap + k + 40);
            nb0 = *(const bf16x8 *)(bp + k + 32);
            nb1 = *(const bf16x8 *)(bp + k + 40);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(__builtin_shufflevector(a0, a0, 0, 1, 2, 3),
                                                       __builtin_shufflevector(b0, b0, 0, 1, 2, 3),
                                                       acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(__builtin_shufflevector(a0, a0, 4, 5, 6, 7),
                                                       __builtin_shufflevector(b0, b0, 4, 5, 6, 7),
                                                       acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(__builtin_shufflevector(a1, a1, 0, 1, 2, 3),
                                                       __builtin_shufflevector(b1, b1, 0, 1, 2, 3),
                                                       acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(__builtin_shufflevector(a1, a1, 4, 5, 6, 7),
                                                       __builtin_shufflevector(b1, b1, 4, 5, 6, 7),
                                                       acc, 0, 0, 0);
        a0 = na0;
        a1 = na1;
        b0 = nb0;
        b1 = nb1;
    }
    const int col = tc + (l & 31);
    for (int v = 0; v < 16; ++v) {
        const int row = tr + 8 * (v >> 2) + 4 * (l >> 5) + (v & 3);
        C[(int64_t)row * ldc + col] = acc[v];
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

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

// Z += bias; hidden: A = SiLU(Z) in both layouts
__global__ void bias_act(float *__restrict__ Z, const float *__restrict__ bias, int Bp, int N,
                         int silu, bf16 *__restrict__ Ab, bf16 *__restrict__ AT) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)Bp * N) return;
    int b, n;
    row_col(i, N, b, n);
    const float z = Z[i] + bias[n];
    Z[i] = z;
    if (silu) {
        const bf16 h = to_bf16(z * sigm(z));
        Ab[i] = h;
        if (AT) AT[(int64_t)n * Bp + b] = h;
    }
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

// The same three conversions when the [column][row] copy is wanted too
// (training), over 64 x 64 tiles through LDS so that both copies are
// written with consecutive lanes on consecutive addresses (the per-element
// kernels' transposed stores touched one 2-byte word per lane per line).
// MODE 0: input rows X [B][n] (zero padded) -> [Bp][N] and [N][Bp];
// MODE 1: Z += bias, A = SiLU(Z) -> [Bp][N] and [N][Bp];
// MODE 2: delta dZ -> [Bp][N] and [N][Bp].  Bp is a multiple of 64.
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
            } else if constexpr (MODE == 1) {
                const float z = Z[i] + bias[col];
                Z[i] = z;
                v = z * sigm(z);
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

// dZ = dA * SiLU'(Z)
__global__ void silu_back(const float *__restrict__ dA, const float *__restrict__ Z, int64_t n,
                          float *__restrict__ dZ) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float z = Z[i];
    const float s = sigm(z);
    dZ[i] = dA[i] * (s * (1.0f + z * (1.0f - s)));
}

// db[n] = sum over rows of dZ[b][n]
__global__ void col_sum(const float *__restrict__ dZ, int Bp, int N, float *__restrict__ db) {
    const int n = blockIdx.x;
    float s = 0.0f;
    for (int b = threadIdx.x; b < Bp; b += blockDim.x) s += dZ[(int64_t)b * N + n];
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[n] = red[0];
}

__global__ void adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                     float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                     float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
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
        return hipMalloc(&p, b < 256 ? 256 : b);
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
    Buf Xb, XT, dA, loss;
    int64_t step = 0;
    hipStream_t st = nullptr;
};

namespace {

using winds::set_error;

#define MLP_HIP(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_error(WINDS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

void gemm(hipStream_t st, const bf16 *A, int lda, const bf16 *B, int ldb, float *C, int ldc, int M,
          int N, int K) {
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    gemm_nt<<<grid, 256, 0, st>>>(A, lda, B, ldb, C, ldc, M, N, K);
}

// forward with the caches training needs (keep_t: also the [feature][batch] layouts)
int forward(winds_mlp *m, const float *X, int B, int Bp, bool keep_t, hipStream_t st) {
    Layer &l0 = m->L[0];
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
        gemm(st, in, l.Kp, l.Wb.as<bf16>(), l.Kp, l.Z.as<float>(), l.Np, Bp, l.Np, l.Kp);
        const bool hidden = i + 1 < m->nl;
        if (hidden && keep_t)
            tile_pack<1><<<dim3((l.Np + 63) / 64, Bp / 64), 256, 0, st>>>(
                l.Z.as<float>(), nullptr, l.b.as<float>(), Bp, l.Np, Bp, l.Np, l.Ab.as<bf16>(),
                l.AT.as<bf16>());
        else
            bias_act<<<nblk((int64_t)Bp * l.Np), 256, 0, st>>>(
                l.Z.as<float>(), l.b.as<float>(), Bp, l.Np, hidden ? 1 : 0, l.Ab.as<bf16>(),
                nullptr);
        MLP_HIP(hipGetLastError());
        in = l.Ab.as<bf16>();
    }
    return WINDS_OK;
}

int backward(winds_mlp *m, const float *T, int B, int Bp, hipStream_t st, float *loss) {
    Layer &lo = m->L[m->nl - 1];
    mse_grad<<<nblk((int64_t)Bp * lo.Np), 256, 0, st>>>(lo.Z.as<float>(), T, B, m->n_out, Bp, lo.Np,
                                                       lo.dZ.as<float>(), loss);
    MLP_HIP(hipGetLastError());
    for (int i = m->nl - 1; i >= 0; --i) {
        Layer &l = m->L[i];
        tile_pack<2><<<dim3((l.Np + 63) / 64, Bp / 64), 256, 0, st>>>(
            l.dZ.as<float>(), nullptr, nullptr, Bp, l.Np, Bp, l.Np, l.Db.as<bf16>(),
            l.DT.as<bf16>());
        MLP_HIP(hipGetLastError());
        const bf16 *xt = i == 0 ? m->XT.as<bf16>() : m->L[i - 1].AT.as<bf16>();
        // dW [Np][Kp] = dZ^T [Np][Bp] * X^T [Kp][Bp]^T
        gemm(st, l.DT.as<bf16>(), Bp, xt, Bp, l.gW.as<float>(), l.Kp, l.Np, l.Kp, Bp);
        col_sum<<<l.Np, 256, 0, st>>>(l.dZ.as<float>(), Bp, l.Np, l.gb.as<float>());
        MLP_HIP(hipGetLastError());
        if (i > 0) {
            Layer &p = m->L[i - 1];
            // dA [Bp][Kp] = dZ [Bp][Np] * W [Np][Kp]  (B operand W^T [Kp][Np])
            gemm(st, l.Db.as<bf16>(), l.Np, l.WT.as<bf16>(), l.Np, m->dA.as<float>(), l.Kp, Bp,
                 l.Kp, l.Np);
            silu_back<<<nblk((int64_t)Bp * p.Np), 256, 0, st>>>(m->dA.as<float>(), p.Z.as<float>(),
                                                               (int64_t)Bp * p.Np, p.dZ.as<float>());
            MLP_HIP(hipGetLastError());
        }
    }
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
    if (m->st) {
        (void)hipStreamSynchronize(m->st);
        (void)hipStreamDestroy(m->st);
    }
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
    int maxK = 0;
    for (int i = 0; i < m->nl; ++i) {
        Layer &l = m->L[i];
        l.K = i == 0 ? n_in : n_hidden;
        l.N = i + 1 < m->nl ? n_hidden : n_out;
        l.Kp = rup(l.K, 32);
        l.Np = rup(l.N, 32);
        maxK = std::max(maxK, l.Kp);
        const size_t w = (size_t)l.Np * l.Kp * 4, b = (size_t)l.Np * 4;
        MC(l.W.alloc(w));
        MC(l.b.alloc(b));
        MC(l.Wb.alloc(w / 2));
        MC(l.WT.alloc(w / 2));
        MC(l.gW.alloc(w));
        MC(l.gb.alloc(b));
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
    MC(m->dA.alloc((size_t)Bp * maxK * 4));
    MC(m->loss.alloc(256));
#undef MC
    int rc = refresh(m, m->st);
    if (!rc && hipStreamSynchronize(m->st) != hipSuccess) rc = set_error(WINDS_E_HIP, "sync");
    if (rc) return fail(rc);
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
    const int Bp = rup(B, 64);
    int rc = forward(m, X, B, Bp, false, st);
    if (rc) return rc;
    const Layer &lo = m->L[m->nl - 1];
    take_cols<<<nblk((int64_t)B * m->n_out), 256, 0, st>>>(lo.Z.as<float>(), lo.Np, B, m->n_out, Y);
    MLP_HIP(hipGetLastError());
    return WINDS_OK;
}

int winds_mlp_train_step(winds_mlp *m, const float *X, const float *T, int32_t B, float lr,
                         float *loss, void *stream) {
    winds::clear_error();
    if (!m || !X || !T) return set_error(WINDS_E_ARG, "null argument");
    if (B < 1 || B > m->Bmax) return set_error(WINDS_E_DIMS, "batch out of range");
    hipStream_t st = (hipStream_t)stream;  // NULL: the default stream
    const int Bp = rup(B, 64);
    if (loss) MLP_HIP(hipMemsetAsync(loss, 0, sizeof(float), st));
    int rc = forward(m, X, B, Bp, true, st);
    if (!rc) rc = backward(m, T, B, Bp, st, loss);
    if (rc) return rc;
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
    for (Layer &l : m->L) {
        const int64_t nw = (int64_t)l.Np * l.Kp;
        adam<<<nblk(nw), 256, 0, st>>>(l.W.as<float>(), l.gW.as<float>(), l.mW.as<float>(),
                                       l.vW.as<float>(), nw, lr, b1, b2, 1e-8f, c1, c2);
        adam<<<nblk(l.Np), 256, 0, st>>>(l.b.as<float>(), l.gb.as<float>(), l.mb.as<float>(),
                                         l.vb.as<float>(), l.Np, lr, b1, b2, 1e-8f, c1, c2);
        MLP_HIP(hipGetLastError());
    }
    return refresh(m, st);
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
    for (const Layer &l : m->L) {
        std::vector<float> w((size_t)l.Np * l.Kp), b(l.Np);
        MLP_HIP(hipMemcpy(w.data(), l.gW.p, w.size() * 4, hipMemcpyDeviceToHost));
        MLP_HIP(hipMemcpy(b.data(), l.gb.p, b.size() * 4, hipMemcpyDeviceToHost));
        for (int n = 0; n < l.N; ++n)
            for (int k = 0; k < l.K; ++k) *grads++ = w[(size_t)n * l.Kp + k];
        for (int n = 0; n < l.N; ++n) *grads++ = b[n];
    }
    if (loss) MLP_HIP(hipMemcpy(loss, m->loss.p, sizeof(float), hipMemcpyDeviceToHost));
    return WINDS_OK;
}

}  // extern "C"
