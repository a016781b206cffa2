// bs_kernels.hip -- Biot-Savart direct-sum kernels for gfx950 (CDNA4).
//
// Work decomposition: one target per lane, every lane of a wavefront walks
// the same segment sequence, so each segment record is wave-uniform and is
// fetched with scalar loads into SGPRs (no LDS staging of records); the
// VALU only does per-pair arithmetic.  The FP64 kernel adds LDS only to
// hand terms between the waves that share 64 targets.  Segment-only terms (L, |L|^2,
// 1/|L|^2, gamma/4pi, core constant, near radius) are hoisted into a
// prepared record by prep_seg32.
//
// Kernels:
//   bs_fp64_ordered  FP64, biot_p2p arithmetic with the sequential segment
//                    order of BiotSavart_nogpu (BiotSavartMex.cu:63-93,
//                    268-282), or BiotSavart.m arithmetic (:86-115) with its
//                    nested sum order sum(sum(sum(.,1),3),4); terms of 64
//                    targets computed by many waves, added in that order by
//                    one; bitwise equal to the CPU restatements.
//   bs_fp32c         FP32 pair arithmetic with r = P-B taken in FP64 and
//                    rounded once per filament point and carried to the next
//                    segment of the filament as its r1, r1 x r2 evaluated as
//                    L x r1 (no cancellation), per-tile partial sums merged
//                    into a (hi, lo) two-sum accumulator; pairs whose target
//                    lies within the near radius of the segment are listed
//                    per wave in LDS and evaluated in FP64 with the reference
//                    arithmetic, 64 at a time.
//   bs_reduce_split  sums the per-split partials in a fixed order.
//   pack/unpack      column-major MATLAB arrays <-> device layout, in the
//                    reference's packed index k = m + n*ns + p*ns*nt.

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "bs_common.hpp"
#include "bs_fp32c.hpp"
#include "bs_kernels.hpp"
#include "bs_pair.hpp"

namespace winds {

// ---------------------------------------------------------------------------
// FP64 kernels
// ---------------------------------------------------------------------------

// FP64 direct sum in the reference's order, with the terms of one target
// computed by many waves.  A workgroup takes 64 targets (one per lane); in
// each round its W waves compute kOrdTerms consecutive segments' terms each
// into LDS, and one wave (wave r mod W in round r, so the work falls on
// every SIMD) adds the round's W * kOrdTerms terms in segment order to the
// running sums, which it hands on through LDS.  The sums follow
// BiotSavart.m's nesting sum(sum(sum(., 1), 3), 4): columns of col segments
// summed from zero, column sums added into the sum of their block of blk
// segments, block sums into the total; col = blk = nseg is the plain
// sequential sum of BiotSavart_nogpu (BiotSavartMex.cu:268-282), since a
// sum that starts at +0 never is -0 and 0 + x = x.  Every addition is the
// reference's, so results are bitwise those of one lane walking all
// segments, while W waves per 64 targets keep the FP64 pipeline fed where
// one lane per target would leave ~2 waves per SIMD.  MODE 0 biot_p2p (MEX),
// 1 BiotSavart.m cut-off, 2 Vatistas of order NF (1, 2) or n (NF = 0).
// Dynamic LDS: ord_lds_bytes(W).
constexpr int kOrdTerms = 2;
__host__ __device__ constexpr size_t ord_lds_doubles(int W) {
    return (size_t)2 * W * kOrdTerms * 3 * 64 + 9 * 64;
}
template <int MODE, int NF>
__global__ __launch_bounds__(1024) void bs_fp64_ordered(
    const double *__restrict__ tx, const double *__restrict__ ty,
    const double *__restrict__ tz, int64_t nt, const Seg64 *__restrict__ seg,
    int64_t nseg, int64_t col, int64_t blk, double d, int n, double co,
    double *__restrict__ ux, double *__restrict__ uy, double *__restrict__ uz) {
    extern __shared__ double s_ord[];
    const int W = blockDim.x >> 6;
    const int per = W * kOrdTerms;  // terms per round
    double *s_acc = s_ord + (size_t)2 * per * 3 * 64;  // [9][64] running sums
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 64 + lane;
    const bool live = t < nt;
    const double px = live ? tx[t] : 0.0, py = live ? ty[t] : 0.0, pz = live ? tz[t] : 0.0;
    if (wave == 0)
        for (int i = 0; i < 9; ++i) s_acc[i * 64 + lane] = 0.0;
    const int64_t rounds = (nseg + per - 1) / per;
    for (int64_t r = 0; r < rounds; ++r) {
        double *buf = s_ord + (size_t)(r & 1) * per * 3 * 64;  // double-buffered by round
        {
            const int64_t k0 = r * per + (int64_t)wave * kOrdTerms;
            double tt[kOrdTerms][3];
#pragma unroll
            for (int i = 0; i < kOrdTerms; ++i) {
                tt[i][0] = tt[i][1] = tt[i][2] = 0.0;
                if (k0 + i < nseg)
                    fp64_term<MODE, NF>(seg[k0 + i], px, py, pz, d, n, co, tt[i][0], tt[i][1],
                                        tt[i][2]);
            }
#pragma unroll
            for (int i = 0; i < kOrdTerms; ++i) {
                double *o = buf + (size_t)(wave * kOrdTerms + i) * 3 * 64 + lane;
                o[0] = tt[i][0];
                o[64] = tt[i][1];
                o[128] = tt[i][2];
            }
        }
        __syncthreads();
        if (wave == (int)(r % W)) {
            double cx = s_acc[lane], cy = s_acc[64 + lane], cz = s_acc[128 + lane];
            double bx = s_acc[192 + lane], by = s_acc[256 + lane], bz = s_acc[320 + lane];
            double sx = s_acc[384 + lane], sy = s_acc[448 + lane], sz = s_acc[512 + lane];
            const int64_t k0 = r * per;
            int64_t kb = k0 % blk, kc = kb % col;  // offsets in the block and the column
            const int m = nseg - k0 < per ? (int)(nseg - k0) : per;
            for (int j = 0; j < m; ++j) {
                if (kc == 0 && k0 + j > 0) {  // a column begins: close the previous one
                    bx += cx;
                    by += cy;
                    bz += cz;
                    cx = cy = cz = 0.0;
                    if (kb == 0) {  // and a block
                        sx += bx;
                        sy += by;
                        sz += bz;
                        bx = by = bz = 0.0;
                    }
                }
                const double *q = buf + (size_t)j * 3 * 64 + lane;
                cx += q[0];
                cy += q[64];
                cz += q[128];
                ++kb;
                ++kc;
                if (kb == blk)
                    kb = kc = 0;
                else if (kc == col)
                    kc = 0;
            }
            s_acc[lane] = cx;
            s_acc[64 + lane] = cy;
            s_acc[128 + lane] = cz;
            s_acc[192 + lane] = bx;
            s_acc[256 + lane] = by;
            s_acc[320 + lane] = bz;
            s_acc[384 + lane] = sx;
            s_acc[448 + lane] = sy;
            s_acc[512 + lane] = sz;
        }
    }
    __syncthreads();
    if (wave == 0 && live) {
        ux[t] = (s_acc[384 + lane] + (s_acc[192 + lane] + s_acc[lane]));
        uy[t] = (s_acc[448 + lane] + (s_acc[256 + lane] + s_acc[64 + lane]));
        uz[t] = (s_acc[512 + lane] + (s_acc[320 + lane] + s_acc[128 + lane]));
    }
}

// FP64 over one segment chunk per grid.y (split-K for few targets); the
// partial of chunk y goes to part[y][3][nt].  MODE: 0 biot_p2p (MEX),
// 1 BiotSavart.m cut-off, 2 BiotSavart.m Vatistas order NF (1, 2) or n
// (NF = 0).  Segments are summed sequentially within a chunk; chunks are
// added by bs_reduce_split.
template <int MODE, int NF>
__global__ __launch_bounds__(256) void bs_fp64_split(
    const double *__restrict__ tx, const double *__restrict__ ty,
    const double *__restrict__ tz, int64_t nt, const Seg64 *__restrict__ seg,
    int64_t nseg, int64_t chunk, double d, int n, double co, double *__restrict__ part) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int64_t k0 = (int64_t)blockIdx.y * chunk;
    const int64_t k1 = k0 + chunk < nseg ? k0 + chunk : nseg;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    fp64_range<MODE, NF>(seg, k0, k1, tx[t], ty[t], tz[t], d, n, co, sx, sy, sz);
    double *o = part + (int64_t)blockIdx.y * 3 * nt;
    o[t] = sx;
    o[nt + t] = sy;
    o[2 * nt + t] = sz;
}

// ---------------------------------------------------------------------------
// FP32 compensated kernel
// ---------------------------------------------------------------------------

__global__ void prep_seg32(const Seg64 *__restrict__ seg, int64_t nseg,
                           int64_t nseg_pad, int core_kind, int n, double d,
                           double near_factor, double near_len_frac,
                           Seg32 *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nseg_pad) return;
    out[k] = make_seg32(k < nseg ? seg + k : nullptr, k > 0 && k < nseg ? seg + k - 1 : nullptr,
                        k + 1 < nseg ? seg + k + 1 : nullptr, core_kind, n, d, near_factor,
                        near_len_frac);
}

// One launch over a single segment set: grid.x covers targets, grid.y
// splits the segment range; split partials go to out[split][3][stride].
// 8 waves per SIMD (at most 64 VGPRs; 12-48 bytes per lane spill, stored
// and reloaded once per 16-record tile, outside the pair loop): 5-6% faster
// than 7 waves at 72 VGPRs; the general Vatistas order, 4 waves before its
// near pairs dropped pow, 45% faster.
template <int CORE, int SEM, bool CO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void bs_fp32c(
    const double *__restrict__ tx, const double *__restrict__ ty,
    const double *__restrict__ tz, int64_t nt, const Seg32 *__restrict__ s32,
    const Seg64 *__restrict__ s64, int64_t nseg_pad, int64_t chunk, Fp32Params P,
    double *__restrict__ out, float *__restrict__ est, int64_t out_stride) {
    extern __shared__ NearLds near_lds[];  // one per wave
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // whole waves past the end leave; the lanes past the end of a partial
    // wave take part in the wave's near-pair evaluation and store nothing
    if (t - (int64_t)(threadIdx.x & 63u) >= nt) return;
    const bool active = t < nt;
    const int64_t tc = active ? t : nt - 1;
    const int64_t k0 = (int64_t)blockIdx.y * chunk;
    const int64_t k1 = k0 + chunk < nseg_pad ? k0 + chunk : nseg_pad;
    double sx, sy, sz;
    float sa;
    fp32c_range<CORE, SEM, CO>(tx[tc], ty[tc], tz[tc], s32, s64, 0, k0, k1, P,
                               near_lds + (threadIdx.x >> 6), sx, sy, sz, sa);
    if (!active) return;
    double *o = out + (int64_t)blockIdx.y * 3 * out_stride;
    o[t] = sx;
    o[out_stride + t] = sy;
    o[2 * out_stride + t] = sz;
    if (est) est[(int64_t)blockIdx.y * out_stride + t] = sa;
}

// Split partials of the FP32 kernel summed in split order, as
// bs_reduce_split, plus the guard's test: a target whose FP32 terms cancel
// (sqrt(a2) > kappa |u|, or a result that is not finite) is appended to the
// re-evaluation list.  The list order depends on scheduling; each listed
// target's FP64 sum does not.
__global__ void bs_reduce_guard(const double *__restrict__ part, const float *__restrict__ est,
                                int64_t nt, int split, double kappa, double *__restrict__ ux,
                                double *__restrict__ uy, double *__restrict__ uz,
                                uint32_t *__restrict__ list, int *__restrict__ count) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double sx = 0.0, sy = 0.0, sz = 0.0, a2 = 0.0;
    int s = 0;
    // eight partials' loads issued together, then added in order
    for (; s + 8 <= split; s += 8) {
        double v[8][3];
        float e[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const double *p = part + (int64_t)(s + i) * 3 * nt;
            v[i][0] = p[t];
            v[i][1] = p[nt + t];
            v[i][2] = p[2 * nt + t];
            e[i] = est[(int64_t)(s + i) * nt + t];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            sx += v[i][0];
            sy += v[i][1];
            sz += v[i][2];
            a2 += (double)e[i];
        }
    }
    for (; s < split; ++s) {
        const double *p = part + (int64_t)s * 3 * nt;
        sx += p[t];
        sy += p[nt + t];
        sz += p[2 * nt + t];
        a2 += (double)est[(int64_t)s * nt + t];
    }
    ux[t] = sx;
    uy[t] = sy;
    uz[t] = sz;
    if (!(sqrt(a2) <= kappa * sqrt(sx * sx + sy * sy + sz * sz)))
        list[atomicAdd(count, 1)] = (uint32_t)t;
}

// ---------------------------------------------------------------------------
// Guard: FP64 re-evaluation of listed targets
// ---------------------------------------------------------------------------

__device__ __forceinline__ int find_span(const WakeSpan *wk, int n, int64_t t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (wk[mid].t_off <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Chunk c of kFixSplit of the wake's records at target (px, py, pz): each
// of the 256 threads sums a contiguous sub-range in segment order, then a
// fixed tree adds the threads' sums; thread 0 returns the chunk's sum.
template <int MODE, int NF>
__device__ double3 fix_chunk(const Seg64 *__restrict__ seg, int64_t ns, int c, double px,
                             double py, double pz, double d, int n, double co,
                             double (*red)[256]) {
    const int64_t k0 = ns * c / kFixSplit, k1 = ns * (c + 1) / kFixSplit;
    const int64_t len = k1 - k0;
    const int j = (int)threadIdx.x;
    const int64_t a = k0 + len * j / 256, b = k0 + len * (j + 1) / 256;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    fp64_range<MODE, NF>(seg, a, b, px, py, pz, d, n, co, sx, sy, sz);
    __syncthreads();
    red[0][j] = sx;
    red[1][j] = sy;
    red[2][j] = sz;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (j < h) {
            red[0][j] += red[0][j + h];
            red[1][j] += red[1][j + h];
            red[2][j] += red[2][j + h];
        }
        __syncthreads();
    }
    return make_double3(red[0][0], red[1][0], red[2][0]);
}

// grid = (kFixSplit, G), 256 threads.  List entry i < kFixCap: workgroup
// (c, g) writes chunk c's sum to fix[i][c]; bs_guard_fix_reduce adds them.
// Entries past kFixCap: workgroup (0, g) adds the chunks itself, in the
// same order, so a target's result does not depend on its list position.
template <int MODE, int NF>
__global__ __launch_bounds__(256) void bs_guard_fix(
    const double *__restrict__ tx, const double *__restrict__ ty, const double *__restrict__ tz,
    const Seg64 *__restrict__ seg, const WakeSpan *__restrict__ wk, int n_wk, WakeSpan one,
    double d, int n, double co, const uint32_t *__restrict__ list, const int *__restrict__ count,
    double *__restrict__ fix, double *__restrict__ ux, double *__restrict__ uy,
    double *__restrict__ uz) {
    __shared__ double red[3][256];
    const int64_t cnt = *count;
    const int c = (int)blockIdx.x;
    for (int64_t i = blockIdx.y; i < cnt; i += gridDim.y) {
        const int64_t t = list[i];
        const WakeSpan w = wk ? wk[find_span(wk, n_wk, t)] : one;
        const Seg64 *s = seg + w.s_off;
        const double px = tx[t], py = ty[t], pz = tz[t];
        if (i < kFixCap) {
            const double3 r = fix_chunk<MODE, NF>(s, w.ns, c, px, py, pz, d, n, co, red);
            if (threadIdx.x == 0) {
                double *o = fix + (i * kFixSplit + c) * 3;
                o[0] = r.x;
                o[1] = r.y;
                o[2] = r.z;
            }
        } else if (c == 0) {
            double sx = 0.0, sy = 0.0, sz = 0.0;
            for (int q = 0; q < kFixSplit; ++q) {
                const double3 r = fix_chunk<MODE, NF>(s, w.ns, q, px, py, pz, d, n, co, red);
                sx += r.x;
                sy += r.y;
                sz += r.z;
            }
            if (threadIdx.x == 0) {
                ux[t] = sx;
                uy[t] = sy;
                uz[t] = sz;
            }
        }
    }
}

__global__ void bs_guard_fix_reduce(const uint32_t *__restrict__ list,
                                    const int *__restrict__ count,
                                    const double *__restrict__ fix, double *__restrict__ ux,
                                    double *__restrict__ uy, double *__restrict__ uz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t cnt = *count;
    if (i >= cnt || i >= kFixCap) return;
    const int64_t t = list[i];
    const double *o = fix + i * kFixSplit * 3;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int q = 0; q < kFixSplit; ++q, o += 3) {
        sx += o[0];
        sy += o[1];
        sz += o[2];
    }
    ux[t] = sx;
    uy[t] = sy;
    uz[t] = sz;
}

__global__ void guard_reset(int *count) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *count = 0;
}

__global__ void bs_reduce_split(const double *__restrict__ part, int64_t nt,
                                int split, double *__restrict__ ux,
                                double *__restrict__ uy, double *__restrict__ uz) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    int s = 0;
    // eight partials' loads issued together, then added in order
    for (; s + 8 <= split; s += 8) {
        double v[8][3];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const double *p = part + (int64_t)(s + i) * 3 * nt;
            v[i][0] = p[t];
            v[i][1] = p[nt + t];
            v[i][2] = p[2 * nt + t];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            sx += v[i][0];
            sy += v[i][1];
            sz += v[i][2];
        }
    }
    for (; s < split; ++s) {
        const double *p = part + (int64_t)s * 3 * nt;
        sx += p[t];
        sy += p[nt + t];
        sz += p[2 * nt + t];
    }
    ux[t] = sx;
    uy[t] = sy;
    uz[t] = sz;
}

// The same sums (and, with est, bs_reduce_guard's test) for few targets and
// many partials: one wave per target stages the target's partials in LDS
// with all its lanes, then lanes 0-3 add the components and the estimate in
// split order -- the same additions as one thread per target, without a
// chain of dependent device-memory loads.  Dynamic LDS: 4 * split doubles.
__global__ __launch_bounds__(64) void bs_reduce_wave(const double *__restrict__ part,
                                                     const float *__restrict__ est, int64_t nt,
                                                     int split, double kappa,
                                                     double *__restrict__ ux,
                                                     double *__restrict__ uy,
                                                     double *__restrict__ uz,
                                                     uint32_t *__restrict__ list,
                                                     int *__restrict__ count) {
    extern __shared__ double s_red[];  // [4][split]
    const int64_t t = blockIdx.x;
    const int lane = threadIdx.x;
    for (int s = lane; s < split; s += 64) {
        const double *p = part + (int64_t)s * 3 * nt + t;
        s_red[s] = p[0];
        s_red[split + s] = p[nt];
        s_red[2 * split + s] = p[2 * nt];
        s_red[3 * split + s] = est ? (double)est[(int64_t)s * nt + t] : 0.0;
    }
    __syncthreads();
    double acc = 0.0;
    if (lane < 4) {
        const double *q = s_red + (int64_t)lane * split;
        for (int s = 0; s < split; ++s) acc += q[s];
    }
    // lanes 0-2 hold the components, lane 3 the sum of squares
    const double sx = __shfl(acc, 0), sy = __shfl(acc, 1), sz = __shfl(acc, 2),
                 a2 = __shfl(acc, 3);
    if (lane == 0) {
        ux[t] = sx;
        uy[t] = sy;
        uz[t] = sz;
        if (est && !(sqrt(a2) <= kappa * sqrt(sx * sx + sy * sy + sz * sz)))
            list[atomicAdd(count, 1)] = (uint32_t)t;
    }
}

// ---------------------------------------------------------------------------
// Layout kernels
// ---------------------------------------------------------------------------

__global__ void pack_segments_cm(const double *__restrict__ F1,
                                 const double *__restrict__ F2, int64_t ns,
                                 int64_t ntm, int64_t nb,
                                 const double *__restrict__ gamma,
                                 const double *__restrict__ rc,
                                 Seg64 *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nf = ns * ntm * nb;
    if (k >= nf) return;
    const int64_t m = k % ns, n = (k / ns) % ntm, p = k / (ns * ntm);
    const int64_t base = m + 3 * ns * n + 3 * ns * ntm * p;
    Seg64 s;
    s.ax = F1[base];
    s.ay = F1[base + ns];
    s.az = F1[base + 2 * ns];
    s.bx = F2[base];
    s.by = F2[base + ns];
    s.bz = F2[base + 2 * ns];
    s.g = gamma[k];
    s.rc = rc ? rc[k] : 0.0;
    out[k] = s;
}

__global__ void pack_targets_cm(const double *__restrict__ P, int64_t ns,
                                int64_t ntm, int64_t nb, double *__restrict__ t3) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np = ns * ntm * nb;
    if (j >= np) return;
    const int64_t m = j % ns, n = (j / ns) % ntm, p = j / (ns * ntm);
    const int64_t base = m + 3 * ns * n + 3 * ns * ntm * p;
    t3[j] = P[base];
    t3[np + j] = P[base + ns];
    t3[2 * np + j] = P[base + 2 * ns];
}

__global__ void unpack_uind_cm(const double *__restrict__ u3, int64_t ns,
                               int64_t ntm, int64_t nb, double *__restrict__ U) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np = ns * ntm * nb;
    if (j >= np) return;
    const int64_t m = j % ns, n = (j / ns) % ntm, p = j / (ns * ntm);
    const int64_t base = m + 3 * ns * n + 3 * ns * ntm * p;
    U[base] = u3[j];
    U[base + ns] = u3[np + j];
    U[base + 2 * ns] = u3[2 * np + j];
}

__global__ void segment_length(const Seg64 *__restrict__ seg, int64_t nseg,
                               int squared, double *__restrict__ L) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nseg) return;
    const Seg64 s = seg[k];
    const double x2x1 = s.bx - s.ax, y2y1 = s.by - s.ay, z2z1 = s.bz - s.az;
    const double l2 = x2x1 * x2x1 + y2y1 * y2y1 + z2z1 * z2z1;
    L[k] = squared ? l2 : sqrt(l2);
}

__global__ void fill_zero(double *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline unsigned grid_for(int64_t n, int block) {
    return (unsigned)((n + block - 1) / block);
}

// Split reduction by the kernel that suits the shape: one wave per target
// while the targets are few (16 waves' worth of threads, one per target,
// would leave the chip idle) and the partials fit in LDS, one thread per
// target otherwise; the sums are the same either way.
static hipError_t reduce_split(const double *part, const float *est, int64_t nt, int split,
                               double kappa, double *u3, uint32_t *list, int *count,
                               hipStream_t st) {
    if (nt < 1024 && split <= 2048) {
        bs_reduce_wave<<<(unsigned)nt, 64, (size_t)4 * split * sizeof(double), st>>>(
            part, est, nt, split, kappa, u3, u3 + nt, u3 + 2 * nt, list, count);
    } else if (est) {
        bs_reduce_guard<<<grid_for(nt, 256), 256, 0, st>>>(part, est, nt, split, kappa, u3,
                                                           u3 + nt, u3 + 2 * nt, list, count);
    } else {
        bs_reduce_split<<<grid_for(nt, 256), 256, 0, st>>>(part, nt, split, u3, u3 + nt,
                                                           u3 + 2 * nt);
    }
    return hipGetLastError();
}

hipError_t launch_zero(double *p, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    fill_zero<<<grid_for(n, 256), 256, 0, st>>>(p, n);
    return hipGetLastError();
}

// Lets bs_fp64_ordered<MODE, NF> take its largest dynamic LDS (once per
// process; thread-safe static initialisation).
template <int MODE, int NF>
static hipError_t ord_allow_lds() {
    static const hipError_t rc =
        hipFuncSetAttribute((const void *)bs_fp64_ordered<MODE, NF>,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(ord_lds_doubles(16) * sizeof(double)));
    return rc;
}

// Waves per workgroup of bs_fp64_ordered.  16 while the target groups are
// fewer than two per CU (more waves per group fill the chip); 8 above, where
// two workgroups share a CU's LDS and keep four waves per SIMD (measured on
// the 3 x 31-station wakes: 1,440 ages 245 ms at 8 against 341 ms at 16,
// 72 ages 1.85 ms at 16 against 2.28 ms at 8, Vatistas-2).
static int ordered_waves(int64_t groups) {
    const int64_t cus = device_cu_count() > 0 ? device_cu_count() : 256;
    return groups < 2 * cus ? 16 : 8;
}

hipError_t launch_fp64(const double *t3, int64_t nt, const Seg64 *seg,
                       int64_t nseg, const EvalCfg &c, double *u3, hipStream_t st,
                       double *part) {
    const int block = c.block > 0 ? c.block : 256;
    const int block256 = block < 256 ? block : 256;  // kernels bounded to 256 threads
    if (nt <= 0) return hipSuccess;
    if (nseg <= 0) return launch_zero(u3, 3 * nt, st);
    const double *tx = t3, *ty = t3 + nt, *tz = t3 + 2 * nt;
    double *ux = u3, *uy = u3 + nt, *uz = u3 + 2 * nt;
    if (c.split > 1 && part) {
        // segment range split over grid.y, partials reduced in split order
        const int64_t chunk = (nseg + c.split - 1) / c.split;
        const int nsplit = (int)((nseg + chunk - 1) / chunk);
        dim3 grid(grid_for(nt, block256), nsplit);
        const int n = c.core == 0 ? 0 : c.vat_n;
#define WINDS_SPLIT(M, F)                                                                   \
    bs_fp64_split<M, F><<<grid, block256, 0, st>>>(tx, ty, tz, nt, seg, nseg, chunk, c.d, n,   \
                                                   c.co, part)
        if (c.semantics == 0)
            WINDS_SPLIT(0, 0);
        else if (n == 0)
            WINDS_SPLIT(1, 0);
        else if (n == 1)
            WINDS_SPLIT(2, 1);
        else if (n == 2)
            WINDS_SPLIT(2, 2);
        else
            WINDS_SPLIT(2, 0);
#undef WINDS_SPLIT
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        return reduce_split(part, nullptr, nt, nsplit, 0.0, u3, nullptr, nullptr, st);
    }
    {
        const bool mex = c.semantics == 0;
        const int64_t col = !mex && c.seg_col > 0 ? c.seg_col : nseg;
        const int64_t blk = !mex && c.seg_blk > 0 ? c.seg_blk : nseg;
        const int n = c.core == 0 ? 0 : c.vat_n;
        const int64_t groups = (nt + 63) / 64;
        const int W = ordered_waves(groups);
        const size_t lds = ord_lds_doubles(W) * sizeof(double);
#define WINDS_ORD(M, F)                                                                         \
    do {                                                                                        \
        hipError_t ea = ord_allow_lds<M, F>();                                                  \
        if (ea != hipSuccess) return ea;                                                        \
        bs_fp64_ordered<M, F><<<(unsigned)groups, 64 * W, lds, st>>>(                           \
            tx, ty, tz, nt, seg, nseg, col, blk, c.d, n, c.co, ux, uy, uz);                     \
    } while (0)
        if (mex)
            WINDS_ORD(0, 0);
        else if (n == 0)
            WINDS_ORD(1, 0);
        else if (n == 1)
            WINDS_ORD(2, 1);
        else if (n == 2)
            WINDS_ORD(2, 2);
        else
            WINDS_ORD(2, 0);
#undef WINDS_ORD
    }
    return hipGetLastError();
}

int64_t seg32_padded(int64_t nseg) { return (nseg + kSegTile - 1) / kSegTile * kSegTile; }

template <int CORE, int SEM, bool CO>
static void fp32c_dispatch(dim3 grid, int block, hipStream_t st, const double *tx,
                           const double *ty, const double *tz, int64_t nt,
                           const Seg32 *s32, const Seg64 *s64, int64_t npad,
                           int64_t chunk, const Fp32Params &P, double *out, float *est,
                           int64_t stride) {
    const size_t lds = (size_t)(block / 64) * sizeof(NearLds);
    bs_fp32c<CORE, SEM, CO><<<grid, block, lds, st>>>(tx, ty, tz, nt, s32, s64, npad, chunk, P, out,
                                                      est, stride);
}

static inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t guard_bytes(int64_t n_est, int64_t nt) {
    return al256((size_t)n_est * sizeof(float)) + al256((size_t)nt * sizeof(uint32_t)) + 256 +
           al256((size_t)kFixCap * kFixSplit * 3 * sizeof(double)) + 256;
}

GuardBufs guard_carve(void *base, int64_t n_est, int64_t nt) {
    GuardBufs g;
    char *p = (char *)(((uintptr_t)base + 255) & ~(uintptr_t)255);
    g.est = (float *)p;
    p += al256((size_t)n_est * sizeof(float));
    g.list = (uint32_t *)p;
    p += al256((size_t)nt * sizeof(uint32_t));
    g.count = (int *)p;
    p += 256;
    g.fix = (double *)p;
    return g;
}

hipError_t launch_guard_reset(const GuardBufs &g, hipStream_t st) {
    guard_reset<<<1, 64, 0, st>>>(g.count);
    return hipGetLastError();
}

hipError_t launch_guard_fix(const double *t3, int64_t nt_total, const Seg64 *seg,
                            const WakeSpan *wk, int n_wk, WakeSpan one, const EvalCfg &c,
                            const GuardBufs &g, double *u3, hipStream_t st) {
    const double *tx = t3, *ty = t3 + nt_total, *tz = t3 + 2 * nt_total;
    double *ux = u3, *uy = u3 + nt_total, *uz = u3 + 2 * nt_total;
    const int n = c.core == 0 ? 0 : c.vat_n;
    // enough workgroups for a few listed targets at a time on every CU; the
    // kernel exits at once when the list is empty
    const dim3 grid(kFixSplit, 64);
#define WINDS_FIX(M, F)                                                                        \
    bs_guard_fix<M, F><<<grid, 256, 0, st>>>(tx, ty, tz, seg, wk, n_wk, one, c.d, n, c.co, g.list, \
                                             g.count, g.fix, ux, uy, uz)
    if (c.semantics == 0)
        WINDS_FIX(0, 0);
    else if (n == 0)
        WINDS_FIX(1, 0);
    else if (n == 1)
        WINDS_FIX(2, 1);
    else if (n == 2)
        WINDS_FIX(2, 2);
    else
        WINDS_FIX(2, 0);
#undef WINDS_FIX
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    bs_guard_fix_reduce<<<(unsigned)(kFixCap / 256), 256, 0, st>>>(g.list, g.count, g.fix, ux, uy,
                                                                   uz);
    return hipGetLastError();
}

// Far cut-off parameters of the FP32 kernels.  co < 0 zeroes every term
// (r1 > co and r2 > co hold for every finite pair): co2 = -1, no band.
static void set_co(Fp32Params &P, const EvalCfg &c) {
    const bool has_co = c.semantics == 1 && c.co < 1.0e18;
    if (!has_co) {
        P.co2 = __builtin_inff();
        P.co2_tol = 0.0f;
    } else if (c.co < 0.0) {
        P.co2 = -1.0f;
        P.co2_tol = 0.0f;
    } else {
        P.co2 = (float)(c.co * c.co);
        // r^2 in the FP32 path is within a few 1e-7 of its FP64 value
        P.co2_tol = P.co2 * 1.0e-5f;
    }
}

Fp32Params fp32_params(const EvalCfg &c) {
    Fp32Params P;
    const int n = c.core == 0 ? 0 : c.vat_n;
    set_co(P, c);
    P.neg_inv_n = n > 0 ? -1.0f / (float)n : 0.0f;
    P.vat_n = n;
    // general Vatistas order: h^2n overflows FP64 above 2^(1024/n); there the
    // reference's term is h^2 * inf^(-1/n) = 0, so such pairs take the FP64
    // branch (which reproduces it); below 2^(128/n) FP32 cannot reach it
    P.h2_big = __builtin_inff();
    if (n >= 9) P.h2_big = (float)(exp2(1023.0 / (double)n) * (1.0 - 1e-6));
    P.d = c.d;
    P.co = c.co;
    return P;
}

hipError_t launch_fp32c(const double *t3, int64_t nt, const Seg64 *seg,
                        int64_t nseg, const EvalCfg &c, double *u3, Seg32 *s32,
                        double *part, const GuardBufs &g, hipStream_t st) {
    if (nt <= 0) return hipSuccess;
    if (nseg <= 0) return launch_zero(u3, 3 * nt, st);
    const int64_t npad = seg32_padded(nseg);
    const int core_kind = c.core == 0 ? 0 : 1;
    const int n = c.core == 0 ? 0 : c.vat_n;
    prep_seg32<<<grid_for(npad, 256), 256, 0, st>>>(seg, nseg, npad, core_kind, n, c.d,
                                                    c.near_factor, c.near_len_frac, s32);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int block = c.block > 0 ? c.block : 256;
    const int split = c.split > 0 ? c.split : 1;
    const int64_t tiles = npad / kSegTile;
    const int64_t chunk = (tiles + split - 1) / split * kSegTile;
    const int nsplit = (int)((npad + chunk - 1) / chunk);
    dim3 grid(grid_for(nt, block), nsplit);
    const double *tx = t3, *ty = t3 + nt, *tz = t3 + 2 * nt;
    const bool guard = c.guard >= 0.0;
    double *out = nsplit == 1 && !guard ? u3 : part;
    float *est = guard ? g.est : nullptr;
    const int kind = c.core == 0 ? 0 : (n == 1 ? 1 : (n == 2 ? 2 : 3));
    const Fp32Params P = fp32_params(c);
    if (guard) {
        e = launch_guard_reset(g, st);
        if (e != hipSuccess) return e;
    }
#define WINDS_FP32C(K, S, C)                                                                \
    fp32c_dispatch<K, S, C>(grid, block, st, tx, ty, tz, nt, s32, seg, npad, chunk, P, out, est, \
                            nt)
    const bool has_co = c.semantics == 1 && c.co < 1.0e18;
    if (c.semantics == 0) {
        WINDS_FP32C(0, 0, false);
    } else if (has_co) {
        switch (kind) {
            case 0: WINDS_FP32C(0, 1, true); break;
            case 1: WINDS_FP32C(1, 1, true); break;
            case 2: WINDS_FP32C(2, 1, true); break;
            default: WINDS_FP32C(3, 1, true); break;
        }
    } else {
        switch (kind) {
            case 0: WINDS_FP32C(0, 1, false); break;
            case 1: WINDS_FP32C(1, 1, false); break;
            case 2: WINDS_FP32C(2, 1, false); break;
            default: WINDS_FP32C(3, 1, false); break;
        }
    }
#undef WINDS_FP32C
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (guard) {
        e = reduce_split(part, g.est, nt, nsplit, c.guard, u3, g.list, g.count, st);
        if (e != hipSuccess) return e;
        return launch_guard_fix(t3, nt, seg, nullptr, 0, WakeSpan{0, nt, 0, nseg}, c, g, u3, st);
    }
    if (nsplit > 1) e = reduce_split(part, nullptr, nt, nsplit, 0.0, u3, nullptr, nullptr, st);
    return e;
}

hipError_t launch_pack_segments(const double *F1, const double *F2, const int64_t *fd,
                                const double *gamma, const double *rc, Seg64 *out,
                                hipStream_t st) {
    const int64_t nf = fd[0] * fd[2] * fd[3];
    if (nf <= 0) return hipSuccess;
    pack_segments_cm<<<grid_for(nf, 256), 256, 0, st>>>(F1, F2, fd[0], fd[2], fd[3], gamma, rc, out);
    return hipGetLastError();
}

hipError_t launch_pack_targets(const double *P, const int64_t *pd, double *t3, hipStream_t st) {
    const int64_t np = pd[0] * pd[2] * pd[3];
    if (np <= 0) return hipSuccess;
    pack_targets_cm<<<grid_for(np, 256), 256, 0, st>>>(P, pd[0], pd[2], pd[3], t3);
    return hipGetLastError();
}

hipError_t launch_unpack_uind(const double *u3, const int64_t *pd, double *U, hipStream_t st) {
    const int64_t np = pd[0] * pd[2] * pd[3];
    if (np <= 0) return hipSuccess;
    unpack_uind_cm<<<grid_for(np, 256), 256, 0, st>>>(u3, pd[0], pd[2], pd[3], U);
    return hipGetLastError();
}

hipError_t launch_segment_length(const Seg64 *seg, int64_t nseg, int squared, double *L,
                                 hipStream_t st) {
    if (nseg <= 0) return hipSuccess;
    segment_length<<<grid_for(nseg, 256), 256, 0, st>>>(seg, nseg, squared, L);
    return hipGetLastError();
}

}  // namespace winds
