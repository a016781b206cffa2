// bs_fp32c.hpp -- FP32 compensated pair loop shared by the single-set and
// the batched (persistent) kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "bs_kernels.hpp"
#include "bs_pair.hpp"

namespace winds {

// Smallest near radius as a fraction of |L|: with R >= 0.51|L| every target
// inside the sphere on the segment as diameter (r1.r2 <= 0) is near, so the
// FP32 arithmetic never meets r1 r2 + r1.r2 ~ 0 (0.5 plus rounding margin).
constexpr double kMinNearLenFrac = 0.51;


// Orders the wave's LDS accesses: LDS operations of one wave execute in
// issue order; this keeps the compiler from moving them across each other.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Squared near radius of FP64 record s with |L|^2 = l2: max(near_factor^2
// rc, near_len_frac^2 |L|^2) (near_len_frac at least kMinNearLenFrac); +inf
// (always FP64) when it is not positive or gamma is NaN.
__device__ __forceinline__ double near_radius2(const Seg64 &s, double l2, double near_factor,
                                               double near_len_frac) {
    double nr2 = near_factor * near_factor * s.rc;
    const double nlf = near_len_frac > kMinNearLenFrac ? near_len_frac : kMinNearLenFrac;
    const double nl2 = nlf * nlf * l2;
    if (nl2 > nr2) nr2 = nl2;
    if (!(nr2 > 0.0) || !(s.g == s.g)) nr2 = __builtin_inf();
    return nr2;
}

__device__ __forceinline__ bool seg_chained(const Seg64 *prev, const Seg64 &s) {
    return prev && prev->bx == s.ax && prev->by == s.ay && prev->bz == s.az;
}

// Prepared record from FP64 record s (prev / next: the records before and
// after it in the same set, or null), or a padding record when s is null.
__device__ __forceinline__ Seg32 make_seg32(const Seg64 *s_ptr, const Seg64 *prev,
                                            const Seg64 *next, int core_kind, int n, double d,
                                            double near_factor, double near_len_frac) {
    Seg32 r;
    if (!s_ptr) {
        r.bx = r.by = r.bz = 1.0e6;
        r.lx = r.ly = r.lz = 0.0f;
        r.near2b = -1.0f;
        r.il2 = 0.0f;
        r.g = 0.0f;
        r.core = 1.0f;
        r.near2 = -1.0f;
        r.flags = kSegSkip;
        r.near2a = -1.0f;
        return r;
    }
    const Seg64 s = *s_ptr;
    const double lx = s.bx - s.ax, ly = s.by - s.ay, lz = s.bz - s.az;
    const double l2 = lx * lx + ly * ly + lz * lz;
    r.bx = s.bx;
    r.by = s.by;
    r.bz = s.bz;
    r.lx = (float)lx;
    r.ly = (float)ly;
    r.lz = (float)lz;
    r.il2 = l2 > 0.0 ? (float)(1.0 / l2) : 0.0f;
    r.g = l2 > 0.0 ? (float)(s.g / (4.0 * kMlPi)) : 0.0f;
    if (core_kind == 0) {
        r.core = (float)(d * l2);
    } else {
        // n = 1, 2: r_c^2n; FLT_MIN keeps r_c^4 + h^4 > 0 without a per-pair
        // max.  n >= 3: r_c^2 itself (the scaled form of the factor).
        const float c = (float)(n >= 3 ? s.rc : ipow(s.rc, n));
        r.core = n >= 3 ? c : (c > FLT_MIN ? c : FLT_MIN);
    }
    const double nr2 = near_radius2(s, l2, near_factor, near_len_frac);
    // the cut-off core needs no h^2 = |r1 x r2|^2 / |L|^2 but the near test:
    // it compares |r1 x r2|^2 with R^2 |L|^2 (a zero-length segment: with
    // +inf, as h^2 = 0 < R^2 would)
    r.near2 = (float)(core_kind == 0 && l2 > 0.0 ? nr2 * l2 : nr2);
    if (core_kind == 0 && !(l2 > 0.0)) r.near2 = __builtin_inff();
    double thr = nr2 + l2;
    r.near2a = (float)thr;
    if (next && seg_chained(s_ptr, *next)) {
        const Seg64 q = *next;
        const double qx = q.bx - q.ax, qy = q.by - q.ay, qz = q.bz - q.az;
        const double q2 = qx * qx + qy * qy + qz * qz;
        const double tq = near_radius2(q, q2, near_factor, near_len_frac) + q2;
        if (tq > thr) thr = tq;
    }
    r.near2b = (float)thr;
    r.flags = seg_chained(prev, s) ? kSegChained : 0;
    return r;
}

__device__ __forceinline__ void two_sum_acc(float &hi, float &lo, float v) {
    const float s = hi + v;
    const float bb = s - hi;
    const float err = (hi - (s - bb)) + (v - bb);
    hi = s;
    lo += err;
}

// Parameters of the FP32 kernels that are not per segment.
struct Fp32Params {
    float co2;        // co^2 (CO kernels)
    float co2_tol;    // |r^2 - co^2| at or below which a pair goes to FP64
    float neg_inv_n;  // -1/n (general Vatistas order)
    int vat_n;
    float h2_big;     // general Vatistas order: pairs with h^2 >= h2_big go to FP64
    double d;         // cut-off (near pairs)
    double co;        // far cut-off (near pairs)
};

struct EvalCfg;
// Parameters for configuration c (co, order, overflow threshold).
Fp32Params fp32_params(const EvalCfg &c);

// FP32 view of P - X for a filament point X: the vector, its squared length,
// the reciprocal of its length and its length.
struct RelPoint {
    float x, y, z, s, i, r;
};

__device__ __forceinline__ RelPoint rel_point(double px, double py, double pz, double x, double y,
                                              double z) {
    // formed in FP64 and rounded once: coordinates reach ~1 km while the
    // pairs evaluated in FP32 are metres apart
    RelPoint r;
    r.x = (float)(px - x);
    r.y = (float)(py - y);
    r.z = (float)(pz - z);
    r.s = fmaf(r.z, r.z, fmaf(r.y, r.y, r.x * r.x));
    r.i = __builtin_amdgcn_rsqf(r.s);
    r.r = r.s * r.i;
    return r;
}

// The FP32 fields of a record as the pair arithmetic reads them.
//
// Measured on gfx950 with independent chains at 8 waves per SIMD: a wave64
// FP32 instruction with VGPR, inline-constant or literal operands issues in
// ~2.3 shader clocks, one with an SGPR operand in ~4.1, FP64 adds and
// conversions in ~4.2, v_rsq_f32 in ~9.8.  In the kernel's mixed stream
// those costs do not simply add: for the Vatistas cores, copying L into
// VGPRs once per record made the evaluation 3% slower, and reading L and
// gamma from LDS took 5% fewer cycles but ran at a 9% lower clock; they keep
// every field in SGPRs.  The cut-off core, with fewer values live per
// record, reads L and gamma (nine uses per record) from a per-wave LDS copy
// of the tile, broadcast into VGPRs: 27.7 -> 26.6 ms at 133,920 x 263,427.
struct RecF {
    float lx, ly, lz, near2b, il2, g, core, near2;
};

__device__ __forceinline__ RecF rec_fields_1(const Seg32 &s) {
    return RecF{s.lx, s.ly, s.lz, s.near2b, s.il2, s.g, s.core, s.near2};
}

// L and gamma from the LDS copy (v = its float4), the rest from the record.
__device__ __forceinline__ RecF rec_fields_v(const Seg32 &s, const float4 &v) {
    return RecF{v.x, v.y, v.z, s.near2b, s.il2, v.w, s.core, s.near2};
}

// (lx, ly, lz, g) of the kSegTile records of tile kt copied into buf (one
// float4 per record): lane l loads field (l & 3) of records (l >> 2) + 16 q.
static_assert(kSegTile % 16 == 0, "rec_stage: 16 records per 64 lanes");
__device__ __forceinline__ void rec_stage(const Seg32 *__restrict__ s32, int64_t kt, float *buf) {
    const int lane = (int)(threadIdx.x & 63u);
    const int c = lane & 3;
    float v[kSegTile / 16];
#pragma unroll
    for (int q = 0; q < kSegTile / 16; ++q) {
        const float *r = (const float *)(s32 + kt + 16 * q + (lane >> 2));
        v[q] = r[c < 3 ? 6 + c : 11];  // lx, ly, lz at words 6-8, g at word 11
    }
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < kSegTile / 16; ++q) buf[64 * q + lane] = v[q];
    wave_lds_sync();
}

// Near pairs go to a per-wave list in LDS instead of a branch in the pair
// loop: a branch is paid by the whole wave (about 170 FP64 instructions for
// the reference arithmetic) however few lanes are near.  Every near lane
// appends its (record, lane) pair; when the list is full and at the end of
// the range, the wave evaluates the listed pairs 64 at a time, one per
// lane, and adds each term to its target's FP64 sum in LDS.  The order of
// those additions is fixed (list order, then lane order within one LDS
// instruction), so the result is the same on every run.
constexpr int kNearCap = 256;  // pairs per wave between evaluations

struct alignas(16) NearLds {
    float rec[4 * kSegTile];  // the tile's (lx, ly, lz, g) (cut-off core)
    uint32_t pair[kNearCap];  // (k - k0) * 64 + lane
    double p[3][64];          // the wave's targets
    double acc[3][64];        // their FP64 near-pair sums
};

template <int CORE, int SEM>
struct NearList {
    NearLds *L;
    const Seg64 *s64;  // FP64 record of prepared record k: s64[k + shift]
    int64_t shift;
    int64_t k0;
    int n;             // pairs listed (wave-uniform)

    __device__ __forceinline__ void init(NearLds *l, const Seg64 *q, int64_t sh, int64_t kbase,
                                         double px, double py, double pz) {
        L = l;
        s64 = q;
        shift = sh;
        k0 = kbase;
        n = 0;
        const int lane = (int)(threadIdx.x & 63u);
        wave_lds_sync();
        L->p[0][lane] = px;
        L->p[1][lane] = py;
        L->p[2][lane] = pz;
        L->acc[0][lane] = 0.0;
        L->acc[1][lane] = 0.0;
        L->acc[2][lane] = 0.0;
        wave_lds_sync();
    }

    // Evaluates the listed pairs (called with the whole wave).
    __device__ __forceinline__ void flush(const Fp32Params &P) {
        wave_lds_sync();
        const int lane = (int)(threadIdx.x & 63u);
        for (int base = 0; base < n; base += 64) {
            const int i = base + lane;
            if (i < n) {
                const uint32_t e = L->pair[i];
                const int tl = (int)(e & 63u);
                const Seg64 q = s64[k0 + (int64_t)(e >> 6) + shift];
                const double px = L->p[0][tl], py = L->p[1][tl], pz = L->p[2][tl];
                double ex, ey, ez;
                if constexpr (SEM == 0) {
                    mex_term(q.ax, q.ay, q.az, q.bx, q.by, q.bz, q.g, px, py, pz, P.d, ex, ey,
                             ez);
                } else {
                    matlab_pair<CORE != 0, (CORE == 1 || CORE == 2) ? CORE : -1>(
                        q.ax, q.ay, q.az, q.bx, q.by, q.bz, q.g, q.rc, px, py, pz, P.d, P.vat_n,
                        P.co, ex, ey, ez);
                }
                atomicAdd(&L->acc[0][tl], ex);
                atomicAdd(&L->acc[1][tl], ey);
                atomicAdd(&L->acc[2][tl], ez);
            }
        }
        wave_lds_sync();
        n = 0;
    }

    // Lists the pair (k, lane) of every lane of mask m (m = ballot(mine)).
    // The caller keeps room for 64 pairs per record between flushes.
    __device__ __forceinline__ void add(uint64_t m, bool mine, int64_t k) {
        const int c = __popcll(m);
        if (mine) {
            const int lane = (int)(threadIdx.x & 63u);
            const int below = (int)__builtin_amdgcn_mbcnt_hi(
                (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            L->pair[n + below] = (uint32_t)((k - k0) * 64 + lane);
        }
        n += c;
    }

    // The sums of this lane's near pairs (after a final flush).
    __device__ __forceinline__ void result(const Fp32Params &P, double &ax, double &ay,
                                           double &az) {
        flush(P);
        const int lane = (int)(threadIdx.x & 63u);
        ax = L->acc[0][lane];
        ay = L->acc[1][lane];
        az = L->acc[2][lane];
    }
};

// One record: the term of segment k (start point a, end point b, FP32
// fields f) added to (qx, qy, qz) in FP32, or to (ax, ay, az) by the FP64
// reference arithmetic for a near pair.  Returns P - B for the next record
// of the filament.
// Far cut-off masks of the wave's lanes at the carried point (CO kernels):
// band, |r|^2 within FP32 rounding of co^2 (the pair goes to FP64); far,
// not |r|^2 <= co^2 (beyond co, or NaN).  Carried like the end-point mask,
// so each point is tested once for the two segments that share it.
struct CoMask {
    uint64_t band, far;
};

__device__ __forceinline__ CoMask co_mask(float s2, const Fp32Params &P) {
    return CoMask{__builtin_amdgcn_ballot_w64(!(fabsf(s2 - P.co2) > P.co2_tol)),
                  __builtin_amdgcn_ballot_w64(!(s2 <= P.co2))};
}

// CHAINED32: when the record continues the filament of the record before
// it forms P - B as r1 - L in FP32.
template <int CORE, int SEM, bool CO, bool CHAINED32>
__device__ __forceinline__ RelPoint fp32c_record(double px, double py, double pz, const Seg32 &s,
                                                 const RecF &f, const Seg64 *__restrict__ q64,
                                                 int64_t k, RelPoint a, const Fp32Params &P,
                                                 float &qx, float &qy, float &qz, uint64_t &mc,
                                                 CoMask &cm, NearList<CORE, SEM> &nl) {
    // mc: the wave's lanes passing the end-point test at A.  A chained record
    // takes it from the record before (whose B threshold covers this
    // record's), the others test their own start point.  Kept as a lane mask
    // in SGPRs: a carried per-lane bool is rebuilt from a VGPR each record.
    const bool chained = s.flags & kSegChained;
    if (__builtin_expect(!chained, 0)) {
        a = rel_point(px, py, pz, q64->ax, q64->ay, q64->az);
        mc = __builtin_amdgcn_ballot_w64(!(a.s >= s.near2a));
        if constexpr (CO) cm = co_mask(a.s, P);
    }
    const float Lx = f.lx, Ly = f.ly, Lz = f.lz;
    // P - B: in FP64 and rounded once, or (CHAINED32, chained) as r1 - L in
    // FP32, where r1 was rounded from FP64 one record before.  The
    // pairs that reach the FP32 arithmetic have |r2| >= 0.51|L|, so each
    // subtraction adds a few roundings of |r2| at most.
    RelPoint b;
    if (__builtin_expect(CHAINED32 && chained, 1)) {
        b.x = a.x - Lx;
        b.y = a.y - Ly;
        b.z = a.z - Lz;
        b.s = fmaf(b.z, b.z, fmaf(b.y, b.y, b.x * b.x));
        b.i = __builtin_amdgcn_rsqf(b.s);
        b.r = b.s * b.i;
    } else {
        b = rel_point(px, py, pz, s.bx, s.by, s.bz);
    }
    const float dot = fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x));
    // r1 x r2 = L x r1 with L = B - A.
    const float cx = fmaf(Ly, a.z, -(Lz * a.y));
    const float cy = fmaf(Lz, a.x, -(Lx * a.z));
    const float cz = fmaf(Lx, a.y, -(Ly * a.x));
    const float c2 = fmaf(cz, cz, fmaf(cy, cy, cx * cx));  // |r1 x r2|^2
    // squared distance to the segment's line (the cut-off core only tests
    // it, as c2 against R^2 |L|^2)
    const float h2 = CORE == 0 ? c2 : c2 * f.il2;
    // Conservative near test: every target closer than R to the closed
    // segment has h2 < R^2 and min(r1^2, r2^2) < R^2 + |L|^2 (or a larger
    // end-point threshold).  NaN anywhere selects the FP64 branch.
    // R >= 0.51|L| makes every target inside the sphere on the segment as
    // diameter (the pairs with r1.r2 <= 0) near, so below r1.r2 > 0 and
    // r1 r2 + r1.r2 never cancels.
    const bool n1 = !(h2 >= f.near2), n3 = !(b.s >= f.near2b);
    // the near mask combined from compare masks in SGPRs (a ballot of a
    // combined per-lane condition would go through a VGPR and a compare)
    const uint64_t m3 = __builtin_amdgcn_ballot_w64(n3);
    uint64_t m = __builtin_amdgcn_ballot_w64(n1) & (mc | m3);
    // The far cut-off (zero when r1 > co and r2 > co) is a step: a pair with
    // r1 or r2 within FP32 rounding of co could fall on the other side of it
    // than in FP64 and drop or keep a whole term, so such pairs take the FP64
    // branch, which decides as the reference does.
    CoMask cb{0, 0};
    if constexpr (CO) {
        cb = co_mask(b.s, P);
        m |= cm.band | cb.band;
    }
    if constexpr (CORE == 3) m |= __builtin_amdgcn_ballot_w64(h2 >= P.h2_big);
    const bool near = __builtin_amdgcn_inverse_ballot_w64(m);
    // zeroed: near pairs (their terms come from the FP64 branch) and, with
    // the far cut-off, pairs with both ends beyond co
    uint64_t zm = m;
    if constexpr (CO) zm |= cm.far & cb.far;
    mc = m3;
    if constexpr (CO) cm = cb;
    // r1 r2 + r1.r2: from the carried lengths (one product per point, where
    // s1 s2 / (r1 r2) costs two per record), or, for the cut-off core, whose
    // d|L|^2 term needs 1/(r1 r2) anyway, from that
    const float pp = a.i * b.i;
    const float qa = CORE == 0 ? fmaf(a.s * b.s, pp, dot) : fmaf(a.r, b.r, dot);
    const float gn = f.g * (a.i + b.i);
    float ubar;
    if constexpr (CORE == 0) {
        // 1/(r1r2 + r1.r2 + d|L|^2/(r1r2))
        ubar = gn * __builtin_amdgcn_rcpf(fmaf(f.core, pp, qa));
    } else if constexpr (CORE == 1) {
        // h^2/(r_c^2 + h^2) / (r1r2 + r1.r2)
        const float G = fmaxf((f.core + h2) * qa, FLT_MIN);
        ubar = gn * (h2 * __builtin_amdgcn_rcpf(G));
    } else if constexpr (CORE == 2) {
        // h^2/sqrt(r_c^4 + h^4) / (r1r2 + r1.r2); r_c^4 >= FLT_MIN
        const float w = fmaf(h2, h2, f.core);
        ubar = gn * (h2 * __builtin_amdgcn_rsqf(w * (qa * qa)));
    } else {
        // h^2/(r_c^2n + h^2n)^(1/n) / (r1r2 + r1.r2), in the scaled form
        // (h^2/M) (1 + x^n)^(-1/n) with M = max(h^2, r_c^2) and
        // x = min(h^2, r_c^2)/M <= 1, so that no power leaves the FP32 range
        // whatever n and h (h^2n itself overflows FP32 at n = 10, h = 100 m)
        const float M = fmaxf(fmaxf(h2, f.core), FLT_MIN);
        float iM = __builtin_amdgcn_rcpf(M);
        iM = fmaf(iM, fmaf(-M, iM, 1.0f), iM);
        const float x = fminf(h2, f.core) * iM;
        float xn = x;
        for (int i = 1; i < P.vat_n; ++i) xn = xn * x;
        const float w = 1.0f + xn;  // in [1, 2]
        // w^(-1/n) from exp2/log2, then one Newton step on y^-n = w,
        // y <- y + y (1 - w y^n) / n, to FP32 rounding
        float root = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(w) * P.neg_inv_n);
        float yn = root;
        for (int i = 1; i < P.vat_n; ++i) yn = yn * root;
        root = fmaf(root * fmaf(-w, yn, 1.0f), -P.neg_inv_n, root) * iM;
        // 1/qa refined by one Newton step: v_rcp_f32's error was the larger
        // part of this order's (max per-target error on the 1,440-age wake
        // 2.0e-6 -> 1.2e-6, 1% of the time)
        float rq = __builtin_amdgcn_rcpf(qa);
        rq = fmaf(rq, fmaf(-qa, rq, 1.0f), rq);
        ubar = gn * (h2 * root * rq);
    }
    ubar = __builtin_amdgcn_inverse_ballot_w64(zm) ? 0.0f : ubar;
    qx = fmaf(ubar, cx, qx);
    qy = fmaf(ubar, cy, qy);
    qz = fmaf(ubar, cz, qz);
    if (__builtin_expect(m != 0, 0)) nl.add(m, near, k);
    return b;
}

// Record k with all sixteen words in SGPRs before the first branch on its
// flags: one batch of scalar loads and one wait per record (left alone, the
// compiler loads the flags first and the rest under the branch, paying the
// scalar-cache latency twice).
__device__ __forceinline__ Seg32 load_record(const Seg32 *__restrict__ s32, int64_t k) {
    Seg32 s = s32[k];
    asm volatile("" : "+s"(s.flags)
                 : "s"(s.bx), "s"(s.by), "s"(s.bz), "s"(s.lx), "s"(s.ly), "s"(s.lz),
                   "s"(s.near2b), "s"(s.il2), "s"(s.g), "s"(s.core), "s"(s.near2),
                   "s"(s.near2a));
    return s;
}

// load_record with the record's LDS float4 v (read before the scalar loads)
// in the same wait: one lgkmcnt wait covers both.
__device__ __forceinline__ Seg32 load_record_v(const Seg32 *__restrict__ s32, int64_t k,
                                               const float4 &v) {
    Seg32 s = s32[k];
    asm volatile("" : "+s"(s.flags)
                 : "s"(s.bx), "s"(s.by), "s"(s.bz), "s"(s.lx), "s"(s.ly), "s"(s.lz),
                   "s"(s.near2b), "s"(s.il2), "s"(s.g), "s"(s.core), "s"(s.near2),
                   "s"(s.near2a), "v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w));
    return s;
}

// Record k (slot j of the tile) and the fields its arithmetic reads: from
// the scalar record, or with LF L and gamma from the tile's LDS copy buf.
template <bool LF>
__device__ __forceinline__ Seg32 load_rec(const Seg32 *__restrict__ s32, int64_t k,
                                          const float *buf, int j, RecF &f) {
    if constexpr (LF) {
        const float4 v = ((const float4 *)buf)[j];
        const Seg32 s = load_record_v(s32, k, v);
        f = rec_fields_v(s, v);
        return s;
    } else {
        const Seg32 s = load_record(s32, k);
        f = rec_fields_1(s);
        return s;
    }
}

// Sum of the terms of segments [k0, k1) (prepared records, k1 - k0 a
// multiple of kSegTile) at target (px, py, pz).  s64 + shift is the FP64
// record of prepared record k.  Every lane of the wave must walk the same
// range: the records are wave-uniform scalar loads, and so is every branch
// on their flags.
//
// CORE: 0 cut-off, 1 Vatistas n=1, 2 Vatistas n=2, 3 Vatistas general n.
// SEM:  0 MEX semantics (near pairs by biot_p2p), 1 BiotSavart.m semantics
//       (near pairs by the BiotSavart.m arithmetic with zeroing).
// CO:   apply the far cut-off r1 > co && r2 > co (BiotSavart.m:111).
// CHAIN32: records go in groups of kGroup; the first forms P - B in FP64,
// the others as r1 - L in FP32 when they continue the filament.
// LF:   L and gamma from a per-wave LDS copy of the tile (see RecF).
//
// The branch hints (__builtin_expect) here and in fp32c_record lay the
// common path -- a record that continues the filament, no near lane, no
// flush -- out to fall through, with the rare blocks out of line: taken
// branches cost the wave issue slots (29.19 -> 28.37 ms on the Vatistas-2
// wake of 133,920 targets x 263,427 segments, same results).
//
// Records per FP64 end point.  Longer chains are faster but carry more
// FP32 rounding into r, and the cancellation guard's error estimate assumes
// the rounding of two: with 4 the largest per-target error on 32 seeded
// filament soups (random polylines, targets on and near them) grew from
// 4.1e-7 to 9.8e-7 for 1.7% less time on the 133,920 x 263,427 wake; with
// 8 and 16 the 64-wake batch's grew from 3.0e-7 to 6.1e-7 and 2.8e-6.
constexpr int kGroup = 2;
template <int CORE, int SEM, bool CO, bool CHAIN32 = true, bool LF = (CORE == 0)>
__device__ __forceinline__ void fp32c_range(double px, double py, double pz,
                                            const Seg32 *__restrict__ s32,
                                            const Seg64 *__restrict__ s64, int64_t shift,
                                            int64_t k0, int64_t k1, const Fp32Params &P,
                                            NearLds *nlds, double &ox, double &oy, double &oz,
                                            float &oa) {
    float hx = 0.f, hy = 0.f, hz = 0.f, lx = 0.f, ly = 0.f, lz = 0.f;
    // sum of |tile partial|^2: how far the FP32 terms cancel at this target
    float at = 0.f;
    NearList<CORE, SEM> nl;
    nl.init(nlds, s64, shift, k0, px, py, pz);
    // P - (end point of the record before), carried along the filament; the
    // range's first record forms it from its own start point
    RelPoint e{0.f, 0.f, 0.f, 1.f, 1.f, 1.f};
    uint64_t mc = 0;  // lanes passing the end-point near test at that point
    CoMask cm{0, 0};  // the far cut-off masks at that point (CO)
    if (k0 < k1 && !(s32[k0].flags & kSegSkip)) {
        e = rel_point(px, py, pz, s64[k0 + shift].ax, s64[k0 + shift].ay, s64[k0 + shift].az);
        mc = __builtin_amdgcn_ballot_w64(!(e.s >= s32[k0].near2a));
        if constexpr (CO) cm = co_mask(e.s, P);
    }
    const float *buf = nlds->rec;
    if constexpr (LF) {
        if (k0 < k1) rec_stage(s32, k0, nlds->rec);
    }
    // Each iteration loads its own records: at 8 waves per SIMD the other
    // waves hide the scalar-cache latency (loading the next iteration's
    // records early helped only the general Vatistas order while it ran 4
    // waves, and made the cut-off 6% slower).
    for (int64_t kt = k0; kt < k1; kt += kSegTile) {
        float qx = 0.f, qy = 0.f, qz = 0.f;
        // kGroup records per iteration: the carried point alternates between
        // two sets of registers instead of being copied
#pragma unroll 1
        for (int j = 0; j < kSegTile; j += kGroup) {
#pragma unroll
            for (int q = 0; q < kGroup; q += 2) {
                const int64_t k = kt + j + q;
                RecF fa, fb;
                const Seg32 sa = load_rec<LF>(s32, k, buf, j + q, fa);
                RelPoint f = e;
                if (__builtin_expect(!(sa.flags & kSegSkip), 1)) {
                    if (q == 0)
                        f = fp32c_record<CORE, SEM, CO, false>(px, py, pz, sa, fa, s64 + k + shift,
                                                               k, e, P, qx, qy, qz, mc, cm, nl);
                    else
                        f = fp32c_record<CORE, SEM, CO, CHAIN32>(px, py, pz, sa, fa,
                                                                 s64 + k + shift, k, e, P, qx, qy,
                                                                 qz, mc, cm, nl);
                }
                const Seg32 sb = load_rec<LF>(s32, k + 1, buf, j + q + 1, fb);
                if (__builtin_expect(!(sb.flags & kSegSkip), 1))
                    e = fp32c_record<CORE, SEM, CO, CHAIN32>(px, py, pz, sb, fb,
                                                             s64 + k + 1 + shift, k + 1, f, P, qx,
                                                             qy, qz, mc, cm, nl);
                else
                    e = f;
                // flushed between record pairs, where little else is live
                // (inside the record it pushed the loop to 89 VGPRs, 5 waves)
                if (__builtin_expect(nl.n > kNearCap - 128, 0)) nl.flush(P);
            }
        }
        if constexpr (LF) {
            if (kt + kSegTile < k1) rec_stage(s32, kt + kSegTile, nlds->rec);
        }
        two_sum_acc(hx, lx, qx);
        two_sum_acc(hy, ly, qy);
        two_sum_acc(hz, lz, qz);
        at = fmaf(qz, qz, fmaf(qy, qy, fmaf(qx, qx, at)));
    }
    double ax, ay, az;
    nl.result(P, ax, ay, az);
    ox = ((double)hx + (double)lx) + ax;
    oy = ((double)hy + (double)ly) + ay;
    oz = ((double)hz + (double)lz) + az;
    oa = at;
}

}  // namespace winds
