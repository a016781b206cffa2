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

// Copy of a wave-uniform value into a VGPR.  gfx950 issues FP32 VALU
// operations whose operands are all VGPRs at twice the rate of those that
// read an SGPR, and a move takes one fast slot: a record value feeding
// several products is moved once.
__device__ __forceinline__ float to_vgpr(float s) {
    float v;
    asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
    return v;
}

// Prepared record k from FP64 record s, or a padding record when s is null.
__device__ __forceinline__ Seg32 make_seg32(const Seg64 *s_ptr, int core_kind, int n, double d,
                                            double near_factor, double near_len_frac) {
    Seg32 r;
    if (!s_ptr) {
        // Padding record: far away, zero strength, never near.  Every FP32
        // intermediate stays finite, so its term is exactly zero.
        r.ax = r.ay = r.az = 1.0e6;
        r.lx = r.ly = r.lz = 0.0f;
        r.near2b = -1.0f;
        r.il2 = 0.0f;
        r.g = 0.0f;
        r.core = 1.0f;
        r.near2 = -1.0f;
        return r;
    }
    const Seg64 s = *s_ptr;
    const double lx = s.bx - s.ax, ly = s.by - s.ay, lz = s.bz - s.az;
    const double l2 = lx * lx + ly * ly + lz * lz;
    r.ax = s.ax;
    r.ay = s.ay;
    r.az = s.az;
    r.lx = (float)lx;
    r.ly = (float)ly;
    r.lz = (float)lz;
    r.il2 = l2 > 0.0 ? (float)(1.0 / l2) : 0.0f;
    r.g = l2 > 0.0 ? (float)(s.g / (4.0 * kMlPi)) : 0.0f;
    const double core = core_kind == 0 ? d * l2 : ipow(s.rc, n);
    r.core = (float)core;
    double nr2 = near_factor * near_factor * s.rc;
    const double nlf = near_len_frac > kMinNearLenFrac ? near_len_frac : kMinNearLenFrac;
    const double nl2 = nlf * nlf * l2;
    if (nl2 > nr2) nr2 = nl2;
    if (!(nr2 > 0.0) || !(s.g == s.g)) nr2 = __builtin_inf();
    r.near2 = (float)nr2;
    r.near2b = (float)(nr2 + l2);
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
    float neg_inv_n;  // -1/n (general Vatistas order)
    int vat_n;
    double d;         // cut-off (near pairs)
    double co;        // far cut-off (near pairs)
};

// Sum of the terms of segments [k0, k1) (prepared records, k1 - k0 a
// multiple of kSegTile) at target (px, py, pz).  s64 + shift is the FP64
// record of prepared record k.  Every lane of the wave must walk the same
// range: the records are wave-uniform scalar loads.
//
// CORE: 0 cut-off, 1 Vatistas n=1, 2 Vatistas n=2, 3 Vatistas general n.
// SEM:  0 MEX semantics (near pairs by biot_p2p), 1 BiotSavart.m semantics
//       (near pairs by the BiotSavart.m arithmetic with zeroing).
// CO:   apply the far cut-off r1 > co && r2 > co (BiotSavart.m:111).
template <int CORE, int SEM, bool CO>
__device__ __forceinline__ void fp32c_range(double px, double py, double pz,
                                            const Seg32 *__restrict__ s32,
                                            const Seg64 *__restrict__ s64, int64_t shift,
                                            int64_t k0, int64_t k1, const Fp32Params &P,
                                            double &ox, double &oy, double &oz) {
    const float co2 = P.co2;
    const float neg_inv_n = P.neg_inv_n;
    const int vat_n = P.vat_n;
    const double d = P.d;
    const double co = P.co;
    float hx = 0.f, hy = 0.f, hz = 0.f, lx = 0.f, ly = 0.f, lz = 0.f;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int64_t kt = k0; kt < k1; kt += kSegTile) {
        float qx = 0.f, qy = 0.f, qz = 0.f;
        // not unrolled: measured 9% faster than two records per iteration
#pragma unroll 1
        for (int j = 0; j < kSegTile; ++j) {
            const int64_t k = kt + j;
            const Seg32 s = s32[k];
            // L feeds three products per component: one move to a VGPR each
            const float Lx = to_vgpr(s.lx), Ly = to_vgpr(s.ly), Lz = to_vgpr(s.lz);
            // r1 = P - A in FP64, rounded once (coordinates reach ~1 km while
            // the pairs evaluated here are metres apart).  r2 = r1 - L in
            // FP32: every pair that reaches this arithmetic has
            // |r2| >= |L|/2 (the near test below sends closer ones to FP64),
            // so |r1| <= 3|r2| and r2 carries at most ~6 ulp of relative error.
            const float r1x = (float)(px - s.ax), r1y = (float)(py - s.ay),
                        r1z = (float)(pz - s.az);
            const float r2x = r1x - Lx, r2y = r1y - Ly, r2z = r1z - Lz;
            const float s1 = fmaf(r1z, r1z, fmaf(r1y, r1y, r1x * r1x));
            const float s2 = fmaf(r2z, r2z, fmaf(r2y, r2y, r2x * r2x));
            const float dot = fmaf(r1z, r2z, fmaf(r1y, r2y, r1x * r2x));
            // r1 x r2 = L x r1 with L = B - A.
            const float cx = fmaf(Ly, r1z, -(Lz * r1y));
            const float cy = fmaf(Lz, r1x, -(Lx * r1z));
            const float cz = fmaf(Lx, r1y, -(Ly * r1x));
            const float c2 = fmaf(cz, cz, fmaf(cy, cy, cx * cx));  // |r1 x r2|^2
            const float h2 = c2 * s.il2;  // squared distance to the segment's line
            // Conservative near test: every target closer than R to the
            // closed segment has h2 < R^2 and min(s1, s2) < R^2 + |L|^2.
            // NaN anywhere selects the FP64 branch.  R >= 0.51|L| makes every
            // target inside the sphere on the segment as diameter (the pairs
            // with r1.r2 <= 0) near, so below r1.r2 > 0 and r1 r2 + r1.r2
            // never cancels.
            const bool near = !(h2 >= s.near2) && !(fminf(s1, s2) >= s.near2b);
            const float i1 = __builtin_amdgcn_rsqf(s1);
            const float i2 = __builtin_amdgcn_rsqf(s2);
            const float pp = i1 * i2;
            const float qa = fmaf(s1 * s2, pp, dot);  // r1 r2 + r1.r2
            const float gn = s.g * (i1 + i2);
            float ubar;
            if constexpr (CORE == 0) {
                // 1/(r1r2 + r1.r2 + d|L|^2/(r1r2))
                ubar = gn * __builtin_amdgcn_rcpf(fmaf(s.core, pp, qa));
            } else if constexpr (CORE == 1) {
                // h^2/(r_c^2 + h^2) / (r1r2 + r1.r2)
                const float G = fmaxf((s.core + h2) * qa, FLT_MIN);
                ubar = gn * (h2 * __builtin_amdgcn_rcpf(G));
            } else if constexpr (CORE == 2) {
                // h^2/sqrt(r_c^4 + h^4) / (r1r2 + r1.r2)
                const float w = fmaxf(fmaf(h2, h2, s.core), FLT_MIN);
                ubar = gn * (h2 * __builtin_amdgcn_rsqf(w * (qa * qa)));
            } else {
                // h^2/(r_c^2n + h^2n)^(1/n) / (r1r2 + r1.r2)
                float hn = h2;
                for (int i = 1; i < vat_n; ++i) hn = hn * h2;
                const float w = fmaxf(s.core + hn, FLT_MIN);
                const float root = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(w) * neg_inv_n);
                ubar = gn * (h2 * root * __builtin_amdgcn_rcpf(qa));
            }
            bool zero = near;
            if constexpr (CO) zero = zero || !(fminf(s1, s2) <= co2);
            ubar = zero ? 0.0f : ubar;
            qx = fmaf(ubar, cx, qx);
            qy = fmaf(ubar, cy, qy);
            qz = fmaf(ubar, cz, qz);
            if (near) {
                const Seg64 q = s64[k + shift];
                if constexpr (SEM == 0) {
                    mex_pair(q.ax, q.ay, q.az, q.bx, q.by, q.bz, q.g, px, py, pz, d, ax, ay, az);
                } else {
                    double ex, ey, ez;
                    matlab_pair<CORE != 0, (CORE == 1 || CORE == 2) ? CORE : 0>(q.ax, q.ay, q.az, q.bx, q.by, q.bz, q.g, q.rc, px, py,
                                           pz, d, vat_n, co, ex, ey, ez);
                    ax += ex;
                    ay += ey;
                    az += ez;
                }
            }
        }
        two_sum_acc(hx, lx, qx);
        two_sum_acc(hy, ly, qy);
        two_sum_acc(hz, lz, qz);
    }
    ox = ((double)hx + (double)lx) + ax;
    oy = ((double)hy + (double)ly) + ay;
    oz = ((double)hz + (double)lz) + az;
}

}  // namespace winds
