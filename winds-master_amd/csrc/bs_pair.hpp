// bs_pair.hpp -- one target/segment interaction in FP64, in the reference's
// own operation order.  Shared by the FP64 kernels and by the near-pair
// branch of the FP32 compensated kernel, so both round exactly alike.
#pragma once

#include <hip/hip_runtime.h>

namespace winds {

// Segment record in device memory: A = F1 endpoint, B = F2 endpoint,
// gamma, rc = squared core radius.  64 bytes, one scalar-cache line.
struct __attribute__((aligned(64))) Seg64 {
    double ax, ay, az, bx, by, bz, g, rc;
};

// BiotSavartMex.cu:35 spells pi as this literal.
constexpr double kMexPi = 3.14159265358979324;
// MATLAB's pi (the double nearest pi; the same value).
constexpr double kMlPi = 3.141592653589793;

// biot_p2p / biot_p2p_nogpu (BiotSavartMex.cu:63-93, 236-266): cut-off core
// with d*|L|^2 in the denominator; rc and co unused; nothing zeroed.
// x1,y1,z1 = F1 endpoint, x2,y2,z2 = F2 endpoint (the reference packs them
// as fx = (x1, x2), BiotSavartMex.cu:347-352).
// The term alone; mex_pair adds it.
__device__ __forceinline__ void mex_term(double x1, double y1, double z1,
                                         double x2, double y2, double z2,
                                         double g, double px, double py,
                                         double pz, double cd, double &tx,
                                         double &ty, double &tz) {
    double ldx = x1 - x2;
    double ldy = y1 - y2;
    double ldz = z1 - z2;
    double pxx1 = px - x1;
    double pxx2 = px - x2;
    double pyy1 = py - y1;
    double pyy2 = py - y2;
    double pzz1 = pz - z1;
    double pzz2 = pz - z2;
    double r1 = sqrt(pxx1 * pxx1 + pyy1 * pyy1 + pzz1 * pzz1);
    double r2 = sqrt(pxx2 * pxx2 + pyy2 * pyy2 + pzz2 * pzz2);
    double r1dr2 = pxx1 * pxx2 + pyy1 * pyy2 + pzz1 * pzz2;
    double r1tr2 = r1 * r2;
    double len = ldx * ldx + ldy * ldy + ldz * ldz;
    double den = r1tr2 * (r1tr2 + r1dr2) + cd * len;
    double ubar = (g * (r1 + r2)) * (1.000 / (4.000 * kMexPi));
    ubar = ubar / den;
    tx = ubar * (pyy1 * pzz2 - pzz1 * pyy2);
    ty = ubar * (pzz1 * pxx2 - pxx1 * pzz2);
    tz = ubar * (pxx1 * pyy2 - pyy1 * pxx2);
}

__device__ __forceinline__ void mex_pair(double x1, double y1, double z1,
                                         double x2, double y2, double z2,
                                         double g, double px, double py,
                                         double pz, double cd, double &ux,
                                         double &uy, double &uz) {
    double tx, ty, tz;
    mex_term(x1, y1, z1, x2, y2, z2, g, px, py, pz, cd, tx, ty, tz);
    ux += tx;
    uy += ty;
    uz += tz;
}

__device__ __forceinline__ double ipow(double x, int n) {
    double r = x;
    for (int i = 1; i < n; ++i) r = r * x;
    return r;
}

// y.^(-1/n): 1/y for n = 1, 1/sqrt(y) for n = 2, pow otherwise.
__device__ __forceinline__ double inv_root(double y, int n) {
    if (n == 1) return 1.0 / y;
    if (n == 2) return 1.0 / sqrt(y);
    return pow(y, -1.0 / (double)n);
}

// y^(-1/n) for n >= 1 without pow: an FP32 first guess and three Newton
// steps on r^-n = y in FP64 (quadratic convergence from ~1e-6).  The guess
// is taken from y = m 2^e (m in [0.5, 1)) as 2^k * 2^(f - log2(m)/n) with
// -e/n = k + f (k integer, f in [0, 1)), so that it stays inside the FP32
// range for every FP64 y; the residual 1 - y r^n is formed as a running
// product from y, which moves monotonically towards 1 and cannot overflow.
// pow's special values are kept: y = 0 -> +inf, y = +inf -> 0, y < 0 or
// NaN -> NaN (BiotSavart.m's (rc^n + Cnu^n)^(-1/n) when Cnu^n overflows is
// inf^(-1/n) = 0).
__device__ __forceinline__ double inv_root_newton(double y, int n) {
    if (!(y > 0.0)) return y == 0.0 ? __builtin_inf() : __builtin_nan("");
    if (__builtin_isinf(y)) return 0.0;
    const double inv_n = 1.0 / (double)n;
    int e;
    const double m = frexp(y, &e);
    const double q = -(double)e * inv_n;
    const double k = floor(q);
    const float ex = (float)(q - k) - __builtin_amdgcn_logf((float)m) * (float)inv_n;
    double r = ldexp((double)__builtin_amdgcn_exp2f(ex), (int)k);
    for (int it = 0; it < 3; ++it) {
        double t = y;
        for (int i = 0; i < n; ++i) t = t * r;
        r = fma(r * (1.0 - t), inv_n, r);
    }
    return r;
}

// The same with the order fixed at compile time (NF = 1 or 2) or taken at
// run time (NF = 0: pow, as the reference; NF = -1: Newton steps, for the
// FP32 kernels' near pairs, where pow's registers cost the kernel half its
// waves).
template <int NF>
__device__ __forceinline__ double inv_root_n(double y, int n) {
    if constexpr (NF == 1) return 1.0 / y;
    if constexpr (NF == 2) return 1.0 / sqrt(y);
    if constexpr (NF == -1) return inv_root_newton(y, n);
    return inv_root(y, n);
}

template <int NF>
__device__ __forceinline__ double ipow_n(double x, int n) {
    if constexpr (NF == 1) return x;
    if constexpr (NF == 2) return x * x;
    return ipow(x, n);
}

// BiotSavart.m:86-115 for one pair: cut-off core (n == 0) or Vatistas
// order n (n >= 1, rc = r_c^2), each expression evaluated left to right as
// MATLAB does, ubar zeroed when NaN/Inf or when r1 > co and r2 > co
// (BiotSavart.m:111).  Returns the term ubar * (r1 x r2).  VISC selects the
// Vatistas branch at compile time (n >= 1), otherwise the cut-off (n = 0);
// NF = 1 or 2 fixes the Vatistas order at compile time, NF = 0 reads n.
template <bool VISC, int NF = 0>
__device__ __forceinline__ void matlab_pair(double x1, double y1, double z1,
                                            double x2, double y2, double z2,
                                            double g, double rc, double px,
                                            double py, double pz, double d,
                                            int n, double co, double &tx,
                                            double &ty, double &tz) {
    double x2x1 = x2 - x1, y2y1 = y2 - y1, z2z1 = z2 - z1;
    double L = x2x1 * x2x1 + y2y1 * y2y1 + z2z1 * z2z1;
    double pxx1 = px - x1, pyy1 = py - y1, pzz1 = pz - z1;
    double pxx2 = px - x2, pyy2 = py - y2, pzz2 = pz - z2;
    double r1 = sqrt(pxx1 * pxx1 + pyy1 * pyy1 + pzz1 * pzz1);
    double r2 = sqrt(pxx2 * pxx2 + pyy2 * pyy2 + pzz2 * pzz2);
    double r1dr2 = pxx1 * pxx2 + pyy1 * pyy2 + pzz1 * pzz2;
    double r1tr2 = r1 * r2;
    double ubar;
    if (VISC) {
        double ld = x2x1 * pxx1 + y2y1 * pyy1 + z2z1 * pzz1;
        double Ldr12 = ld * ld;
        double Cnu = (r1 * r1) - (Ldr12 / L);
        Cnu = Cnu * inv_root_n<NF>(ipow_n<NF>(rc, n) + ipow_n<NF>(Cnu, n), n);
        ubar = Cnu * g / (4.0 * kMlPi) * (r1 + r2) / (r1tr2 * (r1tr2 + r1dr2));
    } else {
        ubar = (g / (4.0 * kMlPi)) * (r1 + r2) / (r1tr2 * (r1tr2 + r1dr2) + (d * L));
    }
    if (__builtin_isnan(ubar) || __builtin_isinf(ubar) || (r1 > co && r2 > co)) ubar = 0.0;
    tx = ubar * (pyy1 * pzz2 - pzz1 * pyy2);
    ty = ubar * (pzz1 * pxx2 - pxx1 * pzz2);
    tz = ubar * (pxx1 * pyy2 - pyy1 * pxx2);
}

// Term of segment record s at (px, py, pz).  MODE 0: biot_p2p (MEX);
// 1: BiotSavart.m cut-off; 2: BiotSavart.m Vatistas, order NF (1, 2) or n
// (NF = 0).
template <int MODE, int NF>
__device__ __forceinline__ void fp64_term(const Seg64 &s, double px, double py, double pz,
                                          double d, int n, double co, double &tx, double &ty,
                                          double &tz) {
    if constexpr (MODE == 0)
        mex_term(s.ax, s.ay, s.az, s.bx, s.by, s.bz, s.g, px, py, pz, d, tx, ty, tz);
    else
        matlab_pair<MODE == 2, NF>(s.ax, s.ay, s.az, s.bx, s.by, s.bz, s.g, s.rc, px, py, pz, d,
                                   n, co, tx, ty, tz);
}

// Adds the terms of segments [k0, k1) to (sx, sy, sz) one after another in
// segment order -- the reference's summation, so the result is bit-identical
// to a plain loop -- while the terms of four segments are computed side by
// side: the FP64 pipeline has independent work while one term's square
// roots and divisions are in flight.
template <int MODE, int NF>
__device__ __forceinline__ void fp64_range(const Seg64 *__restrict__ seg, int64_t k0, int64_t k1,
                                           double px, double py, double pz, double d, int n,
                                           double co, double &sx, double &sy, double &sz) {
    int64_t k = k0;
    for (; k + 4 <= k1; k += 4) {
        double t[4][3];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            fp64_term<MODE, NF>(seg[k + i], px, py, pz, d, n, co, t[i][0], t[i][1], t[i][2]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            sx += t[i][0];
            sy += t[i][1];
            sz += t[i][2];
        }
    }
    for (; k < k1; ++k) {
        double tx, ty, tz;
        fp64_term<MODE, NF>(seg[k], px, py, pz, d, n, co, tx, ty, tz);
        sx += tx;
        sy += ty;
        sz += tz;
    }
}

}  // namespace winds
