/*
 * bs_oracle.c -- CPU restatement of the WInDS induced-velocity operator.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * never by the product library.  Nothing in winds-master_amd/ links it.
 *
 * Three restatements of the same sum  u(P_i) = sum_k ubar_ik * (r1 x r2):
 *
 *   oracle_bs_mex     BiotSavartMex CPU path: biot_p2p_nogpu + BiotSavart_nogpu
 *                     (cuda/source/BiotSavartMex.cu:236-282), targets and
 *                     segments packed as crunch_p / crunch_f / crunch_rg do
 *                     (BiotSavartMex.cu:316-379), i.e. linear index
 *                     k = m + n*ns + p*ns*nt, and the result unpacked as
 *                     reassem_uind (BiotSavartMex.cu:383-401).  Cut-off core
 *                     only; rc and co are ignored, nothing is zeroed (the
 *                     reference's own behaviour, BiotSavartMex.cu:87).
 *
 *   oracle_bs_matlab  core/BiotSavart.m:36-119: cut-off (numeric d) or
 *                     Vatistas order n ('viscN'), NaN/Inf zeroing and the far
 *                     cut-off co (BiotSavart.m:111), summed in MATLAB's
 *                     nested order sum(sum(sum(.,1),3),4) (BiotSavart.m:113).
 *
 *   oracle_bs_quad    the same BiotSavart.m arithmetic in __float128, used only
 *                     to bound the FP64 oracle's own rounding on tiny inputs.
 *
 * plus oracle_filament_length (core/FilamentLength.m:36-41).
 *
 * Pinning (tests/test_oracle.py): closed forms for a single straight segment,
 * an N-gon ring and the Vatistas perpendicular-bisector factor; linearity in
 * Gamma; sign flip on endpoint swap; rigid-motion equivariance; superposition
 * of shed+trail; exact zeros at endpoints and on the segment line; the quad
 * brute force.  The reference ships no golden vectors for this path.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -shared -fPIC (oracle/Makefile).
 * -ffp-contract=off keeps every product and sum rounded as written, so that
 * the FP64 GPU kernel can be compared bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <quadmath.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* BiotSavartMex.cu:35 uses this literal for pi. */
#define MEX_PI 3.14159265358979324
/* MATLAB's pi is the double nearest to pi, the same value. */
#define ML_PI 3.141592653589793

/* ------------------------------------------------------------------------ */
/* MEX path                                                                  */
/* ------------------------------------------------------------------------ */

/* One interaction, BiotSavartMex.cu:236-266 (biot_p2p_nogpu).  The segment is
 * given as the pair arrays the reference packs: fx=(x1,x2), fy, fz, and the
 * target as (px,py,pz).  Every operation is written in the reference's order. */
static inline void mex_p2p(double x1, double x2, double y1, double y2,
                           double z1, double z2, double g,
                           double px, double py, double pz, double cd,
                           double *ux, double *uy, double *uz)
{
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
    double ubar = (g * (r1 + r2)) * (1.000 / (4.000 * MEX_PI));
    ubar = ubar / den;
    *ux += ubar * (pyy1 * pzz2 - pzz1 * pyy2);
    *uy += ubar * (pzz1 * pxx2 - pxx1 * pzz2);
    *uz += ubar * (pxx1 * pyy2 - pyy1 * pxx2);
}

/* Column-major [ns,nd,nt,nb] element (m, c, n, p). */
static inline int64_t cm_idx(const int64_t *di, int64_t m, int64_t c,
                             int64_t n, int64_t p)
{
    return m + c * di[0] + n * di[0] * di[1] + p * di[0] * di[1] * di[2];
}

/*
 * BiotSavart_nogpu (BiotSavartMex.cu:268-282) behind biot_calc's packing.
 * F1,F2: [fns,3,fnt,fnb]; P: [pns,3,pnt,pnb]; gamma: [fns,1,fnt,fnb].
 * uind: [pns,3,pnt,pnb], overwritten.  Targets are independent and each is
 * summed over segments in packed order, so nthreads does not change a bit.
 */
void oracle_bs_mex(const double *F1, const double *F2, const int64_t *fdims,
                   const double *P, const int64_t *pdims,
                   const double *gamma, double d, double *uind, int nthreads)
{
    const int64_t fns = fdims[0], fnt = fdims[2], fnb = fdims[3];
    const int64_t pns = pdims[0], pnt = pdims[2], pnb = pdims[3];
    const int64_t nf = fns * fnt * fnb, np = pns * pnt * pnb;
    double *seg = (double *)malloc((size_t)nf * 7 * sizeof(double));
    for (int64_t p = 0, k = 0; p < fnb; p++)
        for (int64_t n = 0; n < fnt; n++)
            for (int64_t m = 0; m < fns; m++, k++) {
                seg[7 * k + 0] = F1[cm_idx(fdims, m, 0, n, p)];
                seg[7 * k + 1] = F2[cm_idx(fdims, m, 0, n, p)];
                seg[7 * k + 2] = F1[cm_idx(fdims, m, 1, n, p)];
                seg[7 * k + 3] = F2[cm_idx(fdims, m, 1, n, p)];
                seg[7 * k + 4] = F1[cm_idx(fdims, m, 2, n, p)];
                seg[7 * k + 5] = F2[cm_idx(fdims, m, 2, n, p)];
                seg[7 * k + 6] = gamma[k];
            }
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t j = 0; j < np; j++) {
        int64_t m = j % pns, n = (j / pns) % pnt, p = j / (pns * pnt);
        double px = P[cm_idx(pdims, m, 0, n, p)];
        double py = P[cm_idx(pdims, m, 1, n, p)];
        double pz = P[cm_idx(pdims, m, 2, n, p)];
        double ux = 0.0, uy = 0.0, uz = 0.0; /* memset, BiotSavartMex.cu:637-638 */
        for (int64_t k = 0; k < nf; k++) {
            const double *s = seg + 7 * k;
            mex_p2p(s[0], s[1], s[2], s[3], s[4], s[5], s[6], px, py, pz, d,
                    &ux, &uy, &uz);
        }
        uind[cm_idx(pdims, m, 0, n, p)] = ux;
        uind[cm_idx(pdims, m, 1, n, p)] = uy;
        uind[cm_idx(pdims, m, 2, n, p)] = uz;
    }
    free(seg);
}

/*
 * Packed-target variant of the same sum, used by the CPU baseline timing:
 * targets[3*np] as (x,y,z) triples, segments as the 7-double records above
 * are built from the same column-major inputs.  Only targets [j0, j1) are
 * evaluated, so a bounded sample of a large workload can be timed.
 */
void oracle_bs_mex_range(const double *F1, const double *F2,
                         const int64_t *fdims, const double *P,
                         const int64_t *pdims, const double *gamma, double d,
                         int64_t j0, int64_t j1, double *out3, int nthreads)
{
    const int64_t fns = fdims[0], fnt = fdims[2], fnb = fdims[3];
    const int64_t pns = pdims[0], pnt = pdims[2];
    const int64_t nf = fns * fnt * fnb;
    double *seg = (double *)malloc((size_t)nf * 7 * sizeof(double));
    for (int64_t p = 0, k = 0; p < fnb; p++)
        for (int64_t n = 0; n < fnt; n++)
            for (int64_t m = 0; m < fns; m++, k++) {
                seg[7 * k + 0] = F1[cm_idx(fdims, m, 0, n, p)];
                seg[7 * k + 1] = F2[cm_idx(fdims, m, 0, n, p)];
                seg[7 * k + 2] = F1[cm_idx(fdims, m, 1, n, p)];
                seg[7 * k + 3] = F2[cm_idx(fdims, m, 1, n, p)];
                seg[7 * k + 4] = F1[cm_idx(fdims, m, 2, n, p)];
                seg[7 * k + 5] = F2[cm_idx(fdims, m, 2, n, p)];
                seg[7 * k + 6] = gamma[k];
            }
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t j = j0; j < j1; j++) {
        int64_t m = j % pns, n = (j / pns) % pnt, p = j / (pns * pnt);
        double px = P[cm_idx(pdims, m, 0, n, p)];
        double py = P[cm_idx(pdims, m, 1, n, p)];
        double pz = P[cm_idx(pdims, m, 2, n, p)];
        double ux = 0.0, uy = 0.0, uz = 0.0;
        for (int64_t k = 0; k < nf; k++) {
            const double *s = seg + 7 * k;
            mex_p2p(s[0], s[1], s[2], s[3], s[4], s[5], s[6], px, py, pz, d,
                    &ux, &uy, &uz);
        }
        out3[3 * (j - j0) + 0] = ux;
        out3[3 * (j - j0) + 1] = uy;
        out3[3 * (j - j0) + 2] = uz;
    }
    free(seg);
}

/* ------------------------------------------------------------------------ */
/* BiotSavart.m path                                                         */
/* ------------------------------------------------------------------------ */

/* x.^n for a positive integer n, as repeated products. */
static inline double ipow(double x, int n)
{
    double r = x;
    for (int i = 1; i < n; i++) r = r * x;
    return r;
}

/* y.^(-1/n): 1/y for n=1, 1/sqrt(y) for n=2, pow otherwise. */
static inline double inv_root(double y, int n)
{
    if (n == 1) return 1.0 / y;
    if (n == 2) return 1.0 / sqrt(y);
    return pow(y, -1.0 / (double)n);
}

/* ubar*(r1 x r2) for one pair, BiotSavart.m:86-115, with MATLAB's
 * left-to-right evaluation of each expression. */
static inline void ml_p2p(double x1, double y1, double z1, double x2,
                          double y2, double z2, double g, double rc,
                          double px, double py, double pz, double d, int n,
                          double co, double *tx, double *ty, double *tz)
{
    double x2x1 = x2 - x1, y2y1 = y2 - y1, z2z1 = z2 - z1;
    double L = x2x1 * x2x1 + y2y1 * y2y1 + z2z1 * z2z1; /* BiotSavart.m:71 */
    double pxx1 = px - x1, pyy1 = py - y1, pzz1 = pz - z1;
    double pxx2 = px - x2, pyy2 = py - y2, pzz2 = pz - z2;
    double r1 = sqrt(pxx1 * pxx1 + pyy1 * pyy1 + pzz1 * pzz1);
    double r2 = sqrt(pxx2 * pxx2 + pyy2 * pyy2 + pzz2 * pzz2);
    double r1dr2 = pxx1 * pxx2 + pyy1 * pyy2 + pzz1 * pzz2;
    double r1tr2 = r1 * r2;
    double ubar;
    if (n != 0) { /* BiotSavart.m:102-106 */
        double ld = x2x1 * pxx1 + y2y1 * pyy1 + z2z1 * pzz1;
        double Ldr12 = ld * ld;
        double Cnu = (r1 * r1) - (Ldr12 / L);
        Cnu = Cnu * inv_root(ipow(rc, n) + ipow(Cnu, n), n);
        ubar = Cnu * g / (4.0 * ML_PI) * (r1 + r2) / (r1tr2 * (r1tr2 + r1dr2));
    } else { /* BiotSavart.m:108 */
        ubar = (g / (4.0 * ML_PI)) * (r1 + r2) / (r1tr2 * (r1tr2 + r1dr2) + (d * L));
    }
    if (isnan(ubar) || isinf(ubar) || (r1 > co && r2 > co)) ubar = 0.0; /* :111 */
    *tx = ubar * (pyy1 * pzz2 - pzz1 * pyy2);
    *ty = ubar * (pzz1 * pxx2 - pxx1 * pzz2);
    *tz = ubar * (pxx1 * pyy2 - pyy1 * pxx2);
}

/*
 * core/BiotSavart.m with type 'full'.  n = 0 selects the cut-off core with d;
 * n >= 1 selects the Vatistas core of order n (d = 'viscN').  rc holds r_c^2.
 * Per target the sum is taken over segment stations first, then ages, then
 * blades (BiotSavart.m:113-115).
 */
void oracle_bs_matlab(const double *F1, const double *F2, const int64_t *fdims,
                      const double *P, const int64_t *pdims,
                      const double *gamma, const double *rc, double d, int n,
                      double co, double *uind, int nthreads)
{
    const int64_t fns = fdims[0], fnt = fdims[2], fnb = fdims[3];
    const int64_t pns = pdims[0], pnt = pdims[2], pnb = pdims[3];
    const int64_t np = pns * pnt * pnb;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t j = 0; j < np; j++) {
        int64_t m = j % pns, nn = (j / pns) % pnt, p = j / (pns * pnt);
        double px = P[cm_idx(pdims, m, 0, nn, p)];
        double py = P[cm_idx(pdims, m, 1, nn, p)];
        double pz = P[cm_idx(pdims, m, 2, nn, p)];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int64_t fb = 0; fb < fnb; fb++) {
            double bx = 0.0, by = 0.0, bz = 0.0;
            for (int64_t ft = 0; ft < fnt; ft++) {
                double cx = 0.0, cy = 0.0, cz = 0.0;
                for (int64_t fs = 0; fs < fns; fs++) {
                    int64_t k = fs + ft * fns + fb * fns * fnt;
                    double tx, ty, tz;
                    ml_p2p(F1[cm_idx(fdims, fs, 0, ft, fb)],
                           F1[cm_idx(fdims, fs, 1, ft, fb)],
                           F1[cm_idx(fdims, fs, 2, ft, fb)],
                           F2[cm_idx(fdims, fs, 0, ft, fb)],
                           F2[cm_idx(fdims, fs, 1, ft, fb)],
                           F2[cm_idx(fdims, fs, 2, ft, fb)], gamma[k], rc[k],
                           px, py, pz, d, n, co, &tx, &ty, &tz);
                    cx += tx; cy += ty; cz += tz;
                }
                bx += cx; by += cy; bz += cz;
            }
            sx += bx; sy += by; sz += bz;
        }
        uind[cm_idx(pdims, m, 0, nn, p)] = sx;
        uind[cm_idx(pdims, m, 1, nn, p)] = sy;
        uind[cm_idx(pdims, m, 2, nn, p)] = sz;
    }
}

/* ------------------------------------------------------------------------ */
/* __float128 brute force of the BiotSavart.m arithmetic                     */
/* ------------------------------------------------------------------------ */

static inline __float128 ipowq(__float128 x, int n)
{
    __float128 r = x;
    for (int i = 1; i < n; i++) r = r * x;
    return r;
}

void oracle_bs_quad(const double *F1, const double *F2, const int64_t *fdims,
                    const double *P, const int64_t *pdims, const double *gamma,
                    const double *rc, double d, int n, double co, double *uind)
{
    const int64_t fns = fdims[0], fnt = fdims[2], fnb = fdims[3];
    const int64_t pns = pdims[0], pnt = pdims[2], pnb = pdims[3];
    const int64_t np = pns * pnt * pnb;
    const __float128 four_pi = 4 * M_PIq;
    for (int64_t j = 0; j < np; j++) {
        int64_t m = j % pns, nn = (j / pns) % pnt, p = j / (pns * pnt);
        __float128 px = P[cm_idx(pdims, m, 0, nn, p)];
        __float128 py = P[cm_idx(pdims, m, 1, nn, p)];
        __float128 pz = P[cm_idx(pdims, m, 2, nn, p)];
        __float128 sx = 0, sy = 0, sz = 0;
        for (int64_t fb = 0; fb < fnb; fb++)
            for (int64_t ft = 0; ft < fnt; ft++)
                for (int64_t fs = 0; fs < fns; fs++) {
                    int64_t k = fs + ft * fns + fb * fns * fnt;
                    __float128 x1 = F1[cm_idx(fdims, fs, 0, ft, fb)];
                    __float128 y1 = F1[cm_idx(fdims, fs, 1, ft, fb)];
                    __float128 z1 = F1[cm_idx(fdims, fs, 2, ft, fb)];
                    __float128 x2 = F2[cm_idx(fdims, fs, 0, ft, fb)];
                    __float128 y2 = F2[cm_idx(fdims, fs, 1, ft, fb)];
                    __float128 z2 = F2[cm_idx(fdims, fs, 2, ft, fb)];
                    __float128 g = gamma[k], r_c = rc[k];
                    __float128 lx = x2 - x1, ly = y2 - y1, lz = z2 - z1;
                    __float128 L = lx * lx + ly * ly + lz * lz;
                    __float128 ax = px - x1, ay = py - y1, az = pz - z1;
                    __float128 bx = px - x2, by = py - y2, bz = pz - z2;
                    __float128 r1 = sqrtq(ax * ax + ay * ay + az * az);
                    __float128 r2 = sqrtq(bx * bx + by * by + bz * bz);
                    __float128 dot = ax * bx + ay * by + az * bz;
                    __float128 rr = r1 * r2;
                    __float128 ubar;
                    if (n != 0) {
                        __float128 ld = lx * ax + ly * ay + lz * az;
                        __float128 C = r1 * r1 - ld * ld / L;
                        C = C * powq(ipowq(r_c, n) + ipowq(C, n), -1 / (__float128)n);
                        ubar = C * g / four_pi * (r1 + r2) / (rr * (rr + dot));
                    } else {
                        ubar = g / four_pi * (r1 + r2) / (rr * (rr + dot) + d * L);
                    }
                    if (isnanq(ubar) || isinfq(ubar) ||
                        (r1 > (__float128)co && r2 > (__float128)co))
                        ubar = 0;
                    sx += ubar * (ay * bz - az * by);
                    sy += ubar * (az * bx - ax * bz);
                    sz += ubar * (ax * by - ay * bx);
                }
        uind[cm_idx(pdims, m, 0, nn, p)] = (double)sx;
        uind[cm_idx(pdims, m, 1, nn, p)] = (double)sy;
        uind[cm_idx(pdims, m, 2, nn, p)] = (double)sz;
    }
}

/* core/FilamentLength.m:36-41: |F2 - F1| per segment, [fns,1,fnt,fnb]. */
void oracle_filament_length(const double *F1, const double *F2,
                            const int64_t *fdims, double *L)
{
    const int64_t nf = fdims[0] * fdims[2] * fdims[3];
    for (int64_t k = 0; k < nf; k++) {
        int64_t m = k % fdims[0], n = (k / fdims[0]) % fdims[2],
                p = k / (fdims[0] * fdims[2]);
        double dx = F2[cm_idx(fdims, m, 0, n, p)] - F1[cm_idx(fdims, m, 0, n, p)];
        double dy = F2[cm_idx(fdims, m, 1, n, p)] - F1[cm_idx(fdims, m, 1, n, p)];
        double dz = F2[cm_idx(fdims, m, 2, n, p)] - F1[cm_idx(fdims, m, 2, n, p)];
        L[k] = sqrt(dx * dx + dy * dy + dz * dz);
    }
}
