// march.hip -- device-resident free-vortex-wake time march (include/winds_march.h).
//
// State layout for a wake of A ages (MATLAB's column-major order of the
// reference's cell arrays, flattened):
//   nodes   SoA x,y,z over j = m + a*nst + b*nst*A      (wake.domain{p})
//   gs, rs  shed circulation / core^2 over m + a*ns + b*ns*A        [ns,1,A,nb]
//   gt, rt  trailing over m + a*nst + b*nst*(A-1)               [nst,1,A-1,nb]
//   len_s/t filament lengths, same layouts (vcore.m:42-46)
// Segments for the induced velocity are rebuilt every evaluation from nodes
// and circulation, shed first (F1 = node m, F2 = node m+1), then trailing
// (F1 = age a+1, F2 = age a), exactly the pairs of WInDS.m:150-154.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "bs_common.hpp"
#include "bs_kernels.hpp"
#include "bs_pair.hpp"
#include "surrogate.hpp"
#include "winds_march.h"
#include "winds_mlp.h"

namespace winds {
namespace mk {

constexpr double kAlphaRL = 1.25643;  // core/constants.m:22
constexpr double kNu = 1e-5;          // :23
constexpr double kDelta = 100.0;      // :24
constexpr double kA1 = 6.5e-5;        // :25
constexpr double kRho = 1.23;         // :28
constexpr double kPi = 3.141592653589793;

// Rotor description on the device.  Blade data per station; the polars of
// every airfoil concatenated (table k: entries [foil_off[k], +foil_len[k]));
// kinematics per time index in the reference's column-major layouts
// (pos.quarter/trail [nst,3,nt,nb], pos.bound, pos.nodes.bxn/byn/bzn,
// vel.blade [ns,3,nt,nb], blade pitch [nt,nb]).
struct Geom {
    const double *r_node, *chord_node, *twist_node, *dr_node;
    const int *foil;                  // table of station m (0-based)
    const int *foil_off, *foil_len;   // per table
    const double *alpha, *cl, *cd;    // concatenated tables
    const double *quarter, *trail;
    const double *bound, *bxn, *byn, *bzn, *vblade;
    const double *ll_extra;           // induced velocity of other sources at the
                                      // lifting line, or null
    const double *pitch, *rot_speed;
    int nb, nst, ns, nt, n_foils, n_alpha;
};

__device__ __forceinline__ double rad(double deg) { return deg * kPi / 180.0; }

// element (i, c, p0, b) of a column-major [n,3,nt,nb] table, p0 0-based
__device__ __forceinline__ double tab4(const double *t, int n, int nt, int i, int c, int p0,
                                       int b) {
    return t[i + (int64_t)n * (c + 3 * (p0 + (int64_t)nt * b))];
}

// Per-step values that a step's kernels read from device memory rather than
// as arguments, so that every step of a steady march enqueues the same
// launches and a recorded graph of them can be replayed.  k_set_step writes
// them before the step's launches.
struct StepDev {
    int p0;             // table index of the step's time index p (p - 1)
    int row;            // performance-history row of the step
    double p_index;     // p
    double r0[2];       // vcore's r0 at p and p + 1
    double T[2];        // fastout.Time at p and p + 1
    double wind[2][3];  // wind.infty at p and p + 1
};

__global__ void k_set_step(StepDev *__restrict__ sd, StepDev v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *sd = v;
}

// wake ages 0 and 1 of time index p0+1: pos.quarter and pos.trail
// (initials.m:87-88, kinematics.m:49-50)
// (sd: table index sd->p0 + dp instead of p0)
__global__ void k_nodes01(Geom g, int p0, int A, double *__restrict__ X,
                          const StepDev *__restrict__ sd, int dp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nst * g.nb) return;
    if (sd) p0 = sd->p0 + dp;
    const int m = i % g.nst, b = i / g.nst;
    const int64_t N = (int64_t)g.nst * A * g.nb;
    const int64_t j0 = m + (int64_t)b * g.nst * A;
    const int64_t j1 = j0 + g.nst;
    for (int c = 0; c < 3; ++c) {
        X[c * N + j0] = tab4(g.quarter, g.nst, g.nt, m, c, p0, b);
        X[c * N + j1] = tab4(g.trail, g.nst, g.nt, m, c, p0, b);
    }
}

// lifting-line point of station m of blade b at time index p0+1: bound
// point, blade frame bxn/byn/bzn and the kinematic inflow vel.blade
// (velocity.m:43-70)
struct LL {
    double px, py, pz;
    double bx[3], by[3], bz[3];
    double vblade[3];
};

__device__ __forceinline__ LL ll_frame(const Geom &g, int m, int b, int p0) {
    LL L;
    L.px = tab4(g.bound, g.ns, g.nt, m, 0, p0, b);
    L.py = tab4(g.bound, g.ns, g.nt, m, 1, p0, b);
    L.pz = tab4(g.bound, g.ns, g.nt, m, 2, p0, b);
    for (int c = 0; c < 3; ++c) {
        L.bx[c] = tab4(g.bxn, g.ns, g.nt, m, c, p0, b);
        L.by[c] = tab4(g.byn, g.ns, g.nt, m, c, p0, b);
        L.bz[c] = tab4(g.bzn, g.ns, g.nt, m, c, p0, b);
        L.vblade[c] = tab4(g.vblade, g.ns, g.nt, m, c, p0, b);
    }
    return L;
}

// interp1(alpha, tab, x) linear, NaN outside the table
__device__ double interp1(const double *xp, const double *fp, int n, double x) {
    if (!(x >= xp[0] && x <= xp[n - 1])) return __builtin_nan("");
    if (x == xp[n - 1]) return fp[n - 1];
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (xp[mid] <= x)
            lo = mid;
        else
            hi = mid;
    }
    const double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
    return slope * (x - xp[lo]) + fp[lo];
}

// interp1 of two tables over the same abscissae (MATLAB interp1 'linear':
// NaN outside [xp[0], xp[n-1]]), with one search for both.  The search
// starts from the interval of the previous call (hint; -1 for none): the
// interval with xp[lo] <= x < xp[lo + 1] is unique, so a hint that brackets
// x gives the binary search's answer; otherwise the search runs.  sa, sb
// hold the hinted interval's slopes (set whenever the hint changes).
__device__ __forceinline__ void interp1x2_hint(const double *xp, const double *fa,
                                               const double *fb, int n, double x, double &ya,
                                               double &yb, int &hint, double &sa, double &sb) {
    if (!(x >= xp[0] && x <= xp[n - 1])) {
        ya = yb = __builtin_nan("");
        return;
    }
    if (x == xp[n - 1]) {
        ya = fa[n - 1];
        yb = fb[n - 1];
        return;
    }
    int lo = hint;
    if (!(lo >= 0 && lo < n - 1 && xp[lo] <= x && x < xp[lo + 1])) {
        lo = 0;
        int hi = n - 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (xp[mid] <= x)
                lo = mid;
            else
                hi = mid;
        }
        hint = lo;
        // the interval's slopes, kept while the interval stays
        const double dx = xp[lo + 1] - xp[lo];
        sa = (fa[lo + 1] - fa[lo]) / dx;
        sb = (fb[lo + 1] - fb[lo]) / dx;
    }
    ya = sa * (x - xp[lo]) + fa[lo];
    yb = sb * (x - xp[lo]) + fb[lo];
}

// initials.m:110-125: bound circulation 0.5*|U(1)|*chord*cl, cl the
// initial lift coefficient of each station (BEM.m or the elliptic wing's
// closed form, initials.m:111-119), the same on every blade
__global__ void k_init_bound(Geom g, double umag, const double *__restrict__ cl0, int A,
                             double *__restrict__ gs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.ns * g.nb) return;
    const int m = i % g.ns, b = i / g.ns;
    gs[m + (int64_t)b * g.ns * A] = 0.5 * umag * g.chord_node[m] * cl0[m];
}

// Kelvin (WInDS.m:131-140), trailing part: age 0 = diff([0; bound; 0]),
// ages 1.. = previous trailing ages 0.. (the oldest dropped when truncated)
__global__ void k_kelvin_trail(int nst, int nb, const double *__restrict__ gs_old, int A_old,
                               const double *__restrict__ gt_old, int A_new,
                               double *__restrict__ gt_new) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ns = nst - 1;
    const int64_t n = (int64_t)nst * (A_new - 1) * nb;
    if (i >= n) return;
    const int m = (int)(i % nst);
    const int a = (int)((i / nst) % (A_new - 1));
    const int b = (int)(i / ((int64_t)nst * (A_new - 1)));
    double v;
    if (a == 0) {
        const double *gb = gs_old + (int64_t)b * ns * A_old;  // bound = shed age 0
        const double hi = m < ns ? gb[m] : 0.0;
        const double lo = m > 0 ? gb[m - 1] : 0.0;
        v = hi - lo;
    } else {
        v = gt_old[m + (int64_t)(a - 1) * nst + (int64_t)b * nst * (A_old - 1)];
    }
    gt_new[i] = v;
}

// Kelvin, shed part: age 0 = bound; age a>=1 = Cz(a) - C(a-1), C = span
// cumulative sum of trailing ages, Cz(A-1) = 0 (WInDS.m:139-140).  A block
// takes kShedAges ages of one blade: the trailing rows it reads and the
// shed rows it writes are contiguous, so they pass through LDS with
// coalesced accesses, and each thread runs its age's cumulative sums in
// station order.  Dynamic LDS: kelvin_shed_lds(nst) bytes.
constexpr int kShedAges = 32;
inline size_t kelvin_shed_lds(int nst) {
    return ((size_t)(kShedAges + 1) * nst + (size_t)kShedAges * (nst - 1)) * sizeof(double);
}
__global__ __launch_bounds__(kShedAges) void k_kelvin_shed(int nst, int nb,
                                                           const double *__restrict__ gs_old,
                                                           int A_old,
                                                           const double *__restrict__ gt_new,
                                                           int A_new,
                                                           double *__restrict__ gs_new) {
    extern __shared__ double s_kel[];
    const int ns = nst - 1;
    const int nta = A_new - 1;  // trailing ages
    const int blocks_per_blade = (A_new + kShedAges - 1) / kShedAges;
    const int b = blockIdx.x / blocks_per_blade;
    const int a0 = (blockIdx.x % blocks_per_blade) * kShedAges;
    const int na = A_new - a0 < kShedAges ? A_new - a0 : kShedAges;  // ages of this block
    double *tr = s_kel;                               // trailing rows a0-1 .. a0+na-1
    double *sh = s_kel + (size_t)(kShedAges + 1) * nst;  // shed rows a0 .. a0+na-1
    const double *gtb = gt_new + (int64_t)b * nst * nta;
    const int r0 = a0 - 1 > 0 ? a0 - 1 : 0;           // first trailing row read
    const int r1 = a0 + na < nta ? a0 + na : nta;     // one past the last
    for (int k = threadIdx.x; k < (r1 - r0) * nst; k += kShedAges)
        tr[(r0 - (a0 - 1)) * nst + k] = gtb[(int64_t)r0 * nst + k];
    __syncthreads();
    const int i = threadIdx.x;
    if (i < na) {
        const int a = a0 + i;
        double *out = sh + (size_t)i * ns;
        if (a == 0) {
            const double *gb = gs_old + (int64_t)b * ns * A_old;
            for (int m = 0; m < ns; ++m) out[m] = gb[m];
        } else {
            const double *tp = tr + (size_t)i * nst;  // trailing age a-1
            const double *tc = a <= nta - 1 ? tp + nst : nullptr;
            double cprev = 0.0, ccur = 0.0;
            for (int m = 0; m < ns; ++m) {
                cprev += tp[m];
                if (tc) ccur += tc[m];
                out[m] = (tc ? ccur : 0.0) - cprev;
            }
        }
    }
    __syncthreads();
    double *dst = gs_new + (int64_t)b * ns * A_new + (int64_t)a0 * ns;
    for (int k = threadIdx.x; k < na * ns; k += kShedAges) dst[k] = sh[k];
}

// vcore.m:30-51 + FilamentLength.m + filamentmod.m:27-41 for one time index
// (sd: r0 and T of step slot `slot`, 0 for p, 1 for p + 1)
__global__ void k_vcore(int nst, int nb, int A, const double *__restrict__ X,
                        const double *__restrict__ gs, const double *__restrict__ gt, double r0,
                        double T, int fmod, const double *__restrict__ ls_prev,
                        const double *__restrict__ lt_prev, int A_prev,
                        double *__restrict__ ls, double *__restrict__ lt,
                        double *__restrict__ rs, double *__restrict__ rt,
                        const StepDev *__restrict__ sd, int slot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (sd) {
        r0 = sd->r0[slot];
        T = sd->T[slot];
    }
    const int ns = nst - 1;
    const int64_t N = (int64_t)nst * A * nb;
    const int64_t n_s = (int64_t)ns * A * nb;
    const int64_t n_t = (int64_t)nst * (A - 1) * nb;
    if (i >= n_s + n_t) return;
    const bool shed = i < n_s;
    const int64_t k = shed ? i : i - n_s;
    int m, a, b;
    int64_t j1, j2;  // F1, F2 node indices
    if (shed) {
        m = (int)(k % ns);
        a = (int)((k / ns) % A);
        b = (int)(k / ((int64_t)ns * A));
        j1 = m + (int64_t)a * nst + (int64_t)b * nst * A;
        j2 = j1 + 1;
    } else {
        m = (int)(k % nst);
        a = (int)((k / nst) % (A - 1));
        b = (int)(k / ((int64_t)nst * (A - 1)));
        j2 = m + (int64_t)a * nst + (int64_t)b * nst * A;
        j1 = j2 + nst;
    }
    const double dx = X[j2] - X[j1], dy = X[N + j2] - X[N + j1], dz = X[2 * N + j2] - X[2 * N + j1];
    const double len = sqrt(dx * dx + dy * dy + dz * dz);
    const double gam = shed ? gs[k] : gt[k];
    const double Re = fabs(gam / kNu);
    const double rc = r0 * r0 + 4.0 * kAlphaRL * kNu * (1.0 + kA1 * Re) * T;
    double rce = rc;
    if (fmod) {
        // interior ages 2..end-1 (1-based), compared at the same age index
        const bool inner = shed ? (a >= 1 && a <= A - 2) : (a >= 1 && a <= A - 3);
        if (inner) {
            const int64_t kp = shed ? m + (int64_t)a * ns + (int64_t)b * ns * A_prev
                                    : m + (int64_t)a * nst + (int64_t)b * nst * (A_prev - 1);
            const double lold = sqrt(shed ? ls_prev[kp] : lt_prev[kp]);
            const double lnew = sqrt(len);
            const double strain = (lnew - lold) / lold;
            rce = rc * (1.0 / (1.0 + strain));
        }
    }
    if (shed) {
        ls[k] = len;
        rs[k] = rce;
    } else {
        lt[k] = len;
        rt[k] = rce;
    }
}

// Segment records from nodes and circulation (shed first, in the reference's
// order; then trailing, line by line), for shed ages [s0, s1) and trailing
// ages [t0, t1) of every blade.
// zero_changing zeroes the circulation of the filaments the Kutta-Joukowski
// iteration changes: shed ages 0 and 1, trailing age 0.
__global__ void k_segments(int nst, int nb, int A, const double *__restrict__ X,
                           const double *__restrict__ gs, const double *__restrict__ gt,
                           const double *__restrict__ rs, const double *__restrict__ rt,
                           int zero_changing, int s0, int s1, int t0, int t1,
                           Seg64 *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ns = nst - 1;
    const int ws = s1 - s0, wt = t1 - t0;
    const int64_t N = (int64_t)nst * A * nb;
    const int64_t n_s = (int64_t)ns * ws * nb;
    const int64_t n_t = (int64_t)nst * wt * nb;
    if (i >= n_s + n_t) return;
    Seg64 s;
    int64_t j1, j2;
    if (i < n_s) {
        const int m = (int)(i % ns);
        const int a = s0 + (int)((i / ns) % ws);
        const int b = (int)(i / ((int64_t)ns * ws));
        const int64_t k = m + (int64_t)a * ns + (int64_t)b * ns * A;
        j1 = m + (int64_t)a * nst + (int64_t)b * nst * A;
        j2 = j1 + 1;
        s.g = (zero_changing && a <= 1) ? 0.0 : gs[k];
        s.rc = rs[k];
    } else {
        // line by line (blade, station, age from the oldest down):
        // consecutive records are neighbouring filaments
        const int64_t q = i - n_s;
        const int b = (int)(q / ((int64_t)nst * wt));
        const int64_t r = q % ((int64_t)nst * wt);
        const int m = (int)(r / wt);
        const int a = t1 - 1 - (int)(r % wt);
        const int64_t k = m + (int64_t)a * nst + (int64_t)b * nst * (A - 1);
        j2 = m + (int64_t)a * nst + (int64_t)b * nst * A;
        j1 = j2 + nst;
        s.g = (zero_changing && a == 0) ? 0.0 : gt[k];
        s.rc = rt[k];
    }
    s.ax = X[j1];
    s.ay = X[N + j1];
    s.az = X[2 * N + j1];
    s.bx = X[j2];
    s.by = X[N + j2];
    s.bz = X[2 * N + j2];
    out[i] = s;
}

// Near-wake targets: the nodes of age < St, compacted blade-major into
// SoA [3][nst*St*nb].
__global__ void k_gather_near(int nst, int nb, int A, int St, const double *__restrict__ X,
                              double *__restrict__ T) {
    const int64_t n = (int64_t)nst * St * nb;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = (int)(i % nst);
    const int a = (int)((i / nst) % St);
    const int b = (int)(i / ((int64_t)nst * St));
    const int64_t N = (int64_t)nst * A * nb;
    const int64_t j = m + (int64_t)a * nst + (int64_t)b * nst * A;
    T[i] = X[j];
    T[n + i] = X[N + j];
    T[2 * n + i] = X[2 * N + j];
}

// Induced velocity at every node: nodes of age < St from unear (compacted as
// k_gather_near), older nodes from the store taken at the last direct
// evaluation, kept in the frame turning with the rotor (cos/sin of the
// blade-0 azimuth c, s) and indexed by age (ages beyond the store's Aref use
// its oldest age).
__global__ void k_compose(int nst, int nb, int A, int St, const double *__restrict__ unear,
                          const double *__restrict__ ufz, int Aref, double c, double s,
                          double *__restrict__ u) {
    const int64_t N = (int64_t)nst * A * nb;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const int m = (int)(j % nst);
    const int a = (int)((j / nst) % A);
    const int b = (int)(j / ((int64_t)nst * A));
    if (a < St) {
        const int64_t n = (int64_t)nst * St * nb;
        const int64_t i = m + (int64_t)a * nst + (int64_t)b * nst * St;
        u[j] = unear[i];
        u[N + j] = unear[n + i];
        u[2 * N + j] = unear[2 * n + i];
    } else {
        const int ar = min(a, Aref - 1);
        const int64_t Nr = (int64_t)nst * Aref * nb;
        const int64_t k = m + (int64_t)ar * nst + (int64_t)b * nst * Aref;
        const double vy = ufz[Nr + k], vz = ufz[2 * Nr + k];
        u[j] = ufz[k];
        u[N + j] = vy * c - vz * s;
        u[2 * N + j] = vy * s + vz * c;
    }
}

// ufz = the induced velocity u in the frame turning with the rotor
__global__ void k_store_turning(int64_t N, double c, double s, const double *__restrict__ u,
                                double *__restrict__ ufz) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const double uy = u[N + j], uz = u[2 * N + j];
    ufz[j] = u[j];
    ufz[N + j] = uy * c + uz * s;
    ufz[2 * N + j] = -uy * s + uz * c;
}

__global__ void k_accumulate(int64_t n, const double *__restrict__ a, double *__restrict__ u) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] += a[i];
}

__global__ void k_zero(int64_t n, double *__restrict__ u) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = 0.0;
}

// out[0] = max_i |a_i - b_i|, out[1] = max_i |b_i| over points (SoA [3][n]);
// non-negative doubles order like their bit patterns.
__global__ void k_max_err(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                          unsigned long long *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0, r = 0.0;
    if (i < n) {
        const double dx = a[i] - b[i], dy = a[n + i] - b[n + i], dz = a[2 * n + i] - b[2 * n + i];
        e = sqrt(dx * dx + dy * dy + dz * dz);
        r = sqrt(b[i] * b[i] + b[n + i] * b[n + i] + b[2 * n + i] * b[2 * n + i]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        e = fmax(e, __shfl_down(e, off));
        r = fmax(r, __shfl_down(r, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(out, (unsigned long long)__double_as_longlong(e));
        atomicMax(out + 1, (unsigned long long)__double_as_longlong(r));
    }
}

__global__ void k_add_uinf(int64_t n, double ux, double uy, double uz,
                           const double *__restrict__ u, double *__restrict__ v,
                           const StepDev *__restrict__ sd, int slot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (sd) {
        ux = sd->wind[slot][0];
        uy = sd->wind[slot][1];
        uz = sd->wind[slot][2];
    }
    v[i] = ux + u[i];
    v[n + i] = uy + u[n + i];
    v[2 * n + i] = uz + u[2 * n + i];
}

// New ages 2.. from the old nodes: forward Euler (fe.m:31), PCC corrector
// (pcc.m:74-75, vpred != null) or Adams-Bashforth 4 (ab4.m:38-42, h1 != null)
// ab: 4 = ab4.m (h1..h3 the velocities of the three earlier time levels),
// 2 = ab2.m (h1 only), 0 = fe.m, or pcc.m's corrector when vpred is given
__global__ void k_advect(int nst, int nb, int A_old, int A_new, double dt,
                         const double *__restrict__ Xo, const double *__restrict__ vo,
                         const double *__restrict__ vpred, int ab, const double *__restrict__ h1,
                         int A1, const double *__restrict__ h2, int A2,
                         const double *__restrict__ h3, int A3, double *__restrict__ Xn) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nadv = A_new - 2;
    if (nadv <= 0) return;
    const int64_t n = (int64_t)nst * nadv * nb;
    if (i >= n) return;
    const int m = (int)(i % nst);
    const int ai = (int)((i / nst) % nadv);  // slice index, new age ai+2
    const int b = (int)(i / ((int64_t)nst * nadv));
    const int64_t No = (int64_t)nst * A_old * nb;
    const int64_t Nn = (int64_t)nst * A_new * nb;
    const int64_t jo = m + (int64_t)(ai + 1) * nst + (int64_t)b * nst * A_old;
    const int64_t jn = m + (int64_t)(ai + 2) * nst + (int64_t)b * nst * A_new;
    for (int c = 0; c < 3; ++c) {
        const double src = Xo[c * No + jo];
        const double v0 = vo[c * No + jo];
        double x;
        if (vpred) {
            x = src + dt / 2 * (vpred[c * Nn + jn] + v0);
        } else if (ab == 2) {
            // ab2.m:38-39: the earlier level's velocity one age younger
            double sh = 0.0;
            if (ai >= 1) {
                const int64_t N1 = (int64_t)nst * A1 * nb;
                sh = h1[c * N1 + m + (int64_t)ai * nst + (int64_t)b * nst * A1];
            }
            x = src + dt / 2 * (3 * v0 - sh);
        } else if (ab == 4) {
            double acc = 55 * v0;
            const double *hs[3] = {h1, h2, h3};
            const int As[3] = {A1, A2, A3};
            const double coef[3] = {-59, 37, -9};
            for (int k = 1; k <= 3; ++k) {
                double sh = 0.0;
                if (ai >= k) {
                    const int age = 1 + ai - k;
                    const int64_t Nk = (int64_t)nst * As[k - 1] * nb;
                    sh = hs[k - 1][c * Nk + m + (int64_t)age * nst + (int64_t)b * nst * As[k - 1]];
                }
                acc = acc + coef[k - 1] * sh;
            }
            x = src + dt / 24 * acc;
        } else {
            x = src + dt * v0;
        }
        Xn[c * Nn + jn] = x;
    }
}

__global__ void k_ll_points(Geom g, int p0, double *__restrict__ P,
                            const StepDev *__restrict__ sd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = g.ns * g.nb;
    if (i >= n) return;
    if (sd) p0 = sd->p0;
    const LL L = ll_frame(g, i % g.ns, i / g.ns, p0);
    P[i] = L.px;
    P[n + i] = L.py;
    P[2 * n + i] = L.pz;
}

// Unit-circulation influence of the filaments the KJ iteration changes on
// the lifting-line points: K[(t*NCH + s)*3 + c].  Per blade the changing
// filaments are shed age 0 (ns), shed age 1 (ns), trailing age 0 (nst).
__global__ void k_influence(Geom g, int A, const double *__restrict__ X,
                            const double *__restrict__ rs, const double *__restrict__ rt,
                            const double *__restrict__ P, int sem, int vat_n, double d, double co,
                            double *__restrict__ K) {
    const int nst = g.nst, ns = g.ns, nb = g.nb;
    const int per = 2 * ns + nst;
    const int NCH = per * nb;
    const int NLL = ns * nb;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)NLL * NCH) return;
    const int t = (int)(i / NCH);
    const int s = (int)(i % NCH);
    const int b = s / per, r = s % per;
    const int64_t N = (int64_t)nst * A * nb;
    int64_t j1, j2;
    double rc;
    if (r < 2 * ns) {
        const int a = r / ns, m = r % ns;
        j1 = m + (int64_t)a * nst + (int64_t)b * nst * A;
        j2 = j1 + 1;
        rc = rs[m + (int64_t)a * ns + (int64_t)b * ns * A];
    } else {
        const int m = r - 2 * ns;
        j2 = m + (int64_t)b * nst * A;
        j1 = j2 + nst;
        rc = rt[m + (int64_t)b * nst * (A - 1)];
    }
    const double px = P[t], py = P[NLL + t], pz = P[2 * NLL + t];
    double ux = 0.0, uy = 0.0, uz = 0.0;
    if (sem == 0) {
        mex_pair(X[j1], X[N + j1], X[2 * N + j1], X[j2], X[N + j2], X[2 * N + j2], 1.0, px, py, pz,
                 d, ux, uy, uz);
    } else if (vat_n == 0) {
        matlab_pair<false>(X[j1], X[N + j1], X[2 * N + j1], X[j2], X[N + j2], X[2 * N + j2], 1.0,
                           rc, px, py, pz, d, 0, co, ux, uy, uz);
    } else {
        matlab_pair<true>(X[j1], X[N + j1], X[2 * N + j1], X[j2], X[N + j2], X[2 * N + j2], 1.0,
                          rc, px, py, pz, d, vat_n, co, ux, uy, uz);
    }
    K[i * 3 + 0] = ux;
    K[i * 3 + 1] = uy;
    K[i * 3 + 2] = uz;
}

// The changing filaments' velocity at the lifting line as a function of the
// bound circulation alone.  Per blade b the changing circulation is, from
// the bound values G[m] (KuttaJoukowski.m:285-288): shed age 0 G[m], shed
// age 1 c1[m] - G[m] (c1 the span cumsum of trailing age 1, fixed during
// the iteration), trailing age 0 G[m] - G[m-1] (G[-1] = G[ns] = 0).  So
// u_i = u_frozen,i + sum_m K_i(shed 1, m) c1[m] + sum_{b,m} M_i(b,m) G[b][m]
// with M_i(b,m) = K_i(shed 0, m) - K_i(shed 1, m) + K_i(trail 0, m)
// - K_i(trail 0, m+1).  M [NLL][NLL][3] is a third of K, and the iteration
// reads only it.
// M is stored [j][component][i], so that the points i of one bound value j
// and component are adjacent.
__global__ void k_kj_matrix(Geom g, const double *__restrict__ K, double *__restrict__ M) {
    const int ns = g.ns, nst = g.nst, nb = g.nb;
    const int per = 2 * ns + nst, NCH = per * nb, NLL = ns * nb;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)NLL * NLL) return;
    const int j = (int)(k / NLL), i = (int)(k % NLL);
    const int b = j / ns, m = j % ns;
    const double *Ki = K + (int64_t)i * NCH * 3;
    const int s0 = b * per + m, s1 = b * per + ns + m, t0 = b * per + 2 * ns + m;
    for (int c = 0; c < 3; ++c)
        M[((int64_t)j * 3 + c) * NLL + i] =
            ((Ki[3 * s0 + c] - Ki[3 * s1 + c]) + Ki[3 * t0 + c]) - Ki[3 * (t0 + 1) + c];
}

constexpr int kKjThreads = 1024;
constexpr int kKjWaves = kKjThreads / 64;
constexpr int kMaxLL = 512;      // lifting-line points handled by one workgroup
constexpr int kMaxNodesBlade = 256;
constexpr int kMaxAlpha = 1024;  // polar table entries (all tables) held in LDS by the KJ kernel
constexpr int kMaxFoils = 64;    // airfoil tables

// The KJ kernel's LDS, sized to the rotor at launch (offsets in doubles):
// bound circulation, span cumsum of trailing age 1, the frozen velocity,
// the per-point constants, partial sums, the polars and, when it fits,
// the changing filaments' matrix in each point's chord frame.
struct KjLds {
    int kparts, g, c1, fz, ll, part, al, tcl, tcd, mp, total;
};
__host__ __device__ inline KjLds kj_lds(int NLL, int n_alpha, bool with_m) {
    KjLds L;
    L.kparts = kKjThreads / ((NLL + 63) & ~63);
    int o = 0;
    L.g = o;
    o += NLL;
    L.c1 = o;
    o += NLL;
    L.fz = o;
    o += 3 * NLL;
    L.ll = o;
    o += 10 * NLL;
    L.part = o;
    o += 2 * L.kparts * NLL;
    L.al = o;
    o += n_alpha;
    L.tcl = o;
    o += n_alpha;
    L.tcd = o;
    o += n_alpha;
    L.mp = o;
    if (with_m) o += 2 * NLL * NLL;
    L.total = o;
    return L;
}

// Kutta-Joukowski fixed point (KuttaJoukowski.m:42-49 with kj() :204-296)
// for time index p0+1, in FP64, in one workgroup, followed by
// perform.m:16-62.  frozen: induced velocity at the lifting line of every
// filament except the changing ones; K their unit influence; other sources
// (the tower, KuttaJoukowski.m:232-238) come from g.ll_extra.  Each station
// reads the polar of its airfoil (blade.NFoil, KuttaJoukowski.m:262-269).
// Writes the bound circulation, the trailing age 0 and shed age 1
// circulation, cl/cd/aoa and one history row.
template <bool MLDS>
__global__ __launch_bounds__(kKjThreads) void k_kj(Geom g, int p0, int A, double relax,
                                                  double tol, int maxiter,
                                                  const double *__restrict__ frozen,
                                                  const double *__restrict__ K,
                                                  const double *__restrict__ M,
                                                  double *__restrict__ gs, double *__restrict__ gt,
                                                  double *__restrict__ cl_out,
                                                  double *__restrict__ cd_out,
                                                  double *__restrict__ aoa_out,
                                                  double *__restrict__ hist_row, double p_index,
                                                  int *__restrict__ flag,
                                                  const StepDev *__restrict__ sd) {
    // sd: the table index, the history row (hist_row is then the history's
    // start) and p from the step's device values
    if (sd) {
        p0 = sd->p0;
        hist_row += 5 * (int64_t)sd->row;
        p_index = sd->p_index;
    }
    const int nst = g.nst, ns = g.ns, nb = g.nb;
    const int NLL = ns * nb;
    const int per = 2 * ns + nst;
    const int NCH = per * nb;
    extern __shared__ double s_dyn[];
    const KjLds L = kj_lds(NLL, g.n_alpha, MLDS);
    double *s_g = s_dyn + L.g;        // bound circulation
    double *s_c1 = s_dyn + L.c1;      // span cumsum of trailing age 1
    double *s_part = s_dyn + L.part;  // partial sums of the changing filaments
    double *s_al = s_dyn + L.al, *s_tcl = s_dyn + L.tcl, *s_tcd = s_dyn + L.tcd;  // the polars
    double *s_fz = s_dyn + L.fz;  // the frozen filaments' velocity at the points
    double *s_ll = s_dyn + L.ll;  // per point: chord frame, inflow, |inflow|, chord
    double *s_mp = s_dyn + L.mp;  // M in the chord frames (MLDS)
    __shared__ double s_red[kKjWaves];
    __shared__ double s_red2[kKjWaves];
    __shared__ int s_foff[kMaxFoils], s_flen[kMaxFoils];
    __shared__ int s_nan;
    const int tid = threadIdx.x;
    for (int i = tid; i < NLL; i += kKjThreads) {
        const int m = i % ns, b = i / ns;
        s_g[i] = gs[m + (int64_t)b * ns * A];
    }
    for (int i = tid; i < g.n_alpha; i += kKjThreads) {
        s_al[i] = g.alpha[i];
        s_tcl[i] = g.cl[i];
        s_tcd[i] = g.cd[i];
    }
    for (int i = tid; i < g.n_foils; i += kKjThreads) {
        s_foff[i] = g.foil_off[i];
        s_flen[i] = g.foil_len[i];
    }
    // Thread i < NLL owns lifting-line point i (NLL <= kMaxLL < kKjThreads).
    // Its frame, inflow, chord and frozen velocity do not change during the
    // iteration: computed once, kept in LDS (in registers they pushed the
    // kernel past its 128 VGPRs into spills).
    if (tid < NLL) {
        const int m = tid % ns, b = tid / ns;
        const LL L = ll_frame(g, m, b, p0);
        double *q = s_ll + 10 * tid;
        q[0] = L.bx[0];
        q[1] = L.bx[1];
        q[2] = L.bx[2];
        q[3] = L.by[0];
        q[4] = L.by[1];
        q[5] = L.by[2];
        q[6] = L.vblade[0];
        q[7] = L.vblade[1];
        q[8] = sqrt(L.vblade[0] * L.vblade[0] + L.vblade[1] * L.vblade[1] +
                    L.vblade[2] * L.vblade[2]);
        q[9] = g.chord_node[m];
        double ex = 0.0, ey = 0.0, ez = 0.0;
        if (g.ll_extra) {
            ex = tab4(g.ll_extra, ns, g.nt, m, 0, p0, b);
            ey = tab4(g.ll_extra, ns, g.nt, m, 1, p0, b);
            ez = tab4(g.ll_extra, ns, g.nt, m, 2, p0, b);
        }
        s_fz[tid] = frozen[tid] + ex;
        s_fz[NLL + tid] = frozen[NLL + tid] + ey;
        s_fz[2 * NLL + tid] = frozen[2 * NLL + tid] + ez;
    }
    if (tid < nb) {
        const int b = tid;
        const double *t1 = gt + (int64_t)b * nst * (A - 1) + nst;  // trailing age 1
        double c = 0.0;
        for (int m = 0; m < ns; ++m) {
            if (A - 1 >= 2) c += t1[m];
            s_c1[b * ns + m] = A - 1 >= 2 ? c : 0.0;
        }
    }
    if (tid == 0) s_nan = 0;
    // split of the bound values for the changing filaments' sum: kparts
    // ranges, each over NLL points rounded up to a wave
    const int nllw = (NLL + 63) & ~63;
    const int kparts = L.kparts;
    const int kpart = tid / nllw, ki = tid % nllw;
    __syncthreads();
    // the part that does not change during the iteration: the frozen
    // filaments and shed age 1's c1 share (shed age 1 = c1 - G), in the
    // chord frame of the point (the iteration needs only the two in-plane
    // components)
    double rx0 = 0.0, ry0 = 0.0;
    if (tid < NLL) {
        const double *Ki = K + (int64_t)tid * NCH * 3;
        double cx = 0.0, cy = 0.0, cz = 0.0;
        for (int b = 0; b < nb; ++b)
            for (int m = 0; m < ns; ++m) {
                const double c1 = s_c1[b * ns + m];
                const int sj = b * per + ns + m;
                cx += Ki[3 * sj] * c1;
                cy += Ki[3 * sj + 1] * c1;
                cz += Ki[3 * sj + 2] * c1;
            }
        const double ux = s_fz[tid] + cx, uy = s_fz[NLL + tid] + cy, uz = s_fz[2 * NLL + tid] + cz;
        const double *q = s_ll + 10 * tid;
        rx0 = q[0] * ux + q[1] * uy + q[2] * uz;
        ry0 = q[3] * ux + q[4] * uy + q[5] * uz;
    }
    // The changing filaments' share from the bound values: they split into
    // kparts ranges, one thread per (range, point ki); each term is M's
    // column for (ki, j) turned into ki's chord frame.  With MLDS the terms
    // are formed once into LDS; otherwise they are formed from M at every
    // iteration, with the same arithmetic.
    const int jlo = (int)((int64_t)kpart * NLL / kparts);
    const int jhi = (int)((int64_t)(kpart + 1) * NLL / kparts);
    const bool active = kpart < kparts && ki < NLL;
    double bxk[3] = {0.0, 0.0, 0.0}, byk[3] = {0.0, 0.0, 0.0};
    if (active) {
        const double *q = s_ll + 10 * ki;
        bxk[0] = q[0];
        bxk[1] = q[1];
        bxk[2] = q[2];
        byk[0] = q[3];
        byk[1] = q[4];
        byk[2] = q[5];
        if constexpr (MLDS)
            for (int jj = jlo; jj < jhi; ++jj) {
                const double *Mj = M + (int64_t)jj * 3 * NLL + ki;
                const double a = Mj[0], b = Mj[NLL], c = Mj[2 * NLL];
                s_mp[2 * jj * NLL + ki] = bxk[0] * a + bxk[1] * b + bxk[2] * c;
                s_mp[(2 * jj + 1) * NLL + ki] = byk[0] * a + byk[1] * b + byk[2] * c;
            }
    }
    __syncthreads();
    double dgmax = 1.0;
    int j = 0;
    int hint = -1;  // the polar interval of point tid at the last iteration
    double sl_a = 0.0, sl_b = 0.0;  // and its cl, cd slopes
    while (dgmax > tol && j < maxiter) {
        if (active) {
            double px = 0.0, py = 0.0;
            if constexpr (MLDS) {
                // four bound values' LDS reads issued together, added in order
                int jj = jlo;
                for (; jj + 4 <= jhi; jj += 4) {
                    double mx[4], my[4], gj[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        gj[q] = s_g[jj + q];
                        mx[q] = s_mp[2 * (jj + q) * NLL + ki];
                        my[q] = s_mp[(2 * (jj + q) + 1) * NLL + ki];
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        px += mx[q] * gj[q];
                        py += my[q] * gj[q];
                    }
                }
                for (; jj < jhi; ++jj) {
                    const double gj = s_g[jj];
                    px += s_mp[2 * jj * NLL + ki] * gj;
                    py += s_mp[(2 * jj + 1) * NLL + ki] * gj;
                }
            } else {
                for (int jj = jlo; jj < jhi; ++jj) {
                    const double *Mj = M + (int64_t)jj * 3 * NLL + ki;
                    const double a = Mj[0], b = Mj[NLL], c = Mj[2 * NLL];
                    const double gj = s_g[jj];
                    px += (bxk[0] * a + bxk[1] * b + bxk[2] * c) * gj;
                    py += (byk[0] * a + byk[1] * b + byk[2] * c) * gj;
                }
            }
            s_part[kpart * NLL + ki] = px;
            s_part[(kparts + kpart) * NLL + ki] = py;
        }
        __syncthreads();
        double gnew = 0.0;
        double local_max = 0.0;
        if (tid < NLL) {
            const int i = tid;
            const double *q = s_ll + 10 * i;
            double rx = rx0, ry = ry0;
            for (int r = 0; r < kparts; ++r) {
                rx += s_part[r * NLL + i];
                ry += s_part[(kparts + r) * NLL + i];
            }
            const double tu = q[6] + rx, tv = q[7] + ry;
            const double Vinf = q[8], chord = q[9];
            const double aoa = atan2(-tv, tu) * (180.0 / kPi);
            double cl, cd;
            const int f = g.foil[i % ns];
            const int o = s_foff[f];
            interp1x2_hint(s_al + o, s_tcl + o, s_tcd + o, s_flen[f], aoa, cl, cd, hint, sl_a,
                           sl_b);
            if (cl != cl) s_nan = 1;
            const double gn = 0.5 * Vinf * chord * cl;
            const double dg = gn - s_g[i];
            gnew = s_g[i] + relax * dg;
            local_max = fabs(dg / (fabs(gn) + 1.0));
            cl_out[i] = cl;  // the last iteration's values are the step's result
            cd_out[i] = cd;
            aoa_out[i] = aoa;
        }
        // largest relative change: wave maxima, then the 16 of them
        for (int off = 32; off > 0; off >>= 1) local_max = fmax(local_max, __shfl_xor(local_max, off));
        if ((tid & 63) == 0) s_red[tid >> 6] = local_max;
        __syncthreads();
        if (tid < NLL) s_g[tid] = gnew;
        dgmax = s_red[0];
        for (int w = 1; w < kKjWaves; ++w) dgmax = fmax(dgmax, s_red[w]);
        if (s_nan) dgmax = 0.0;  // diverging: stop, the host reports it
        ++j;
        __syncthreads();
    }
    // write back the bound, trailing age 0 and shed age 1 circulation
    if (tid < nb) {
        const int b = tid;
        double c0 = 0.0;
        double *gtb = gt + (int64_t)b * nst * (A - 1);
        double *gsb = gs + (int64_t)b * ns * A;
        for (int m = 0; m < nst; ++m) {
            const double hi = m < ns ? s_g[b * ns + m] : 0.0;
            const double lo = m > 0 ? s_g[b * ns + m - 1] : 0.0;
            const double tr0 = hi - lo;
            gtb[m] = tr0;
            if (m < ns) {
                c0 += tr0;
                gsb[m] = s_g[b * ns + m];
                if (A >= 2) gsb[ns + m] = (A - 1 >= 2 ? s_c1[b * ns + m] : 0.0) - c0;
            }
        }
    }
    __syncthreads();
    // perform.m:16-62 from this time's cl, cd, aoa: velX = the blade-frame
    // inflow vel.blade turned back by twist and pitch (DCMRot with
    // [-90 AeroTwst BldPitch] about z, negated; perform.m:19-24)
    double dT = 0.0, dQ = 0.0;
    for (int i = tid; i < NLL; i += kKjThreads) {
        const int m = i % ns, b = i / ns;
        const double pitch = g.pitch[p0 + (int64_t)g.nt * b];
        const double th = rad(-90.0 + g.twist_node[m] + pitch);
        const double ct = cos(th), st = sin(th);
        const double vb0 = tab4(g.vblade, ns, g.nt, m, 0, p0, b);
        const double vb1 = tab4(g.vblade, ns, g.nt, m, 1, p0, b);
        const double v1 = -(ct * vb0 - st * vb1);
        const double v2 = -(st * vb0 + ct * vb1);
        const double tsr = v2 / v1;
        const double phi = aoa_out[i] - (pitch + g.twist_node[m]);
        const double tphi = tan(rad(phi));
        const double aap = tsr / tphi;
        const double a = aap * (v1 - tphi * v2) / (aap * v1 + tphi * v2);
        const double ap = a / aap;
        const double cph = cos(rad(phi)), sph = sin(rad(phi));
        const double dCZ = cl_out[i] * cph + cd_out[i] * sph;
        const double dCX = cl_out[i] * sph - cd_out[i] * cph;
        const double w1 = v1 * (1 - a), w2 = v2 * (1 + ap);
        const double W2 = w1 * w1 + w2 * w2;
        const double c = g.chord_node[m];
        dT += 0.5 * kRho * c * dCZ * W2 * g.dr_node[m];
        dQ += 0.5 * kRho * c * dCX * W2 * g.r_node[m] * g.dr_node[m];
    }
    for (int off = 32; off > 0; off >>= 1) {
        dT += __shfl_xor(dT, off);
        dQ += __shfl_xor(dQ, off);
    }
    if ((tid & 63) == 0) {
        s_red[tid >> 6] = dT;
        s_red2[tid >> 6] = dQ;
    }
    __syncthreads();
    if (tid == 0) {
        double T = 0.0, Q = 0.0;
        for (int w = 0; w < kKjWaves; ++w) {
            T += s_red[w];
            Q += s_red2[w];
        }
        hist_row[0] = p_index;
        hist_row[1] = T;
        hist_row[2] = Q;
        hist_row[3] = (kPi / 30.0) * g.rot_speed[p0] * Q;  // perform.m:52-56
        hist_row[4] = (double)j;
        if (s_nan) *flag = 1;
    }
}

}  // namespace mk
}  // namespace winds

using namespace winds::mk;

namespace {

struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t b) {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = winds::dev_alloc((void **)&p, b < 256 ? 256 : b);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    double *d() const { return (double *)p; }
};

unsigned blocks(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

struct winds_march {
    winds_march_cfg cfg{};
    winds::EvalCfg wake, ll;
    Geom g{};
    hipStream_t st = nullptr;
    int p = 1;        // time index of nodes[cur] (the next step to run)
    int A = 2;        // ages of nodes[cur]
    int cur = 0;      // nodes[cur]: time p
    int gprev = 0;    // gs/gt[gprev]: circulation of time p-1 after its KJ solve
    int A_gprev = 1;  // ages of gs/gt[gprev]
    int lcur = 0;     // ls/lt[lcur]: filament lengths of time p-1
    int nsteps_done = 0;
    // per time index on the host (index p-1): fastout.Time, wind.infty,
    // fastout.TipSpdRat, fastout.Azimuth, pos.hub
    std::vector<double> h_time, h_wind, h_tsr, h_azi, h_hub;
    DBuf geom, nodes[2], vel, velp, uind, hist[3];
    int hist_A[3] = {0, 0, 0};
    int n_hist = 0;
    // the KJ kernel's LDS bytes, and whether its matrix is held there
    size_t kj_lds_bytes = 0;
    int kj_m_lds = 0;
    DBuf gs[2], gt[2], gs1, gt1, rs, rt, rs1, rt1, ls[3], lt[3];
    int ls_A[3] = {0, 0, 0};
    DBuf segs, ws, llp, frozen, K, Mkj, kjout, perf, flag;
    // per-step device values (StepDev) and the recorded steps of a steady
    // march, one graph per phase of the buffer rotations
    DBuf stepd;
    bool in_step = false;
    bool use_graphs = false;
    int steady_eager = 0;
    std::map<int, hipGraphExec_t> graphs;
    // far-wake surrogate (winds_march_set_surrogate); net is not owned
    winds_mlp *net = nullptr;
    winds_march_surr sc{};
    bool surr = false;
    bool fallback = true;  // use the direct far sum until a check passes
    int A_ref = 0;         // ages of the turning-frame store ufz (0: empty)
    int64_t since_check = 0;
    const double *cl0 = nullptr;  // initial lift coefficient per station (device)
    DBuf norm, tgt, unear, ufar, umlp, ufz, llfar, llmlp, feat, lab, yout, desc, err;
    double stats[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
};

namespace {

using winds::set_error;

#define WM_HIP(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_error(WINDS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int64_t nodes_of(const winds_march *m, int A) { return (int64_t)m->g.nst * A * m->g.nb; }
int64_t segs_of(const winds_march *m, int A) {
    return (int64_t)m->g.ns * A * m->g.nb + (int64_t)m->g.nst * (A - 1) * m->g.nb;
}

int ensure_ws(winds_march *m, int64_t nt, int64_t nseg, const winds::EvalCfg &c) {
    const size_t need = winds::eval_ws_bytes(nt, nseg, c);
    if (need > m->ws.bytes) {
        WM_HIP(hipStreamSynchronize(m->st));
        WM_HIP(m->ws.alloc(need + (need >> 3)));
    }
    return WINDS_OK;
}

enum Part { kAll = 0, kNear = 1, kFar = 2 };

// Pack the segments of one part of the wake into m->segs; returns their count.
// Near: both endpoints younger than the split age S (shed ages < S, trailing
// ages < S-1); far: the others.
int64_t pack_segments(winds_march *m, const double *X, int A, const double *gsp,
                      const double *gtp, const double *rsp, const double *rtp, int zero_changing,
                      Part part) {
    const int S = m->sc.split;
    int s0 = 0, s1 = A, t0 = 0, t1 = A - 1;
    if (part == kNear) {
        s1 = std::min(S, A);
        t1 = std::min(S - 1, A - 1);
    } else if (part == kFar) {
        s0 = std::min(S, A);
        t0 = std::min(S - 1, A - 1);
    }
    const int64_t n = (int64_t)m->g.ns * (s1 - s0) * m->g.nb + (int64_t)m->g.nst * (t1 - t0) * m->g.nb;
    if (n > 0)
        k_segments<<<blocks(n), 256, 0, m->st>>>(m->g.nst, m->g.nb, A, X, gsp, gtp, rsp, rtp,
                                                 zero_changing, s0, s1, t0, t1,
                                                 (winds::Seg64 *)m->segs.p);
    return n;
}

// Induced velocity at targets t3 (nt) of the segments last packed (ns)
int eval_packed(winds_march *m, int64_t ns, const double *t3, int64_t nt, double *u3,
                const winds::EvalCfg &c) {
    WM_HIP(hipGetLastError());
    if (ns == 0) {
        k_zero<<<blocks(3 * nt), 256, 0, m->st>>>(3 * nt, u3);
        WM_HIP(hipGetLastError());
        return WINDS_OK;
    }
    int rc = ensure_ws(m, nt, ns, c);
    if (rc) return rc;
    return winds::eval_dev(t3, nt, (const winds::Seg64 *)m->segs.p, ns, c, u3, m->ws.p,
                           m->ws.bytes, m->st);
}

// Induced velocity at targets t3 (nt) of the wake (nodes, circulation, cores)
int induced(winds_march *m, const double *X, int A, const double *gsp, const double *gtp,
            const double *rsp, const double *rtp, const double *t3, int64_t nt, double *u3,
            const winds::EvalCfg &c, int zero_changing, Part part = kAll) {
    const int64_t ns = pack_segments(m, X, A, gsp, gtp, rsp, rtp, zero_changing, part);
    return eval_packed(m, ns, t3, nt, u3, c);
}

int kelvin(winds_march *m, const double *gs_old, int A_old, const double *gt_old, int A_new,
           double *gs_new, double *gt_new) {
    const int64_t nt = (int64_t)m->g.nst * (A_new - 1) * m->g.nb;
    k_kelvin_trail<<<blocks(nt), 256, 0, m->st>>>(m->g.nst, m->g.nb, gs_old, A_old, gt_old, A_new,
                                                  gt_new);
    WM_HIP(hipGetLastError());
    const int shed_blocks = (A_new + kShedAges - 1) / kShedAges * m->g.nb;
    k_kelvin_shed<<<shed_blocks, kShedAges, kelvin_shed_lds(m->g.nst), m->st>>>(
        m->g.nst, m->g.nb, gs_old, A_old, gt_new, A_new, gs_new);
    WM_HIP(hipGetLastError());
    return WINDS_OK;
}

double umag_of(const winds_march *m, int p) {
    const double *w = m->h_wind.data();
    const int nt = m->cfg.n_times;
    const double x = w[p - 1], y = w[nt + p - 1], z = w[2 * nt + p - 1];
    return sqrt(x * x + y * y + z * z);
}

// initials.m:134-135: r0(p) = sqrt(4 alpha nu delta T0), T0 = 2 pi R / (12 TSR |U|)
double r0_of(const winds_march *m, int p) {
    const double T0 = 2 * kPi * m->cfg.tip_radius / (12 * m->h_tsr[p - 1] * umag_of(m, p));
    return sqrt(4 * kAlphaRL * kNu * kDelta * T0);
}

int vcore(winds_march *m, int p, const double *X, int A, const double *gsp, const double *gtp,
          int lprev, int lout, double *rsp, double *rtp) {
    const double T = m->h_time[p - 1];  // vcore.m:34-37: fastout.Time(time)
    const int fmod = p > 3 ? 1 : 0;
    const int64_t n = segs_of(m, A);
    const StepDev *sd = m->in_step ? (const StepDev *)m->stepd.p : nullptr;
    k_vcore<<<blocks(n), 256, 0, m->st>>>(m->g.nst, m->g.nb, A, X, gsp, gtp, r0_of(m, p), T, fmod,
                                          m->ls[lprev].d(), m->lt[lprev].d(),
                                          std::max(2, m->ls_A[lprev]), m->ls[lout].d(),
                                          m->lt[lout].d(), rsp, rtp, sd, p == m->p ? 0 : 1);
    WM_HIP(hipGetLastError());
    m->ls_A[lout] = A;
    return WINDS_OK;
}

int advect(winds_march *m, const double *Xo, int A_old, const double *vo, const double *vpred,
           int ab, int p_next, double *Xn, int A_new) {
    // in a step, p_next - 1 = p: the step's table index + 1
    k_nodes01<<<blocks((int64_t)m->g.nst * m->g.nb), 256, 0, m->st>>>(
        m->g, p_next - 1, A_new, Xn, m->in_step ? (const StepDev *)m->stepd.p : nullptr, 1);
    WM_HIP(hipGetLastError());
    const int64_t n = (int64_t)m->g.nst * std::max(0, A_new - 2) * m->g.nb;
    if (n > 0) {
        const double *h1 = nullptr, *h2 = nullptr, *h3 = nullptr;
        int a1 = 0, a2 = 0, a3 = 0;
        if (ab) {
            // the last pushed history entry is time p-1
            auto slot = [&](int k) { return (m->n_hist - k) % 3; };
            h1 = m->hist[slot(1)].d();
            a1 = m->hist_A[slot(1)];
            if (ab == 4) {
                h2 = m->hist[slot(2)].d();
                a2 = m->hist_A[slot(2)];
                h3 = m->hist[slot(3)].d();
                a3 = m->hist_A[slot(3)];
            }
        }
        k_advect<<<blocks(n), 256, 0, m->st>>>(m->g.nst, m->g.nb, A_old, A_new, m->cfg.dt, Xo, vo,
                                               vpred, ab, h1, a1, h2, a2, h3, a3, Xn);
        WM_HIP(hipGetLastError());
    }
    return WINDS_OK;
}

int push_hist(winds_march *m, const double *vel, int A) {
    const int slot = m->n_hist % 3;
    WM_HIP(hipMemcpyAsync(m->hist[slot].p, vel, (size_t)3 * nodes_of(m, A) * sizeof(double),
                          hipMemcpyDeviceToDevice, m->st));
    m->hist_A[slot] = A;
    m->n_hist++;
    return WINDS_OK;
}

// ---- far-wake surrogate -------------------------------------------------

bool surr_on(const winds_march *m, int A) { return m->surr && A > m->sc.split; }

// Angle of the frame turning with the rotor at time index p: the azimuth of
// blade 0 (fastout.Azimuth) about the rotor axis, offset so that blade 0
// points along +y at psi = 0 (kinematics.m: blade b lies along z before the
// rotation by its azimuth about x).
double psi_of(const winds_march *m, int p) { return m->h_azi[p - 1] * kPi / 180.0 + 0.5 * kPi; }

winds::surr::Geo geo_of(const winds_march *m, int A, int p) {
    const int nt = m->cfg.n_times;
    const double *h = m->h_hub.data();
    return winds::surr::Geo{m->g.nb, m->g.nst, A, m->sc.split, m->cfg.max_ages, m->cfg.tip_radius,
                            umag_of(m, p), psi_of(m, p), {h[p - 1], h[nt + p - 1], h[2 * nt + p - 1]}};
}

int64_t near_targets(const winds_march *m) {
    return (int64_t)m->g.nst * m->sc.target_ages * m->g.nb;
}

// feat rows [row0, row0+np) for the points P; network output rows likewise
int mlp_features(winds_march *m, const double *X, int A, const double *gtp, int t,
                 const double *P, int64_t np, int64_t row0) {
    const float *nin = (const float *)m->norm.p;
    WM_HIP(winds::surr::launch_features(geo_of(m, A, t), X, gtp, P, np, nin, (float *)m->desc.p,
                                        (float *)m->feat.p + row0 * winds::surr::kNIn, m->st));
    return WINDS_OK;
}

int mlp_apply(winds_march *m, int A, int t, int64_t np, int64_t row0, double *u3) {
    const float *nout = (const float *)m->norm.p + 2 * winds::surr::kNIn;
    WM_HIP(winds::surr::launch_apply(geo_of(m, A, t), (const float *)m->yout.p + row0 * 3, np,
                                     nout, u3, m->st));
    return WINDS_OK;
}

// u3 += the far wake's induced velocity at P by the configured method
int far_part(winds_march *m, const double *X, int A, const double *gsp, const double *gtp,
             const double *rsp, const double *rtp, int t, const double *P, int64_t np,
             double *u3, const winds::EvalCfg &c, double *scratch) {
    if (m->sc.mode == WINDS_SURR_DROP || np == 0) return WINDS_OK;
    if (m->sc.mode == WINDS_SURR_MLP && !m->fallback) {
        int rc = mlp_features(m, X, A, gtp, t, P, np, 0);
        if (!rc) rc = winds_mlp_forward(m->net, (const float *)m->feat.p, (int32_t)np,
                                        (float *)m->yout.p, m->st);
        if (!rc) rc = mlp_apply(m, A, t, np, 0, u3);
        return rc;
    }
    int rc = induced(m, X, A, gsp, gtp, rsp, rtp, P, np, scratch, c, 0, kFar);
    if (rc) return rc;
    k_accumulate<<<blocks(3 * np), 256, 0, m->st>>>(3 * np, scratch, u3);
    WM_HIP(hipGetLastError());
    return WINDS_OK;
}

// Direct far sum at the near-wake targets and the lifting line of time t,
// the network's prediction at the same points before it is trained on them,
// the fallback decision and the online training step.  u is the direct
// induced velocity of the whole wake at its nodes (this check's step).
int surr_check(winds_march *m, const double *X, int A, const double *gsp, const double *gtp,
               const double *rsp, const double *rtp, int t, const double *u) {
    const int St = m->sc.target_ages;
    const int64_t Nt = near_targets(m);
    const int NLL = m->g.ns * m->g.nb;
    k_gather_near<<<blocks(Nt), 256, 0, m->st>>>(m->g.nst, m->g.nb, A, St, X, m->tgt.d());
    WM_HIP(hipGetLastError());
    k_ll_points<<<blocks(NLL), 256, 0, m->st>>>(m->g, t - 1, m->llp.d(), nullptr);
    WM_HIP(hipGetLastError());
    m->stats[0] += 1;
    if (m->sc.mode != WINDS_SURR_MLP) return WINDS_OK;
    const int64_t nf = pack_segments(m, X, A, gsp, gtp, rsp, rtp, 0, kFar);
    int rc = eval_packed(m, nf, m->tgt.d(), Nt, m->ufar.d(), m->wake);
    if (!rc) rc = eval_packed(m, nf, m->llp.d(), NLL, m->llfar.d(), m->ll);
    if (!rc) rc = mlp_features(m, X, A, gtp, t, m->tgt.d(), Nt, 0);
    if (!rc) rc = mlp_features(m, X, A, gtp, t, m->llp.d(), NLL, Nt);
    const int32_t rows = (int32_t)(Nt + NLL);
    if (!rc) rc = winds_mlp_forward(m->net, (const float *)m->feat.p, rows, (float *)m->yout.p,
                                    m->st);
    if (rc) return rc;
    k_zero<<<blocks(3 * Nt), 256, 0, m->st>>>(3 * Nt, m->umlp.d());
    k_zero<<<blocks(3 * NLL), 256, 0, m->st>>>(3 * NLL, m->llmlp.d());
    WM_HIP(hipGetLastError());
    rc = mlp_apply(m, A, t, Nt, 0, m->umlp.d());
    if (!rc) rc = mlp_apply(m, A, t, NLL, Nt, m->llmlp.d());
    if (rc) return rc;
    // the whole wake's induced velocity at the same points: the near
    // targets' share of u, and a direct sum at the lifting line
    k_gather_near<<<blocks(Nt), 256, 0, m->st>>>(m->g.nst, m->g.nb, A, St, u, m->unear.d());
    WM_HIP(hipGetLastError());
    rc = induced(m, X, A, gsp, gtp, rsp, rtp, m->llp.d(), NLL, m->frozen.d(), m->ll, 0);
    if (rc) return rc;
    // e = {max|net - far|, max|far|, -, max|u|} for the nodes, then the lifting line
    unsigned long long *e = (unsigned long long *)m->err.p;
    WM_HIP(hipMemsetAsync(e, 0, 8 * sizeof(unsigned long long), m->st));
    k_max_err<<<blocks(Nt), 256, 0, m->st>>>(Nt, m->umlp.d(), m->ufar.d(), e);
    k_max_err<<<blocks(Nt), 256, 0, m->st>>>(Nt, m->unear.d(), m->unear.d(), e + 2);
    k_max_err<<<blocks(NLL), 256, 0, m->st>>>(NLL, m->llmlp.d(), m->llfar.d(), e + 4);
    k_max_err<<<blocks(NLL), 256, 0, m->st>>>(NLL, m->frozen.d(), m->frozen.d(), e + 6);
    WM_HIP(hipGetLastError());
    double h[8];
    WM_HIP(hipMemcpyAsync(h, e, sizeof(h), hipMemcpyDeviceToHost, m->st));
    WM_HIP(hipStreamSynchronize(m->st));
    auto ratio = [](double a, double b) { return b > 0.0 ? a / b : 0.0; };
    const double en = ratio(h[0], h[3]), el = ratio(h[4], h[7]);
    m->stats[4] = en;
    m->stats[5] = el;
    m->stats[6] = std::max(m->stats[6], std::max(en, el));
    m->stats[8] = ratio(h[0], h[1]);
    m->stats[9] = ratio(h[4], h[5]);
    m->fallback = !(en <= m->sc.tol && el <= m->sc.tol);
    if (m->fallback) m->stats[1] += 1;
    if (m->sc.train_iters > 0) {
        float *lab = (float *)m->lab.p;
        const float *nout = (const float *)m->norm.p + 2 * winds::surr::kNIn;
        WM_HIP(winds::surr::launch_labels(geo_of(m, A, t), m->ufar.d(), Nt, nout, lab, m->st));
        WM_HIP(winds::surr::launch_labels(geo_of(m, A, t), m->llfar.d(), NLL, nout, lab + 3 * Nt,
                                          m->st));
        for (int i = 0; i < m->sc.train_iters; ++i) {
            rc = winds_mlp_train_step(m->net, (const float *)m->feat.p, lab, rows,
                                      (float)m->sc.lr, nullptr, m->st);
            if (rc) return rc;
            m->stats[7] += 1;
        }
    }
    return WINDS_OK;
}

enum How { kDirect = 0, kDirectCheck = 1, kSurrogate = 2 };

// Induced velocity of the wake at its own nodes, into u (SoA [3][N]).
int wake_velocity(winds_march *m, const double *X, int A, const double *gsp, const double *gtp,
                  const double *rsp, const double *rtp, int t, double *u, How how) {
    const int64_t N = nodes_of(m, A);
    if (how != kSurrogate) {
        int rc = induced(m, X, A, gsp, gtp, rsp, rtp, X, N, u, m->wake, 0);
        if (rc || how == kDirect) return rc;
        const double ps = psi_of(m, t);
        k_store_turning<<<blocks(N), 256, 0, m->st>>>(N, cos(ps), sin(ps), u, m->ufz.d());
        WM_HIP(hipGetLastError());
        m->A_ref = A;
        return surr_check(m, X, A, gsp, gtp, rsp, rtp, t, u);
    }
    const int St = m->sc.target_ages;
    const int64_t Nt = near_targets(m);
    k_gather_near<<<blocks(Nt), 256, 0, m->st>>>(m->g.nst, m->g.nb, A, St, X, m->tgt.d());
    WM_HIP(hipGetLastError());
    int rc = induced(m, X, A, gsp, gtp, rsp, rtp, m->tgt.d(), Nt, m->unear.d(), m->wake, 0, kNear);
    if (!rc)
        rc = far_part(m, X, A, gsp, gtp, rsp, rtp, t, m->tgt.d(), Nt, m->unear.d(), m->wake,
                      m->ufar.d());
    if (rc) return rc;
    const double ps = psi_of(m, t);
    k_compose<<<blocks(N), 256, 0, m->st>>>(m->g.nst, m->g.nb, A, St, m->unear.d(), m->ufz.d(),
                                            m->A_ref, cos(ps), sin(ps), u);
    WM_HIP(hipGetLastError());
    return WINDS_OK;
}

// Free stream of time index p added to the induced velocity u (vel.domain,
// initials.m:90 and WInDS.m:160)
int add_wind(winds_march *m, int p, int64_t N, const double *u, double *v) {
    const int nt = m->cfg.n_times;
    const double *w = m->h_wind.data();
    k_add_uinf<<<blocks(N), 256, 0, m->st>>>(N, w[p - 1], w[nt + p - 1], w[2 * nt + p - 1], u, v,
                                             m->in_step ? (const StepDev *)m->stepd.p : nullptr,
                                             p == m->p ? 0 : 1);
    WM_HIP(hipGetLastError());
    return WINDS_OK;
}

// initials.m:44-152 (time index 1)
int setup(winds_march *m) {
    const int A = 2;
    double *X = m->nodes[0].d();
    k_nodes01<<<blocks((int64_t)m->g.nst * m->g.nb), 256, 0, m->st>>>(m->g, 0, A, X, nullptr, 0);
    WM_HIP(hipGetLastError());
    // the bound circulation as a one-age shed array feeds the Kelvin step
    k_init_bound<<<blocks((int64_t)m->g.ns * m->g.nb), 256, 0, m->st>>>(
        m->g, umag_of(m, 1), m->cl0, 1, m->gs1.d());
    WM_HIP(hipGetLastError());
    int rc = kelvin(m, m->gs1.d(), 1, m->gt1.d(), A, m->gs[0].d(), m->gt[0].d());
    if (rc) return rc;
    rc = vcore(m, 1, X, A, m->gs[0].d(), m->gt[0].d(), 1, 0, m->rs.d(), m->rt.d());
    if (rc) return rc;
    const int64_t N = nodes_of(m, A);
    rc = induced(m, X, A, m->gs[0].d(), m->gt[0].d(), m->rs.d(), m->rt.d(), X, N, m->uind.d(),
                 m->wake, 0);
    if (rc) return rc;
    rc = add_wind(m, 1, N, m->uind.d(), m->vel.d());
    if (rc) return rc;
    rc = push_hist(m, m->vel.d(), A);
    if (rc) return rc;
    const int A_new = std::min(A + 1, m->cfg.max_ages);
    rc = advect(m, X, A, m->vel.d(), nullptr, 0, 2, m->nodes[1].d(), A_new);
    if (rc) return rc;
    m->cur = 1;
    m->A = A_new;
    m->gprev = 0;
    m->A_gprev = A;
    m->lcur = 0;
    m->p = 2;
    return WINDS_OK;
}

// Per-step values of time index p (and p + 1) into device memory
// (StepDev), enqueued before the step's other launches.
int set_step(winds_march *m) {
    const int p = m->p;
    const int nt = m->cfg.n_times;
    const double *w = m->h_wind.data();
    StepDev v{};
    v.p0 = p - 1;
    v.row = m->nsteps_done;
    v.p_index = (double)p;
    for (int k = 0; k < 2; ++k) {
        const int q = p + k;  // q <= n_times (step_check)
        v.r0[k] = r0_of(m, q);
        v.T[k] = m->h_time[q - 1];
        v.wind[k][0] = w[q - 1];
        v.wind[k][1] = w[nt + q - 1];
        v.wind[k][2] = w[2 * nt + q - 1];
    }
    k_set_step<<<1, 64, 0, m->st>>>((StepDev *)m->stepd.p, v);
    WM_HIP(hipGetLastError());
    return WINDS_OK;
}

int step_check(const winds_march *m) {
    const int p = m->p;
    if (p + 1 > m->cfg.n_times)
        return set_error(WINDS_E_DIMS, "the kinematic tables end at time index " +
                                           std::to_string(m->cfg.n_times) +
                                           "; stepping time index " + std::to_string(p) +
                                           " needs the next one");
    if (m->nsteps_done >= m->cfg.n_times)
        return set_error(WINDS_E_NOMEM, "performance history is full");
    return WINDS_OK;
}

// What a step changes on the host besides its launches.
struct StepPlan {
    int lnow, lpred, A_new;
    bool pcc;  // the corrector ran (vcore of p + 1 into lpred)
};

// WInDS.m:127-183, the launches of time index p (after set_step), with the
// host bookkeeping they imply except the rotation step_finish does
// LDS for the KJ kernel: the matrix goes in when the workgroup limit allows
// it, else the kernel reads it from device memory at every iteration.
int kj_lds_plan(winds_march *m, int NLL, int n_alpha) {
    int dev = 0, lim = 0;
    WM_HIP(hipGetDevice(&dev));
    WM_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    hipFuncAttributes fa, fb;
    WM_HIP(hipFuncGetAttributes(&fa, (const void *)k_kj<true>));
    WM_HIP(hipFuncGetAttributes(&fb, (const void *)k_kj<false>));
    fa.sharedSizeBytes = std::max(fa.sharedSizeBytes, fb.sharedSizeBytes);
    const size_t avail = (size_t)lim > fa.sharedSizeBytes ? (size_t)lim - fa.sharedSizeBytes : 0;
    const size_t with_m = (size_t)kj_lds(NLL, n_alpha, true).total * sizeof(double);
    const size_t without = (size_t)kj_lds(NLL, n_alpha, false).total * sizeof(double);
    m->kj_m_lds = with_m <= avail;
    m->kj_lds_bytes = m->kj_m_lds ? with_m : without;
    if (m->kj_lds_bytes > avail)
        return set_error(WINDS_E_ARG, "march: the lifting line needs more LDS than a workgroup has");
    WM_HIP(hipFuncSetAttribute(m->kj_m_lds ? (const void *)k_kj<true> : (const void *)k_kj<false>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)m->kj_lds_bytes));
    return 0;
}

// The Kelvin shed kernel's LDS for nst stations (up to 130 KB at the
// largest blade).
int kelvin_lds_plan(int nst) {
    WM_HIP(hipFuncSetAttribute((const void *)k_kelvin_shed,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kelvin_shed_lds(nst)));
    return 0;
}

int step_body(winds_march *m, StepPlan &sp) {
    const int p = m->p;
    const int t = p;
    const int A = m->A;
    const int nc = m->cur, nn = 1 - m->cur;        // nodes: time p / time p+1
    const int gp = m->gprev, gc = 1 - m->gprev;    // circulation: time p-1 / time p
    double *X = m->nodes[nc].d();
    double *gs = m->gs[gc].d(), *gt = m->gt[gc].d();
    int rc = kelvin(m, m->gs[gp].d(), m->A_gprev, m->gt[gp].d(), A, gs, gt);
    if (rc) return rc;
    const int lprev = m->lcur, lnow = (m->lcur + 1) % 3, lpred = (m->lcur + 2) % 3;
    sp.lnow = lnow;
    sp.lpred = lpred;
    sp.pcc = false;
    rc = vcore(m, p, X, A, gs, gt, lprev, lnow, m->rs.d(), m->rt.d());
    if (rc) return rc;
    const int64_t N = nodes_of(m, A);
    // surrogate steps: a direct sum with a check every check_every steps
    const bool sa = surr_on(m, A);
    const bool chk = sa && (m->A_ref == 0 || m->since_check % m->sc.check_every == 0);
    const How how = !sa ? kDirect : (chk ? kDirectCheck : kSurrogate);
    rc = wake_velocity(m, X, A, gs, gt, m->rs.d(), m->rt.d(), t, m->uind.d(), how);
    if (rc) return rc;
    rc = add_wind(m, p, N, m->uind.d(), m->vel.d());
    if (rc) return rc;
    const int A_new = std::min(A + 1, m->cfg.max_ages);
    sp.A_new = A_new;
    double *Xn = m->nodes[nn].d();
    const int integ = m->cfg.integrator;
    if (integ == WINDS_INTEG_AB4 && p >= 4) {
        rc = advect(m, X, A, m->vel.d(), nullptr, 4, p + 1, Xn, A_new);
    } else if (integ == WINDS_INTEG_AB2 && p >= 2) {
        rc = advect(m, X, A, m->vel.d(), nullptr, 2, p + 1, Xn, A_new);
    } else {
        rc = advect(m, X, A, m->vel.d(), nullptr, 0, p + 1, Xn, A_new);
        if (!rc && integ == WINDS_INTEG_PCC && p > 2) {
            // pcc.m:40-75: Kelvin and vcore for p+1 at the predicted nodes,
            // induced velocity there, trapezoidal corrector
            sp.pcc = true;
            rc = kelvin(m, gs, A, gt, A_new, m->gs1.d(), m->gt1.d());
            if (!rc)
                rc = vcore(m, p + 1, Xn, A_new, m->gs1.d(), m->gt1.d(), lnow, lpred, m->rs1.d(),
                           m->rt1.d());
            const int64_t Nn = nodes_of(m, A_new);
            // a check step's corrector goes through the surrogate too (the
            // check has just refreshed the held far-wake velocities and
            // decided between the network and the direct far sum), unless
            // every step is a check: that march stays the plain one
            const How how_c = how == kDirect || (how == kDirectCheck && m->sc.check_every <= 1)
                                  ? kDirect
                                  : kSurrogate;
            if (!rc)
                rc = wake_velocity(m, Xn, A_new, m->gs1.d(), m->gt1.d(), m->rs1.d(), m->rt1.d(),
                                   t + 1, m->uind.d(), how_c);
            if (!rc) rc = add_wind(m, p + 1, Nn, m->uind.d(), m->velp.d());
            if (!rc) rc = advect(m, X, A, m->vel.d(), m->velp.d(), 0, p + 1, Xn, A_new);
        }
    }
    if (rc) return rc;
    // Kutta-Joukowski at time p on the wake of time p (KuttaJoukowski.m:215-219)
    const StepDev *sd = (const StepDev *)m->stepd.p;
    const int NLL = m->g.ns * m->g.nb;
    k_ll_points<<<blocks(NLL), 256, 0, m->st>>>(m->g, p - 1, m->llp.d(), sd);
    WM_HIP(hipGetLastError());
    if (how == kSurrogate) {
        rc = induced(m, X, A, gs, gt, m->rs.d(), m->rt.d(), m->llp.d(), NLL, m->frozen.d(), m->ll,
                     1, kNear);
        if (!rc)
            rc = far_part(m, X, A, gs, gt, m->rs.d(), m->rt.d(), t, m->llp.d(), NLL,
                          m->frozen.d(), m->ll, m->llfar.d());
    } else {
        rc = induced(m, X, A, gs, gt, m->rs.d(), m->rt.d(), m->llp.d(), NLL, m->frozen.d(), m->ll,
                     1);
    }
    if (rc) return rc;
    const int NCH = (2 * m->g.ns + m->g.nst) * m->g.nb;
    k_influence<<<blocks((int64_t)NLL * NCH), 256, 0, m->st>>>(
        m->g, A, X, m->rs.d(), m->rt.d(), m->llp.d(), m->ll.semantics,
        m->ll.core == 0 ? 0 : m->ll.vat_n, m->ll.d, m->ll.co, m->K.d());
    WM_HIP(hipGetLastError());
    double *kj = m->kjout.d();
    // the history row and p come from sd
    k_kj_matrix<<<blocks((int64_t)NLL * NLL), 256, 0, m->st>>>(m->g, m->K.d(), m->Mkj.d());
    WM_HIP(hipGetLastError());
    if (m->kj_m_lds)
        k_kj<true><<<1, kKjThreads, m->kj_lds_bytes, m->st>>>(
            m->g, p - 1, A, m->cfg.relax, m->cfg.tol, m->cfg.maxiter, m->frozen.d(), m->K.d(),
            m->Mkj.d(), gs, gt, kj, kj + NLL, kj + 2 * NLL, m->perf.d(), (double)p,
            (int *)m->flag.p, sd);
    else
        k_kj<false><<<1, kKjThreads, m->kj_lds_bytes, m->st>>>(
            m->g, p - 1, A, m->cfg.relax, m->cfg.tol, m->cfg.maxiter, m->frozen.d(), m->K.d(),
            m->Mkj.d(), gs, gt, kj, kj + NLL, kj + 2 * NLL, m->perf.d(), (double)p,
            (int *)m->flag.p, sd);
    WM_HIP(hipGetLastError());
    rc = push_hist(m, m->vel.d(), A);
    if (rc) return rc;
    if (sa) {
        if (how == kSurrogate)
            m->stats[(m->sc.mode == WINDS_SURR_MLP && !m->fallback) ? 2 : 3] += 1;
        m->since_check++;
    }
    return WINDS_OK;
}

// The buffer rotation that ends a step.
void step_finish(winds_march *m, const StepPlan &sp) {
    m->nsteps_done++;
    m->lcur = sp.lnow;
    m->gprev = 1 - m->gprev;
    m->A_gprev = m->A;
    m->cur = 1 - m->cur;
    m->A = sp.A_new;
    m->p = m->p + 1;
}

int step_once(winds_march *m) {
    int rc = step_check(m);
    if (!rc) rc = set_step(m);
    if (rc) return rc;
    StepPlan sp{};
    m->in_step = true;
    rc = step_body(m, sp);
    m->in_step = false;
    if (rc) return rc;
    step_finish(m, sp);
    return WINDS_OK;
}

void drop_graphs(winds_march *m) {
    for (auto &kv : m->graphs) (void)hipGraphExecDestroy(kv.second);
    m->graphs.clear();
    m->steady_eager = 0;
}

// A step of a steady direct march (the wake at max_ages, no surrogate):
// every launch is the same as in the steps six before and after it (the
// node, circulation, length and history buffers rotate with periods 2, 2, 3
// and 3), and everything that varies is read from StepDev, so each phase
// of the rotation is recorded once into a graph and replayed.
int step_graph(winds_march *m) {
    int rc = step_check(m);
    if (!rc) rc = set_step(m);
    if (rc) return rc;
    const int key = m->cur + 2 * m->gprev + 4 * m->lcur + 12 * (m->n_hist % 3);
    StepPlan sp{};
    auto it = m->graphs.find(key);
    if (it == m->graphs.end()) {
        WM_HIP(hipStreamBeginCapture(m->st, hipStreamCaptureModeThreadLocal));
        m->in_step = true;
        rc = step_body(m, sp);
        m->in_step = false;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(m->st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess)
            return set_error(WINDS_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiateWithFlags(&ex, g, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess)
            return set_error(WINDS_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        it = m->graphs.emplace(key, ex).first;
    } else {
        // the host bookkeeping of the recorded launches (vcore's lengths,
        // the velocity history)
        const int A = m->A;
        sp.lnow = (m->lcur + 1) % 3;
        sp.lpred = (m->lcur + 2) % 3;
        sp.A_new = std::min(A + 1, m->cfg.max_ages);
        sp.pcc = m->cfg.integrator == WINDS_INTEG_PCC && m->p > 2;
        m->ls_A[sp.lnow] = A;
        if (sp.pcc) m->ls_A[sp.lpred] = sp.A_new;
        m->hist_A[m->n_hist % 3] = A;
        m->n_hist++;
    }
    WM_HIP(hipGraphLaunch(it->second, m->st));
    step_finish(m, sp);
    return WINDS_OK;
}

bool steady_direct(const winds_march *m) {
    return m->use_graphs && !m->surr && m->A == m->cfg.max_ages && m->A_gprev == m->A;
}

}  // namespace

extern "C" {

void winds_march_destroy(winds_march *m) {
    if (!m) return;
    if (m->st) (void)hipStreamSynchronize(m->st);
    drop_graphs(m);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

int winds_march_create(const winds_march_cfg *cfg, winds_march **out) {
    winds::clear_error();
    if (!cfg || !out) return set_error(WINDS_E_ARG, "null argument");
    *out = nullptr;
    const winds_march_cfg &c = *cfg;
    if (c.nb < 1 || c.nst < 2 || c.max_ages < 3 || c.n_times < 2)
        return set_error(WINDS_E_DIMS, "bad march sizes");
    if ((int64_t)(c.nst - 1) * c.nb > kMaxLL || c.nst > kMaxNodesBlade)
        return set_error(WINDS_E_DIMS, "too many lifting-line points for the KJ solver");
    if (c.n_foils < 1 || c.n_foils > kMaxFoils || c.n_alpha < 2 || c.n_alpha > kMaxAlpha || !c.foil_len)
        return set_error(WINDS_E_DIMS, "airfoil tables: 1..64 tables of 2..1024 entries in all");
    {
        int64_t tot = 0;
        for (int k = 0; k < c.n_foils; ++k) {
            if (c.foil_len[k] < 2) return set_error(WINDS_E_DIMS, "an airfoil table has < 2 entries");
            tot += c.foil_len[k];
        }
        if (tot != c.n_alpha) return set_error(WINDS_E_DIMS, "airfoil table lengths do not add up");
    }
    if (!c.r_node || !c.chord_node || !c.twist_node || !c.dr_node || !c.foil || !c.cl_init ||
        !c.alpha || !c.cl || !c.cd)
        return set_error(WINDS_E_ARG, "null blade or airfoil array");
    if (!c.time || !c.wind || !c.tsr || !c.rot_speed || !c.azimuth || !c.pitch || !c.hub ||
        !c.quarter || !c.trail || !c.bound || !c.bxn || !c.byn || !c.bzn || !c.vblade)
        return set_error(WINDS_E_ARG, "null kinematic table");
    for (int i = 0; i < c.nst - 1; ++i)
        if (c.foil[i] < 1 || c.foil[i] > c.n_foils)
            return set_error(WINDS_E_ARG, "blade.NFoil names a table that does not exist");
    if (c.integrator < WINDS_INTEG_PCC || c.integrator > WINDS_INTEG_AB2)
        return set_error(WINDS_E_MODE, "unknown integrator");
    if (!(c.dt > 0.0) || !(c.tip_radius > 0.0)) return set_error(WINDS_E_ARG, "bad dt/TipRad");
    winds_march *m = new winds_march();
    m->cfg = c;
    int rc = winds::opts_to_cfg(&c.wake, m->wake);
    if (!rc) rc = winds::opts_to_cfg(&c.ll, m->ll);
    if (rc) {
        delete m;
        return rc;
    }
    // lifting-line sums have few targets: FP64 splits the segment range
    if (m->ll.precision == WINDS_PREC_FP64 && m->ll.split <= 1) m->ll.split = 512;
    rc = kj_lds_plan(m, (c.nst - 1) * c.nb, c.n_alpha);
    if (!rc) rc = kelvin_lds_plan(c.nst);
    if (rc) {
        delete m;
        return rc;
    }
    const int nst = c.nst, ns = c.nst - 1, nb = c.nb, na = c.n_alpha, nt = c.n_times;
    const int nf = c.n_foils;
    m->h_time.assign(c.time, c.time + nt);
    m->h_wind.assign(c.wind, c.wind + 3 * (size_t)nt);
    m->h_tsr.assign(c.tsr, c.tsr + nt);
    m->h_azi.assign(c.azimuth, c.azimuth + nt);
    m->h_hub.assign(c.hub, c.hub + 3 * (size_t)nt);
    // device copy: doubles first, then the integer arrays
    const size_t n_st = (size_t)nst * 3 * nt * nb, n_ns = (size_t)ns * 3 * nt * nb;
    const size_t nd = 5 * (size_t)ns + 3 * (size_t)na + (size_t)nt * nb + nt + 2 * n_st +
                      5 * n_ns + (c.ll_extra ? n_ns : 0);
    const size_t ni = (size_t)ns + 2 * (size_t)nf;
    std::vector<double> h(nd);
    std::vector<int> hi(ni);
    double *q = h.data();
    auto put = [&](const double *src, size_t n) {
        double *at = q;
        std::copy(src, src + n, q);
        q += n;
        return at - h.data();
    };
    const size_t o_rn = put(c.r_node, ns), o_cn = put(c.chord_node, ns), o_tn = put(c.twist_node, ns),
                 o_dr = put(c.dr_node, ns), o_cl0 = put(c.cl_init, ns), o_al = put(c.alpha, na),
                 o_cl = put(c.cl, na), o_cd = put(c.cd, na), o_pi = put(c.pitch, (size_t)nt * nb),
                 o_rs = put(c.rot_speed, nt), o_qu = put(c.quarter, n_st),
                 o_tr = put(c.trail, n_st), o_bo = put(c.bound, n_ns), o_bx = put(c.bxn, n_ns),
                 o_by = put(c.byn, n_ns), o_bz = put(c.bzn, n_ns), o_vb = put(c.vblade, n_ns);
    const size_t o_ex = c.ll_extra ? put(c.ll_extra, n_ns) : 0;
    for (int i = 0; i < ns; ++i) hi[i] = c.foil[i] - 1;
    for (int k = 0, off = 0; k < nf; off += c.foil_len[k], ++k) {
        hi[ns + k] = off;
        hi[ns + nf + k] = c.foil_len[k];
    }
    auto fail = [&](int code) {
        winds_march_destroy(m);
        return code;
    };
#define WMC(expr)                                                                           \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(set_error(WINDS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
    } while (0)
    WMC(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
    const size_t ibytes = ((ni * sizeof(int) + 255) / 256) * 256;
    WMC(m->geom.alloc(nd * sizeof(double) + ibytes));
    WMC(hipMemcpy(m->geom.p, h.data(), nd * sizeof(double), hipMemcpyHostToDevice));
    int *GI = (int *)((char *)m->geom.p + nd * sizeof(double));
    WMC(hipMemcpy(GI, hi.data(), ni * sizeof(int), hipMemcpyHostToDevice));
    const double *G = m->geom.d();
    m->g.r_node = G + o_rn;
    m->g.chord_node = G + o_cn;
    m->g.twist_node = G + o_tn;
    m->g.dr_node = G + o_dr;
    m->cl0 = G + o_cl0;
    m->g.alpha = G + o_al;
    m->g.cl = G + o_cl;
    m->g.cd = G + o_cd;
    m->g.pitch = G + o_pi;
    m->g.rot_speed = G + o_rs;
    m->g.quarter = G + o_qu;
    m->g.trail = G + o_tr;
    m->g.bound = G + o_bo;
    m->g.bxn = G + o_bx;
    m->g.byn = G + o_by;
    m->g.bzn = G + o_bz;
    m->g.vblade = G + o_vb;
    m->g.ll_extra = c.ll_extra ? G + o_ex : nullptr;
    m->g.foil = GI;
    m->g.foil_off = GI + ns;
    m->g.foil_len = GI + ns + nf;
    m->g.nb = nb;
    m->g.nst = nst;
    m->g.ns = ns;
    m->g.nt = nt;
    m->g.n_foils = nf;
    m->g.n_alpha = na;
    const int AM = c.max_ages;
    const size_t NB = (size_t)3 * nst * AM * nb * sizeof(double);
    const size_t SS = (size_t)ns * AM * nb * sizeof(double);
    const size_t ST = (size_t)nst * AM * nb * sizeof(double);
    const size_t SEG = ((size_t)ns * AM * nb + (size_t)nst * AM * nb) * sizeof(winds::Seg64);
    const int NLL = ns * nb;
    const int NCH = (2 * ns + nst) * nb;
    for (int i = 0; i < 2; ++i) {
        WMC(m->nodes[i].alloc(NB));
        WMC(m->gs[i].alloc(SS));
        WMC(m->gt[i].alloc(ST));
    }
    for (int i = 0; i < 3; ++i) {
        WMC(m->hist[i].alloc(NB));
        WMC(m->ls[i].alloc(SS));
        WMC(m->lt[i].alloc(ST));
    }
    WMC(m->vel.alloc(NB));
    WMC(m->velp.alloc(NB));
    WMC(m->uind.alloc(NB));
    WMC(m->gs1.alloc(SS));
    WMC(m->gt1.alloc(ST));
    WMC(m->rs.alloc(SS));
    WMC(m->rt.alloc(ST));
    WMC(m->rs1.alloc(SS));
    WMC(m->rt1.alloc(ST));
    WMC(m->segs.alloc(SEG));
    WMC(m->llp.alloc((size_t)3 * NLL * sizeof(double)));
    WMC(m->frozen.alloc((size_t)3 * NLL * sizeof(double)));
    WMC(m->K.alloc((size_t)3 * NLL * NCH * sizeof(double)));
    WMC(m->Mkj.alloc((size_t)3 * NLL * NLL * sizeof(double)));
    WMC(m->kjout.alloc((size_t)3 * NLL * sizeof(double)));
    WMC(m->perf.alloc((size_t)5 * nt * sizeof(double)));
    WMC(m->flag.alloc(256));
    WMC(m->stepd.alloc(sizeof(StepDev)));
    WMC(hipMemsetAsync(m->flag.p, 0, 256, m->st));
    WMC(hipMemsetAsync(m->gt1.p, 0, ST, m->st));
    for (int i = 0; i < 3; ++i) {
        WMC(hipMemsetAsync(m->ls[i].p, 0, SS, m->st));
        WMC(hipMemsetAsync(m->lt[i].p, 0, ST, m->st));
    }
#undef WMC
    rc = setup(m);
    if (rc) return fail(rc);
    *out = m;
    return WINDS_OK;
}

int winds_march_step(winds_march *m, int32_t nsteps) {
    winds::clear_error();
    if (!m) return set_error(WINDS_E_ARG, "null march");
    for (int i = 0; i < nsteps; ++i) {
        int rc;
        if (!steady_direct(m)) {
            m->steady_eager = 0;
            rc = step_once(m);
        } else if (m->steady_eager < 2) {
            // steady steps run eagerly twice first (the workspace is then
            // sized for them) before any is recorded
            rc = step_once(m);
            m->steady_eager++;
        } else {
            rc = step_graph(m);
        }
        if (rc) return rc;
    }
    return WINDS_OK;
}

int winds_march_set_graphs(winds_march *m, int32_t on) {
    winds::clear_error();
    if (!m) return set_error(WINDS_E_ARG, "null march");
    WM_HIP(hipStreamSynchronize(m->st));
    drop_graphs(m);
    m->use_graphs = on != 0;
    return WINDS_OK;
}

int winds_march_sync(winds_march *m) {
    winds::clear_error();
    if (!m) return set_error(WINDS_E_ARG, "null march");
    WM_HIP(hipStreamSynchronize(m->st));
    int flag = 0;
    WM_HIP(hipMemcpy(&flag, m->flag.p, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) return set_error(WINDS_E_ARG, "Diverging soln!!! (NaN lift coefficient in KJ)");
    return WINDS_OK;
}

int winds_march_info(const winds_march *m, int64_t out[6]) {
    if (!m || !out) return set_error(WINDS_E_ARG, "null argument");
    out[0] = m->p;
    out[1] = m->A;
    out[2] = nodes_of(m, m->A);
    out[3] = segs_of(m, m->A);
    out[4] = m->A_gprev;
    out[5] = m->nsteps_done;
    return WINDS_OK;
}

int winds_march_get_state(winds_march *m, double *nodes, double *gshed, double *gtrail) {
    winds::clear_error();
    if (!m) return set_error(WINDS_E_ARG, "null march");
    WM_HIP(hipStreamSynchronize(m->st));
    const int A = m->A, nst = m->g.nst, ns = m->g.ns, nb = m->g.nb;
    const int64_t N = nodes_of(m, A);
    if (nodes) {
        std::vector<double> h((size_t)3 * N);
        WM_HIP(hipMemcpy(h.data(), m->nodes[m->cur].p, h.size() * sizeof(double),
                         hipMemcpyDeviceToHost));
        // SoA over j = s + a*nst + b*nst*A  ->  [nst,3,A,nb] column-major
        for (int b = 0; b < nb; ++b)
            for (int a = 0; a < A; ++a)
                for (int s = 0; s < nst; ++s) {
                    const int64_t j = s + (int64_t)a * nst + (int64_t)b * nst * A;
                    for (int c = 0; c < 3; ++c)
                        nodes[s + (int64_t)c * nst + (int64_t)a * 3 * nst + (int64_t)b * 3 * nst * A] =
                            h[c * N + j];
                }
    }
    // circulation of the last completed time index (after its KJ solve)
    const int Ag = m->A_gprev;
    if (gshed)
        WM_HIP(hipMemcpy(gshed, m->gs[m->gprev].p, (size_t)ns * Ag * nb * sizeof(double),
                         hipMemcpyDeviceToHost));
    if (gtrail)
        WM_HIP(hipMemcpy(gtrail, m->gt[m->gprev].p, (size_t)nst * (Ag - 1) * nb * sizeof(double),
                         hipMemcpyDeviceToHost));
    return WINDS_OK;
}

int winds_march_set_state(winds_march *m, int32_t p, int32_t ages, const double *nodes,
                          const double *gshed, const double *gtrail) {
    winds::clear_error();
    if (!m || !nodes || !gshed || !gtrail) return set_error(WINDS_E_ARG, "null argument");
    if (ages < 3 || ages > m->cfg.max_ages || p < 4 || p >= m->cfg.n_times)
        return set_error(WINDS_E_DIMS, "bad ages/p");
    WM_HIP(hipStreamSynchronize(m->st));
    drop_graphs(m);
    const int nst = m->g.nst, ns = m->g.ns, nb = m->g.nb, A = ages;
    const int64_t N = nodes_of(m, A);
    std::vector<double> h((size_t)3 * N);
    for (int b = 0; b < nb; ++b)
        for (int a = 0; a < A; ++a)
            for (int s = 0; s < nst; ++s) {
                const int64_t j = s + (int64_t)a * nst + (int64_t)b * nst * A;
                for (int c = 0; c < 3; ++c)
                    h[c * N + j] =
                        nodes[s + (int64_t)c * nst + (int64_t)a * 3 * nst + (int64_t)b * 3 * nst * A];
            }
    // the wake given is time p; its circulation stands in for time p-1 (the
    // Kelvin step of time p regenerates it from the bound values)
    WM_HIP(hipMemcpy(m->nodes[m->cur].p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    WM_HIP(hipMemcpy(m->gs[m->gprev].p, gshed, (size_t)ns * (A - 1) * nb * sizeof(double),
                     hipMemcpyHostToDevice));
    WM_HIP(hipMemcpy(m->gt[m->gprev].p, gtrail, (size_t)nst * (A - 2) * nb * sizeof(double),
                     hipMemcpyHostToDevice));
    m->A_gprev = A - 1;
    m->A = A;
    m->p = p;
    // lengths of time p-1: this wake's own lengths (zero strain on the first step)
    const int64_t ns_ = segs_of(m, A);
    k_vcore<<<blocks(ns_), 256, 0, m->st>>>(nst, nb, A, m->nodes[m->cur].d(), m->gs[m->gprev].d(),
                                            m->gt[m->gprev].d(), r0_of(m, p), 0.0, 0, m->ls[0].d(),
                                            m->lt[0].d(), A, m->ls[m->lcur].d(), m->lt[m->lcur].d(),
                                            m->rs.d(), m->rt.d(), nullptr, 0);
    WM_HIP(hipGetLastError());
    m->ls_A[m->lcur] = A;
    m->A_ref = 0;
    m->since_check = 0;
    return WINDS_OK;
}

int winds_march_get_perf(winds_march *m, int32_t first, int32_t n, double *rows) {
    winds::clear_error();
    if (!m || !rows) return set_error(WINDS_E_ARG, "null argument");
    if (first < 0 || n < 0 || first + n > m->nsteps_done)
        return set_error(WINDS_E_DIMS, "rows out of range");
    WM_HIP(hipStreamSynchronize(m->st));
    WM_HIP(hipMemcpy(rows, m->perf.d() + 5 * (int64_t)first, (size_t)5 * n * sizeof(double),
                     hipMemcpyDeviceToHost));
    return WINDS_OK;
}

int winds_march_get_kj(winds_march *m, double *cl, double *cd, double *aoa) {
    winds::clear_error();
    if (!m) return set_error(WINDS_E_ARG, "null march");
    WM_HIP(hipStreamSynchronize(m->st));
    const size_t n = (size_t)m->g.ns * m->g.nb;
    const double *k = m->kjout.d();
    if (cl) WM_HIP(hipMemcpy(cl, k, n * sizeof(double), hipMemcpyDeviceToHost));
    if (cd) WM_HIP(hipMemcpy(cd, k + n, n * sizeof(double), hipMemcpyDeviceToHost));
    if (aoa) WM_HIP(hipMemcpy(aoa, k + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost));
    return WINDS_OK;
}

const double *winds_march_uind_dev(const winds_march *m) { return m ? m->uind.d() : nullptr; }

int winds_march_set_surrogate(winds_march *m, winds_mlp *net, const winds_march_surr *s) {
    winds::clear_error();
    if (!m) return set_error(WINDS_E_ARG, "null march");
    WM_HIP(hipStreamSynchronize(m->st));
    drop_graphs(m);
    if (!s) {
        m->surr = false;
        m->net = nullptr;
        return WINDS_OK;
    }
    const winds_march_surr &c = *s;
    if (c.mode < WINDS_SURR_MLP || c.mode > WINDS_SURR_DROP)
        return set_error(WINDS_E_MODE, "unknown surrogate mode");
    if (c.split < 2 || c.split >= m->cfg.max_ages || c.target_ages < 1 || c.target_ages > c.split ||
        c.check_every < 1 || c.train_iters < 0)
        return set_error(WINDS_E_DIMS, "bad surrogate split/targets/check interval");
    if (!(c.tol > 0.0) || !(c.lr >= 0.0)) return set_error(WINDS_E_ARG, "bad tol/lr");
    const int nst = m->g.nst, nb = m->g.nb, NLL = m->g.ns * m->g.nb;
    const int64_t Nt = (int64_t)nst * c.target_ages * nb;
    const int64_t rows = Nt + NLL;
    constexpr int NIN = winds::surr::kNIn;
    if (c.mode == WINDS_SURR_MLP) {
        if (!net || !c.in_norm || !c.out_norm)
            return set_error(WINDS_E_ARG, "network mode needs a network and normalisation");
        int64_t info[6];
        int rc = winds_mlp_info(net, info);
        if (rc) return rc;
        if (info[0] != NIN || info[3] != winds::surr::kNOut)
            return set_error(WINDS_E_DIMS, "network inputs/outputs do not match the surrogate");
        if (info[4] < rows)
            return set_error(WINDS_E_DIMS, "network max_batch below near targets + lifting line");
    }
    const size_t D = sizeof(double);
    const size_t NB = (size_t)3 * nst * m->cfg.max_ages * nb * D;
    WM_HIP(m->norm.alloc((2 * NIN + 6) * sizeof(float)));
    WM_HIP(m->tgt.alloc(3 * Nt * D));
    WM_HIP(m->unear.alloc(3 * Nt * D));
    WM_HIP(m->ufar.alloc(3 * Nt * D));
    WM_HIP(m->umlp.alloc(3 * Nt * D));
    WM_HIP(m->ufz.alloc(NB));
    WM_HIP(m->llfar.alloc(3 * NLL * D));
    WM_HIP(m->llmlp.alloc(3 * NLL * D));
    WM_HIP(m->feat.alloc(rows * NIN * sizeof(float)));
    WM_HIP(m->lab.alloc(rows * 3 * sizeof(float)));
    WM_HIP(m->yout.alloc(rows * 3 * sizeof(float)));
    WM_HIP(m->desc.alloc(64 * sizeof(float)));
    WM_HIP(m->err.alloc(8 * sizeof(unsigned long long)));
    if (c.mode == WINDS_SURR_MLP) {
        std::vector<float> h(2 * NIN + 6);
        std::copy(c.in_norm, c.in_norm + 2 * NIN, h.begin());
        std::copy(c.out_norm, c.out_norm + 6, h.begin() + 2 * NIN);
        WM_HIP(hipMemcpy(m->norm.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    m->net = c.mode == WINDS_SURR_MLP ? net : nullptr;
    m->sc = c;
    m->sc.in_norm = nullptr;  // copied to the device
    m->sc.out_norm = nullptr;
    m->surr = true;
    m->fallback = true;
    m->A_ref = 0;
    m->since_check = 0;
    for (double &v : m->stats) v = 0.0;
    return WINDS_OK;
}

int winds_march_surr_stats(const winds_march *m, double out[10]) {
    if (!m || !out) return set_error(WINDS_E_ARG, "null argument");
    for (int i = 0; i < 10; ++i) out[i] = m->stats[i];
    return WINDS_OK;
}

}  // extern "C"
