// ground.hip -- ground-effect vortex-panel plane (include/winds_ground.h).
//
// Vortex-ring panels on z = 0 (ground/VP_Influence_Coefficient_Matrix_v7.m,
// ground/VP_Induced_Velocity_LM.m), all in FP64 in the reference's order of
// operations: the four segment terms of a ring added 1-2-3-4, divided by 4
// and by pi, rings accumulated in panel order.  The influence matrix is built
// once; the circulation solve runs conjugate gradients on the device (A is
// symmetric positive definite for the flat grid) with fixed-order
// reductions, so every run gives the same bits.

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "bs_common.hpp"
#include "winds_ground.h"

namespace winds {
namespace gnd {

constexpr double kPi = 3.141592653589793;

struct Grid {
    int np;                // panels per side
    double x0, dx, y0, dy;  // node k along x: x0 + k dx; along y: y0 + k dy
    double co;             // absolute cut-off co_frac (dx + dy) / 2
};

__device__ __forceinline__ double xn(const Grid &g, int k) { return g.x0 + k * g.dx; }
__device__ __forceinline__ double yn(const Grid &g, int k) { return g.y0 + k * g.dy; }

// centre of panel i (row n = i / np along y, column m = i % np along x),
// VP_Influence_Coefficient_Matrix_v7.m:11-18
__device__ __forceinline__ void centre(const Grid &g, int i, double &px, double &py, double &pz) {
    const int n = i / g.np, m = i % g.np;
    px = xn(g, m) + 0.5 * (xn(g, m + 1) - xn(g, m));
    py = yn(g, n) + 0.5 * (yn(g, n + 1) - yn(g, n));
    pz = 0.0 + 0.5 * (0.0 - 0.0);
}

// velocity of ring j per unit circulation at P, (A1+A2+A3+A4)/4/pi with
// r_k = P - corner k, n_k = |r_k|, and segment a->b contributing
// r_a x r_b (n_a + n_b) / (n_a n_b (n_a n_b + r_a . r_b) + co)
// (VP_Induced_Velocity_LM.m:47-73); each n_k is computed once, as there
__device__ __forceinline__ void ring(const Grid &g, int j, double px, double py, double pz,
                                     double &ux, double &uy, double &uz) {
    const int n = j / g.np, m = j % g.np;
    const double x1 = xn(g, m), x2 = xn(g, m + 1), y1 = yn(g, n), y2 = yn(g, n + 1);
    const double rx[4] = {px - x1, px - x2, px - x2, px - x1};
    const double ry[4] = {py - y1, py - y1, py - y2, py - y2};
    const double rz = pz - 0.0;
    double nn[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) nn[k] = sqrt(rx[k] * rx[k] + ry[k] * ry[k] + rz * rz);
    double ax[4], ay[4], az[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int a = k, b = (k + 1) & 3;
        const double dot = rx[a] * rx[b] + ry[a] * ry[b] + rz * rz;
        const double t = (nn[a] + nn[b]) / (nn[a] * nn[b] * (nn[a] * nn[b] + dot) + g.co);
        ax[k] = (ry[a] * rz - rz * ry[b]) * t;
        ay[k] = (rz * rx[b] - rx[a] * rz) * t;
        az[k] = (rx[a] * ry[b] - ry[a] * rx[b]) * t;
    }
    ux = (((ax[0] + ax[1]) + ax[2]) + ax[3]) / 4.0 / kPi;
    uy = (((ay[0] + ay[1]) + ay[2]) + ay[3]) / 4.0 / kPi;
    uz = (((az[0] + az[1]) + az[2]) + az[3]) / 4.0 / kPi;
}

// panel normal n_j = c2 / |c2| with c2 = r2 x r3 at control point 1 (:93)
__global__ void k_normals(Grid g, double *__restrict__ nrm) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = g.np * g.np;
    if (j >= N) return;
    double px, py, pz;
    centre(g, 0, px, py, pz);
    const int n = j / g.np, m = j % g.np;
    const double r2x = px - xn(g, m + 1), r2y = py - yn(g, n), r2z = pz;
    const double r3x = px - xn(g, m + 1), r3y = py - yn(g, n + 1), r3z = pz;
    const double cx = r2y * r3z - r2z * r3y, cy = r2z * r3x - r2x * r3z, cz = r2x * r3y - r2y * r3x;
    const double l = sqrt(cx * cx + cy * cy + cz * cz);
    nrm[j] = cx / l;
    nrm[N + j] = cy / l;
    nrm[2 * N + j] = cz / l;
}

__global__ void k_centres(Grid g, double *__restrict__ C) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = g.np * g.np;
    if (i >= N) return;
    centre(g, i, C[i], C[N + i], C[2 * N + i]);
}

// A[i*N + j] = ring j at centre i, dotted with n_j (:95)
__global__ void k_matrix(Grid g, const double *__restrict__ nrm, double *__restrict__ A) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int N = g.np * g.np;
    if (k >= (int64_t)N * N) return;
    const int i = (int)(k / N), j = (int)(k % N);
    double px, py, pz, ux, uy, uz;
    centre(g, i, px, py, pz);
    ring(g, j, px, py, pz, ux, uy, uz);
    A[k] = ux * nrm[j] + uy * nrm[N + j] + uz * nrm[2 * N + j];
}

// V_n(j) = -V_ext(j) . n_j (:94), the right-hand side; also x = 0, r = b, p = b
__global__ void k_rhs(int N, const double *__restrict__ v, const double *__restrict__ nrm,
                      double *__restrict__ b, double *__restrict__ x, double *__restrict__ r,
                      double *__restrict__ p) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const double bj = -(v[j] * nrm[j] + v[N + j] * nrm[N + j] + v[2 * N + j] * nrm[2 * N + j]);
    b[j] = bj;
    x[j] = 0.0;
    r[j] = bj;
    p[j] = bj;
}

// scalars of the iteration (device): 0 rs = r.r, 1 pAp, 2 rs_new, 3 bb = b.b,
// 4 tol^2 bb, 5 done flag, 6 iterations
constexpr int kS = 8;

// y = A p, one wave per row, lanes over columns in order, fixed wave tree
__global__ void k_matvec(int N, const double *__restrict__ A, const double *__restrict__ p,
                         double *__restrict__ y, const double *__restrict__ s) {
    if (s[5] != 0.0) return;
    const int row = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= N) return;
    const double *a = A + (int64_t)row * N;
    double acc = 0.0;
    for (int j = lane; j < N; j += 64) acc += a[j] * p[j];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
    if (lane == 0) y[row] = acc;
}

// s[slot] = a . b in one workgroup of 1024 threads: strided partial sums in
// order, then a fixed tree
__global__ __launch_bounds__(1024) void k_dot(int N, const double *__restrict__ a,
                                              const double *__restrict__ b,
                                              double *__restrict__ s, int slot, int force) {
    if (!force && s[5] != 0.0) return;
    __shared__ double red[1024];
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int j = t; j < N; j += 1024) acc += a[j] * b[j];
    red[t] = acc;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) s[slot] = red[0];
}

__global__ void k_init_scalars(double *__restrict__ s, double tol) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        s[0] = s[3];
        s[4] = tol * tol * s[3];
        s[5] = s[3] == 0.0 ? 1.0 : 0.0;
        s[6] = 0.0;
    }
}

// x += alpha p, r -= alpha Ap with alpha = rs / pAp
__global__ void k_update(int N, double *__restrict__ x, double *__restrict__ r,
                         const double *__restrict__ p, const double *__restrict__ Ap,
                         const double *__restrict__ s) {
    if (s[5] != 0.0) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const double alpha = s[0] / s[1];
    x[j] += alpha * p[j];
    r[j] -= alpha * Ap[j];
}

// p = r + beta p with beta = rs_new / rs (k_advance then sets rs = rs_new
// and the done flag, in a kernel of its own so no workgroup reads a scalar
// another one is writing)
__global__ void k_direction(int N, const double *__restrict__ r, double *__restrict__ p,
                            const double *__restrict__ s) {
    if (s[5] != 0.0) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const double beta = s[2] / s[0];
    p[j] = r[j] + beta * p[j];
}

__global__ void k_advance(double *__restrict__ s, int max_iter) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || s[5] != 0.0) return;
    s[0] = s[2];
    s[6] += 1.0;
    if (!(s[0] > s[4]) || s[6] >= max_iter) s[5] = 1.0;
}

// u = sum_j gamma_j ring_j(P) in panel order (VP_Induced_Velocity_LM.m:37-83)
__global__ void k_induced(Grid g, const double *__restrict__ gam, const double *__restrict__ P,
                          int64_t n, double *__restrict__ u) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = P[i], py = P[n + i], pz = P[2 * n + i];
    const int N = g.np * g.np;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int j = 0; j < N; ++j) {
        double ux, uy, uz;
        ring(g, j, px, py, pz, ux, uy, uz);
        const double gj = gam[j];
        sx = sx + gj * ux;
        sy = sy + gj * uy;
        sz = sz + gj * uz;
    }
    u[i] = sx;
    u[n + i] = sy;
    u[2 * n + i] = sz;
}

}  // namespace gnd
}  // namespace winds

struct winds_ground {
    winds::gnd::Grid g{};
    int N = 0;
    double *A = nullptr;    // N x N
    double *nrm = nullptr;  // [3][N]
    double *work = nullptr; // b, x, r, p, Ap [N] each, scalars
    hipStream_t st = nullptr;
};

namespace {

using winds::set_error;

#define WG_HIP(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_error(WINDS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

unsigned blocks(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

extern "C" {

void winds_ground_destroy(winds_ground *g) {
    if (!g) return;
    if (g->st) (void)hipStreamSynchronize(g->st);
    if (g->A) (void)hipFree(g->A);
    if (g->nrm) (void)hipFree(g->nrm);
    if (g->work) (void)hipFree(g->work);
    if (g->st) (void)hipStreamDestroy(g->st);
    delete g;
}

int winds_ground_create(int32_t np_side, double x_min, double x_max, double y_min, double y_max,
                        double co_frac, winds_ground **out) {
    winds::clear_error();
    if (!out) return set_error(WINDS_E_ARG, "null output");
    *out = nullptr;
    if (np_side < 1 || np_side > 128)
        return set_error(WINDS_E_DIMS, "panels per side must be 1..128");
    if (!(x_max > x_min) || !(y_max > y_min) || !(co_frac >= 0.0))
        return set_error(WINDS_E_ARG, "bad ground plane extent or cut-off");
    winds_ground *g = new winds_ground();
    g->g.np = np_side;
    g->g.x0 = x_min;
    g->g.dx = (x_max - x_min) / np_side;
    g->g.y0 = y_min;
    g->g.dy = (y_max - y_min) / np_side;
    g->g.co = co_frac * (g->g.dx + g->g.dy) / 2;
    const int N = np_side * np_side;
    g->N = N;
    auto fail = [&](int code) {
        winds_ground_destroy(g);
        return code;
    };
#define WGC(expr)                                                                           \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(set_error(WINDS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
    } while (0)
    WGC(hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking));
    WGC(winds::dev_alloc((void **)&g->A, (size_t)N * N * sizeof(double)));
    WGC(winds::dev_alloc((void **)&g->nrm, (size_t)3 * N * sizeof(double)));
    WGC(winds::dev_alloc((void **)&g->work, ((size_t)5 * N + winds::gnd::kS) * sizeof(double)));
    winds::gnd::k_normals<<<blocks(N), 256, 0, g->st>>>(g->g, g->nrm);
    WGC(hipGetLastError());
    winds::gnd::k_matrix<<<blocks((int64_t)N * N), 256, 0, g->st>>>(g->g, g->nrm, g->A);
    WGC(hipGetLastError());
    WGC(hipStreamSynchronize(g->st));
#undef WGC
    *out = g;
    return WINDS_OK;
}

int64_t winds_ground_panels(const winds_ground *g) { return g ? g->N : 0; }

int winds_ground_geometry_dev(const winds_ground *g, double *centres, double *A, void *stream) {
    winds::clear_error();
    if (!g) return set_error(WINDS_E_ARG, "null ground");
    hipStream_t st = (hipStream_t)stream;
    WG_HIP(hipStreamSynchronize(g->st));
    if (centres) {
        winds::gnd::k_centres<<<blocks(g->N), 256, 0, st>>>(g->g, centres);
        WG_HIP(hipGetLastError());
    }
    if (A)
        WG_HIP(hipMemcpyAsync(A, g->A, (size_t)g->N * g->N * sizeof(double),
                              hipMemcpyDeviceToDevice, st));
    return WINDS_OK;
}

int winds_ground_solve_dev(winds_ground *g, const double *v_ext, double *gamma, double tol,
                           int32_t max_iter, void *stream, int32_t *iterations, double *residual) {
    winds::clear_error();
    if (!g || !v_ext || !gamma) return set_error(WINDS_E_ARG, "null argument");
    using namespace winds::gnd;
    hipStream_t st = (hipStream_t)stream;
    WG_HIP(hipStreamSynchronize(g->st));
    const int N = g->N;
    if (!(tol > 0.0)) tol = 1e-13;
    if (max_iter <= 0) max_iter = N;
    double *b = g->work, *r = b + N, *p = r + N, *Ap = p + N, *s = Ap + N + N;
    double *x = gamma;
    k_rhs<<<blocks(N), 256, 0, st>>>(N, v_ext, g->nrm, b, x, r, p);
    WG_HIP(hipGetLastError());
    k_dot<<<1, 1024, 0, st>>>(N, b, b, s, 3, 1);
    k_init_scalars<<<1, 64, 0, st>>>(s, tol);
    WG_HIP(hipGetLastError());
    // iterations in batches; between batches one read of the done flag
    const int batch = 16;
    for (int done_it = 0; done_it < max_iter; done_it += batch) {
        for (int k = 0; k < batch && done_it + k < max_iter; ++k) {
            k_matvec<<<blocks((int64_t)N * 64), 256, 0, st>>>(N, g->A, p, Ap, s);
            k_dot<<<1, 1024, 0, st>>>(N, p, Ap, s, 1, 0);
            k_update<<<blocks(N), 256, 0, st>>>(N, x, r, p, Ap, s);
            k_dot<<<1, 1024, 0, st>>>(N, r, r, s, 2, 0);
            k_direction<<<blocks(N), 256, 0, st>>>(N, r, p, s);
            k_advance<<<1, 64, 0, st>>>(s, max_iter);
        }
        WG_HIP(hipGetLastError());
        double flag = 0.0;
        WG_HIP(hipMemcpyAsync(&flag, s + 5, sizeof(double), hipMemcpyDeviceToHost, st));
        WG_HIP(hipStreamSynchronize(st));
        if (flag != 0.0) break;
    }
    if (iterations || residual) {
        double hs[kS];
        WG_HIP(hipMemcpyAsync(hs, s, sizeof hs, hipMemcpyDeviceToHost, st));
        WG_HIP(hipStreamSynchronize(st));
        if (iterations) *iterations = (int32_t)hs[6];
        if (residual) *residual = hs[3] > 0.0 ? std::sqrt(hs[0] / hs[3]) : 0.0;
    }
    return WINDS_OK;
}

int winds_ground_induced_dev(const winds_ground *g, const double *gamma, const double *P,
                             int64_t n, double *u, void *stream) {
    winds::clear_error();
    if (!g) return set_error(WINDS_E_ARG, "null ground");
    if (n < 0) return set_error(WINDS_E_DIMS, "negative size");
    if (n == 0) return WINDS_OK;
    if (!gamma || !P || !u) return set_error(WINDS_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    WG_HIP(hipStreamSynchronize(g->st));
    winds::gnd::k_induced<<<blocks(n), 256, 0, st>>>(g->g, gamma, P, n, u);
    WG_HIP(hipGetLastError());
    return WINDS_OK;
}

}  // extern "C"
