// surrogate.hip -- inputs, targets and outputs of the far-wake network.
//
// The far wake (every segment with an endpoint of age >= split) is reduced
// to a small descriptor, one workgroup per age bin; each query point then
// becomes one FP32 input row: its position in the frame turning with blade 0,
// followed by the descriptor.  Network outputs are induced velocities in that frame
// divided by the free-stream speed.

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "bs_common.hpp"
#include "surrogate.hpp"

namespace winds {
namespace surr {

namespace {

constexpr int kThreads = 256;

// One workgroup per age bin; workgroup 0 also writes the global entries.
__global__ __launch_bounds__(kThreads) void k_desc(Geo g, const double *__restrict__ X,
                                                   const double *__restrict__ gt,
                                                   float *__restrict__ desc) {
    const int nst = g.nst, nb = g.nb, A = g.ages, S = g.split;
    const int64_t N = (int64_t)nst * A * nb;
    const int w = max(1, (g.max_ages - S + kBins - 1) / kBins);
    const int k = blockIdx.x;
    __shared__ double red[kPerBin][kThreads];
    const int lo = S + k * w;
    const int hi = min(S + (k + 1) * w, A - 1);
    const int na = max(0, hi - lo);
    const int count = na * nb;
    double acc[kPerBin] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int idx = threadIdx.x; idx < count; idx += kThreads) {
        const int b = idx / na, a = lo + idx % na;
        const int64_t j = (nst - 1) + (int64_t)a * nst + (int64_t)b * nst * A;
        const int64_t t0 = (int64_t)a * nst + (int64_t)b * nst * (A - 1);
        acc[0] += gt[t0 + nst - 1];
        acc[1] += gt[t0];
        const double y = X[N + j] - g.hub[1], z = X[2 * N + j] - g.hub[2];
        acc[2] += X[j] - g.hub[0];
        acc[3] += sqrt(y * y + z * z);
        acc[4] += y;
    }
    for (int q = 0; q < kPerBin; ++q) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < kPerBin; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double inv = count > 0 ? 1.0 / count : 0.0;
        const double gs = 10.0 / (g.U * g.R);
        desc[k * kPerBin + 0] = (float)(red[0][0] * inv * gs);
        desc[k * kPerBin + 1] = (float)(red[1][0] * inv * gs);
        desc[k * kPerBin + 2] = (float)(red[2][0] * inv / g.R);
        desc[k * kPerBin + 3] = (float)(red[3][0] * inv / g.R);
        desc[k * kPerBin + 4] = (float)(red[4][0] * inv / g.R);
    }
    if (k == 0 && threadIdx.x == 0) {
        float *gl = desc + kBins * kPerBin;
        gl[0] = (float)cos(g.psi);
        gl[1] = (float)sin(g.psi);
        gl[2] = (float)cos(g.nb * g.psi);
        gl[3] = (float)sin(g.nb * g.psi);
        const double span = (double)max(1, g.max_ages - 1 - S);
        gl[4] = (float)fmin(1.0, fmax(0.0, (double)(A - 1 - S) / span));
    }
}

__global__ void k_features(Geo g, const float *__restrict__ desc, const double *__restrict__ P,
                           int64_t np, const float *__restrict__ norm, float *__restrict__ feat) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double x = P[i] - g.hub[0], y = P[np + i] - g.hub[1], z = P[2 * np + i] - g.hub[2];
    const double c = cos(g.psi), s = sin(g.psi);
    float f[4];
    f[0] = (float)(x / g.R);
    f[1] = (float)((y * c + z * s) / g.R);
    f[2] = (float)((-y * s + z * c) / g.R);
    f[3] = (float)(sqrt(y * y + z * z) / g.R);
    float *o = feat + i * kNIn;
    for (int k = 0; k < kNIn; ++k) {
        float v = k < 4 ? f[k] : desc[k - 4];
        if (norm) v = (v - norm[k]) * norm[kNIn + k];
        o[k] = v;
    }
}

__global__ void k_labels(Geo g, const double *__restrict__ u3, int64_t np,
                         const float *__restrict__ onorm, float *__restrict__ T) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double c = cos(g.psi), s = sin(g.psi);
    const double ux = u3[i], uy = u3[np + i], uz = u3[2 * np + i];
    const double r[3] = {ux / g.U, (uy * c + uz * s) / g.U, (-uy * s + uz * c) / g.U};
    for (int k = 0; k < 3; ++k) {
        double v = r[k];
        if (onorm) v = (v - onorm[k]) / onorm[3 + k];
        T[i * 3 + k] = (float)v;
    }
}

__global__ void k_apply(Geo g, const float *__restrict__ Y, int64_t np,
                        const float *__restrict__ onorm, double *__restrict__ u3) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double r[3];
    for (int k = 0; k < 3; ++k) {
        double v = Y[i * 3 + k];
        if (onorm) v = v * onorm[3 + k] + onorm[k];
        r[k] = v * g.U;
    }
    const double c = cos(g.psi), s = sin(g.psi);
    u3[i] += r[0];
    u3[np + i] += r[1] * c - r[2] * s;
    u3[2 * np + i] += r[1] * s + r[2] * c;
}

unsigned nblk(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

hipError_t launch_features(const Geo &g, const double *X, const double *gt, const double *P,
                           int64_t np, const float *norm, float *desc, float *feat,
                           hipStream_t st) {
    k_desc<<<kBins, kThreads, 0, st>>>(g, X, gt, desc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || np == 0) return e;
    k_features<<<nblk(np), kThreads, 0, st>>>(g, desc, P, np, norm, feat);
    return hipGetLastError();
}

hipError_t launch_labels(const Geo &g, const double *u3, int64_t np, const float *onorm, float *T,
                         hipStream_t st) {
    if (np == 0) return hipSuccess;
    k_labels<<<nblk(np), kThreads, 0, st>>>(g, u3, np, onorm, T);
    return hipGetLastError();
}

hipError_t launch_apply(const Geo &g, const float *Y, int64_t np, const float *onorm, double *u3,
                        hipStream_t st) {
    if (np == 0) return hipSuccess;
    k_apply<<<nblk(np), kThreads, 0, st>>>(g, Y, np, onorm, u3);
    return hipGetLastError();
}

int check_geo(const Geo &g) {
    if (g.nb < 1 || g.nst < 2 || g.ages < 2 || g.split < 1 || g.max_ages <= g.split)
        return set_error(WINDS_E_DIMS, "bad surrogate wake sizes");
    if (!(g.R > 0.0) || !(g.U > 0.0)) return set_error(WINDS_E_ARG, "R and U must be positive");
    return WINDS_OK;
}

}  // namespace surr
}  // namespace winds

using namespace winds::surr;

namespace {
int hip_status(hipError_t e, const char *what) {
    if (e == hipSuccess) return WINDS_OK;
    return winds::set_error(WINDS_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

extern "C" {

int winds_surr_features_dev(const winds_surr_geom *g, const double *X, const double *gt,
                            const double *P, int64_t np, const float *in_norm, float *desc,
                            float *feat, void *stream) {
    winds::clear_error();
    if (!g || !X || !gt || !desc) return winds::set_error(WINDS_E_ARG, "null argument");
    if (np < 0 || (np > 0 && (!P || !feat))) return winds::set_error(WINDS_E_ARG, "bad points");
    const Geo geo = from_abi(*g);
    int rc = check_geo(geo);
    if (rc) return rc;
    return hip_status(launch_features(geo, X, gt, P, np, in_norm, desc, feat, (hipStream_t)stream),
                      "surrogate features");
}

int winds_surr_labels_dev(const winds_surr_geom *g, const double *u3, int64_t np,
                          const float *out_norm, float *T, void *stream) {
    winds::clear_error();
    if (!g || np < 0 || (np > 0 && (!u3 || !T))) return winds::set_error(WINDS_E_ARG, "bad argument");
    const Geo geo = from_abi(*g);
    int rc = check_geo(geo);
    if (rc) return rc;
    return hip_status(launch_labels(geo, u3, np, out_norm, T, (hipStream_t)stream),
                      "surrogate labels");
}

int winds_surr_apply_dev(const winds_surr_geom *g, const float *Y, int64_t np,
                         const float *out_norm, double *u3, void *stream) {
    winds::clear_error();
    if (!g || np < 0 || (np > 0 && (!u3 || !Y))) return winds::set_error(WINDS_E_ARG, "bad argument");
    const Geo geo = from_abi(*g);
    int rc = check_geo(geo);
    if (rc) return rc;
    return hip_status(launch_apply(geo, Y, np, out_norm, u3, (hipStream_t)stream),
                      "surrogate apply");
}

}  // extern "C"
