// bs_api.cpp -- the C ABI declared in include/winds_bs.h.
//
// winds_bs_eval is the drop-in for the MEX gateway: argument checks of
// mexFunction (BiotSavartMex.cu:702-783) become status codes, biot_calc's
// per-call malloc/pack/cudaMalloc/copy/launch/copy/free (:452-685) becomes a
// per-device context with grow-only device buffers, device-side packing in
// the same index order, one launch, and one copy back.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "bs_common.hpp"
#include "bs_kernels.hpp"
#include "winds_bs.h"

namespace {

thread_local std::string g_err;

// Default near radius of the FP32 kernel: pairs whose target lies within
// max(40 r_c, 0.51 |L|) of the segment are evaluated in FP64.  Chosen by
// measurement on the 133,920-target single-wake workload and on smaller
// wakes of other TSR, yaw and size: per-target relative error below 1e-6
// against FP64 for the cut-off and Vatistas cores, for about 9% of the
// evaluation time (20 r_c: 6% faster, max error 7.0e-7 on the single wake;
// 10 r_c: 1.2e-6).  0.51 |L| is also the least fraction the FP32 arithmetic
// allows (every target inside the sphere on the segment as diameter must be
// near).
constexpr double kNearFactor = 40.0;
constexpr double kNearLenFrac = 0.51;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

}  // namespace

namespace winds {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
void clear_error() { g_err.clear(); }
int device_cu_count();
}  // namespace winds

namespace {

#define WINDS_HIP(expr)                                                             \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                       \
            return fail(WINDS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            bytes = 0;
            if (e != hipSuccess) return e;
        }
        size_t want = need < 4096 ? 4096 : need;
        hipError_t e = winds::dev_alloc(&p, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// Per-device state of the host-pointer entry point.
struct HostCtx {
    bool ready = false;
    hipStream_t stream = nullptr;
    DevBuf f1, f2, p, gamma, rc, seg, t3, u3, ucm, len, ws;
};

constexpr int kMaxDev = 64;
std::mutex g_mu;
HostCtx g_ctx[kMaxDev];

int check_device(int dev) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(WINDS_E_DEVICE, "no HIP device visible");
    if (dev < 0 || dev >= n || dev >= kMaxDev)
        return fail(WINDS_E_DEVICE, "device index " + std::to_string(dev) + " out of range");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return fail(WINDS_E_DEVICE, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(WINDS_E_DEVICE, std::string("device arch ") + prop.gcnArchName +
                                        " is not gfx950; this library carries gfx950 code only");
    return WINDS_OK;
}

int cu_count() { return winds::device_cu_count(); }

}  // namespace

hipError_t winds::dev_alloc(void **p, size_t bytes) {
    if (const char *v = std::getenv("WINDS_FAIL_ALLOC_BYTES")) {
        const long long lim = std::atoll(v);
        if (lim > 0 && bytes >= (unsigned long long)lim) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
    }
    return hipMalloc(p, bytes);
}

int winds::device_cu_count() {
    static int cached = 0;
    if (cached) return cached;
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
        cached = n;
    else
        cached = 256;
    return cached;
}

namespace {

bool dims_ok(const int64_t *d) {
    for (int i = 0; i < 4; ++i)
        if (d[i] < 0) return false;
    return true;
}

int to_cfg(const winds_bs_opts *o, winds::EvalCfg &c) {
    if (!o) return fail(WINDS_E_ARG, "null options");
    if (o->core != WINDS_CORE_CUTOFF && o->core != WINDS_CORE_VATISTAS)
        return fail(WINDS_E_MODE, "unknown core model");
    if (o->semantics != WINDS_SEM_MEX && o->semantics != WINDS_SEM_MATLAB)
        return fail(WINDS_E_MODE, "unknown semantics");
    if (o->precision != WINDS_PREC_FP64 && o->precision != WINDS_PREC_FP32C)
        return fail(WINDS_E_MODE, "unknown precision");
    if (o->core == WINDS_CORE_VATISTAS && o->semantics != WINDS_SEM_MATLAB)
        return fail(WINDS_E_MODE, "the Vatistas core exists only with BiotSavart.m semantics");
    if (o->core == WINDS_CORE_VATISTAS && (o->vat_n < 1 || o->vat_n > 64))
        return fail(WINDS_E_MODE, "Vatistas order must be an integer in [1, 64]");
    if (o->block != 0 && (o->block % 64 != 0 || o->block < 64 || o->block > 1024))
        return fail(WINDS_E_ARG, "block must be 0 or a multiple of 64 in [64, 1024]");
    if (o->precision == WINDS_PREC_FP32C && o->block > 256)
        return fail(WINDS_E_ARG, "FP32 compensated kernel takes at most 256 threads per block");
    if (o->split < 0 || o->split > 4096) return fail(WINDS_E_ARG, "split out of range");
    if (std::isnan(o->co)) return fail(WINDS_E_ARG, "co is NaN");
    c.core = o->core;
    c.vat_n = o->vat_n;
    c.semantics = o->semantics;
    c.precision = o->precision;
    c.d = o->d;
    c.co = o->semantics == WINDS_SEM_MATLAB ? o->co : __builtin_inf();
    c.near_factor = o->near_factor > 0 ? o->near_factor : kNearFactor;
    c.near_len_frac = o->near_len_frac > 0 ? o->near_len_frac : kNearLenFrac;
    c.block = o->block;
    c.split = o->split;
    c.seg_col = o->seg_col;
    c.seg_blk = o->seg_blk;
    if (std::isnan(o->guard_kappa)) return fail(WINDS_E_ARG, "guard_kappa is NaN");
    c.guard = o->guard_kappa == 0.0 ? winds::kGuardKappa : o->guard_kappa;
    return WINDS_OK;
}

// Split of the segment range for the FP32 kernel: enough waves to keep every
// SIMD busy at the kernel's occupancy, chosen so the last round of waves is
// as full as possible, and chunks of at most ~3k records (short chunks even
// out the waves that take the FP64 near-pair branch; measured on the
// 1,440-age wake, Vatistas-2: 32 splits 35.0 ms, 64-96 splits 34.0-34.1 ms,
// 129 splits 34.6 ms, 258 splits 35.9 ms).
int pick_split(int64_t nt, int64_t nseg, const winds::EvalCfg &c) {
    const int64_t waves = ((nt + 63) / 64);
    const int64_t tiles = winds::seg32_padded(nseg) / winds::kSegTile;
    // SIMDs x resident waves: the kernel runs 8 waves per SIMD (sized for 4
    // waves where 7 ran, the surrogate's near-wake sum, 6,696 targets x
    // 26,259 segments, took 324 us at 39 splits against 243 us at 68)
    (void)c;
    const int64_t slots = (int64_t)cu_count() * 4 * 8;
    int best = 1;
    double best_cost = 1e300;
    for (int s = 1; s <= 1024; ++s) {
        if (s > 1 && tiles / s < 4) break;
        const double w = (double)(waves * s);
        const double rounds = std::ceil(w / (double)slots);
        // time ~ rounds * (work per wave ~ 1/s); small bias toward fewer splits
        const double cost = rounds / s * (1.0 + 0.002 * s);
        if (cost < best_cost) {
            best_cost = cost;
            best = s;
        }
    }
    const int64_t by_chunk = std::min<int64_t>((nseg + 3071) / 3072, std::min<int64_t>(tiles / 4, 1024));
    return (int)std::max<int64_t>(best, by_chunk);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t ws_bytes(int64_t nt, int64_t nseg, const winds::EvalCfg &c, int *split_out) {
    if (c.precision != WINDS_PREC_FP32C) {
        // FP64 splits only when asked to (the sequential sum is the reference order)
        const int split = c.split > 1 ? c.split : 1;
        if (split_out) *split_out = split;
        return split > 1 ? align256((size_t)split * 3 * (size_t)nt * sizeof(double)) + 256 : 256;
    }
    const int split = c.split > 0 ? c.split : pick_split(nt, nseg, c);
    if (split_out) *split_out = split;
    const bool guard = c.guard >= 0.0;
    const size_t s32 = align256((size_t)winds::seg32_padded(nseg) * sizeof(winds::Seg32));
    const size_t part = split > 1 || guard ? align256((size_t)split * 3 * (size_t)nt * sizeof(double)) : 0;
    const size_t gb = guard ? winds::guard_bytes((int64_t)split * nt, nt) : 0;
    return s32 + part + gb + 256;
}

// The guard's buffers inside an FP32C workspace laid out by ws_bytes.
winds::GuardBufs guard_of(void *ws, int64_t nt, int64_t nseg, int split) {
    char *base = (char *)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    char *g = base + align256((size_t)winds::seg32_padded(nseg) * sizeof(winds::Seg32)) +
              align256((size_t)split * 3 * (size_t)nt * sizeof(double));
    return winds::guard_carve(g, (int64_t)split * nt, nt);
}

int eval_dev_impl(const double *t3, int64_t nt, const winds::Seg64 *seg, int64_t nseg,
                  winds::EvalCfg c, double *u3, void *ws, size_t wsb, hipStream_t st) {
    int split = 1;
    const size_t need = ws_bytes(nt, nseg, c, &split);
    if (c.precision == WINDS_PREC_FP64) {
        double *part = nullptr;
        if (split > 1) {
            if (!ws || wsb < need)
                return fail(WINDS_E_NOMEM, "workspace of " + std::to_string(wsb) + " bytes; need " +
                                               std::to_string(need));
            part = (double *)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
        }
        WINDS_HIP(winds::launch_fp64(t3, nt, seg, nseg, c, u3, st, part));
        return WINDS_OK;
    }
    if (!ws || wsb < need)
        return fail(WINDS_E_NOMEM, "workspace of " + std::to_string(wsb) + " bytes; need " +
                                       std::to_string(need));
    c.split = split;
    char *base = (char *)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    winds::Seg32 *s32 = (winds::Seg32 *)base;
    double *part = (double *)(base + align256((size_t)winds::seg32_padded(nseg) * sizeof(winds::Seg32)));
    winds::GuardBufs g;
    if (c.guard >= 0.0) g = guard_of(ws, nt, nseg, split);
    WINDS_HIP(winds::launch_fp32c(t3, nt, seg, nseg, c, u3, s32, part, g, st));
    return WINDS_OK;
}

bool starts(const char *s, const char *p) { return s && std::strncmp(s, p, std::strlen(p)) == 0; }

}  // namespace

extern "C" {

void winds_bs_opts_default(winds_bs_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->core = WINDS_CORE_CUTOFF;
    o->vat_n = 2;
    o->semantics = WINDS_SEM_MEX;
    o->precision = WINDS_PREC_FP64;
    o->d = 1e-4;
    o->co = __builtin_inf();
    o->near_factor = kNearFactor;
    o->near_len_frac = kNearLenFrac;
    o->guard_kappa = winds::kGuardKappa;
}

const char *winds_last_error(void) { return g_err.c_str(); }

const char *winds_version(void) { return "winds-mi355x 0.1 (gfx950)"; }

int winds_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

size_t winds_bs_workspace_bytes(int64_t n_targets, int64_t n_segments, const winds_bs_opts *o) {
    winds::EvalCfg c;
    if (to_cfg(o, c) != WINDS_OK) return 0;
    return ws_bytes(n_targets, n_segments, c, nullptr);
}

int winds_bs_eval_dev(const double *targets, int64_t n_targets, const double *segments,
                      int64_t n_segments, const winds_bs_opts *o, double *uind, void *workspace,
                      size_t workspace_bytes, void *stream) {
    g_err.clear();
    winds::EvalCfg c;
    int rc = to_cfg(o, c);
    if (rc) return rc;
    if (n_targets < 0 || n_segments < 0) return fail(WINDS_E_DIMS, "negative size");
    if (n_targets > 0 && (!targets || !uind)) return fail(WINDS_E_ARG, "null target/output pointer");
    if (n_segments > 0 && !segments) return fail(WINDS_E_ARG, "null segment pointer");
    if (n_targets > (int64_t)1 << 31 || n_segments > (int64_t)1 << 31)
        return fail(WINDS_E_DIMS, "more than 2^31 targets or segments");
    return eval_dev_impl(targets, n_targets, (const winds::Seg64 *)segments, n_segments, c, uind,
                         workspace, workspace_bytes, (hipStream_t)stream);
}

}  // extern "C"

namespace {

// Workspace of winds_bs_eval_dev_noalloc: a file-scope device array (one
// copy per GPU, held until the process ends).  In host code `scratch` names
// a host-side stand-in, not device memory: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
constexpr size_t kScratchBytes = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDev] = {};

int scratch_address(void **out) {
    int dev = 0;
    WINDS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDev) return fail(WINDS_E_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(WINDS_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                         hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return WINDS_OK;
}

}  // namespace

extern "C" {

int winds_bs_eval_dev_noalloc(const double *targets, int64_t n_targets, const double *segments,
                              int64_t n_segments, const winds_bs_opts *o, double *uind,
                              void *stream) {
    g_err.clear();
    winds::EvalCfg c;
    int rc = to_cfg(o, c);
    if (rc) return rc;
    if (n_targets < 0 || n_segments < 0) return fail(WINDS_E_DIMS, "negative size");
    if (n_targets > 0 && (!targets || !uind)) return fail(WINDS_E_ARG, "null target/output pointer");
    if (n_segments > 0 && !segments) return fail(WINDS_E_ARG, "null segment pointer");
    if (n_targets > (int64_t)1 << 31 || n_segments > (int64_t)1 << 31)
        return fail(WINDS_E_DIMS, "more than 2^31 targets or segments");
    // the tuned split of the FP32C segment range, or the largest that fits
    // in the scratch array; a split asked for explicitly must fit as it is
    int split = 1;
    size_t need = ws_bytes(n_targets, n_segments, c, &split);
    if (need > kScratchBytes && c.precision == WINDS_PREC_FP32C && o->split == 0) {
        while (split > 1 && need > kScratchBytes) {
            c.split = --split;
            need = ws_bytes(n_targets, n_segments, c, &split);
        }
    }
    if (need > kScratchBytes)
        return fail(WINDS_E_NOMEM, "evaluation needs " + std::to_string(need) +
                                       " bytes of workspace; the no-allocation entry point has " +
                                       std::to_string(kScratchBytes));
    void *ws = nullptr;
    rc = scratch_address(&ws);
    if (rc) return rc;
    if (c.precision == WINDS_PREC_FP32C) c.split = split;
    return eval_dev_impl(targets, n_targets, (const winds::Seg64 *)segments, n_segments, c, uind,
                         ws, kScratchBytes, (hipStream_t)stream);
}

int winds_bs_guard_count(const void *workspace, int64_t n_targets, int64_t n_segments,
                         const winds_bs_opts *o, int64_t *count, void *stream) {
    g_err.clear();
    if (!count) return fail(WINDS_E_ARG, "null count");
    *count = 0;
    winds::EvalCfg c;
    int rc = to_cfg(o, c);
    if (rc) return rc;
    if (c.precision != WINDS_PREC_FP32C || c.guard < 0.0 || n_targets <= 0 || n_segments <= 0)
        return WINDS_OK;
    if (!workspace) return fail(WINDS_E_ARG, "null workspace");
    int split = 1;
    (void)ws_bytes(n_targets, n_segments, c, &split);
    const winds::GuardBufs g = guard_of(const_cast<void *>(workspace), n_targets, n_segments, split);
    int v = 0;
    hipStream_t st = (hipStream_t)stream;
    WINDS_HIP(hipMemcpyAsync(&v, g.count, sizeof(int), hipMemcpyDeviceToHost, st));
    WINDS_HIP(hipStreamSynchronize(st));
    *count = v;
    return WINDS_OK;
}

int winds_pack_segments_dev(const double *F1, const double *F2, const int64_t fdims[4],
                            const double *gamma, const double *rc, double *segments,
                            void *stream) {
    g_err.clear();
    if (!fdims || !dims_ok(fdims) || fdims[1] != 3) return fail(WINDS_E_DIMS, "F dims must be [ns,3,nt,nb]");
    const int64_t nf = fdims[0] * fdims[2] * fdims[3];
    if (nf > 0 && (!F1 || !F2 || !gamma || !segments)) return fail(WINDS_E_ARG, "null pointer");
    WINDS_HIP(winds::launch_pack_segments(F1, F2, fdims, gamma, rc, (winds::Seg64 *)segments,
                                          (hipStream_t)stream));
    return WINDS_OK;
}

int winds_pack_targets_dev(const double *P, const int64_t pdims[4], double *targets, void *stream) {
    g_err.clear();
    if (!pdims || !dims_ok(pdims) || pdims[1] != 3) return fail(WINDS_E_DIMS, "P dims must be [ns,3,nt,nb]");
    const int64_t np = pdims[0] * pdims[2] * pdims[3];
    if (np > 0 && (!P || !targets)) return fail(WINDS_E_ARG, "null pointer");
    WINDS_HIP(winds::launch_pack_targets(P, pdims, targets, (hipStream_t)stream));
    return WINDS_OK;
}

int winds_unpack_uind_dev(const double *uind_soa, const int64_t pdims[4], double *uind_cm,
                          void *stream) {
    g_err.clear();
    if (!pdims || !dims_ok(pdims) || pdims[1] != 3) return fail(WINDS_E_DIMS, "P dims must be [ns,3,nt,nb]");
    const int64_t np = pdims[0] * pdims[2] * pdims[3];
    if (np > 0 && (!uind_soa || !uind_cm)) return fail(WINDS_E_ARG, "null pointer");
    WINDS_HIP(winds::launch_unpack_uind(uind_soa, pdims, uind_cm, (hipStream_t)stream));
    return WINDS_OK;
}

int winds_segment_length_dev(const double *segments, int64_t n_segments, int squared, double *L,
                             void *stream) {
    g_err.clear();
    if (n_segments < 0) return fail(WINDS_E_DIMS, "negative size");
    if (n_segments > 0 && (!segments || !L)) return fail(WINDS_E_ARG, "null pointer");
    WINDS_HIP(winds::launch_segment_length((const winds::Seg64 *)segments, n_segments, squared, L,
                                           (hipStream_t)stream));
    return WINDS_OK;
}

int winds_bs_eval(const double *F1, const double *F2, const int64_t fdims[4], const double *P,
                  const int64_t pdims[4], const double *gamma, const double *rc, double d,
                  const char *cmod, double co, const char *type, const char *gpu,
                  const double gpuhw[3], double *uind, double *L) {
    g_err.clear();
    (void)gpu;  // evaluation always runs on the GPU; the flag is accepted as given
    if (!fdims || !pdims) return fail(WINDS_E_ARG, "null dims");
    if (!dims_ok(fdims) || !dims_ok(pdims)) return fail(WINDS_E_DIMS, "negative dimension");
    if (fdims[1] != 3) return fail(WINDS_E_DIMS, "F1/F2 must be [ns,3,nt,nb]");
    if (pdims[1] != 3) return fail(WINDS_E_DIMS, "P must be [ns,3,nt,nb]");
    if (!type) return fail(WINDS_E_ARG, "null type");
    if (!cmod) return fail(WINDS_E_ARG, "null cmod");

    bool length_only = false;
    winds_bs_opts o;
    winds_bs_opts_default(&o);
    if (starts(type, "leng")) {
        length_only = true;
    } else if (starts(type, "full") || starts(type, "fuln") || starts(type, "fuls")) {
        o.precision = WINDS_PREC_FP64;
    } else if (starts(type, "fulc")) {
        o.precision = WINDS_PREC_FP32C;
    } else {
        return fail(WINDS_E_MODE, std::string("unknown type '") + type + "'");
    }
    if (std::strcmp(cmod, "none") == 0 || cmod[0] == 0) {
        o.core = WINDS_CORE_CUTOFF;
        o.semantics = WINDS_SEM_MEX;
        o.d = d;
    } else if (std::strcmp(cmod, "cutm") == 0) {
        o.core = WINDS_CORE_CUTOFF;
        o.semantics = WINDS_SEM_MATLAB;
        o.d = d;
        o.co = co;
    } else if (std::strcmp(cmod, "visc") == 0) {
        if (!(d >= 1.0 && d <= 64.0 && std::floor(d) == d))
            return fail(WINDS_E_MODE, "cmod 'visc' needs d = the integer Vatistas order n");
        o.core = WINDS_CORE_VATISTAS;
        o.vat_n = (int)d;
        o.semantics = WINDS_SEM_MATLAB;
        o.co = co;
    } else {
        return fail(WINDS_E_MODE, std::string("unknown cmod '") + cmod + "'");
    }
    int device = 0;
    if (gpuhw) {
        const double th = gpuhw[0];
        if (!(th == 0.0 || (th >= 64 && th <= 1024 && std::fmod(th, 64.0) == 0.0)))
            return fail(WINDS_E_ARG, "gpuhw(1) threads must be 0 or a multiple of 64 up to 1024");
        o.block = (int)th;
        if (o.precision == WINDS_PREC_FP32C && o.block > 256) o.block = 256;
        if (!(gpuhw[1] <= 1.0)) return fail(WINDS_E_ARG, "gpuhw(2): one GPU per call");
        device = (int)gpuhw[2];
    }
    const int64_t nf = fdims[0] * fdims[2] * fdims[3];
    const int64_t np = pdims[0] * pdims[2] * pdims[3];
    if (nf > 0 && (!F1 || !F2 || !gamma)) return fail(WINDS_E_ARG, "null F1/F2/gamma");
    if (np > 0 && (!P || !uind)) return fail(WINDS_E_ARG, "null P/uind");
    if (o.core == WINDS_CORE_VATISTAS && nf > 0 && !rc) return fail(WINDS_E_ARG, "null rc");
    if (nf > (int64_t)1 << 31 || np > (int64_t)1 << 31) return fail(WINDS_E_DIMS, "too many elements");
    o.seg_col = (int32_t)fdims[0];
    o.seg_blk = (int32_t)(fdims[0] * fdims[2]);

    winds::EvalCfg c;
    int rcode = to_cfg(&o, c);
    if (rcode) return rcode;

    std::lock_guard<std::mutex> lock(g_mu);
    rcode = check_device(device);
    if (rcode) return rcode;
    WINDS_HIP(hipSetDevice(device));
    HostCtx &x = g_ctx[device];
    if (!x.ready) {
        WINDS_HIP(hipStreamCreate(&x.stream));
        x.ready = true;
    }
    hipStream_t st = x.stream;
    const size_t fbytes = (size_t)nf * 3 * sizeof(double);
    const size_t pbytes = (size_t)np * 3 * sizeof(double);
    const size_t gbytes = (size_t)nf * sizeof(double);
    if (nf > 0) {
        WINDS_HIP(x.f1.ensure(fbytes));
        WINDS_HIP(x.f2.ensure(fbytes));
        WINDS_HIP(x.gamma.ensure(gbytes));
        WINDS_HIP(x.rc.ensure(gbytes));
        WINDS_HIP(x.seg.ensure((size_t)nf * sizeof(winds::Seg64)));
        WINDS_HIP(hipMemcpyAsync(x.f1.p, F1, fbytes, hipMemcpyHostToDevice, st));
        WINDS_HIP(hipMemcpyAsync(x.f2.p, F2, fbytes, hipMemcpyHostToDevice, st));
        WINDS_HIP(hipMemcpyAsync(x.gamma.p, gamma, gbytes, hipMemcpyHostToDevice, st));
        if (rc) WINDS_HIP(hipMemcpyAsync(x.rc.p, rc, gbytes, hipMemcpyHostToDevice, st));
        WINDS_HIP(winds::launch_pack_segments(x.f1.as<double>(), x.f2.as<double>(), fdims,
                                              x.gamma.as<double>(), rc ? x.rc.as<double>() : nullptr,
                                              x.seg.as<winds::Seg64>(), st));
    }
    if (L && nf > 0) {
        WINDS_HIP(x.len.ensure(gbytes));
        WINDS_HIP(winds::launch_segment_length(x.seg.as<winds::Seg64>(), nf, length_only ? 0 : 1,
                                               x.len.as<double>(), st));
        WINDS_HIP(hipMemcpyAsync(L, x.len.p, gbytes, hipMemcpyDeviceToHost, st));
    }
    if (length_only) {
        if (np > 0 && uind) std::memset(uind, 0, pbytes);  // BiotSavart.m:75
        WINDS_HIP(hipStreamSynchronize(st));
        return WINDS_OK;
    }
    if (np > 0) {
        WINDS_HIP(x.p.ensure(pbytes));
        WINDS_HIP(x.t3.ensure(pbytes));
        WINDS_HIP(x.u3.ensure(pbytes));
        WINDS_HIP(x.ucm.ensure(pbytes));
        WINDS_HIP(hipMemcpyAsync(x.p.p, P, pbytes, hipMemcpyHostToDevice, st));
        WINDS_HIP(winds::launch_pack_targets(x.p.as<double>(), pdims, x.t3.as<double>(), st));
        const size_t wsb = ws_bytes(np, nf, c, nullptr);
        WINDS_HIP(x.ws.ensure(wsb));
        rcode = eval_dev_impl(x.t3.as<double>(), np, x.seg.as<winds::Seg64>(), nf, c,
                              x.u3.as<double>(), x.ws.p, x.ws.bytes, st);
        if (rcode) return rcode;
        WINDS_HIP(winds::launch_unpack_uind(x.u3.as<double>(), pdims, x.ucm.as<double>(), st));
        WINDS_HIP(hipMemcpyAsync(uind, x.ucm.p, pbytes, hipMemcpyDeviceToHost, st));
    }
    WINDS_HIP(hipStreamSynchronize(st));
    return WINDS_OK;
}

}  // extern "C"

int winds::opts_to_cfg(const winds_bs_opts *o, winds::EvalCfg &c) { return to_cfg(o, c); }

int winds::eval_dev(const double *t3, int64_t nt, const winds::Seg64 *seg, int64_t nseg,
                    const winds::EvalCfg &c, double *u3, void *ws, size_t wsb, hipStream_t st) {
    return eval_dev_impl(t3, nt, seg, nseg, c, u3, ws, wsb, st);
}

size_t winds::eval_ws_bytes(int64_t nt, int64_t nseg, const winds::EvalCfg &c) {
    return ws_bytes(nt, nseg, c, nullptr);
}
