// bs_batch.hip -- many independent wakes of uneven size in one persistent launch.
//
// A plan splits every wake into work items (wake, tile of `block` targets,
// chunk of `chunk` segments).  Items are ordered longest first and, among
// equal items, chunk-major so that concurrently running workgroups read the
// same prepared segment records from L2.  One persistent grid (workgroups
// per CU x CUs) pulls items from an atomic counter; each item writes its
// partial sums to its own slot, and a reduce kernel adds the slots of each
// target in chunk order, so the result is independent of scheduling.  The
// whole evaluation (record preparation, counter reset, persistent kernel,
// reduce) can be recorded into a hipGraph once and replayed.
//
// FP32C: partial sums per (target tile, segment chunk) item, reduced per
// target in chunk order, then the cancellation guard (targets whose FP32
// terms cancel are summed again in FP64).  FP64: one item per target tile
// with the wake's whole segment range, each lane summing its target
// sequentially in segment order -- the order of BiotSavart_nogpu
// (BiotSavartMex.cu:268-282) and of the single-wake FP64 kernels -- so every
// wake's result is bitwise equal to its own FP64 evaluation.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "bs_common.hpp"
#include "bs_fp32c.hpp"
#include "bs_kernels.hpp"
#include "winds_bs.h"

namespace winds {

struct __attribute__((aligned(16))) WorkItem {
    int64_t t0;      // first target (global index)
    int64_t k0, k1;  // prepared-record range (global padded index)
    int64_t shift;   // FP64 record index = prepared index + shift
    int32_t tn;      // targets in this item (<= block)
    int32_t slot;    // partial-sum slot
};

struct BatchWake {
    int64_t t_off, nt;   // targets
    int64_t s_off, ns;   // FP64 records
    int64_t p_off, np;   // prepared records (padded to kSegTile)
    int64_t slot0;       // first slot of the wake
    int32_t nchunk;      // chunks per target tile
    int32_t ntile;       // target tiles
};

__device__ __forceinline__ int find_wake_by(const BatchWake *w, int n, int64_t k, bool by_p) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t off = by_p ? w[mid].p_off : w[mid].t_off;
        if (off <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void prep_seg32_batch(const Seg64 *__restrict__ seg, const BatchWake *__restrict__ wk,
                                 int n_wakes, int64_t np_total, int core_kind, int n, double d,
                                 double near_factor, double near_len_frac,
                                 Seg32 *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np_total) return;
    const int w = find_wake_by(wk, n_wakes, k, true);
    const int64_t j = k - wk[w].p_off;
    const Seg64 *src = j < wk[w].ns ? seg + wk[w].s_off + j : nullptr;
    const Seg64 *prev = j > 0 && j < wk[w].ns ? src - 1 : nullptr;
    const Seg64 *next = j + 1 < wk[w].ns ? src + 1 : nullptr;
    out[k] = make_seg32(src, prev, next, core_kind, n, d, near_factor, near_len_frac);
}

__global__ void reset_counter(int *c, int *guard_count) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *c = 0;
        if (guard_count) *guard_count = 0;
    }
}

// 8 waves per SIMD (at most 64 VGPRs), with the second record of each
// pair formed in FP32 (r1 - L), which is what fits the loop in 64 VGPRs:
// 1,071 G int/s on the 64-wake batch against 992 at 7 waves (72 VGPRs) and
// 943 at 8 waves with FP64 end points throughout (spills).  The compiler's
// own choice, 74 VGPRs and 6 waves, measured 9% slower than 7 waves.
template <int CORE, int SEM, bool CO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void bs_batch_persistent(
    const double *__restrict__ tx, const double *__restrict__ ty, const double *__restrict__ tz,
    const Seg32 *__restrict__ s32, const Seg64 *__restrict__ s64,
    const WorkItem *__restrict__ items, int n_items, int *__restrict__ counter, Fp32Params P,
    double *__restrict__ part, float *__restrict__ est) {
    __shared__ int s_item;
    extern __shared__ NearLds near_lds[];  // one per wave
    const int B = blockDim.x;
    const int lane = threadIdx.x;
    for (;;) {
        if (lane == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        // the index is the same for the whole workgroup; readfirstlane keeps
        // it (and the item fields and segment records) in scalar registers
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= n_items) break;
        const WorkItem w = items[it];
        const bool active = lane < w.tn;
        const int64_t t = w.t0 + (active ? lane : 0);
        double sx, sy, sz;
        float sa;
        fp32c_range<CORE, SEM, CO>(tx[t], ty[t], tz[t], s32, s64, w.shift, w.k0, w.k1, P,
                                   near_lds + (lane >> 6), sx, sy, sz, sa);
        if (active) {
            double *o = part + (int64_t)w.slot * 3 * B;
            o[lane] = sx;
            o[B + lane] = sy;
            o[2 * B + lane] = sz;
            if (est) est[(int64_t)w.slot * B + lane] = sa;
        }
    }
}

// FP64 items: every lane sums its target over [k0, k1) of the FP64 records
// in segment order (fp64_range) and stores the result itself.  MODE 0
// biot_p2p, 1 BiotSavart.m cut-off, 2 BiotSavart.m Vatistas of order NF (1,
// 2) or n (NF = 0).  1-2 waves per SIMD, as the single-set FP64 kernel.
template <int MODE, int NF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void bs_batch_fp64(
    const double *__restrict__ tx, const double *__restrict__ ty, const double *__restrict__ tz,
    const Seg64 *__restrict__ s64, const WorkItem *__restrict__ items, int n_items,
    int *__restrict__ counter, double d, int n, double co, double *__restrict__ ux,
    double *__restrict__ uy, double *__restrict__ uz) {
    __shared__ int s_item;
    const int lane = threadIdx.x;
    for (;;) {
        if (lane == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= n_items) break;
        const WorkItem w = items[it];
        if (lane < w.tn) {
            const int64_t t = w.t0 + lane;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            fp64_range<MODE, NF>(s64, w.k0, w.k1, tx[t], ty[t], tz[t], d, n, co, sx, sy, sz);
            ux[t] = sx;
            uy[t] = sy;
            uz[t] = sz;
        }
    }
}

// Slots of each target added in chunk order; with est, the guard's test
// (as bs_reduce_guard) appends the targets whose FP32 terms cancel to list.
__global__ void bs_batch_reduce(const double *__restrict__ part, const float *__restrict__ est,
                                const BatchWake *__restrict__ wk, int n_wakes, int64_t nt_total,
                                int B, double kappa, double *__restrict__ ux,
                                double *__restrict__ uy, double *__restrict__ uz,
                                uint32_t *__restrict__ list, int *__restrict__ count) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nt_total) return;
    const int w = find_wake_by(wk, n_wakes, g, false);
    const int64_t t = g - wk[w].t_off;
    const int64_t tile = t / B;
    const int lane = (int)(t % B);
    const int nc = wk[w].nchunk;
    const int64_t s0 = wk[w].slot0 + tile * nc;
    const double *p = part + s0 * 3 * B;
    double sx = 0.0, sy = 0.0, sz = 0.0, a2 = 0.0;
    for (int c = 0; c < nc; ++c, p += 3 * B) {
        sx += p[lane];
        sy += p[B + lane];
        sz += p[2 * B + lane];
        if (est) a2 += (double)est[(s0 + c) * B + lane];
    }
    ux[g] = sx;
    uy[g] = sy;
    uz[g] = sz;
    if (est && !(sqrt(a2) <= kappa * sqrt(sx * sx + sy * sy + sz * sz)))
        list[atomicAdd(count, 1)] = (uint32_t)g;
}

}  // namespace winds

using winds::BatchWake;
using winds::WorkItem;

struct winds_batch {
    int device = 0;
    winds::EvalCfg cfg;
    winds_batch_tune tune{};
    int n_wakes = 0;
    std::vector<int64_t> nt, ns;
    std::vector<BatchWake> hw;
    int64_t nt_total = 0, ns_total = 0, np_total = 0, inter = 0;
    int64_t n_items = 0, n_slots = 0;
    BatchWake *d_wakes = nullptr;
    winds::WakeSpan *d_spans = nullptr;
    WorkItem *d_items = nullptr;
    int *d_counter = nullptr;
    int grid = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *gkey[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace {

using winds::set_error;

#define WB_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return set_error(WINDS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

void free_plan_device(winds_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    if (b->d_wakes) (void)hipFree(b->d_wakes);
    if (b->d_spans) (void)hipFree(b->d_spans);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_counter) (void)hipFree(b->d_counter);
    b->d_wakes = nullptr;
    b->d_spans = nullptr;
    b->d_items = nullptr;
    b->d_counter = nullptr;
    for (auto &k : b->gkey) k = nullptr;
}

template <int K, int S, bool C>
int occ_blocks(int block) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, winds::bs_batch_persistent<K, S, C>, block,
                                                     (size_t)(block / 64) *
                                                         sizeof(winds::NearLds)) != hipSuccess ||
        n < 1)
        n = std::max(1, 16 / std::max(1, block / 64));
    return n;
}

template <int M, int F>
int occ_blocks64(int block) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, winds::bs_batch_fp64<M, F>, block, 0) !=
            hipSuccess ||
        n < 1)
        n = std::max(1, 8 / std::max(1, block / 64));
    return n;
}

// Resident workgroups per CU of the persistent kernel instance used by cfg.
int default_wg_per_cu(const winds::EvalCfg &c, int block) {
    const int n = c.core == 0 ? 0 : c.vat_n;
    if (c.precision == WINDS_PREC_FP64) {
        if (c.semantics == 0) return occ_blocks64<0, 0>(block);
        if (n == 0) return occ_blocks64<1, 0>(block);
        if (n == 1) return occ_blocks64<2, 1>(block);
        if (n == 2) return occ_blocks64<2, 2>(block);
        return occ_blocks64<2, 0>(block);
    }
    const bool co = c.semantics == 1 && c.co < 1.0e18;
    const int kind = c.core == 0 ? 0 : (n == 1 ? 1 : (n == 2 ? 2 : 3));
    if (c.semantics == 0) return occ_blocks<0, 0, false>(block);
    if (co) {
        switch (kind) {
            case 0: return occ_blocks<0, 1, true>(block);
            case 1: return occ_blocks<1, 1, true>(block);
            case 2: return occ_blocks<2, 1, true>(block);
            default: return occ_blocks<3, 1, true>(block);
        }
    }
    switch (kind) {
        case 0: return occ_blocks<0, 1, false>(block);
        case 1: return occ_blocks<1, 1, false>(block);
        case 2: return occ_blocks<2, 1, false>(block);
        default: return occ_blocks<3, 1, false>(block);
    }
}

int build_plan(winds_batch *b) {
    const int B = b->tune.block;
    const bool fp64 = b->cfg.precision == WINDS_PREC_FP64;
    // FP64: the whole segment range per item (the reference's sequential order)
    const int64_t chunk = fp64 ? INT64_MAX / 4 : b->tune.chunk;
    b->hw.assign(b->n_wakes, BatchWake{});
    int64_t toff = 0, soff = 0, poff = 0, slot = 0, inter = 0;
    for (int w = 0; w < b->n_wakes; ++w) {
        BatchWake &x = b->hw[w];
        x.t_off = toff;
        x.nt = b->nt[w];
        x.s_off = soff;
        x.ns = b->ns[w];
        x.p_off = poff;
        x.np = fp64 ? x.ns : winds::seg32_padded(x.ns);
        x.ntile = (int32_t)((x.nt + B - 1) / B);
        x.nchunk = x.np > 0 ? (int32_t)((x.np + chunk - 1) / chunk) : 0;
        x.slot0 = slot;
        slot += (int64_t)x.ntile * x.nchunk;
        toff += x.nt;
        soff += x.ns;
        poff += x.np;
        inter += x.nt * x.ns;
    }
    b->nt_total = toff;
    b->ns_total = soff;
    b->np_total = poff;
    b->inter = inter;
    b->n_slots = slot;
    if (slot > INT32_MAX) return set_error(WINDS_E_DIMS, "too many work items");
    std::vector<WorkItem> items;
    items.reserve((size_t)slot);
    for (int w = 0; w < b->n_wakes; ++w) {
        const BatchWake &x = b->hw[w];
        for (int c = 0; c < x.nchunk; ++c) {
            const int64_t k0 = x.p_off + (int64_t)c * std::min(chunk, x.np);
            const int64_t k1 = std::min(k0 + chunk, x.p_off + x.np);
            for (int t = 0; t < x.ntile; ++t) {
                WorkItem it;
                it.t0 = x.t_off + (int64_t)t * B;
                it.tn = (int32_t)std::min<int64_t>(B, x.nt - (int64_t)t * B);
                it.k0 = k0;
                it.k1 = k1;
                it.shift = x.s_off - x.p_off;
                it.slot = (int32_t)(x.slot0 + (int64_t)t * x.nchunk + c);
                items.push_back(it);
            }
        }
    }
    // longest first; equal items keep their chunk-major order
    std::stable_sort(items.begin(), items.end(), [](const WorkItem &a, const WorkItem &c) {
        return (int64_t)a.tn * (a.k1 - a.k0) > (int64_t)c.tn * (c.k1 - c.k0);
    });
    b->n_items = (int64_t)items.size();
    free_plan_device(b);
    std::vector<winds::WakeSpan> spans(b->n_wakes);
    for (int w = 0; w < b->n_wakes; ++w)
        spans[w] = winds::WakeSpan{b->hw[w].t_off, b->hw[w].nt, b->hw[w].s_off, b->hw[w].ns};
    WB_HIP(winds::dev_alloc((void **)&b->d_wakes, sizeof(BatchWake) * std::max(1, b->n_wakes)));
    WB_HIP(winds::dev_alloc((void **)&b->d_spans, sizeof(winds::WakeSpan) * std::max(1, b->n_wakes)));
    WB_HIP(winds::dev_alloc((void **)&b->d_items, sizeof(WorkItem) * std::max<size_t>(1, items.size())));
    WB_HIP(winds::dev_alloc((void **)&b->d_counter, 256));
    WB_HIP(hipMemcpy(b->d_wakes, b->hw.data(), sizeof(BatchWake) * b->n_wakes, hipMemcpyHostToDevice));
    WB_HIP(hipMemcpy(b->d_spans, spans.data(), sizeof(winds::WakeSpan) * b->n_wakes,
                     hipMemcpyHostToDevice));
    if (!items.empty())
        WB_HIP(hipMemcpy(b->d_items, items.data(), sizeof(WorkItem) * items.size(),
                         hipMemcpyHostToDevice));
    const int wg = b->tune.wg_per_cu > 0 ? b->tune.wg_per_cu : default_wg_per_cu(b->cfg, B);
    b->grid = std::max(1, winds::device_cu_count() * wg);
    return WINDS_OK;
}

bool guarded(const winds_batch *b) {
    return b->cfg.precision == WINDS_PREC_FP32C && b->cfg.guard >= 0.0;
}

size_t ws_need(const winds_batch *b) {
    if (b->cfg.precision == WINDS_PREC_FP64) return 256;
    const size_t n_est = (size_t)b->n_slots * b->tune.block;
    return a256((size_t)b->np_total * sizeof(winds::Seg32)) +
           a256((size_t)b->n_slots * 3 * b->tune.block * sizeof(double)) +
           (guarded(b) ? winds::guard_bytes((int64_t)n_est, b->nt_total) : 0) + 256;
}

winds::GuardBufs guard_of(const winds_batch *b, void *ws) {
    char *base = (char *)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    char *g = base + a256((size_t)b->np_total * sizeof(winds::Seg32)) +
              a256((size_t)b->n_slots * 3 * b->tune.block * sizeof(double));
    return winds::guard_carve(g, (int64_t)b->n_slots * b->tune.block, b->nt_total);
}

template <int K, int S, bool C>
void launch_persistent(const winds_batch *b, const double *t3, const winds::Seg32 *s32,
                       const winds::Seg64 *s64, const winds::Fp32Params &P, double *part,
                       float *est, hipStream_t st) {
    winds::bs_batch_persistent<K, S, C><<<b->grid, b->tune.block,
                                          (size_t)(b->tune.block / 64) * sizeof(winds::NearLds),
                                          st>>>(
        t3, t3 + b->nt_total, t3 + 2 * b->nt_total, s32, s64, b->d_items, (int)b->n_items,
        b->d_counter, P, part, est);
}

int enqueue_fp64(const winds_batch *b, const double *t3, const double *segs, double *u3,
                 hipStream_t st) {
    const winds::EvalCfg &c = b->cfg;
    const winds::Seg64 *s64 = (const winds::Seg64 *)segs;
    const int n = c.core == 0 ? 0 : c.vat_n;
    const int64_t N = b->nt_total;
    // wakes without segments keep u = 0: zero everything first (a kernel,
    // not a memset, so that graph capture records no memset node)
    WB_HIP(winds::launch_zero(u3, 3 * N, st));
    winds::reset_counter<<<1, 64, 0, st>>>(b->d_counter, nullptr);
    WB_HIP(hipGetLastError());
    if (b->n_items == 0) return WINDS_OK;
#define WB_FP64(M, F)                                                                          \
    winds::bs_batch_fp64<M, F><<<b->grid, b->tune.block, 0, st>>>(                               \
        t3, t3 + N, t3 + 2 * N, s64, b->d_items, (int)b->n_items, b->d_counter, c.d, n, c.co, u3, \
        u3 + N, u3 + 2 * N)
    if (c.semantics == 0)
        WB_FP64(0, 0);
    else if (n == 0)
        WB_FP64(1, 0);
    else if (n == 1)
        WB_FP64(2, 1);
    else if (n == 2)
        WB_FP64(2, 2);
    else
        WB_FP64(2, 0);
#undef WB_FP64
    WB_HIP(hipGetLastError());
    return WINDS_OK;
}

int enqueue(const winds_batch *b, const double *t3, const double *segs, double *u3, void *ws,
            hipStream_t st) {
    const winds::EvalCfg &c = b->cfg;
    if (b->nt_total == 0) return WINDS_OK;
    if (c.precision == WINDS_PREC_FP64) return enqueue_fp64(b, t3, segs, u3, st);
    char *base = (char *)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    winds::Seg32 *s32 = (winds::Seg32 *)base;
    double *part = (double *)(base + a256((size_t)b->np_total * sizeof(winds::Seg32)));
    const winds::Seg64 *s64 = (const winds::Seg64 *)segs;
    const int n = c.core == 0 ? 0 : c.vat_n;
    const bool guard = guarded(b);
    winds::GuardBufs g;
    if (guard) g = guard_of(b, ws);
    if (b->np_total > 0) {
        winds::prep_seg32_batch<<<(unsigned)((b->np_total + 255) / 256), 256, 0, st>>>(
            s64, b->d_wakes, b->n_wakes, b->np_total, c.core == 0 ? 0 : 1, n, c.d, c.near_factor,
            c.near_len_frac, s32);
        WB_HIP(hipGetLastError());
    }
    winds::reset_counter<<<1, 64, 0, st>>>(b->d_counter, guard ? g.count : nullptr);
    WB_HIP(hipGetLastError());
    const winds::Fp32Params P = winds::fp32_params(c);
    float *est = guard ? g.est : nullptr;
    const bool has_co = c.semantics == 1 && c.co < 1.0e18;
    const int kind = c.core == 0 ? 0 : (n == 1 ? 1 : (n == 2 ? 2 : 3));
    if (b->n_items > 0) {
        if (c.semantics == 0) {
            launch_persistent<0, 0, false>(b, t3, s32, s64, P, part, est, st);
        } else if (has_co) {
            switch (kind) {
                case 0: launch_persistent<0, 1, true>(b, t3, s32, s64, P, part, est, st); break;
                case 1: launch_persistent<1, 1, true>(b, t3, s32, s64, P, part, est, st); break;
                case 2: launch_persistent<2, 1, true>(b, t3, s32, s64, P, part, est, st); break;
                default: launch_persistent<3, 1, true>(b, t3, s32, s64, P, part, est, st); break;
            }
        } else {
            switch (kind) {
                case 0: launch_persistent<0, 1, false>(b, t3, s32, s64, P, part, est, st); break;
                case 1: launch_persistent<1, 1, false>(b, t3, s32, s64, P, part, est, st); break;
                case 2: launch_persistent<2, 1, false>(b, t3, s32, s64, P, part, est, st); break;
                default: launch_persistent<3, 1, false>(b, t3, s32, s64, P, part, est, st); break;
            }
        }
        WB_HIP(hipGetLastError());
    }
    winds::bs_batch_reduce<<<(unsigned)((b->nt_total + 255) / 256), 256, 0, st>>>(
        part, est, b->d_wakes, b->n_wakes, b->nt_total, b->tune.block, c.guard, u3,
        u3 + b->nt_total, u3 + 2 * b->nt_total, g.list, g.count);
    WB_HIP(hipGetLastError());
    if (guard)
        WB_HIP(winds::launch_guard_fix(t3, b->nt_total, s64, b->d_spans, b->n_wakes,
                                       winds::WakeSpan{0, 0, 0, 0}, c, g, u3, st));
    return WINDS_OK;
}

int check_tune(const winds_batch_tune &t) {
    if (t.block != 64 && t.block != 128 && t.block != 256)
        return set_error(WINDS_E_ARG, "batch block must be 64, 128 or 256");
    if (t.chunk < winds::kSegTile || t.chunk % winds::kSegTile != 0)
        return set_error(WINDS_E_ARG, "batch chunk must be a positive multiple of the segment tile");
    if (t.wg_per_cu < 0 || t.wg_per_cu > 64) return set_error(WINDS_E_ARG, "wg_per_cu out of range");
    return WINDS_OK;
}

}  // namespace

extern "C" {

int winds_batch_create(int32_t n_wakes, const int64_t *n_targets, const int64_t *n_segments,
                       const winds_bs_opts *o, const winds_batch_tune *tune, winds_batch **out) {
    winds::clear_error();
    if (!out) return set_error(WINDS_E_ARG, "null output");
    *out = nullptr;
    if (n_wakes < 0 || (n_wakes > 0 && (!n_targets || !n_segments)))
        return set_error(WINDS_E_ARG, "bad wake list");
    winds::EvalCfg c;
    int rc = winds::opts_to_cfg(o, c);
    if (rc) return rc;
    for (int w = 0; w < n_wakes; ++w)
        if (n_targets[w] < 0 || n_segments[w] < 0) return set_error(WINDS_E_DIMS, "negative size");
    winds_batch_tune t{256, 16384, 0, 0};
    if (tune) t = *tune;
    rc = check_tune(t);
    if (rc) return rc;
    winds_batch *b = new winds_batch();
    b->cfg = c;
    b->tune = t;
    b->n_wakes = n_wakes;
    b->nt.assign(n_targets, n_targets + n_wakes);
    b->ns.assign(n_segments, n_segments + n_wakes);
    if (hipGetDevice(&b->device) != hipSuccess) {
        delete b;
        return set_error(WINDS_E_DEVICE, "no HIP device");
    }
    rc = build_plan(b);
    if (rc) {
        free_plan_device(b);
        delete b;
        return rc;
    }
    if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess) {
        free_plan_device(b);
        delete b;
        return set_error(WINDS_E_HIP, "stream creation failed");
    }
    *out = b;
    return WINDS_OK;
}

void winds_batch_destroy(winds_batch *b) {
    if (!b) return;
    free_plan_device(b);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    delete b;
}

size_t winds_batch_workspace_bytes(const winds_batch *b) { return b ? ws_need(b) : 0; }

int winds_batch_guard_count(const winds_batch *b, const void *workspace, int64_t *count,
                            void *stream) {
    winds::clear_error();
    if (!b || !count) return set_error(WINDS_E_ARG, "null argument");
    *count = 0;
    if (!guarded(b) || b->nt_total == 0) return WINDS_OK;
    if (!workspace) return set_error(WINDS_E_ARG, "null workspace");
    const winds::GuardBufs g = guard_of(b, const_cast<void *>(workspace));
    int v = 0;
    hipStream_t st = (hipStream_t)stream;
    WB_HIP(hipMemcpyAsync(&v, g.count, sizeof(int), hipMemcpyDeviceToHost, st));
    WB_HIP(hipStreamSynchronize(st));
    *count = v;
    return WINDS_OK;
}

int winds_batch_info(const winds_batch *b, int64_t out[4]) {
    if (!b || !out) return set_error(WINDS_E_ARG, "null argument");
    out[0] = b->nt_total;
    out[1] = b->ns_total;
    out[2] = b->inter;
    out[3] = b->n_items;
    return WINDS_OK;
}

int winds_batch_eval_dev(winds_batch *b, const double *targets, const double *segments,
                         double *uind, void *workspace, size_t workspace_bytes, void *stream,
                         int use_graph) {
    winds::clear_error();
    if (!b) return set_error(WINDS_E_ARG, "null plan");
    if (b->nt_total > 0 && (!targets || !uind)) return set_error(WINDS_E_ARG, "null target/output");
    if (b->ns_total > 0 && !segments) return set_error(WINDS_E_ARG, "null segments");
    if (!workspace || workspace_bytes < ws_need(b))
        return set_error(WINDS_E_NOMEM, "workspace too small: need " + std::to_string(ws_need(b)));
    hipStream_t st = (hipStream_t)stream;
    if (!use_graph) return enqueue(b, targets, segments, uind, workspace, st);
    const void *key[4] = {targets, segments, uind, workspace};
    if (!b->exec || std::memcmp(key, b->gkey, sizeof key) != 0) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        WB_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
        int rc = enqueue(b, targets, segments, uind, workspace, b->cap);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_error(WINDS_E_HIP, std::string("capture: ") + hipGetErrorString(e));
        b->graph = g;
        WB_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        std::memcpy(b->gkey, key, sizeof key);
    }
    WB_HIP(hipGraphLaunch(b->exec, st));
    return WINDS_OK;
}

int winds_batch_autotune(winds_batch *b, const double *targets, const double *segments,
                         double *uind, void *stream, winds_batch_tune *best, double *best_ms) {
    winds::clear_error();
    if (!b) return set_error(WINDS_E_ARG, "null plan");
    hipStream_t st = (hipStream_t)stream;
    const int blocks[] = {256, 128, 64};
    const int chunks[] = {2048, 4096, 8192, 16384, 32768};
    const winds_batch_tune start = b->tune;
    winds_batch_tune win = start;
    double win_ms = 1e300;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    WB_HIP(hipEventCreate(&e0));
    WB_HIP(hipEventCreate(&e1));
    int rc = WINDS_OK;
    bool first = true;
    // FP64 items take whole wakes: the chunk does not apply
    const int n_chunks = b->cfg.precision == WINDS_PREC_FP64 ? 1 : (int)(sizeof chunks / sizeof *chunks);
    for (int B : blocks) {
        for (int ci = 0; ci < n_chunks; ++ci) {
            const int ch = chunks[ci];
            b->tune = winds_batch_tune{B, ch, 0, 0};
            rc = build_plan(b);
            if (rc) break;
            const size_t need = ws_need(b);
            if (need > ws_bytes) {
                if (ws) (void)hipFree(ws);
                ws = nullptr;
                if (winds::dev_alloc((void **)&ws, need) != hipSuccess) {
                    ws_bytes = 0;
                    continue;  // candidate does not fit in memory: skip it
                }
                ws_bytes = need;
            }
            if (first) {  // warm the code and caches once
                rc = enqueue(b, targets, segments, uind, ws, st);
                if (rc) break;
                first = false;
            }
            (void)hipEventRecord(e0, st);
            rc = enqueue(b, targets, segments, uind, ws, st);
            if (rc) break;
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = set_error(WINDS_E_HIP, "autotune synchronisation failed");
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < win_ms) {
                win_ms = ms;
                win = b->tune;
            }
        }
        if (rc) break;
    }
    if (ws) (void)hipFree(ws);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tune = rc ? start : win;
    int rc2 = build_plan(b);
    if (rc) return rc;
    if (rc2) return rc2;
    if (best) *best = win;
    if (best_ms) *best_ms = win_ms;
    return WINDS_OK;
}

}  // extern "C"
