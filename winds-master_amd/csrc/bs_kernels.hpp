// bs_kernels.hpp -- host-side view of the direct-sum kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bs_pair.hpp"

namespace winds {

// Segments per compensated partial sum in the FP32 kernel; segment tables
// are padded to a multiple of it.
#ifndef WINDS_SEG_TILE
#define WINDS_SEG_TILE 16
#endif
constexpr int kSegTile = WINDS_SEG_TILE;

// Prepared segment record of the FP32 kernel (64 bytes, scalar loads).
// The end point B stays FP64 so that r2 = P-B is formed in FP64 and rounded
// once.  Filaments are chains (a trailing line, a shed line): when segment
// k starts where segment k-1 ends, r1 of segment k is r2 of segment k-1 and
// the kernel carries it (with |r|^2 and 1/|r|) from one record to the next;
// otherwise it forms r1 from the Seg64 record's start point.  The FP64
// near-pair branch reads the Seg64 record.
struct __attribute__((aligned(64))) Seg32 {
    double bx, by, bz;   // F2 endpoint
    float lx, ly, lz;    // L = B - A, rounded once
    float near2b;        // end-point threshold at B: near2 + |L|^2, raised to
                         // the next record's near2 + |L|^2 when that record
                         // is chained, so one compare at B serves both
    float il2;           // 1/|L|^2 (0 for a zero-length segment)
    float g;             // gamma / (4 pi) (0 for a zero-length segment)
    float core;          // cut-off: d |L|^2; Vatistas: max(rc^n, FLT_MIN) (rc = r_c^2)
    float near2;         // squared near radius R^2 (cut-off core: R^2 |L|^2,
                         // compared with |r1 x r2|^2); +inf: always FP64
    int32_t flags;       // kSegChained, kSegSkip
    float near2a;        // end-point threshold at A: near2 + |L|^2 (read when
                         // the record is not chained)
};
constexpr int32_t kSegChained = 1;  // starts at the previous record's end point
constexpr int32_t kSegSkip = 2;     // padding: no term
static_assert(sizeof(Seg32) == 64, "Seg32 layout");
static_assert(sizeof(Seg64) == 64, "Seg64 layout");

// Cancellation guard of the FP32 path.  Each target also accumulates
// a2 = sum over its 16-record tiles of |tile partial|^2 (FP32 terms only).
// The FP32 error of a target stays below ~4e-7 * sqrt(a2) (measured by
// emulating the kernel's arithmetic on rotor wakes of 87k-134k targets,
// Vatistas-2 and cut-off cores), so a target with sqrt(a2) > kappa |u| has
// FP32 terms that cancel to a small result, and its sum is evaluated again
// in FP64 over every segment.  kappa = 1 keeps the FP32 error below ~4e-7
// |u| and selects ~0.1% of the targets of those wakes.
constexpr double kGuardKappa = 1.0;
// Targets re-evaluated with the segment range split across kFixSplit
// workgroups (partials reduced in a fixed order); targets past kFixCap in
// the list are summed by one workgroup each.
constexpr int kFixSplit = 64;
constexpr int64_t kFixCap = 2048;

// Targets [t_off, t_off + nt) of one wake and its FP64 records
// [s_off, s_off + ns).
struct WakeSpan {
    int64_t t_off, nt, s_off, ns;
};

struct EvalCfg {
    int core = 0;        // 0 cut-off, 1 Vatistas
    int vat_n = 2;
    int semantics = 0;   // 0 MEX, 1 MATLAB
    int precision = 0;   // 0 FP64, 1 FP32C
    double d = 1e-4;
    double co = __builtin_inf();
    double near_factor = 40.0;
    double near_len_frac = 0.51;
    int block = 0;
    int split = 0;
    int64_t seg_col = 0;
    int64_t seg_blk = 0;
    double guard = kGuardKappa;  // FP32C guard threshold kappa; < 0: off
};

// Device buffers of the FP32 guard: n_est a2 partials (one per target and
// partial sum), the list of targets to re-evaluate (room for nt) and its
// length, and the partials of the split re-evaluation
// ([kFixCap][kFixSplit][3]).
struct GuardBufs {
    float *est = nullptr;
    uint32_t *list = nullptr;
    int *count = nullptr;
    double *fix = nullptr;
};
size_t guard_bytes(int64_t n_est, int64_t nt);
GuardBufs guard_carve(void *base, int64_t n_est, int64_t nt);

int64_t seg32_padded(int64_t nseg);

hipError_t launch_zero(double *p, int64_t n, hipStream_t st);
// FP64 sum.  split > 1 (with part holding split*3*nt doubles) splits the
// segment range across workgroups; otherwise one sequential sum per target.
hipError_t launch_fp64(const double *t3, int64_t nt, const Seg64 *seg, int64_t nseg,
                       const EvalCfg &c, double *u3, hipStream_t st, double *part = nullptr);
// FP32 compensated sum.  part holds max(split, 1) * 3 * nt doubles; with
// c.guard >= 0, g carries the guard's buffers (split = the split in c).
hipError_t launch_fp32c(const double *t3, int64_t nt, const Seg64 *seg, int64_t nseg,
                        const EvalCfg &c, double *u3, Seg32 *s32, double *part,
                        const GuardBufs &g, hipStream_t st);
// FP64 re-evaluation of the targets in g.list (count at g.count): the
// guard's second pass.  wk (device, n_wk entries) locates each target's
// wake; with wk null the single wake `one` holds every target.
hipError_t launch_guard_fix(const double *t3, int64_t nt_total, const Seg64 *seg,
                            const WakeSpan *wk, int n_wk, WakeSpan one, const EvalCfg &c,
                            const GuardBufs &g, double *u3, hipStream_t st);
// Zero the guard's counter (a kernel, so that graph capture records no memset).
hipError_t launch_guard_reset(const GuardBufs &g, hipStream_t st);
hipError_t launch_pack_segments(const double *F1, const double *F2, const int64_t *fd,
                                const double *gamma, const double *rc, Seg64 *out,
                                hipStream_t st);
hipError_t launch_pack_targets(const double *P, const int64_t *pd, double *t3, hipStream_t st);
hipError_t launch_unpack_uind(const double *u3, const int64_t *pd, double *U, hipStream_t st);
hipError_t launch_segment_length(const Seg64 *seg, int64_t nseg, int squared, double *L,
                                 hipStream_t st);

}  // namespace winds
