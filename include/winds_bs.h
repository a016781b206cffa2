/*
 * winds_bs.h -- C ABI of the MI355X free-vortex-wake induced-velocity engine.
 *
 * Every entry point takes plain pointers and sizes and returns a status code
 * (0 = success, negative = error; winds_last_error() holds the message).
 * Nothing throws or longjmps across the boundary.
 *
 * Two layers:
 *
 *  1. winds_bs_eval -- the drop-in.  Same 11 inputs and 2 outputs as the
 *     MATLAB gateway [uind,L] = BiotSavartMex(F1,F2,P,gamma,rc,d,cmod,co,
 *     type,gpu,gpuhw) (cuda/source/BiotSavartMex.cu:14-28, mexFunction at
 *     :688-807), with MATLAB column-major host arrays in and out.  A
 *     mexFunction shim (winds-master_amd/mex/BiotSavartMex.cpp) forwards to it.
 *
 *  2. Device-resident calls for the time march and benchmarks: inputs already
 *     in device memory, caller-provided workspace, caller's HIP stream.
 *
 * Device data layout:
 *   targets   3 x n doubles, SoA: x[0..n), y[0..n), z[0..n)
 *   segments  n records of 8 doubles: ax ay az bx by bz gamma rc
 *             (A = F1 endpoint, B = F2 endpoint, rc = squared core radius)
 *   uind      3 x n doubles, SoA, like targets
 */
#ifndef WINDS_BS_H
#define WINDS_BS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define WINDS_OK 0
#define WINDS_E_ARG (-1)      /* bad argument (null pointer, bad string)   */
#define WINDS_E_DIMS (-2)     /* array dimensions inconsistent             */
#define WINDS_E_MODE (-3)     /* unsupported cmod / type combination       */
#define WINDS_E_HIP (-4)      /* HIP runtime error                         */
#define WINDS_E_NOMEM (-5)    /* device allocation or workspace too small  */
#define WINDS_E_DEVICE (-6)   /* no usable gfx950 device                   */

/* core models (BiotSavart.m:49-53, 102-108) */
#define WINDS_CORE_CUTOFF 0   /* den += d*|L|^2                            */
#define WINDS_CORE_VATISTAS 1 /* factor h^2 (rc^n + h^2n)^(-1/n)           */

/* semantics */
#define WINDS_SEM_MEX 0    /* BiotSavartMex: co ignored, nothing zeroed (BiotSavartMex.cu:63-93) */
#define WINDS_SEM_MATLAB 1 /* BiotSavart.m: NaN/Inf zeroed, r1>co & r2>co zeroed (BiotSavart.m:111) */

/* arithmetic */
#define WINDS_PREC_FP64 0  /* FP64, reference operation order, sequential sum  */
#define WINDS_PREC_FP32C 1 /* FP32 with compensated sums; FP64 for near pairs   */

typedef struct winds_bs_opts {
    int32_t core;          /* WINDS_CORE_*                                   */
    int32_t vat_n;         /* Vatistas order n >= 1 (core == VATISTAS)       */
    int32_t semantics;     /* WINDS_SEM_*                                    */
    int32_t precision;     /* WINDS_PREC_*                                   */
    double d;              /* squared cut-off fraction (core == CUTOFF)      */
    double co;             /* far cut-off distance (SEM_MATLAB); +inf = none */
    double near_factor;    /* FP32C: FP64 for pairs with the target within   */
    double near_len_frac;  /*   max(near_factor*r_c, near_len_frac*|L|) of   */
                           /*   the segment; <= 0 selects 40 and 0.51;       */
                           /*   fractions below 0.51 act as 0.51             */
    int32_t block;         /* threads per block; 0 = tuned; the FP32C and    */
                           /*   MEX-semantics FP64 kernels use at most 256   */
    int32_t split;         /* segment-range split; FP32C: 0 = tuned; FP64:  */
                           /*   0/1 = one sequential sum (reference order),  */
                           /*   > 1 = split sums (for few targets)           */
    int32_t seg_col;       /* SEM_MATLAB+FP64: segments per station column   */
    int32_t seg_blk;       /*   and per blade, for the nested MATLAB sum     */
    double guard_kappa;    /* FP32C cancellation guard: a target whose FP32  */
                           /*   terms' 16-segment partial sums have a root   */
                           /*   sum of squares above guard_kappa * |u| is    */
                           /*   evaluated again in FP64 over all segments;   */
                           /*   0 selects 1.0, negative turns the guard off  */
} winds_bs_opts;

/* Fill o with the defaults: cut-off, d = 1e-4, MEX semantics, FP64. */
void winds_bs_opts_default(winds_bs_opts *o);

/* Message of the last failed call on this thread ("" if none). */
const char *winds_last_error(void);

/* Library version string. */
const char *winds_version(void);

/*
 * Drop-in for BiotSavartMex (cuda/source/BiotSavartMex.cu:688-807).
 *
 *  F1, F2  [fdims] = [ns,3,nt,nb] column-major doubles, segment endpoints
 *  P       [pdims] = [ns',3,nt',nb'] column-major doubles, targets
 *          (callers pad 2-D/3-D P to 4-D, BiotSavartMex.cu:745-764)
 *  gamma   [ns,1,nt,nb] circulation (ns*nt*nb values are read)
 *  rc      [ns,1,nt,nb] squared core radius (ns*nt*nb values are read), or
 *          NULL: then every r_c is 0 (allowed for the cut-off cores only;
 *          the FP32C near radius is then max(0.51|L|, 0))
 *  d       cut-off (cmod "none"/"cutm") or Vatistas order n (cmod "visc")
 *  cmod    "none": cut-off core, MEX semantics (what BiotSavartMex computes)
 *          "cutm": cut-off core, BiotSavart.m semantics (zeroing, co)
 *          "visc": Vatistas core of order n = d, BiotSavart.m semantics;
 *                  this is how WInDS.m:84-91 splits d = 'viscN' for the MEX
 *  co      far cut-off distance (BiotSavart.m:111); ignored for "none"
 *  type    "full", "fuln", "fuls": FP64 direct sum (BiotSavartMex.cu:591-608)
 *          "fulc": FP32 compensated direct sum with FP64 near pairs
 *          "leng", "length": segment lengths only (BiotSavartMex.cu:582-587,
 *                  BiotSavart.m:74-76)
 *  gpu     "true" or "false"; the engine evaluates on the GPU either way
 *  gpuhw   [threads, ncards, device]: threads per block (0 or a multiple of
 *          64 up to 1024), ncards must be 0 or 1, device = HIP device index
 *  uind    out, [pdims] column-major
 *  L       out, [ns,1,nt,nb] of the F dims: |F2-F1| for "leng"/"length",
 *          |F2-F1|^2 otherwise (BiotSavart.m:71-76); may be NULL
 */
int winds_bs_eval(const double *F1, const double *F2, const int64_t fdims[4],
                  const double *P, const int64_t pdims[4], const double *gamma,
                  const double *rc, double d, const char *cmod, double co,
                  const char *type, const char *gpu, const double gpuhw[3],
                  double *uind, double *L);

/*
 * Same sum on device memory (targets/segments/uind laid out as above) on the
 * given HIP stream (NULL = default stream).  The call only enqueues work.
 * workspace must hold winds_bs_workspace_bytes(...) bytes of device memory,
 * 256-byte aligned.  Replaces biot_calc's per-call malloc/copy/launch/copy
 * (BiotSavartMex.cu:452-685) for callers that keep the wake on the device.
 */
size_t winds_bs_workspace_bytes(int64_t n_targets, int64_t n_segments,
                                const winds_bs_opts *o);
int winds_bs_eval_dev(const double *targets, int64_t n_targets,
                      const double *segments, int64_t n_segments,
                      const winds_bs_opts *o, double *uind, void *workspace,
                      size_t workspace_bytes, void *stream);

/*
 * winds_bs_eval_dev without a workspace argument and without any device
 * allocation: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The workspace is a 64 MiB device array
 * inside the library (one per GPU, held until the process ends); the FP32C
 * segment-range split is lowered until the evaluation fits in it, and an
 * evaluation that does not fit even unsplit (about 1M segments, or 2.4M
 * targets at 1 split) returns WINDS_E_NOMEM before launching anything.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (different streams, different threads) race and must be serialised by the
 * caller.  Targets, segments and uind are as for winds_bs_eval_dev.
 */
int winds_bs_eval_dev_noalloc(const double *targets, int64_t n_targets,
                              const double *segments, int64_t n_segments,
                              const winds_bs_opts *o, double *uind, void *stream);

/*
 * Number of targets the FP32C cancellation guard evaluated again in FP64 in
 * the last evaluation that used this workspace with these sizes and options
 * (0 for FP64 or with the guard off).  Synchronises stream.  No counterpart
 * in the reference (it has no FP32 path).
 */
int winds_bs_guard_count(const void *workspace, int64_t n_targets, int64_t n_segments,
                         const winds_bs_opts *o, int64_t *count, void *stream);

/*
 * Column-major [ns,3,nt,nb] endpoint arrays (device) to segment records
 * (device), in BiotSavartMex's packed order k = m + n*ns + p*ns*nt
 * (crunch_f/crunch_rg, BiotSavartMex.cu:316-357).
 */
int winds_pack_segments_dev(const double *F1, const double *F2,
                            const int64_t fdims[4], const double *gamma,
                            const double *rc, double *segments, void *stream);

/* Column-major [ns,3,nt,nb] targets (device) to SoA (crunch_p, :361-379). */
int winds_pack_targets_dev(const double *P, const int64_t pdims[4],
                           double *targets, void *stream);

/* SoA uind (device) back to column-major [ns,3,nt,nb] (reassem_uind, :383-401). */
int winds_unpack_uind_dev(const double *uind_soa, const int64_t pdims[4],
                          double *uind_cm, void *stream);

/* Segment lengths |B-A| (squared = 0) or |B-A|^2 (squared = 1) per record
 * (calcLengthOnly, BiotSavartMex.cu:197-206; FilamentLength.m:36-41). */
int winds_segment_length_dev(const double *segments, int64_t n_segments,
                             int squared, double *L, void *stream);

/* Number of visible HIP devices (0 when no GPU); never fails. */
int winds_device_count(void);

/*
 * Batched evaluation of many independent wakes of uneven size in one
 * persistent launch.  No counterpart exists in the reference, which evaluates
 * one wake per BiotSavartMex call (BiotSavartMex.cu:452-685); this is the
 * same operator applied to wake w's own targets and segments for every w.
 *
 * Wake w owns targets [t_off, t_off + n_targets[w]) of one concatenated SoA
 * target array (t_off = sum of earlier n_targets) and segment records
 * [s_off, s_off + n_segments[w]) of one concatenated record array.  uind is
 * SoA over all targets.  The plan holds the work list (wake, target tile,
 * segment chunk) in longest-first order; a persistent kernel pulls items
 * from an atomic counter; per-item partial sums are reduced per target in a
 * fixed order, so results do not depend on scheduling.
 *
 * WINDS_PREC_FP32C: FP32 compensated sums with FP64 near pairs and the
 * cancellation guard (guard_kappa).  WINDS_PREC_FP64: one item per target
 * tile covering the wake's whole segment range, each target summed in
 * segment order with the reference arithmetic of the semantics selected
 * (biot_p2p, BiotSavartMex.cu:63-93, 268-282; BiotSavart.m:86-115), so each
 * wake's result is bitwise equal to winds_bs_eval_dev in FP64 on that wake
 * alone (seg_col = seg_blk = 0).  FP64 needs no workspace beyond 256 bytes.
 */
typedef struct winds_batch winds_batch;

typedef struct winds_batch_tune {
    int32_t block;      /* targets per work item = threads per workgroup (64, 128, 256) */
    int32_t chunk;      /* segments per work item (multiple of 16)                      */
    int32_t wg_per_cu;  /* persistent workgroups per compute unit; 0 = by occupancy     */
    int32_t reserved;
} winds_batch_tune;

/* tune == NULL selects block 256, chunk 16384, workgroups by occupancy. */
int winds_batch_create(int32_t n_wakes, const int64_t *n_targets,
                       const int64_t *n_segments, const winds_bs_opts *o,
                       const winds_batch_tune *tune, winds_batch **out);
void winds_batch_destroy(winds_batch *b);
size_t winds_batch_workspace_bytes(const winds_batch *b);
/* Enqueue one evaluation on stream.  use_graph != 0 records the evaluation
 * into a hipGraph on first use (per set of pointers) and replays it after. */
int winds_batch_eval_dev(winds_batch *b, const double *targets, const double *segments,
                         double *uind, void *workspace, size_t workspace_bytes,
                         void *stream, int use_graph);
/* Time each candidate launch shape on these inputs, keep the fastest in b,
 * and report it (and its time in ms) through best / best_ms. */
int winds_batch_autotune(winds_batch *b, const double *targets, const double *segments,
                         double *uind, void *stream, winds_batch_tune *best,
                         double *best_ms);
/* Totals: targets, segments, interactions (sum over wakes of nt*ns), items. */
int winds_batch_info(const winds_batch *b, int64_t out[4]);
/* Targets the FP32C guard evaluated again in FP64 in the last evaluation
 * with this workspace (0 for FP64 or with the guard off).  Synchronises
 * stream. */
int winds_batch_guard_count(const winds_batch *b, const void *workspace, int64_t *count,
                            void *stream);

#ifdef __cplusplus
}
#endif

#endif /* WINDS_BS_H */
