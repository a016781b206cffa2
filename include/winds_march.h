/*
 * winds_march.h -- device-resident free-vortex-wake time march (C ABI).
 *
 * Replaces the per-step MATLAB loop of the reference with one HIP stream of
 * kernels on wake state that never leaves the GPU:
 *
 *   setup             core/initials.m:44-152 (two ages, bound circulation
 *                     0.5*|U(1)|*chord*cl_init, r0, vcore, induced velocity,
 *                     forward Euler)
 *   per step p        WInDS.m:127-183: Kelvin circulation update
 *                     (WInDS.m:131-140), vcore + filamentmod (core/vcore.m:
 *                     30-51, core/filamentmod.m:27-41), shed + trailing
 *                     induced velocity at every node (the BiotSavartMex calls
 *                     of WInDS.m:150-154), integrator (numerical/pcc.m:31-75,
 *                     fe.m:30-31, ab2.m:34-40, ab4.m:34-42), Kutta-Joukowski fixed point
 *                     in FP64 with each station's own airfoil table
 *                     (core/KuttaJoukowski.m:42-49, 173-298) and rotor
 *                     thrust/torque/power (postproc/perform.m:16-62).
 *
 * The motion of the rotor comes in as tables per time index -- what the
 * reference computes before its time loop (core/kinematics.m, core/
 * velocity.m; WInDS.m:112-116) -- so any platform, nacelle, shaft and blade
 * motion the caller states drives the march: the first two rows of wake
 * nodes (pos.quarter, pos.trail), the lifting-line points (pos.bound), the
 * blade frames there (pos.nodes.bxn/byn/bzn) and the kinematic inflow in
 * that frame (vel.blade), with fastout's time, azimuth, tip speed ratio,
 * rotor speed and blade pitch and the free stream wind.infty.  The wake keeps
 * at most max_ages ages; older ages are dropped.
 */
#ifndef WINDS_MARCH_H
#define WINDS_MARCH_H

#include <stdint.h>

#include "winds_bs.h"
#include "winds_mlp.h"

#ifdef __cplusplus
extern "C" {
#endif

#define WINDS_INTEG_PCC 0 /* numerical/pcc.m */
#define WINDS_INTEG_FE 1  /* numerical/fe.m  */
#define WINDS_INTEG_AB4 2 /* numerical/ab4.m */
#define WINDS_INTEG_AB2 3 /* numerical/ab2.m (forward Euler at the first step) */

typedef struct winds_march winds_march;

/* Arrays are column-major (MATLAB's layout); [ns,3,nt,nb] means element
 * (m, c, p, b) at m + ns*(c + 3*(p + nt*b)).  ns = nst - 1 stations. */
typedef struct winds_march_cfg {
    int32_t nb;            /* blades (turbine.NumBl)                              */
    int32_t nst;           /* trailing nodes per blade (length(blade.RTrail))     */
    int32_t max_ages;      /* wake truncation (>= 3)                              */
    int32_t n_times;       /* time indices the tables hold (length(fastout.Time)) */
    /* blade (NRELrotor.m:34-55 names), [ns] each */
    const double *r_node;      /* blade.RNodes                                    */
    const double *chord_node;  /* blade.Chord                                     */
    const double *twist_node;  /* blade.AeroTwst, degrees                         */
    const double *dr_node;     /* blade.DRNodes                                   */
    const int32_t *foil;       /* blade.NFoil, 1-based airfoil table per station  */
    const double *cl_init;     /* initial lift coefficient (initials.m:111-119:   */
                               /*   BEM.m, or the elliptic wing's closed form)    */
    double tip_radius;         /* blade.TipRad (initials.m:134)                   */
    /* airfoil tables (airfoils.profiles): n_foils tables, concatenated */
    int32_t n_foils;           /* 1..64                                           */
    int32_t n_alpha;           /* total entries, 2..1024                          */
    const int32_t *foil_len;   /* [n_foils] entries of each table (>= 2)          */
    const double *alpha;       /* [n_alpha] degrees, increasing within a table    */
    const double *cl;          /* [n_alpha]                                       */
    const double *cd;          /* [n_alpha]                                       */
    /* per time index */
    const double *time;        /* [nt] fastout.Time, s                            */
    const double *wind;        /* [nt,3] wind.infty, m/s                          */
    const double *tsr;         /* [nt] fastout.TipSpdRat                          */
    const double *rot_speed;   /* [nt] fastout.RotSpeed, rpm                      */
    const double *azimuth;     /* [nt] fastout.Azimuth, degrees                   */
    const double *pitch;       /* [nt,nb] fastout.BldPitch1..nb, degrees          */
    const double *hub;         /* [nt,3] pos.hub                                  */
    const double *quarter;     /* [nst,3,nt,nb] pos.quarter (wake age 0)          */
    const double *trail;       /* [nst,3,nt,nb] pos.trail (wake age 1)            */
    const double *bound;       /* [ns,3,nt,nb] pos.bound (lifting line)           */
    const double *bxn;         /* [ns,3,nt,nb] pos.nodes.bxn                      */
    const double *byn;         /* [ns,3,nt,nb] pos.nodes.byn                      */
    const double *bzn;         /* [ns,3,nt,nb] pos.nodes.bzn                      */
    const double *vblade;      /* [ns,3,nt,nb] vel.blade                          */
    const double *ll_extra;    /* [ns,3,nt,nb] induced velocity of other sources  */
                               /*   at the lifting line (vel.uind_tower,          */
                               /*   KuttaJoukowski.m:232-238), or NULL            */
    double dt;             /* time step 1/user.t(3), s (fe.m:30)                  */
    double relax;          /* KJ relaxation (WInDS.m:53)                          */
    double tol;            /* KJ tolerance (WInDS.m:41)                           */
    int32_t maxiter;       /* KJ iterations (WInDS.m:47)                          */
    int32_t integrator;    /* WINDS_INTEG_*                                       */
    winds_bs_opts wake;    /* evaluation of the wake on its own nodes             */
    winds_bs_opts ll;      /* evaluation at the lifting line (KJ)                 */
} winds_march_cfg;

/* Build the state and run the reference's setup (time index 1).  Enqueued on
 * an internal stream; later calls keep the order. */
int winds_march_create(const winds_march_cfg *cfg, winds_march **out);
void winds_march_destroy(winds_march *m);

/* Advance nsteps time steps (WInDS.m loop body), asynchronously.  Stepping
 * time index p needs the tables of p+1 (the integrator places the new first
 * two rows of nodes there); beyond that the call fails. */
int winds_march_step(winds_march *m, int32_t nsteps);

/* on != 0: once the wake is at max_ages, without the surrogate, each step
 * is recorded into a HIP graph the first time its phase of the buffer
 * rotation (six phases) comes round, after two steady steps run eagerly,
 * and replayed from then on (the step's time index, history row, r0, Time
 * and free stream are read from device memory that a small kernel writes
 * before each step).  The results are those of the eager steps, bit for
 * bit.  Off by default; changing it, set_state and set_surrogate drop the
 * recorded graphs. */
int winds_march_set_graphs(winds_march *m, int32_t on);

/* Wait for queued work; report a diverging Kutta-Joukowski solve
 * (NaN lift coefficient, KuttaJoukowski.m:272-274) as an error. */
int winds_march_sync(winds_march *m);

/* out = {next time index p, ages, nodes, segments, ages of the circulation
 * of time p-1 (the last solved one), steps taken}. */
int winds_march_info(const winds_march *m, int64_t out[6]);

/* Copies to host, MATLAB layouts: nodes [nst,3,ages,nb]; shed circulation
 * [ns,1,ages,nb]; trailing circulation [nst,1,ages-1,nb]; each may be NULL. */
int winds_march_get_state(winds_march *m, double *nodes, double *gshed, double *gtrail);

/* Start from a prepared wake: nodes [nst,3,ages,nb] of time index p (>= 4,
 * p < n_times)
 * and the circulation of time p-1, shed [ns,1,ages-1,nb] (bound = age 1)
 * and trailing [nst,1,ages-2,nb], as the previous step would leave them. */
int winds_march_set_state(winds_march *m, int32_t p, int32_t ages, const double *nodes,
                          const double *gshed, const double *gtrail);

/* Performance history rows {p, T [N], Q [N m], P [W], KJ iterations} for the
 * steps [first, first + n) counted from the first stepped time index. */
int winds_march_get_perf(winds_march *m, int32_t first, int32_t n, double *rows);

/* Last Kutta-Joukowski result at the lifting line, [ns, nb] each. */
int winds_march_get_kj(winds_march *m, double *cl, double *cd, double *aoa);

/* Total induced velocity of the last wake evaluation at every node, SoA
 * [3][nodes] in device memory (valid until the next step). */
const double *winds_march_uind_dev(const winds_march *m);

/* Far-wake surrogate (no counterpart in the reference, which sums the whole
 * wake at every node, WInDS.m:150-154).  Once the wake is longer than split
 * ages, a step evaluates
 *   near targets (nodes of age < target_ages) and the lifting line:
 *       direct sum over the near segments (both endpoints younger than split)
 *       + the far segments' velocity from the network (WINDS_SURR_MLP), the
 *       direct far sum (WINDS_SURR_DIRECT) or nothing (WINDS_SURR_DROP);
 *   older nodes: the induced velocity of the node of the same age at the
 *       last check, held in the frame turning with the rotor.
 * Every check_every steps (and on the first such step) the step's wake
 * evaluation at time p is a full direct sum (with PCC the corrector's
 * evaluation at the predicted nodes then goes through the surrogate as on
 * other steps, except with check_every = 1, which is the plain march); it
 * refreshes the held velocities and, in WINDS_SURR_MLP,
 * compares the network with the direct far sum at the near targets and the
 * lifting line: error = max |u_net - u_far| / max |u_ind| over each point
 * set, u_ind being the whole wake's induced velocity there (the velocity the
 * network's error perturbs).  A check above tol switches the following steps
 * to the direct far sum until a check passes; train_iters Adam steps (rate
 * lr) on the check's direct sums follow every check.  The network needs
 * WINDS_SURR_NIN inputs, 3 outputs and max_batch >= nst*target_ages*nb +
 * (nst-1)*nb; it is not owned and must outlive its use.  s == NULL turns the
 * surrogate off. */
#define WINDS_SURR_MLP 0
#define WINDS_SURR_DIRECT 1
#define WINDS_SURR_DROP 2

typedef struct winds_march_surr {
    int32_t mode;
    int32_t split;          /* first far age, 2 <= split < max_ages          */
    int32_t target_ages;    /* 1 <= target_ages <= split                      */
    int32_t check_every;    /* >= 1                                           */
    int32_t train_iters;    /* >= 0                                           */
    int32_t reserved;
    double tol;             /* fallback threshold, e.g. 0.01                  */
    double lr;              /* Adam rate of the online training               */
    const float *in_norm;   /* host [2][WINDS_SURR_NIN]: mean, 1/std          */
    const float *out_norm;  /* host [2][3]: mean, std                         */
} winds_march_surr;

int winds_march_set_surrogate(winds_march *m, winds_mlp *net, const winds_march_surr *s);

/* out = {checks, failed checks, network steps, direct-far steps, last node
 * error, last lifting-line error, largest error, Adam steps, last node and
 * lifting-line errors relative to max |u_far| instead of max |u_ind|}. */
int winds_march_surr_stats(const winds_march *m, double out[10]);

#ifdef __cplusplus
}
#endif

#endif /* WINDS_MARCH_H */
