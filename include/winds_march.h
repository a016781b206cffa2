/*
 * winds_march.h -- device-resident free-vortex-wake time march (C ABI).
 *
 * Replaces the per-step MATLAB loop of the reference with one HIP stream of
 * kernels on wake state that never leaves the GPU:
 *
 *   setup             core/initials.m:44-152 (two ages, bound circulation
 *                     0.5*|U|*chord*cl(geometric aoa), r0, vcore, induced
 *                     velocity, forward Euler)
 *   per step p        WInDS.m:127-183: Kelvin circulation update
 *                     (WInDS.m:131-140), vcore + filamentmod (core/vcore.m:
 *                     30-51, core/filamentmod.m:27-41), shed + trailing
 *                     induced velocity at every node (the BiotSavartMex calls
 *                     of WInDS.m:150-154), integrator (numerical/pcc.m:31-75,
 *                     fe.m:30-31, ab4.m:34-42), Kutta-Joukowski fixed point
 *                     in FP64 (core/KuttaJoukowski.m:42-49, 173-298) and rotor
 *                     thrust/torque/power (postproc/perform.m:42-62).
 *
 * The rotor is rigid (constant Omega, no platform motion): blade b at
 * azimuth 2 pi b/nb + Omega t, quarter chord r e_r, trailing edge
 * r e_r + 0.75 c e_c, blade frame bxn = e_c, bzn = e_r, byn = bzn x bxn with
 * e_c = -cos(theta) e_t + sin(theta) e_x, theta = twist + pitch.  The wake
 * keeps at most max_ages ages; older ages are dropped.
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

typedef struct winds_march winds_march;

typedef struct winds_march_cfg {
    int32_t nb;            /* blades                                           */
    int32_t nst;           /* trailing nodes per blade (ns = nst-1 stations)   */
    int32_t max_ages;      /* wake truncation (>= 3)                           */
    int32_t n_alpha;       /* polar table length, 2..1024                      */
    const double *r_trail;      /* [nst] node radii (blade.RTrail)             */
    const double *chord_trail;  /* [nst] (blade.ChordTrail)                    */
    const double *twist_trail;  /* [nst] degrees (blade.AeroTwstTrail)         */
    const double *alpha;        /* [n_alpha] degrees, increasing               */
    const double *cl;           /* [n_alpha]                                   */
    const double *cd;           /* [n_alpha]                                   */
    double R;              /* tip radius (initials.m:134)                      */
    double tsr;            /* tip speed ratio (initials.m:134)                 */
    double omega;          /* rad/s                                            */
    double dt;             /* time step, s                                     */
    double pitch_deg;      /* blade pitch                                      */
    double uinf[3];        /* free stream, m/s                                 */
    double relax;          /* KJ relaxation (WInDS.m:53)                       */
    double tol;            /* KJ tolerance (WInDS.m:41)                        */
    int32_t maxiter;       /* KJ iterations (WInDS.m:47)                       */
    int32_t integrator;    /* WINDS_INTEG_*                                    */
    int32_t max_steps;     /* capacity of the performance history             */
    int32_t reserved;
    winds_bs_opts wake;    /* evaluation of the wake on its own nodes          */
    winds_bs_opts ll;      /* evaluation at the lifting line (KJ)              */
} winds_march_cfg;

/* Build the state and run the reference's setup (time index 1).  Enqueued on
 * an internal stream; later calls keep the order. */
int winds_march_create(const winds_march_cfg *cfg, winds_march **out);
void winds_march_destroy(winds_march *m);

/* Advance nsteps time steps (WInDS.m loop body), asynchronously. */
int winds_march_step(winds_march *m, int32_t nsteps);

/* Wait for queued work; report a diverging Kutta-Joukowski solve
 * (NaN lift coefficient, KuttaJoukowski.m:272-274) as an error. */
int winds_march_sync(winds_march *m);

/* out = {next time index p, ages, nodes, segments, ages of the circulation
 * of time p-1 (the last solved one), steps taken}. */
int winds_march_info(const winds_march *m, int64_t out[6]);

/* Copies to host, MATLAB layouts: nodes [nst,3,ages,nb]; shed circulation
 * [ns,1,ages,nb]; trailing circulation [nst,1,ages-1,nb]; each may be NULL. */
int winds_march_get_state(winds_march *m, double *nodes, double *gshed, double *gtrail);

/* Start from a prepared wake: nodes [nst,3,ages,nb] of time index p (>= 4)
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
 * Every check_every steps (and on the first such step) the step is a full
 * direct sum; it refreshes the held velocities and, in WINDS_SURR_MLP,
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
