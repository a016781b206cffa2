/*
 * winds_ground.h -- ground-effect vortex-panel plane (C ABI).
 *
 * Device version of the reference's ground model (ground/Main_Driver.m,
 * ground/VP_Influence_Coefficient_Matrix_v7.m, ground/VP_Induced_Velocity_LM.m):
 * an Np x Np grid of square vortex-ring panels on the plane z = 0 between
 * [x_min, x_max] x [y_min, y_max] (meshgrid nodes, Main_Driver.m:8-22).  Each
 * ring is four straight segments (corners 1-2-3-4 counter-clockwise seen from
 * +z), whose term at a point P is r_a x r_b (n_a + n_b) /
 * (n_a n_b (n_a n_b + r_a . r_b) + co) / (4 pi), r = P - corner, n = |r|, with
 * the absolute cut-off co = co_frac (dx + dy) / 2 (VP_*.m:66-95).
 *
 *   create   grid, panel centres (the control points), panel normals and the
 *            influence matrix A(i, j) = (velocity of ring j per unit
 *            circulation at centre i) . n_j, in FP64 on the device
 *   solve    circulation Gamma with A Gamma = -V_ext . n at every centre (the
 *            reference's inv(A) * V_n, VP_Influence_Coefficient_Matrix_v7.m:
 *            96-98), by conjugate gradients (A is symmetric positive definite
 *            for this flat grid), deterministic reductions
 *   induced  the panels' induced velocity at any points (VP_Induced_Velocity_LM.m)
 *
 * Arrays: panel-major in the reference's order j = m + n*Np (m along x, n
 * along y); points and velocities SoA [3][n] doubles in device memory.
 */
#ifndef WINDS_GROUND_H
#define WINDS_GROUND_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct winds_ground winds_ground;

/* 1 <= Np <= 128 panels per side (A takes 8 Np^4 bytes of device memory:
 * 50 MB at the reference's Np = 50, Main_Driver.m:7); x_max > x_min,
 * y_max > y_min; co_frac >= 0 (Main_Driver.m:23 uses 0.05).  Builds A on the
 * current device.  One object serves one solve at a time: solves that
 * overlap on one object share its work vectors and race. */
int winds_ground_create(int32_t np_side, double x_min, double x_max, double y_min,
                        double y_max, double co_frac, winds_ground **out);
void winds_ground_destroy(winds_ground *g);

/* Number of panels (Np^2). */
int64_t winds_ground_panels(const winds_ground *g);

/* Panel centres (SoA [3][N]) and the influence matrix (row-major N x N,
 * A[i*N + j]) into device buffers; either may be NULL. */
int winds_ground_geometry_dev(const winds_ground *g, double *centres, double *A, void *stream);

/* Gamma [N] (device) from the external velocity at the panel centres
 * (v_ext SoA [3][N], device): free stream plus every other induced velocity
 * there.  Stops when |A Gamma - V_n| <= tol |V_n| (tol <= 0: 1e-13) or after
 * max_iter iterations (<= 0: N); iterations and the final relative residual
 * are returned through the last two pointers (either may be NULL; reading
 * them synchronises stream). */
int winds_ground_solve_dev(winds_ground *g, const double *v_ext, double *gamma, double tol,
                           int32_t max_iter, void *stream, int32_t *iterations,
                           double *residual);

/* u [3][n] (device) = the panels' induced velocity at points P [3][n]
 * (device) for circulation gamma [N] (device). */
int winds_ground_induced_dev(const winds_ground *g, const double *gamma, const double *P,
                             int64_t n, double *u, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* WINDS_GROUND_H */
