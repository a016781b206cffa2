/*
 * winds_surrogate.h -- inputs and outputs of the far-wake surrogate network
 * (include/winds_mlp.h) for a wake in the device march layout.
 *
 * No counterpart exists in the reference (core/BiotSavart.m sums the whole
 * wake directly).  The network predicts the velocity that the far wake
 * induces at a point of the near wake or the lifting line:
 *
 *   far sources    every segment with an endpoint of age >= split
 *   input          the point in the frame turning with blade 0 (x/R,
 *                  y'/R, z'/R, r/R) and a descriptor of the far wake:
 *                  WINDS_SURR_BINS age bins over [split, max_ages), each
 *                  with the mean tip and root trailing circulation
 *                  (/(U R), x10), the mean axial, radial and lateral
 *                  position of the tip nodes (/R); and cos/sin of the blade-0
 *                  azimuth psi and of nb*psi and the filled fraction of the
 *                  far range.
 *   output         the induced velocity in the turning frame divided by U.
 *
 * Wake layout (as in winds_march.h): nodes SoA [3][N] with node
 * j = m + a*nst + b*nst*ages, trailing circulation [nst, ages-1, nb].
 * in_norm = {mean[NIN], 1/std[NIN]} and out_norm = {mean[3], std[3]}
 * (device FP32) normalise inputs and outputs; NULL means identity.
 */
#ifndef WINDS_SURROGATE_H
#define WINDS_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WINDS_SURR_BINS 8
#define WINDS_SURR_PER_BIN 5
#define WINDS_SURR_NGLOBAL 5
#define WINDS_SURR_NDESC (WINDS_SURR_BINS * WINDS_SURR_PER_BIN + WINDS_SURR_NGLOBAL)
#define WINDS_SURR_NIN (4 + WINDS_SURR_NDESC)
#define WINDS_SURR_NOUT 3

typedef struct winds_surr_geom {
    int32_t nb, nst, ages;   /* wake sizes                                  */
    int32_t split;           /* first far age                               */
    int32_t max_ages;        /* end of the descriptor's age range           */
    int32_t reserved;
    double R;                /* tip radius                                  */
    double U;                /* free-stream speed                           */
    double psi;              /* azimuth of blade 0 (rad), about +x          */
    double hub[3];           /* rotor centre the positions are taken from   */
} winds_surr_geom;

/* feat [np][WINDS_SURR_NIN] for the points P (SoA [3][np]); desc is a device
 * scratch of WINDS_SURR_NDESC floats that receives the raw descriptor. */
int winds_surr_features_dev(const winds_surr_geom *g, const double *X, const double *gt,
                            const double *P, int64_t np, const float *in_norm, float *desc,
                            float *feat, void *stream);

/* Normalised training targets T [np][3] from induced velocities u3 (SoA). */
int winds_surr_labels_dev(const winds_surr_geom *g, const double *u3, int64_t np,
                          const float *out_norm, float *T, void *stream);

/* u3 (SoA [3][np]) += the velocity a network output Y [np][3] stands for. */
int winds_surr_apply_dev(const winds_surr_geom *g, const float *Y, int64_t np,
                         const float *out_norm, double *u3, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* WINDS_SURROGATE_H */
