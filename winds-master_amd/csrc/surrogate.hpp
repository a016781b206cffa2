// surrogate.hpp -- launchers shared by the surrogate C ABI and the march.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "winds_surrogate.h"

namespace winds {
namespace surr {

constexpr int kBins = WINDS_SURR_BINS;
constexpr int kPerBin = WINDS_SURR_PER_BIN;
constexpr int kNDesc = WINDS_SURR_NDESC;
constexpr int kNIn = WINDS_SURR_NIN;
constexpr int kNOut = WINDS_SURR_NOUT;

struct Geo {
    int nb, nst, ages, split, max_ages;
    double R, U, psi;
    double hub[3];  // rotor centre: positions are taken relative to it
};

inline Geo from_abi(const winds_surr_geom &g) {
    return Geo{g.nb, g.nst, g.ages, g.split, g.max_ages, g.R, g.U, g.psi,
               {g.hub[0], g.hub[1], g.hub[2]}};
}

int check_geo(const Geo &g);

// the descriptor (one workgroup per age bin), then one input row per point
hipError_t launch_features(const Geo &g, const double *X, const double *gt, const double *P,
                           int64_t np, const float *norm, float *desc, float *feat,
                           hipStream_t st);
hipError_t launch_labels(const Geo &g, const double *u3, int64_t np, const float *onorm, float *T,
                         hipStream_t st);
hipError_t launch_apply(const Geo &g, const float *Y, int64_t np, const float *onorm, double *u3,
                        hipStream_t st);

}  // namespace surr
}  // namespace winds
