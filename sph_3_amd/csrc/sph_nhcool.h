// sph_nhcool.h -- cooling update from a table over (log n_H, log T, r, log N_H)
// (sph_nhcool.hip).
#pragma once
#include "sph_common.h"

#define SPH_NH_GRID_MAX 2048  // grid points of the four axes together (staged in LDS)

// The table of HCoolMu.bin (layout read at CR/v3t.cu:45-122): the four grids, their clamp
// bounds, and log10 heating, log10 cooling, mu and log10 u on [nH][T][r][NH].
struct NHTable {
    int dims[4];      // n_nH, n_T, n_r, n_NH
    float bounds[8];  // nH, T, r, NH: low, high
    const float *grids;  // the four grids in that order (sum of dims floats)
    const float *heat, *cool, *mu, *logu;
};
// Code units and step: rho_cgs = rho unit_density, u_cgs = u unit_u, dt_s = dt unit_time.
struct NHCoolArgs {
    float unit_density, unit_u, unit_time, XH, gamma, dt;
    double u_floor;
};

namespace sph {

// u_out[i] of doCloudyHCooling for n particles (positions x, y, z, density rho, energy u,
// log10 N_H); typ (may be null) skips particles with typ != 0 (u unchanged).  floors: one
// spread counter (SPH_SPREAD_WORDS words, cleared by the caller) += floored particles.
template <typename RHO>
hipError_t nh_cool_update(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const RHO *rho,
                          const double *u, const double *log_nh, const int32_t *typ, const NHTable &t,
                          const NHCoolArgs &a, double *u_out, unsigned long long *floors);

}  // namespace sph
