// sph_agn.h -- host entry points of the AGN source terms (sph_agn.hip).
#pragma once
#include "sph_common.h"

namespace sph {

// a_i += -2 sigma^2 x_i / |x_i|^2: the isothermal halo's pull G M(r) x / r^3 with
// M(r) = 2 sigma^2 r / G (G cancels); a particle at the origin gets nothing.
hipError_t halo_accel(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, double sigma,
                      double *ax, double *ay, double *az);

// The state arrays an injection appends to (sorted-order arrays of the resident state;
// zmet / j21 may be null).
struct InjectTarget {
    double *x, *y, *z, *vx, *vy, *vz, *m, *h, *u, *duprev, *ax, *ay, *az;
    int32_t *orig;
    double *zmet, *j21;
};
struct InjectParams {
    double r_inject, v_in, m_inject, u_inject, Z_inject, J21_inject;
    uint64_t seed, step;
};
// *hsrc (device) = h of the particle whose caller index is `caller` (one match among n).
hipError_t h_of_caller(hipStream_t st, int64_t n, const int32_t *orig, const double *h, int64_t caller,
                       double *hsrc);
// Appends `pairs` particle pairs at slots [n, n + 2 pairs) with caller indices n, n + 1, ...:
// pair j at +-r, |r| = r_inject U0, direction (theta = 2 pi U1, phi = acos(2 U2 - 1)), velocity
// +-v_in along r, mass m_inject, u = u_inject, h = *hsrc, zero acceleration and du/dt history.
// U_k = inject_uniform(seed, step, j, k).
hipError_t inject_pairs(hipStream_t st, int64_t n, int64_t pairs, const InjectParams &p, const double *hsrc,
                        const InjectTarget &t);

}  // namespace sph

// Counter-based uniform draw in [0, 1): SplitMix64 finalisers (Steele, Lea & Flood 2014) of
// (seed, step, pair, component); the top 53 bits scaled by 2^-53.
__host__ __device__ inline uint64_t sph_mix64(uint64_t v)
{
    v += 0x9E3779B97F4A7C15ull;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}
__host__ __device__ inline double inject_uniform(uint64_t seed, uint64_t step, uint64_t pair, int k)
{
    const uint64_t v = sph_mix64(seed ^ sph_mix64(step ^ sph_mix64(pair * 4 + (uint64_t)k)));
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}
