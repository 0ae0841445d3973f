// sph_grav.h -- host entry point of the tree gravity walk (sph_grav.hip).
#pragma once
#include "sph_common.h"

namespace sph {

// a_i = G sum m_j f(r_ij) (r_j - r_i) over the walk's interaction set (self
// excluded).  soft_kind 0: spline with eps_ij = (eps_i + eps_j)/2 (node:
// epsmax), eps may be null (= 0); 1: Plummer with eps_plummer.  theta = 0
// gives the exact direct sum.  counters[0] += particle-particle,
// counters[1] += particle-node interactions, counters[2] += stack overflows
// (must stay 0).
template <typename R>
hipError_t gravity(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const double *m,
                   const double *eps, const WNode *nodes, int bucket, double theta, double G, int soft_kind,
                   double eps_plummer, double *ax, double *ay, double *az, unsigned long long *counters);

}  // namespace sph
