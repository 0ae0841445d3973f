// sph_grav.h -- host entry points of the tree gravity walk (sph_grav.hip).
#pragma once
#include "sph_common.h"
#include "sph_order.h"

namespace sph {

// Target groups: the maximal tree nodes with at most 64 particles.  Writes
// their node ids to groups[] (capacity 2N-1) and the count to *ngroups
// (synchronises the stream).
hipError_t gravity_groups(hipStream_t st, int64_t n, const int32_t *ncount, const int32_t *parent, int32_t *groups,
                          unsigned int *ngroups_dev, int64_t *ngroups);

// a_i = G sum m_j f(r_ij) (r_j - r_i) over the group's interaction list (self
// excluded).  soft_kind 0: spline with eps_ij = (eps_i + eps_j)/2 (node:
// epsmax), eps may be null (= 0); 1: Plummer with eps_plummer.  theta = 0
// gives the exact direct sum.  counters: spread array (sph_common.h) of three
// counters, 0 += particle-particle, 1 += particle-node interactions, 2 += stack
// overflows
// (must stay 0).  m_max: an upper bound of the particle masses (INFINITY if unknown), which
// decides whether a Plummer softening > 0 keeps each particle's weight on itself finite.
// phi (may be null): also the potential phi_i = G sum_j m_j psi(r_ij) over the same list
// (psi = 1/sqrt(r^2 + eps^2) or the spline potential of getPE; coincident pairs excluded).
template <typename R>
hipError_t gravity(hipStream_t st, int64_t ngroups, const int32_t *groups, const double *x, const double *y,
                   const double *z, const double *m, const double *eps, const WNode *nodes, int bucket, double theta,
                   double G, int soft_kind, double eps_plummer, double m_max, double *ax, double *ay, double *az,
                   unsigned long long *counters, double *phi = nullptr);

}  // namespace sph
