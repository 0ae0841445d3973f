// sph_tree.h -- host entry points of the tree kernels (sph_tree.hip).
#pragma once
#include "sph_common.h"
#include "sph_order.h"

namespace sph {

// Topology from sorted keys: nodes[].first/count/left/escape, right[N-1],
// parent[2N-1] (root parent = -1).
hipError_t tree_topology(hipStream_t st, int64_t n, const uint64_t *skeys, WNode *nodes, int32_t *right,
                         int32_t *parent);
// FP64 moments (box, mass, centre of mass, size, hmax, epsmax) bottom-up.
// h / eps may be null (then 0).  Scratch: flags N uint32, frontier
// moments_scratch_entries(N) int32.
hipError_t tree_moments(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                        const double *m, const double *h, const double *eps, WNode *nodes,
                        const int32_t *right, const int32_t *parent, uint32_t *flags, int32_t *frontier);
int64_t moments_scratch_entries(int64_t n);

}  // namespace sph
