// sph_tree.h -- host entry points of the tree kernels (sph_tree.hip).
#pragma once
#include "sph_common.h"
#include "sph_order.h"

namespace sph {

// Topology from sorted keys: nodes[].first/count/left/right, right[N-1],
// parent[2N-1] (root parent = -1), ncount[2N-1] (= nodes[].count, packed for scans).
hipError_t tree_topology(hipStream_t st, int64_t n, const uint64_t *skeys, WNode *nodes, int32_t *right,
                         int32_t *parent, int32_t *ncount);
// nodes[].escape (pre-order successor) of every node, for the stackless walks.
hipError_t tree_escape(hipStream_t st, int64_t n, WNode *nodes);
// FP64 moments (box, mass, centre of mass, size, hmax, epsmax) bottom-up, for the buckets
// (nodes of at most `bucket` particles whose parent holds more, and the root) and every
// node above them; nodes below the buckets keep stale records, so walks must stop at
// buckets of at least this size.  h / eps may be null (then 0).  Scratch: flags N uint32,
// frontier moments_scratch_entries(N) int32.
hipError_t tree_moments(hipStream_t st, int64_t n, int bucket, const double *x, const double *y, const double *z,
                        const double *m, const double *h, const double *eps, WNode *nodes,
                        const int32_t *right, const int32_t *parent, const int32_t *ncount, uint32_t *flags,
                        int32_t *frontier);
int64_t moments_scratch_entries(int64_t n);

}  // namespace sph
