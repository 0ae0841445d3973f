// sph_ngb.h -- host entry points of the neighbour kernels (sph_ngb.hip).
#pragma once
#include "sph_common.h"
#include "sph_order.h"

namespace sph {

struct HSolveParams {
    int64_t m;                   // targets
    const double *tx, *ty, *tz;  // target positions (device)
    const double *x, *y, *z;     // sorted sources (device)
    const WNode *nodes;
    int bucket;
    int mode;                    // 0 Newton, 1 reference count stepping, 2 k-th distance
    double n_ngb, tol;
    int n_low, n_up, max_steps;
    int kth;
    int cap;                     // candidate slots per target
    double r_fac;                // initial radius = 2 * h_in * r_fac
    const double *h_in;          // start values (device)
    double *h_out;               // result (device)
    double *cand, *R;            // scratch: cap*m, m
    int32_t *steps, *state;      // scratch: m, m
    unsigned long long *counters;  // scratch: 2
    int64_t *unconverged_total;  // host accumulator (may be null)
};

// Runs rounds until every target is decided; *n_unconverged = targets still
// pending after max_rounds (0 on success).  Synchronises the stream.
hipError_t hsolve(hipStream_t st, const HSolveParams &p, int max_rounds, int64_t *n_unconverged);

// CSR of S_i = { j : r_ij^2 < (h_i+h_j)^2 } over sorted particles.  Counts,
// scans (off[0..n]), reads back *total; fills idx if idx_cap >= *total.
hipError_t neighbours(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                      const double *h, const WNode *nodes, int bucket, int32_t *cnt, int64_t *off,
                      int64_t *scan_tmp, int32_t *idx, int64_t idx_cap, int64_t *total);
hipError_t neighbours_fill(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                           const double *h, const WNode *nodes, int bucket, const int64_t *off, int32_t *idx);

}  // namespace sph
