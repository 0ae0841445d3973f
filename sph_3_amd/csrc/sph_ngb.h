// sph_ngb.h -- host entry points of the neighbour kernels (sph_ngb.hip).
#pragma once
#include "sph_common.h"
#include "sph_order.h"

#include <algorithm>

// Largest candidate capacity per target of the h solves (register slots of the solve).
#define SPH_HCAP_SLOTS 256

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
    int cap;                     // candidate slots per target (<= SPH_HCAP_SLOTS)
    double r_fac;                // initial radius = 2 * h_in * r_fac
    const double *h_in;          // start values (device)
    double *h_out;               // result (device)
    double *R;                   // scratch: m
    int32_t *cidx = nullptr;     // scratch of modes 0 and 2: cap*m candidate rows (source indices)
    double *cand = nullptr;      // optional: cap*m squared distances beside them; null: the solve
                                 //   recomputes them from the positions (the same bits, less memory)
    int32_t *ccnt = nullptr;     // m candidate counts (required by modes 0 and 2)
    int32_t *steps, *state;      // scratch: m, m
    unsigned long long *counters;  // scratch: spread array of 5 counters (5 * SPH_SPREAD_WORDS words)
    unsigned long long *totals;    // scratch: 5 words
    int64_t *work;                 // host (may be null): += distance tests of the walks, kernel terms of Newton
    int64_t *unconverged_total;  // host accumulator (may be null)
    const int32_t *groups;       // tree-node groups when the targets are the sorted particles, else null
    int64_t ngroups;
    int64_t *rounds;             // host: rounds used (may be null)
    // hsolve_begin / hsolve_end: pinned host words (5) that receive the first round's totals
    // and an event recorded after their copy
    unsigned long long *host_totals = nullptr;
    hipEvent_t done = nullptr;
    // mode 0 (may be null): m int32 and one device word for the targets whose solution needs
    // more candidates than a row holds; they are solved by a candidate-free walk solver after
    // the rounds and their candidate count is set to cap + 1 (incomplete row)
    int32_t *big_list = nullptr;
    unsigned long long *big_count = nullptr;
    int64_t *n_big = nullptr;  // host (may be null): targets given to that solver
    double per = 0.0;          // side of a periodic box: minimum-image distances (modes 0 and 2; 0: open)
    // modes 0 and 2 (may be null): hsolve_plist_ints(m, ngroups) int32 and 2 device words for the
    // lists of pending groups and targets that the rounds after the first launch
    int32_t *plist = nullptr;
    unsigned long long *plist_counts = nullptr;
};
inline int64_t hsolve_plist_ints(int64_t m, int64_t ngroups) { return m + std::max<int64_t>(ngroups, (m + 63) / 64); }

// Runs rounds until every target is decided; *n_unconverged = targets still
// pending after max_rounds (0 on success).  Synchronises the stream.  Returns
// hipErrorIllegalState if a walk stack overflowed.
hipError_t hsolve(hipStream_t st, const HSolveParams &p, int max_rounds, int64_t *n_unconverged);
// The same in two halves (Newton, count-free modes 0 and 2): hsolve_begin enqueues the first
// round and the asynchronous copy of its totals to p.host_totals (pinned, 6 words) followed by
// p.done, without waiting; the caller may enqueue independent work; hsolve_end waits for
// p.done (not the stream) and runs further rounds synchronously if targets are pending
// (*more_rounds = true: the results of anything enqueued from the first round's output
// must be rebuilt).
hipError_t hsolve_begin(hipStream_t st, const HSolveParams &p);
hipError_t hsolve_end(hipStream_t st, const HSolveParams &p, int max_rounds, int64_t *n_unconverged,
                      bool *more_rounds);

// CSR of S_i = { j : r_ij^2 < (h_i+h_j)^2 } over sorted particles, one
// wavefront per tree-node group (per > 0: minimum-image distances in a periodic box).  hscale (device, may be null): the node h_max
// were built from an earlier h and are scaled by *hscale in the prune.  Counts, scans (off[0..n]), reads back
// *total; fills idx if idx_cap >= *total.  overflow: 1 scratch counter.
hipError_t neighbours(hipStream_t st, int64_t n, int64_t ngroups, const int32_t *groups, const double *x,
                      const double *y, const double *z, const double *h, const WNode *nodes, int bucket, int32_t *cnt,
                      int64_t *off, int64_t *scan_tmp, unsigned long long *overflow, int32_t *idx, int64_t idx_cap,
                      int64_t *total, const double *hscale = nullptr, double per = 0.0);
hipError_t neighbours_fill(hipStream_t st, int64_t ngroups, const int32_t *groups, const double *x, const double *y,
                           const double *z, const double *h, const WNode *nodes, int bucket, const int64_t *off,
                           int32_t *idx, unsigned long long *overflow, const double *hscale = nullptr,
                           double per = 0.0);
// The same sets in one walk as padded rows: idx[i*stride + c], c < cnt[i] (sorted
// particles).  *n_long = rows with more than stride entries (those rows are
// truncated: rebuild as CSR), *total = entries.  spread: 2 * SPH_SPREAD_WORDS
// scratch words, totals: 2 scratch words, overflow: 1 scratch counter.
// Synchronises the stream.
hipError_t neighbour_rows(hipStream_t st, int64_t ngroups, const int32_t *groups, const double *x, const double *y,
                          const double *z, const double *h, const WNode *nodes, int bucket, int stride, int32_t *cnt,
                          int32_t *idx, unsigned long long *overflow, unsigned long long *spread,
                          unsigned long long *totals, int64_t *n_long, int64_t *total,
                          const double *hscale = nullptr, double per = 0.0);
// Padded neighbour rows (as neighbour_rows) built from the h solve's final candidate rows
// (targets = the sorted particles, same positions and the solved h; per: the solve's box) instead of
// a tree walk; exact by the containment argument at k_cand_rows.  ext: ext_cap int2 scratch
// for the pairs found from the other side (SPH_SPREAD regions), next: SPH_SPREAD_WORDS
// scratch words (one counter per region), cnt0: n int32 scratch (row lengths before the
// appended pairs).  Targets the candidate-free h solver handled (incomplete candidate rows)
// get their rows from one walk each (big).  *n_long > 0: some row is too long or an extra
// was lost -- rebuild with a walk.  In two halves: _begin enqueues the kernels and the asynchronous copy of the
// totals and region counters into host_words (pinned, 2 + SPH_SPREAD words) followed by
// done; _end waits for done (not the stream) and evaluates them.
// big: the targets of the candidate-free h solver (their rows are built by a walk each).
// cand: the solve's squared-distance rows, or null when it kept indices only; rec: scratch of
// n double2 (h_j, R_j) with cand, else n CandRec (position, h and R per particle, packed so
// that a candidate's distance and both tests take one 48-byte gather).
struct __align__(16) CandRec {
    double x, y, z, h, R, pad;
};
struct BigRows {
    const int32_t *list = nullptr;
    int64_t count = 0;
    const double *x = nullptr, *y = nullptr, *z = nullptr;
    const WNode *nodes = nullptr;
    int bucket = 0;
    const double *hscale = nullptr;  // device: node h_max belong to an earlier h (may be null)
    double per = 0.0;                // side of a periodic box (0: open boundaries)
};
hipError_t neighbour_rows_from_candidates_begin(hipStream_t st, int64_t n, int cap, const double *x,
                                               const double *y, const double *z, double per, const double *cand,
                                               const int32_t *cidx, const int32_t *ccnt, const double *h,
                                               const double *R, int stride,
                                               int32_t *cnt, int32_t *cnt0, int32_t *idx, int2 *ext,
                                               int64_t ext_cap, unsigned long long *next,
                                               unsigned long long *spread, unsigned long long *totals,
                                               unsigned long long *host_words, hipEvent_t done,
                                               const BigRows &big, void *rec);
hipError_t neighbour_rows_from_candidates_end(hipEvent_t done, const unsigned long long *host_words, int64_t ext_cap,
                                             int64_t *n_long, int64_t *total, int64_t *n_extra);
// *out (device) = max_i h_new[i] / h_old[i], rounded up by 8 ulp: the factor that
// keeps node h_max of h_old a valid bound for the new h.
hipError_t h_ratio_max(hipStream_t st, int64_t n, const double *h_new, const double *h_old, double *out);

}  // namespace sph
