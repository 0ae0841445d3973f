// sph_order.h -- host entry points of the ordering kernels (sph_order.hip).
#pragma once
#include "sph_common.h"

namespace sph {

// box[0..5] = lo xyz, hi xyz; box[6] = 65536 / max extent.  part: 6*1024 doubles.
hipError_t bbox(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                double *part, double *box);
// 48-bit Morton keys (16 bits per axis) and identity values.
hipError_t morton(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                  const double *box, uint64_t *keys, int32_t *vals);
int64_t radix_hist_entries(int64_t n);
// Stable LSD sort of (k0, v0) on the low nbits (multiple of 8) bits.  Result
// lands in (k0, v0) when nbits/8 is even, else in (k1, v1).
hipError_t radix_sort(hipStream_t st, int64_t n, int nbits, uint64_t *k0, int32_t *v0,
                      uint64_t *k1, int32_t *v1, uint32_t *hist);
int64_t scan_tmp_entries(int64_t n);
// off[0..n] = exclusive prefix of cnt[0..n-1] (off[n] = total).
hipError_t scan_counts(hipStream_t st, int64_t n, const int32_t *cnt, int64_t *off, int64_t *tmp);
// dst[i] = src[perm[i]*stride + comp]
template <typename T, typename U>
hipError_t gather(hipStream_t st, int64_t n, const int32_t *perm, const T *src, int stride, int comp, U *dst);
// dst[j][i] = src[j][perm[i]] for j < k <= SPH_GATHER_MAX FP64 arrays and, when idst is not
// null, idst[i] = isrc[perm[i]], in one pass.
#define SPH_GATHER_MAX 16
hipError_t gather_set(hipStream_t st, int64_t n, const int32_t *perm, int k, const double *const *src,
                      double *const *dst, const int32_t *isrc, int32_t *idst);
// dst[perm[i]*stride + comp] = src[i]
template <typename T>
hipError_t scatter(hipStream_t st, int64_t n, const int32_t *perm, const T *src, int stride, int comp, double *dst);
// out[i] = a[b[i]]
hipError_t compose(hipStream_t st, int64_t n, const int32_t *a, const int32_t *b, int32_t *out);
template <typename T>
hipError_t fill(hipStream_t st, int64_t n, T *a, T v);

}  // namespace sph
