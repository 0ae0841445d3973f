// sph_hydro.h -- host entry points of the CSR hydro passes (sph_hydro.hip).
// R = float (mixed form) or double (all-FP64 form).  All arrays are sorted
// order; rows holds the neighbour sets S_i = { j : r_ij^2 < (h_i+h_j)^2 }.
#pragma once
#include "sph_common.h"

// Neighbour rows of the sorted particles: CSR (stride == 0: row i is
// idx[off[i] .. off[i+1])) or padded (stride > 0: row i is
// idx[i*stride .. i*stride + cnt[i])), with the per-particle inputs the passes
// gather, packed as 32-byte records: p4[i] = (x, y, z, h) and v4[i] = (vx, vy, vz, m)
// (one gather per neighbour each instead of four; sph::pack4 builds them).
struct NgbRows {
    const int64_t *off;
    const int32_t *cnt;
    const int32_t *idx;
    int stride;
    double L;  // side of the periodic box the sets were built in (0: open boundaries)
    const double4 *p4;
    const double4 *v4;
};

__device__ __forceinline__ void row_range(const NgbRows &r, int64_t i, int64_t &b, int64_t &e)
{
    if (r.stride > 0) {
        b = i * (int64_t)r.stride;
        e = b + r.cnt[i];
    } else {
        b = r.off[i];
        e = r.off[i + 1];
    }
}

namespace sph {

// out[i] = (a[i], b[i], c[i], d[i]).
hipError_t pack4(hipStream_t st, int64_t n, const double *a, const double *b, const double *c, const double *d,
                 double4 *out);
// rho = sum m_j W(r_ij, hbar_ij); fused EOS: P = (gamma-1) rho u,
// c = sqrt(gamma (gamma-1) u), prho2 = P / rho^2 (skipped when u is null).
template <typename R>
hipError_t density(hipStream_t st, int64_t n, const double *u, double gamma, const NgbRows &rows, R *rho, R *P,
                   R *c, R *prho2);
// |div v|, |curl v| and (bal non-null) Balsara f = |div v| / (|div v| + |curl v| + 1e-4 c/h).
template <typename R>
hipError_t divcurl(hipStream_t st, int64_t n, const R *rho, const R *c, const NgbRows &rows, R *divV, R *curlV,
                   R *bal);
// density() and divcurl() in one pass (u required; bal may be null).
template <typename R>
hipError_t density_divcurl(hipStream_t st, int64_t n, const double *u, double gamma, const NgbRows &rows, R *rho,
                           R *P, R *c, R *prho2, R *divV, R *curlV, R *bal);
// a = a_grav (optional, may be null) + pressure/AV acceleration; bal null = no Balsara.
template <typename R>
hipError_t accsph(hipStream_t st, int64_t n, const R *rho, const R *prho2, const R *c, const R *bal, double alpha,
                  const NgbRows &rows, const double *gx, const double *gy, const double *gz, double *ax, double *ay,
                  double *az);
template <typename R>
hipError_t dudt(hipStream_t st, int64_t n, const R *rho, const R *prho2, const R *c, const R *bal, double alpha,
                const NgbRows &rows, R *out);

}  // namespace sph
