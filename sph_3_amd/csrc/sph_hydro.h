// sph_hydro.h -- host entry points of the CSR hydro passes (sph_hydro.hip).
// R = float (mixed form) or double (all-FP64 form).  All arrays are sorted
// order; off/idx is the CSR of S_i = { j : r_ij^2 < (h_i+h_j)^2 }.
#pragma once
#include "sph_common.h"

namespace sph {

// rho = sum m_j W(r_ij, hbar_ij); fused EOS: P = (gamma-1) rho u,
// c = sqrt(gamma (gamma-1) u), prho2 = P / rho^2 (skipped when u is null).
template <typename R>
hipError_t density(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const double *m,
                   const double *h, const double *u, double gamma, const int64_t *off, const int32_t *idx, R *rho, R *P,
                   R *c, R *prho2);
// |div v|, |curl v| and (bal non-null) Balsara f = |div v| / (|div v| + |curl v| + 1e-4 c/h).
template <typename R>
hipError_t divcurl(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const double *vx,
                   const double *vy, const double *vz, const double *m, const double *h, const R *rho, const R *c,
                   const int64_t *off, const int32_t *idx, R *divV, R *curlV, R *bal);
// a = a_grav (optional, may be null) + pressure/AV acceleration; bal null = no Balsara.
template <typename R>
hipError_t accsph(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const double *vx,
                  const double *vy, const double *vz, const double *m, const double *h, const R *rho,
                  const R *prho2, const R *c, const R *bal, double alpha, const int64_t *off, const int32_t *idx,
                  const double *gx, const double *gy, const double *gz, double *ax, double *ay, double *az);
template <typename R>
hipError_t dudt(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const double *vx,
                const double *vy, const double *vz, const double *m, const double *h, const R *rho, const R *prho2,
                const R *c, const R *bal, double alpha, const int64_t *off, const int32_t *idx, R *out);

}  // namespace sph
