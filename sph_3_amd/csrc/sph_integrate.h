// sph_integrate.h -- host entry points of sph_integrate.hip.
#pragma once
#include "sph_common.h"

namespace sph {

// v += a*dtk; if x != null: x += v*dtd (with the kicked v).
hipError_t kick_drift(hipStream_t st, int64_t n, double *vx, double *vy, double *vz, const double *ax,
                      const double *ay, const double *az, double dtk, double dtd, double *x, double *y, double *z,
                      double L = 0.0);
// first: u += du*dt; else u += dt/2 (du + du_prev).  Then du_prev = du.
template <typename R>
hipError_t u_update(hipStream_t st, int64_t n, double *u, const R *du, double *du_prev, double dt, int first);
// prho2 = P/rho^2; if bal != null: bal = divV / (divV + curlV + 1e-4 c/h).
template <typename R>
hipError_t prep(hipStream_t st, int64_t n, const R *rho, const R *P, const R *c, const R *divV, const R *curlV,
                const double *h, R *prho2, R *bal);
// out10 = [KE, U, PE, px, py, pz, Lx, Ly, Lz, M] in FP64 with KE = sum m v^2/2, U = sum m u,
// PE = -1/2 sum m phi, p = sum m v, L = sum m x cross v; null arrays count as zeros;
// part: 10*1024 doubles.
hipError_t totals(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const double *vx,
                  const double *vy, const double *vz, const double *m, const double *u, const double *phi,
                  double *part, double *out10);
// out5 = [KE, px, py, pz, M]; part: 5*1024 doubles.  Velocity component k of particle i
// is vk[i * vstride] (1: three arrays; 3: vx = v, vy = v + 1, vz = v + 2 of one n x 3 array).
hipError_t kinetics(hipStream_t st, int64_t n, const double *vx, const double *vy, const double *vz,
                    const double *m, double *part, double *out5, int64_t vstride = 1);

}  // namespace sph
