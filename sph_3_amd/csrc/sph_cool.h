// sph_cool.h -- host entry points of the cooling kernels (sph_cool.hip).
#pragma once
#include "sph_common.h"

// Grid of the cooling table: axis a in {log10 n_H, log10 T, log10 Z, log10 J21}
// has n[a] points lo[a] + i*d[a].
struct CoolAxes {
    double lo[4], d[4];
    int n[4];
};

namespace sph {

// Photo-ionisation (rf1[0..2]) and heating (rf1[3..5]) integrals for J21 = 1.
hipError_t cool_radfield(hipStream_t st, double *rf1);
// Exact FP64 heating / cooling per n_H^2 at (T, n_H, J21, Z); null arrays take the constants.
hipError_t cool_rates_exact(hipStream_t st, int64_t n, const double *T, const double *nH, const double *J21,
                            double J21c, const double *Z, double Zc, const double *rf1, double *heat, double *lam);
// what = 0: T(u); what = 1: Gamma - Lambda at T(u) clamped to [1e4, 1e8] K.
hipError_t cool_u_exact(hipStream_t st, int64_t n, int what, const double *u, const double *nH, double XH,
                        const double *J21, double J21c, const double *Z, double Zc, const double *rf1, double *out);
// Implicit u update over dt (cgs); -1 where the downward bracket reached u_floor.
hipError_t do_cooling_exact(hipStream_t st, int64_t n, const double *rho, const double *u_old, const double *dt,
                            double dtc, double XH, const double *J21, double J21c, const double *Z, double Zc,
                            double u_floor, const double *rf1, double *out);
// log10 Lambda, log10 Gamma on the grid (FP32, layout ((nH*nT + T)*nZ + Z)*nU + U).
hipError_t cool_table_build(hipStream_t st, const CoolAxes &ax, const double *rf1, float *logL, float *logG);
// 4D multilinear lookup of log10 Lambda / log10 Gamma at log10 inputs.
hipError_t cool_table_lookup(hipStream_t st, int64_t n, const CoolAxes &ax, const float *logL, const float *logG,
                             const float *lnH, const float *lT, const float *lZ, const float *lU, float *outL,
                             float *outG);

}  // namespace sph
