// sph_cool.h -- host entry points of the cooling kernels (sph_cool.hip).
#pragma once
#include "sph_common.h"

// Grid of the cooling table: axis a in {log10 n_H, log10 T, log10 Z, log10 J21}
// has n[a] points lo[a] + i*d[a].
struct CoolAxes {
    double lo[4], d[4];
    int n[4];
};

// Cell index and fraction of v on one table axis (clamped into the grid).
__device__ __forceinline__ void cool_table_pos(float v, float lo, float d, int n, int &i, float &f)
{
    float x = (v - lo) / d;
    x = fminf(fmaxf(x, 0.0f), (float)(n - 1) - 1e-4f);
    i = (int)x;
    if (i > n - 2) i = n - 2;
    f = x - (float)i;
}

// 4D multilinear interpolation of one or two tables (tabG may be null) at the
// log10 point v = (log n_H, log T, log Z, log J21); layout ((nH*nT+T)*nZ+Z)*nU+U.
__device__ __forceinline__ void cool_table_interp(const CoolAxes &ax, const float *__restrict__ tabL,
                                                  const float *__restrict__ tabG, const float v[4], float &outL,
                                                  float &outG)
{
    int i[4];
    float f[4];
#pragma unroll
    for (int a = 0; a < 4; a++) cool_table_pos(v[a], (float)ax.lo[a], (float)ax.d[a], ax.n[a], i[a], f[a]);
    const int64_t s3 = 1, s2 = ax.n[3], s1 = (int64_t)ax.n[2] * ax.n[3], s0 = (int64_t)ax.n[1] * s1;
    const int64_t base = i[0] * s0 + i[1] * s1 + i[2] * s2 + i[3] * s3;
    float accL = 0.0f, accG = 0.0f;
#pragma unroll
    for (int c = 0; c < 16; c++) {
        const int b0 = (c >> 3) & 1, b1 = (c >> 2) & 1, b2 = (c >> 1) & 1, b3 = c & 1;
        const float w = (b0 ? f[0] : 1.0f - f[0]) * (b1 ? f[1] : 1.0f - f[1]) * (b2 ? f[2] : 1.0f - f[2]) *
                        (b3 ? f[3] : 1.0f - f[3]);
        const int64_t o = base + b0 * s0 + b1 * s1 + b2 * s2 + b3 * s3;
        accL += w * tabL[o];
        if (tabG) accG += w * tabG[o];
    }
    outL = accL;
    outG = accG;
}

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
// 4D multilinear lookup of log10 Lambda / log10 Gamma at log10 inputs (logG, outG may be null).
hipError_t cool_table_lookup(hipStream_t st, int64_t n, const CoolAxes &ax, const float *logL, const float *logG,
                             const float *lnH, const float *lT, const float *lZ, const float *lU, float *outL,
                             float *outG);
// Exact log10 Lambda at n points given as FP32 log10 rows (log n_H, log T, log Z, log J21).
hipError_t cool_log_lambda_points(hipStream_t st, int64_t n, const float *x4, const double *rf1, float *logL);

}  // namespace sph
