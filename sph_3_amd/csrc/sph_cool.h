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

// Grid of the temperature table: log10 T on axes {log10 n_H, log10 u [erg/g],
// log10 J21}, n[a] points lo[a] + i*d[a], built for one hydrogen fraction XH.
struct CoolTAxes {
    double lo[3], d[3];
    int n[3];
};

// Trilinear lookup of log10 T at (log n_H, log u, log J21), inputs clamped to the grid;
// layout (nH*nu + u)*nJ + J.
__device__ __forceinline__ float cool_ttable_interp(const CoolTAxes &ax, const float *__restrict__ tabT,
                                                    const float v[3])
{
    int i[3];
    float f[3];
#pragma unroll
    for (int a = 0; a < 3; a++) cool_table_pos(v[a], (float)ax.lo[a], (float)ax.d[a], ax.n[a], i[a], f[a]);
    const int64_t s1 = ax.n[2], s0 = (int64_t)ax.n[1] * s1;
    const int64_t base = i[0] * s0 + i[1] * s1 + i[2];
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int b0 = (c >> 2) & 1, b1 = (c >> 1) & 1, b2 = c & 1;
        const float w = (b0 ? f[0] : 1.0f - f[0]) * (b1 ? f[1] : 1.0f - f[1]) * (b2 ? f[2] : 1.0f - f[2]);
        acc += w * tabT[base + b0 * s0 + b1 * s1 + b2];
    }
    return acc;
}

// Explicit cooling update of one step (CR/v4t.h:1768-1884 doCloudyHCooling
// semantics with per-n_H^2 rates): rho_cgs = rho * rho_cgs, n_H = XH rho_cgs / m_H,
// u_cgs = u * u_cgs; (n_H, T, Z, J21) are clamped into the cooling table's box
// [lo, hi] (log10) before the rates are evaluated; the net rate per volume is
// (Gamma - Lambda) n_H^2 with the unclamped n_H; u_cgs' = max(u_cgs + rate / rho_cgs
// * dt_s, u_floor) and u' = u_cgs' / u_cgs.
struct CoolUpdate {
    double rho_cgs, u_cgs, dt_s, XH, u_floor;
    double lo[4], hi[4];     // log10 box: n_H, T, Z, J21
    double lin_lo[4], lin_hi[4];  // the same box in linear units (10^lo, 10^hi)
};

enum { SPH_COOLK_EXACT = 0, SPH_COOLK_TABLE = 1, SPH_COOLK_PREP = 2 };

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
// log10 T(u) (the exact u -> T iteration) on the temperature grid.
hipError_t cool_ttable_build(hipStream_t st, const CoolTAxes &tax, double XH, const double *rf1, float *logT);
// T at (u [erg/g], n_H, J21) by the temperature table (J21 null: the constant J21c).
hipError_t cool_ttable_lookup(hipStream_t st, int64_t n, const CoolTAxes &tax, const float *logT, const double *u,
                              const double *nH, const double *J21, double J21c, double *T);
// Explicit cooling update (CoolUpdate) of n particles, u_out may alias u; RHO = float or
// double.  kind SPH_COOLK_EXACT: exact FP64 T(u) and rates; SPH_COOLK_TABLE: T from the
// temperature table, log Lambda and log Gamma from the 4D table; SPH_COOLK_PREP: writes
// the clamped log10 inputs as four FP32 columns cols[k*n + i] and the table's log Gamma
// (logG) for cool_apply.  nfloor (spread counter 0) += particles set to the floor.
template <typename RHO>
hipError_t cool_update(hipStream_t st, int64_t n, int kind, const RHO *rho, const double *u, const double *Z,
                       const double *J21, const CoolUpdate &cu, const CoolAxes &ax, const float *tabL,
                       const float *tabG, const CoolTAxes &tax, const float *tabT, const double *rf1, double *u_out,
                       float *cols, float *logG, unsigned long long *nfloor);
// The update of cool_update with log10 Lambda and log10 Gamma given per particle.
template <typename RHO>
hipError_t cool_apply(hipStream_t st, int64_t n, const RHO *rho, const double *u, const float *logL,
                      const float *logG, const CoolUpdate &cu, double *u_out, unsigned long long *nfloor);

}  // namespace sph
