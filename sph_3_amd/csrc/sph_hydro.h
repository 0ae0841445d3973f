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
    // Compact records, read instead of p4/v4 by the mixed form's density, acceleration and
    // du/dt passes (null: not packed; sph::pack_compact): pf[i] = (x, y, z as 32-bit fixed
    // point in steps of grid[7], the bits of h as FP32) and vf[i] = (vx, vy, vz, m) as FP32.
    const uint4 *pf = nullptr;
    const float4 *vf = nullptr;
    const double *grid = nullptr;
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

// Four-wide record of the pass type: float4 (mixed form) or double4 (FP64 form).
template <typename R> struct R4T;
template <> struct R4T<float> { using type = float4; };
template <> struct R4T<double> { using type = double4; };
template <typename R> using R4 = typename R4T<R>::type;

namespace sph {

// out[i] = (a[i], b[i], c[i], d[i]).
hipError_t pack4(hipStream_t st, int64_t n, const double *a, const double *b, const double *c, const double *d,
                 double4 *out);
// Compact records of the mixed form (NgbRows::pf, vf): coordinates as 32-bit fixed point,
// rounded to the nearest step sc.  Open boundaries (per = 0): from the low corner of the box
// grid[0..5] (lo xyz, hi xyz; the Morton bounding box), sc = 2^(k-32) with 2^k the smallest
// power of two above the largest extent; periodic box of side per: from the origin,
// sc = per / 2^32, modulo 2^32 (a pair's 32-bit difference is then its minimum image).
// sc is written to grid[7].  vel_only: only vf is rewritten (positions, h and the grid as at
// the last full call).
hipError_t pack_compact(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                        const double *h, const double *vx, const double *vy, const double *vz, const double *m,
                        double per, double *grid, uint4 *pf, float4 *vf, bool vel_only);
// s4[i] = (rho, P/rho^2, c, Balsara f or 1 when bal is null): the packed record the
// acceleration and du/dt passes gather.
template <typename R>
hipError_t pack_s4(hipStream_t st, int64_t n, const R *rho, const R *prho2, const R *c, const R *bal, R4<R> *s4);
// rho = sum m_j W(r_ij, hbar_ij); fused EOS: P = (gamma-1) rho u,
// c = sqrt(gamma (gamma-1) u), prho2 = P / rho^2 (skipped when u is null); s4 (may be
// null, needs u): the packed record with f = 1.
template <typename R>
hipError_t density(hipStream_t st, int64_t n, const double *u, double gamma, const NgbRows &rows, R *rho, R *P,
                   R *c, R *prho2, R4<R> *s4 = nullptr);
// |div v|, |curl v| and (bal non-null) Balsara f = |div v| / (|div v| + |curl v| + 1e-4 c/h).
template <typename R>
hipError_t divcurl(hipStream_t st, int64_t n, const R *rho, const R *c, const NgbRows &rows, R *divV, R *curlV,
                   R *bal);
// density() and divcurl() in one pass (u required; bal may be null; s4 may be null).
template <typename R>
hipError_t density_divcurl(hipStream_t st, int64_t n, const double *u, double gamma, const NgbRows &rows, R *rho,
                           R *P, R *c, R *prho2, R *divV, R *curlV, R *bal, R4<R> *s4 = nullptr);
// a = a_grav (optional, may be null) + pressure/AV acceleration from the packed records
// s4 (pack_s4); balsara = false: no Balsara switch (f unused).
template <typename R>
hipError_t accsph(hipStream_t st, int64_t n, const R4<R> *s4, bool balsara, double alpha, const NgbRows &rows,
                  const double *gx, const double *gy, const double *gz, double *ax, double *ay, double *az);
template <typename R>
hipError_t dudt(hipStream_t st, int64_t n, const R4<R> *s4, bool balsara, double alpha, const NgbRows &rows, R *out);

// Column density from the density gradient (the reference's grad_rho_norm_ngb and the
// host loop after it): g_i = |sum_j m_j nW(r_ij, hbar_ij) (x_j - x_i)| over the rows
// (pair arithmetic in R, the kernel gradient's q + 1e-10 as in that kernel), then in FP64
// N_H = 0.5 (h_i / cbrt(n_ngb) + rho_i / g_i) unit_length_cm * rho_i unit_density_cgs XH / m_H,
// log_nh[i] = log10 N_H (cm^-2) and grad[i] = g_i (code units; may be null).
// Per-particle time-step criteria over the rows (dt_array_indiv_dt,
// Particles_in_BH_Tree_III/Evrard_Collapse_2024/hfvCLibs_v7_hfv_v8.1.h:1191-1270, with the signal
// velocity's maximum over S_i instead of over all particles):
//   vsig_i = max_j (c_i + c_j - 3 w_ij), w_ij = v_ij . r_ij / (r_ij + 1e-5)   (pairs in R)
//   dt_hyd = C h / vsig, dt_cour = C h / (h |div v| + max(c, |v|)(1 + 0.6 alpha)),
//   dt_G1 = 0.4 sqrt(eps / |a_tot|), dt_G2 = 0.4 eps / |v|, dt_f = sqrt(h / |a_g|),
//   dt_kin = sqrt(h / |a_tot|); dt_i = the smallest (FP64), dh/dt = h div v / 3.
// s4.z = c; divv in R; a_tot = (ax, ay, az); a_g = (gx, gy, gz) or null (then a_tot);
// eps per particle or, when null, eps_const.  dt (required), dhdt (may be null); *dt_min_bits
// (device, initialise to ~0) receives the minimum's bit pattern (positive doubles order as integers).
struct StepCriteria {
    double C_CFL, alpha, eps_const;
};
template <typename R>
hipError_t timestep(hipStream_t st, int64_t n, const R4<R> *s4, const R *divv, const NgbRows &rows,
                    const double *ax, const double *ay, const double *az, const double *gx, const double *gy,
                    const double *gz, const double *eps, const StepCriteria &sc, double *dt, double *dhdt,
                    unsigned long long *dt_min_bits);

struct ColumnParams {
    double unit_density_cgs, unit_length_cm, XH, n_ngb;
};
template <typename R>
hipError_t column_density(hipStream_t st, int64_t n, const R *rho, const NgbRows &rows, const ColumnParams &cp,
                          double *grad, double *log_nh);

}  // namespace sph
