// sph_cool.hip -- per-particle radiative cooling/heating: the exact FP64
// function, a 4D table built from it on the device, and the table lookup.
//
// Exact function: restates the closed-form primordial H/He network of
// Particles_in_BH_Tree_III/hcool_with_unit_tests_Here/photolibs3.py (Cen 1992
// rates :140-165, ionisation equilibrium by damped fixed-point iteration
// :169-212, Verner 1996 photo-rates for a J_nu ~ nu^-1 field :215-329,
// cooling terms :332-376, u -> T :379-420, implicit u update :489-547) with
// the field amplitude J21 per particle, plus a metal-line term Z*Lambda_RB95(T)
// (Rosen & Bregman 1995 solar CIE fit) scaled by the metallicity Z [Z_sun].
//
// Table: log10 Lambda and log10 Gamma (per n_H^2) on a grid uniform in
// (log n_H, log T, log Z, log J21), FP32; lookup is 4D multilinear in log
// space with inputs clamped to the grid (as CR/v4t.h:1768-1884
// doCloudyHCooling clamps to its table bounds), O(1) index arithmetic.
#include "sph_cool.h"

// One rounding per operation in this file's FP64 physics and update (as the
// oracle is compiled), so an update computes the same bits in every kernel that
// performs it; the FP32 table interpolation (sph_cool.h, above this line) keeps
// the default contraction it has in every other file that uses it.
#pragma clang fp contract(off)

#define CL_C 29979245800.0
#define CL_HEV 4.135667696e-15
#define CL_HPL 6.626196e-27
#define CL_KB 1.3807e-16
#define CL_MH 1.6726e-24

struct CoolRates {
    double aHp, aHep, aHepp, ad, geH0, geHe0, geHep;
};

__device__ __forceinline__ CoolRates cool_rates(double T)
{
    CoolRates r;
    const double Tfact = 1.0 / (1.0 + sqrt(T / 1e5));
    r.aHp = 8.41e-11 * pow(T / 1000.0, -0.2) / (1.0 + pow(T / 1e6, 0.7)) / sqrt(T);
    r.aHep = 1.5e-10 * pow(T, -0.6353);
    r.aHepp = 4.0 * r.aHp;
    r.ad = 1.9e-3 * pow(T, -1.5) * exp(-470000.0 / T) * (1.0 + 0.3 * exp(-94000.0 / T));
    r.geH0 = 5.85e-11 * sqrt(T) * exp(-157809.1 / T) * Tfact;
    r.geHe0 = 2.38e-11 * sqrt(T) * exp(-285335.4 / T) * Tfact;
    r.geHep = 5.68e-12 * sqrt(T) * exp(-631515.0 / T) * Tfact;
    return r;
}

// n_X / n_H: nH0, nHe0, nHp, ne, nHep, nHepp (gJ: photo-ionisation rates).
__device__ __forceinline__ void cool_abund(const CoolRates &r, double nH, const double *gJ, double *ab)
{
    const double y = 0.24 / (4.0 - 4.0 * 0.24);
    double ne = 1.0, ne_old, dne = 1.0;
    double nH0 = 0, nHp = 0, nHep = 0, nHe0 = 0, nHepp = 0;
    int niter = 1;
    while (dne > 1e-4 && niter < 100) {
        ne_old = ne;
        const double a = gJ[0] / (ne * nH), b = gJ[1] / (ne * nH), c = gJ[2] / (ne * nH);
        nH0 = r.aHp / (r.aHp + r.geH0 + a);
        nHp = 1.0 - nH0;
        nHep = y / (1.0 + (r.aHep + r.ad) / (r.geHe0 + b) + (r.geHep + c) / r.aHepp);
        nHe0 = nHep * (r.aHep + r.ad) / (r.geHe0 + b);
        nHepp = nHep * (r.geHep + c) / r.aHepp;
        ne = nHp + nHep + 2.0 * nHepp;
        ne = 0.5 * (ne + ne_old);
        dne = fabs(ne - ne_old);
        niter++;
    }
    ab[0] = nH0; ab[1] = nHe0; ab[2] = nHp; ab[3] = ne; ab[4] = nHep; ab[5] = nHepp;
}

__device__ __forceinline__ double rb95(double T)
{
    if (T < 300.0) return 0.0;
    if (T < 2000.0) return 2.2380e-32 * T * T;
    if (T < 8000.0) return 1.0012e-30 * pow(T, 1.5);
    if (T < 1e5) return 4.6240e-36 * pow(T, 2.867);
    if (T < 4e7) return 1.7800e-18 * pow(T, -0.65);
    return 3.2217e-27 * sqrt(T);
}

// Heating and cooling per n_H^2 at (T, n_H); rf1 = field integrals for J21 = 1.
__device__ __forceinline__ void cool_heat_lambda(double T, double nH, double J21, double Z, const double *rf1,
                                                 double &heat, double &lam)
{
    const CoolRates r = cool_rates(T);
    double gJ[3] = {rf1[0] * J21, rf1[1] * J21, rf1[2] * J21};
    double ab[6];
    cool_abund(r, nH, gJ, ab);
    const double nH0 = ab[0], nHe0 = ab[1], nHp = ab[2], ne = ab[3], nHep = ab[4], nHepp = ab[5];
    heat = nH0 * (rf1[3] * J21) / nH + nHe0 * (rf1[4] * J21) / nH + nHep * (rf1[5] * J21) / nH;
    const double Tfact = 1.0 / (1.0 + sqrt(T / 1e5));
    const double BetaH0 = 7.5e-19 * exp(-118348.0 / T) * Tfact;
    const double BetaHep = 5.54e-17 * pow(T, -0.397) * exp(-473638.0 / T) * Tfact;
    const double lt = 5.5 - log10(T);
    const double Betaff = 1.43e-27 * sqrt(T) * (1.0 + 0.34 * exp(-lt * lt / 3.0));
    const double Lexc = BetaH0 * ne * nH0 + BetaHep * ne * nHep;
    const double Lion = 2.18e-11 * r.geH0 * ne * nH0 + 3.94e-11 * r.geHe0 * ne * nHe0 + 8.72e-11 * r.geHep * ne * nHep;
    const double Lrec = 1.036e-16 * T * ne * (r.aHp * nHp) + 1.036e-16 * T * ne * (r.aHep * nHep) +
                        1.036e-16 * T * ne * (r.aHepp * nHepp) + 6.526e-11 * r.ad * ne * nHep;
    const double Lff = Betaff * (nHp + nHep + 4.0 * nHepp) * ne;
    lam = Lexc + Lion + Lrec + Lff + Z * rb95(T);
}

__device__ __forceinline__ double u_to_T(double u, double nH, double XH, double J21, const double *rf1)
{
    const double gamma = 5.0 / 3.0;
    const double yHe = (1.0 - XH) / (4.0 * XH);
    double mu = (1.0 + 4.0 * yHe) / (1.0 + yHe + 1.0);
    double temp = (gamma - 1.0) * CL_MH / CL_KB * mu * u;
    const double gJ[3] = {rf1[0] * J21, rf1[1] * J21, rf1[2] * J21};
    double dT = temp, ab[6];
    int niter = 1;
    while (dT / temp > 1e-4 && niter < 100) {
        const double temp_old = temp;
        cool_abund(cool_rates(temp_old), nH, gJ, ab);
        mu = (1.0 + 4.0 * yHe) / (1.0 + yHe + ab[3]);
        temp = (gamma - 1.0) * CL_MH / CL_KB * mu * u;
        temp = 0.5 * (temp + temp_old);
        dT = fabs(temp - temp_old);
        niter++;
    }
    return temp;
}

__device__ __forceinline__ double rate_from_u(double u, double nH, double XH, double J21, double Z, const double *rf1)
{
    double T = u_to_T(u, nH, XH, J21, rf1);
    T = T < 1e4 ? 1e4 : (T > 1e8 ? 1e8 : T);
    double h, l;
    cool_heat_lambda(T, nH, J21, Z, rf1, h, l);
    return h - l;
}

// ---------------------------------------------------------------- field integrals
// One thread per species: trapezoid integrals over numpy.arange(nu_edge/1e16,
// 6.0, 0.02) * 1e16 (length ceil((stop-start)/step), values start + i*delta,
// delta = (start + step) - start), amplitude J21 = 1.
__global__ void k_radfield(double *__restrict__ rf1)
{
    const int s = threadIdx.x;
    if (s >= 3) return;
    const double E[3] = {13.60, 24.59, 54.42};
    const double P[3][7] = {{4.298E-1, 5.475E4, 3.288E1, 2.963, 0.0, 0.0, 0.0},
                            {1.361E+1, 9.492E2, 1.469E0, 3.188, 2.039, 4.434E-1, 2.136},
                            {1.720E+0, 1.369E4, 3.288E1, 2.963, 0.0, 0.0, 0.0}};
    const double start = (E[s] / CL_HEV) / 1e16;
    const int n = (int)ceil((6.0 - start) / 0.02);
    const double delta = (start + 0.02) - start;
    const double nu_e = E[s] / CL_HEV;
    const double nu_e2 = CL_C / ((CL_C / nu_e * 1e8) * 1e-8);
    const double dnu = (start + delta) * 1e16 - start * 1e16;
    double g = 0.0, hr = 0.0, pg = 0.0, ph = 0.0;
    for (int i = 0; i < n; i++) {
        const double nu = (start + i * delta) * 1e16;
        const double Ee = CL_HEV * nu;
        const double x = Ee / P[s][0] - P[s][5];
        const double y = sqrt(x * x + P[s][6] * P[s][6]);
        const double Fy = ((x - 1.0) * (x - 1.0) + P[s][4] * P[s][4]) * pow(y, 0.5 * P[s][3] - 5.5) *
                          pow(1.0 + sqrt(y / P[s][2]), -P[s][3]);
        const double sig = P[s][1] * Fy * 1e-18;
        const double J = 1.0 * 1e-21 / (nu / nu_e2);
        const double fg = 4.0 * SPH_PI * J * sig / CL_HPL / nu;
        const double fh = 4.0 * SPH_PI * J * sig * (nu - E[s] / CL_HEV) / nu;
        if (i > 0) {
            g += dnu * (pg + fg) / 2.0;
            hr += dnu * (ph + fh) / 2.0;
        }
        pg = fg;
        ph = fh;
    }
    rf1[s] = g;
    rf1[3 + s] = hr;
}

// ---------------------------------------------------------------- exact kernels
__global__ void __launch_bounds__(256) k_cool_rates(int64_t n, const double *__restrict__ T,
                                                    const double *__restrict__ nH, const double *__restrict__ J21,
                                                    double J21c, const double *__restrict__ Z, double Zc,
                                                    const double *__restrict__ rf1, double *__restrict__ heat,
                                                    double *__restrict__ lam)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double h, l;
    cool_heat_lambda(T[i], nH[i], J21 ? J21[i] : J21c, Z ? Z[i] : Zc, rf1, h, l);
    heat[i] = h;
    lam[i] = l;
}

__global__ void __launch_bounds__(256) k_cool_u(int64_t n, int what, const double *__restrict__ u,
                                                const double *__restrict__ nH, double XH,
                                                const double *__restrict__ J21, double J21c,
                                                const double *__restrict__ Z, double Zc,
                                                const double *__restrict__ rf1, double *__restrict__ out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double j = J21 ? J21[i] : J21c, z = Z ? Z[i] : Zc;
    out[i] = what == 0 ? u_to_T(u[i], nH[i], XH, j, rf1) : rate_from_u(u[i], nH[i], XH, j, z, rf1);
}

// Implicit update of HC/photolibs3.py:489-547 (bracket by factors of 1.1,
// bisection to |du/u| <= 1e-4); the downward bracket stops at u_floor (the
// reference has no floor) and then returns -1.
__global__ void __launch_bounds__(256) k_do_cooling(int64_t n, const double *__restrict__ rho,
                                                    const double *__restrict__ u_old, const double *__restrict__ dt,
                                                    double dtc, double XH, const double *__restrict__ J21,
                                                    double J21c, const double *__restrict__ Z, double Zc,
                                                    double u_floor, const double *__restrict__ rf1,
                                                    double *__restrict__ out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = rho[i], uo = u_old[i], d = dt ? dt[i] : dtc, j = J21 ? J21[i] : J21c, z = Z ? Z[i] : Zc;
    const double nH = XH * r / CL_MH;
    const double ratefact = nH * nH / r;
    double u = uo, uu = u, ul = u;
    double net = rate_from_u(u, nH, XH, j, z, rf1);
    if (u - uo - ratefact * net * d > 0.0) {
        while (uu - uo - ratefact * rate_from_u(uu, nH, XH, j, z, rf1) * d > 0.0) {
            uu /= 1.1;
            if (uu < u_floor) { out[i] = -1.0; return; }
        }
        ul = uu;
        uu = ul * 1.1;
    }
    if (u - uo - ratefact * net * d < 0.0) {
        int guard = 0;
        while (uu - uo - ratefact * rate_from_u(uu, nH, XH, j, z, rf1) * d < 0.0 && guard < 2000) {
            uu *= 1.1;
            guard++;
        }
        ul = uu / 1.1;
    }
    int niter = 1;
    double du = u;
    while (fabs(du / u) > 1e-4 && niter < 100) {
        u = 0.5 * (ul + uu);
        net = rate_from_u(u, nH, XH, j, z, rf1);
        if (u - uo - ratefact * net * d > 0.0) uu = u;
        else ul = u;
        du = fabs(uu - ul);
        niter++;
    }
    out[i] = u;
}

// ---------------------------------------------------------------- table
__global__ void __launch_bounds__(256) k_cool_table(CoolAxes ax, const double *__restrict__ rf1,
                                                    float *__restrict__ logL, float *__restrict__ logG)
{
    const int64_t total = (int64_t)ax.n[0] * ax.n[1] * ax.n[2] * ax.n[3];
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= total) return;
    // layout: ((i_nH * nT + i_T) * nZ + i_Z) * nU + i_U
    int64_t r = q;
    const int iU = (int)(r % ax.n[3]); r /= ax.n[3];
    const int iZ = (int)(r % ax.n[2]); r /= ax.n[2];
    const int iT = (int)(r % ax.n[1]); r /= ax.n[1];
    const int iN = (int)r;
    const double lnH = ax.lo[0] + iN * ax.d[0], lT = ax.lo[1] + iT * ax.d[1];
    const double lZ = ax.lo[2] + iZ * ax.d[2], lU = ax.lo[3] + iU * ax.d[3];
    double h, l;
    cool_heat_lambda(pow(10.0, lT), pow(10.0, lnH), pow(10.0, lU), pow(10.0, lZ), rf1, h, l);
    logL[q] = (float)log10(fmax(l, 1e-40));
    logG[q] = (float)log10(fmax(h, 1e-40));
}

// 4D multilinear lookup in log space; inputs are log10 values (FP32).
__global__ void __launch_bounds__(256) k_cool_lookup(int64_t n, CoolAxes ax, const float *__restrict__ logL,
                                                     const float *__restrict__ logG, const float *__restrict__ lnH,
                                                     const float *__restrict__ lT, const float *__restrict__ lZ,
                                                     const float *__restrict__ lU, float *__restrict__ outL,
                                                     float *__restrict__ outG)
{
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float v[4] = {lnH[p], lT[p], lZ[p], lU[p]};
    float L, G;
    cool_table_interp(ax, logL, logG, v, L, G);
    outL[p] = L;
    if (outG) outG[p] = G;
}

// Exact log10 Lambda at FP32 log10 points (training and validation samples).
__global__ void __launch_bounds__(256) k_cool_points(int64_t n, const float *__restrict__ x4,
                                                     const double *__restrict__ rf1, float *__restrict__ logL)
{
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double lnH = x4[4 * p], lT = x4[4 * p + 1], lZ = x4[4 * p + 2], lU = x4[4 * p + 3];
    double h, l;
    cool_heat_lambda(pow(10.0, lT), pow(10.0, lnH), pow(10.0, lU), pow(10.0, lZ), rf1, h, l);
    logL[p] = (float)log10(fmax(l, 1e-40));
}

// ---------------------------------------------------------------- temperature table
__global__ void __launch_bounds__(256) k_cool_ttable(CoolTAxes ax, double XH, const double *__restrict__ rf1,
                                                     float *__restrict__ logT)
{
    const int64_t total = (int64_t)ax.n[0] * ax.n[1] * ax.n[2];
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= total) return;
    int64_t r = q;
    const int iJ = (int)(r % ax.n[2]); r /= ax.n[2];
    const int iu = (int)(r % ax.n[1]); r /= ax.n[1];
    const int iN = (int)r;
    const double lnH = ax.lo[0] + iN * ax.d[0], lu = ax.lo[1] + iu * ax.d[1], lJ = ax.lo[2] + iJ * ax.d[2];
    logT[q] = (float)log10(u_to_T(pow(10.0, lu), pow(10.0, lnH), XH, pow(10.0, lJ), rf1));
}

__global__ void __launch_bounds__(256) k_cool_tlookup(int64_t n, CoolTAxes ax, const float *__restrict__ logT,
                                                      const double *__restrict__ u, const double *__restrict__ nH,
                                                      const double *__restrict__ J21, double J21c,
                                                      double *__restrict__ T)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v[3] = {(float)log10(nH[i]), (float)log10(u[i]), (float)log10(J21 ? J21[i] : J21c)};
    T[i] = pow(10.0, (double)cool_ttable_interp(ax, logT, v));
}

// ---------------------------------------------------------------- explicit update
__device__ __forceinline__ double clampd(double v, double lo, double hi) { return fmin(fmax(v, lo), hi); }

// u_cgs + rate n_H^2 / rho_cgs dt_s, floored; returns the code-unit value and
// whether the floor was applied.
__device__ __forceinline__ double cool_finish(const CoolUpdate &cu, double rho_cgs, double nH, double u_cgs,
                                              double net, bool &floored)
{
    double un = u_cgs + net * (nH * nH / rho_cgs) * cu.dt_s;
    floored = !(un >= cu.u_floor);
    if (floored) un = cu.u_floor;
    return un / cu.u_cgs;
}

__device__ __forceinline__ void count_floor(unsigned long long *nfloor, bool floored)
{
    const uint64_t b = __ballot(floored);
    if ((threadIdx.x & 63) == 0 && b) spread_add(nfloor, 0, (unsigned long long)__popcll(b));
}

template <int KIND, typename RHO>
__global__ void __launch_bounds__(256) k_cool_update(int64_t n, const RHO *__restrict__ rho, const double *u,
                                                     const double *__restrict__ Z, const double *__restrict__ J21,
                                                     CoolUpdate cu, CoolAxes ax, const float *__restrict__ tabL,
                                                     const float *__restrict__ tabG, CoolTAxes tax,
                                                     const float *__restrict__ tabT, const double *__restrict__ rf1,
                                                     double *u_out, float *__restrict__ cols,
                                                     float *__restrict__ logG, unsigned long long *nfloor)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const int64_t k = valid ? i : 0;
    const double rho_cgs = (double)rho[k] * cu.rho_cgs;
    const double nH = cu.XH * rho_cgs / CL_MH;
    const double u_cgs = u[k] * cu.u_cgs;
    bool floored = false;
    if (KIND == SPH_COOLK_EXACT) {
        const double nHc = clampd(nH, cu.lin_lo[0], cu.lin_hi[0]);
        const double zc = clampd(Z[k], cu.lin_lo[2], cu.lin_hi[2]);
        const double jc = clampd(J21[k], cu.lin_lo[3], cu.lin_hi[3]);
        const double Tc = clampd(u_to_T(u_cgs, nHc, cu.XH, jc, rf1), cu.lin_lo[1], cu.lin_hi[1]);
        double heat, lam;
        cool_heat_lambda(Tc, nHc, jc, zc, rf1, heat, lam);
        const double un = cool_finish(cu, rho_cgs, nH, u_cgs, heat - lam, floored);
        if (valid) u_out[i] = un;
    } else {
        const float lnH = (float)clampd(log10(nH), cu.lo[0], cu.hi[0]);
        const float lZ = (float)clampd(log10(Z[k]), cu.lo[2], cu.hi[2]);
        const float lJ = (float)clampd(log10(J21[k]), cu.lo[3], cu.hi[3]);
        const float tv[3] = {lnH, (float)log10(u_cgs), lJ};
        const float lT = fminf(fmaxf(cool_ttable_interp(tax, tabT, tv), (float)cu.lo[1]), (float)cu.hi[1]);
        const float v[4] = {lnH, lT, lZ, lJ};
        float L, G;
        if (KIND == SPH_COOLK_TABLE) {
            cool_table_interp(ax, tabL, tabG, v, L, G);
            const double un = cool_finish(cu, rho_cgs, nH, u_cgs, exp10((double)G) - exp10((double)L), floored);
            if (valid) u_out[i] = un;
        } else {
            cool_table_interp(ax, tabG, nullptr, v, G, L);  // Gamma only
            if (valid) {
#pragma unroll
                for (int a = 0; a < 4; a++) cols[a * n + i] = v[a];
                logG[i] = G;
            }
        }
    }
    if (KIND != SPH_COOLK_PREP) count_floor(nfloor, valid && floored);
}

template <typename RHO>
__global__ void __launch_bounds__(256) k_cool_apply(int64_t n, const RHO *__restrict__ rho, const double *u,
                                                    const float *__restrict__ logL, const float *__restrict__ logG,
                                                    CoolUpdate cu, double *u_out, unsigned long long *nfloor)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const int64_t k = valid ? i : 0;
    const double rho_cgs = (double)rho[k] * cu.rho_cgs;
    const double nH = cu.XH * rho_cgs / CL_MH;
    bool floored = false;
    const double un =
        cool_finish(cu, rho_cgs, nH, u[k] * cu.u_cgs, exp10((double)logG[k]) - exp10((double)logL[k]), floored);
    if (valid) u_out[i] = un;
    count_floor(nfloor, valid && floored);
}

namespace sph {

hipError_t cool_radfield(hipStream_t st, double *rf1)
{
    k_radfield<<<1, 64, 0, st>>>(rf1);
    return hipGetLastError();
}

hipError_t cool_rates_exact(hipStream_t st, int64_t n, const double *T, const double *nH, const double *J21,
                            double J21c, const double *Z, double Zc, const double *rf1, double *heat, double *lam)
{
    k_cool_rates<<<sph_blocks(n, 256), 256, 0, st>>>(n, T, nH, J21, J21c, Z, Zc, rf1, heat, lam);
    return hipGetLastError();
}

hipError_t cool_u_exact(hipStream_t st, int64_t n, int what, const double *u, const double *nH, double XH,
                        const double *J21, double J21c, const double *Z, double Zc, const double *rf1, double *out)
{
    k_cool_u<<<sph_blocks(n, 256), 256, 0, st>>>(n, what, u, nH, XH, J21, J21c, Z, Zc, rf1, out);
    return hipGetLastError();
}

hipError_t do_cooling_exact(hipStream_t st, int64_t n, const double *rho, const double *u_old, const double *dt,
                            double dtc, double XH, const double *J21, double J21c, const double *Z, double Zc,
                            double u_floor, const double *rf1, double *out)
{
    k_do_cooling<<<sph_blocks(n, 256), 256, 0, st>>>(n, rho, u_old, dt, dtc, XH, J21, J21c, Z, Zc, u_floor, rf1, out);
    return hipGetLastError();
}

hipError_t cool_table_build(hipStream_t st, const CoolAxes &ax, const double *rf1, float *logL, float *logG)
{
    const int64_t total = (int64_t)ax.n[0] * ax.n[1] * ax.n[2] * ax.n[3];
    k_cool_table<<<sph_blocks(total, 256), 256, 0, st>>>(ax, rf1, logL, logG);
    return hipGetLastError();
}

hipError_t cool_table_lookup(hipStream_t st, int64_t n, const CoolAxes &ax, const float *logL, const float *logG,
                             const float *lnH, const float *lT, const float *lZ, const float *lU, float *outL,
                             float *outG)
{
    k_cool_lookup<<<sph_blocks(n, 256), 256, 0, st>>>(n, ax, logL, logG, lnH, lT, lZ, lU, outL, outG);
    return hipGetLastError();
}

hipError_t cool_log_lambda_points(hipStream_t st, int64_t n, const float *x4, const double *rf1, float *logL)
{
    k_cool_points<<<sph_blocks(n, 256), 256, 0, st>>>(n, x4, rf1, logL);
    return hipGetLastError();
}

hipError_t cool_ttable_build(hipStream_t st, const CoolTAxes &tax, double XH, const double *rf1, float *logT)
{
    const int64_t total = (int64_t)tax.n[0] * tax.n[1] * tax.n[2];
    k_cool_ttable<<<sph_blocks(total, 256), 256, 0, st>>>(tax, XH, rf1, logT);
    return hipGetLastError();
}

hipError_t cool_ttable_lookup(hipStream_t st, int64_t n, const CoolTAxes &tax, const float *logT, const double *u,
                              const double *nH, const double *J21, double J21c, double *T)
{
    k_cool_tlookup<<<sph_blocks(n, 256), 256, 0, st>>>(n, tax, logT, u, nH, J21, J21c, T);
    return hipGetLastError();
}

template <typename RHO>
hipError_t cool_update(hipStream_t st, int64_t n, int kind, const RHO *rho, const double *u, const double *Z,
                       const double *J21, const CoolUpdate &cu, const CoolAxes &ax, const float *tabL,
                       const float *tabG, const CoolTAxes &tax, const float *tabT, const double *rf1, double *u_out,
                       float *cols, float *logG, unsigned long long *nfloor)
{
    const unsigned nb = sph_blocks(n, 256);
    if (kind == SPH_COOLK_EXACT)
        k_cool_update<SPH_COOLK_EXACT, RHO><<<nb, 256, 0, st>>>(n, rho, u, Z, J21, cu, ax, tabL, tabG, tax, tabT, rf1,
                                                                 u_out, cols, logG, nfloor);
    else if (kind == SPH_COOLK_TABLE)
        k_cool_update<SPH_COOLK_TABLE, RHO><<<nb, 256, 0, st>>>(n, rho, u, Z, J21, cu, ax, tabL, tabG, tax, tabT, rf1,
                                                                 u_out, cols, logG, nfloor);
    else if (kind == SPH_COOLK_PREP)
        k_cool_update<SPH_COOLK_PREP, RHO><<<nb, 256, 0, st>>>(n, rho, u, Z, J21, cu, ax, tabL, tabG, tax, tabT, rf1,
                                                                u_out, cols, logG, nfloor);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename RHO>
hipError_t cool_apply(hipStream_t st, int64_t n, const RHO *rho, const double *u, const float *logL,
                      const float *logG, const CoolUpdate &cu, double *u_out, unsigned long long *nfloor)
{
    k_cool_apply<RHO><<<sph_blocks(n, 256), 256, 0, st>>>(n, rho, u, logL, logG, cu, u_out, nfloor);
    return hipGetLastError();
}

#define INST(RHO)                                                                                                    \
    template hipError_t cool_update<RHO>(hipStream_t, int64_t, int, const RHO *, const double *, const double *,      \
                                         const double *, const CoolUpdate &, const CoolAxes &, const float *,        \
                                         const float *, const CoolTAxes &, const float *, const double *, double *,  \
                                         float *, float *, unsigned long long *);                                    \
    template hipError_t cool_apply<RHO>(hipStream_t, int64_t, const RHO *, const double *, const float *,            \
                                        const float *, const CoolUpdate &, double *, unsigned long long *);
INST(float)
INST(double)
#undef INST

}  // namespace sph
