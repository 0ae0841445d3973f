/*
 * sph_cooling.c -- FP64 CPU restatement of the reference's closed-form
 * primordial cooling/heating (Particles_in_BH_Tree_III/hcool_with_unit_tests_Here/
 * photolibs3.py, "HC/" below) plus the build's metal-line term.
 *
 * TEST INFRASTRUCTURE ONLY (see sph_oracle.c).  Pinned against golden vectors
 * produced by the reference module itself (tests/golden/cooling_photolibs3.json).
 *
 * Generalisation: the reference's radiation field has a fixed amplitude
 * Jcoeff = 1e-7 (HC/photolibs3.py:233); here it is the argument J21 (same
 * spectral shape, J_nu = J21 1e-21 (nu/nu_edge)^-1).  Z (in solar units) adds
 * Z * Lambda_RB95(T), the solar-metallicity CIE cooling fit of Rosen & Bregman
 * (1995, ApJ 440, 634), as the metal-line term; Z = 0 is the reference.
 */
#include <math.h>
#include <stdint.h>

#define CL_C 29979245800.0
#define CL_HEV 4.135667696e-15
#define CL_HPL 6.626196e-27
#define CL_KB 1.3807e-16
#define CL_MH 1.6726e-24

/* RandCIRates, HC/photolibs3.py:140-165 */
void or_cool_rates(double T, double *o)
{
    double Tfact = 1.0 / (1.0 + sqrt(T / 1e5));
    o[0] = 8.41e-11 * pow(T / 1000.0, -0.2) / (1.0 + pow(T / 1e6, 0.7)) / sqrt(T);
    o[1] = 1.5e-10 * pow(T, -0.6353);
    o[2] = 4.0 * o[0];
    o[3] = 1.9e-3 * pow(T, -1.5) * exp(-470000.0 / T) * (1.0 + 0.3 * exp(-94000.0 / T));
    o[4] = 5.85e-11 * sqrt(T) * exp(-157809.1 / T) * Tfact;
    o[5] = 2.38e-11 * sqrt(T) * exp(-285335.4 / T) * Tfact;
    o[6] = 5.68e-12 * sqrt(T) * exp(-631515.0 / T) * Tfact;
}

/* Verner et al. (1996) photoionisation cross section, HC/photolibs3.py:120-135 */
static double ph_sigma(double nu, double E0, double sig0, double ya, double P, double yw, double y0, double y1)
{
    double E = CL_HEV * nu;
    double x = E / E0 - y0;
    double y = sqrt(x * x + y1 * y1);
    double Fy = ((x - 1.0) * (x - 1.0) + yw * yw) * pow(y, 0.5 * P - 5.5) * pow(1.0 + sqrt(y / ya), -P);
    return sig0 * Fy * 1e-18;
}

/* Frequency grid of HC/photolibs3.py:247-249: numpy.arange(nu_edge/1e16, 6.0,
 * 0.02) * 1e16, with numpy's float arange (length ceil((stop-start)/step),
 * values start + i*((start+step)-start)). */
static int nu_grid(double Eedge, double *nu, int cap)
{
    double start = (Eedge / CL_HEV) / 1e16;
    int n = (int)ceil((6.0 - start) / 0.02);
    if (n > cap) n = cap;
    double delta = (start + 0.02) - start;
    for (int i = 0; i < n; i++) nu[i] = (start + i * delta) * 1e16;
    return n;
}

/* RadiationField with amplitude J21, HC/photolibs3.py:230-329: photoionisation
 * rates gJ[0..2] (H0, He0, He+) and heating rates per atom gJ[3..5]. */
void or_cool_radfield(double J21, double *out)
{
    static const double E[3] = {13.60, 24.59, 54.42};
    static const double par[3][7] = {{4.298E-1, 5.475E4, 3.288E1, 2.963, 0.0, 0.0, 0.0},
                                     {1.361E+1, 9.492E2, 1.469E0, 3.188, 2.039, 4.434E-1, 2.136},
                                     {1.720E+0, 1.369E4, 3.288E1, 2.963, 0.0, 0.0, 0.0}};
    double nu[1024];
    for (int s = 0; s < 3; s++) {
        int n = nu_grid(E[s], nu, 1024);
        /* edge frequency through the wavelength round trip of J_neu_func (:215-227) */
        double nu_e = E[s] / CL_HEV;
        double lam = CL_C / nu_e * 1e8;
        double nu_e2 = CL_C / (lam * 1e-8);
        double g = 0.0, hr = 0.0, prev_g = 0.0, prev_h = 0.0;
        double dnu = nu[1] - nu[0];
        for (int i = 0; i < n; i++) {
            double sig = ph_sigma(nu[i], par[s][0], par[s][1], par[s][2], par[s][3], par[s][4], par[s][5], par[s][6]);
            double J = J21 * 1e-21 / (nu[i] / nu_e2);
            double fg = 4.0 * M_PI * J * sig / CL_HPL / nu[i];
            double fh = 4.0 * M_PI * J * sig * (nu[i] - E[s] / CL_HEV) / nu[i];
            if (i > 0) {
                g += dnu * (prev_g + fg) / 2.0;
                hr += dnu * (prev_h + fh) / 2.0;
            }
            prev_g = fg;
            prev_h = fh;
        }
        out[s] = g;
        out[3 + s] = hr;
    }
}

/* Abundance_hX, HC/photolibs3.py:169-212: damped fixed-point iteration on
 * n_e/n_H.  out = nH0, nHe0, nHp, ne, nHep, nHepp (relative to n_H). */
void or_cool_abund(double T, double nH, const double *gJ, double *out)
{
    double r[7];
    or_cool_rates(T, r);
    const double aHp = r[0], aHep = r[1], aHepp = r[2], ad = r[3], geH0 = r[4], geHe0 = r[5], geHep = r[6];
    const double Y = 0.24, y = Y / (4.0 - 4.0 * Y);
    double ne = 1.0, ne_old, dne = 1.0;
    double nH0 = 0, nHp = 0, nHep = 0, nHe0 = 0, nHepp = 0;
    int niter = 1;
    while (dne > 1e-4 && niter < 100) {
        ne_old = ne;
        nH0 = aHp / (aHp + geH0 + gJ[0] / (ne * nH));
        nHp = 1.0 - nH0;
        nHep = y / (1.0 + (aHep + ad) / (geHe0 + gJ[1] / (ne * nH)) + (geHep + gJ[2] / (ne * nH)) / aHepp);
        nHe0 = nHep * (aHep + ad) / (geHe0 + gJ[1] / (ne * nH));
        nHepp = nHep * (geHep + gJ[2] / (ne * nH)) / aHepp;
        ne = nHp + nHep + 2.0 * nHepp;
        ne = 0.5 * (ne + ne_old);
        dne = fabs(ne - ne_old);
        niter++;
    }
    out[0] = nH0; out[1] = nHe0; out[2] = nHp; out[3] = ne; out[4] = nHep; out[5] = nHepp;
}

/* Rosen & Bregman (1995) solar-metallicity cooling function [erg cm^3 s^-1] */
double or_cool_rb95(double T)
{
    if (T < 300.0) return 0.0;
    if (T < 2000.0) return 2.2380e-32 * T * T;
    if (T < 8000.0) return 1.0012e-30 * pow(T, 1.5);
    if (T < 1e5) return 4.6240e-36 * pow(T, 2.867);
    if (T < 4e7) return 1.7800e-18 * pow(T, -0.65);
    return 3.2217e-27 * sqrt(T);
}

/* coolingHeatingRates, HC/photolibs3.py:332-376 (+ Z * RB95):
 * out[0] = heating / nH^2, out[1] = cooling / nH^2. */
void or_cool_heat_lambda(double T, double nH, double J21, double Z, double *out)
{
    double rf[6], r[7], ab[6];
    or_cool_radfield(J21, rf);
    or_cool_rates(T, r);
    or_cool_abund(T, nH, rf, ab);
    const double nH0 = ab[0], nHe0 = ab[1], nHp = ab[2], ne = ab[3], nHep = ab[4], nHepp = ab[5];
    const double aHp = r[0], aHep = r[1], aHepp = r[2], ad = r[3], geH0 = r[4], geHe0 = r[5], geHep = r[6];
    double heat = nH0 * rf[3] / nH + nHe0 * rf[4] / nH + nHep * rf[5] / nH;
    double Tfact = 1.0 / (1.0 + sqrt(T / 1e5));
    double BetaH0 = 7.5e-19 * exp(-118348.0 / T) * Tfact;
    double BetaHep = 5.54e-17 * pow(T, -0.397) * exp(-473638.0 / T) * Tfact;
    double lt = 5.5 - log10(T);
    double Betaff = 1.43e-27 * sqrt(T) * (1.0 + 0.34 * exp(-lt * lt / 3.0));
    double Lexc = BetaH0 * ne * nH0 + BetaHep * ne * nHep;
    double Lion = 2.18e-11 * geH0 * ne * nH0 + 3.94e-11 * geHe0 * ne * nHe0 + 8.72e-11 * geHep * ne * nHep;
    double Lrec = 1.036e-16 * T * ne * (aHp * nHp) + 1.036e-16 * T * ne * (aHep * nHep) +
                  1.036e-16 * T * ne * (aHepp * nHepp) + 6.526e-11 * ad * ne * nHep;
    double Lff = Betaff * (nHp + nHep + 4.0 * nHepp) * ne;
    out[0] = heat;
    out[1] = Lexc + Lion + Lrec + Lff + Z * or_cool_rb95(T);
}

/* convert_u_to_temp_h, HC/photolibs3.py:379-420 */
double or_cool_u_to_T(double u, double nH, double XH, double J21)
{
    const double gamma = 5.0 / 3.0;
    double yHe = (1.0 - XH) / (4.0 * XH);
    double mu = (1.0 + 4.0 * yHe) / (1.0 + yHe + 1.0);
    double temp = (gamma - 1.0) * CL_MH / CL_KB * mu * u;
    double rf[6], ab[6];
    or_cool_radfield(J21, rf);
    double dT = temp, temp_old;
    int niter = 1;
    while (dT / temp > 1e-4 && niter < 100) {
        temp_old = temp;
        or_cool_abund(temp_old, nH, rf, ab);
        mu = (1.0 + 4.0 * yHe) / (1.0 + yHe + ab[3]);
        temp = (gamma - 1.0) * CL_MH / CL_KB * mu * u;
        temp = 0.5 * (temp + temp_old);
        dT = fabs(temp - temp_old);
        niter++;
    }
    return temp;
}

/* coolingRateFromU_h, HC/photolibs3.py:470-486: Gamma - Lambda (per nH^2)
 * at T(u) clamped to [1e4, 1e8] K. */
double or_cool_rate_from_u(double u, double nH, double XH, double J21, double Z)
{
    double T = or_cool_u_to_T(u, nH, XH, J21);
    if (T < 1e4) T = 1e4;
    if (T > 1e8) T = 1e8;
    double hl[2];
    or_cool_heat_lambda(T, nH, J21, Z, hl);
    return hl[0] - hl[1];
}

/* DoCooling_h, HC/photolibs3.py:489-547: implicit u update by bracketing
 * (factors of 1.1) and bisection to |du/u| <= 1e-4.  The reference's downward
 * bracket search has no floor (and overflows for strong cooling); here it
 * stops at u_floor and returns -1 if the bracket could not be formed. */
double or_do_cooling(double rho, double u_old, double dt, double XH, double J21, double Z, double u_floor)
{
    double nH = XH * rho / CL_MH;
    double ratefact = nH * nH / rho;
    double u = u_old, u_upper = u, u_lower = u;
    double net = or_cool_rate_from_u(u, nH, XH, J21, Z);
    if (u - u_old - ratefact * net * dt > 0.0) {
        while (u_upper - u_old - ratefact * or_cool_rate_from_u(u_upper, nH, XH, J21, Z) * dt > 0.0) {
            u_upper /= 1.1;
            if (u_upper < u_floor) return -1.0;
        }
        u_lower = u_upper;
        u_upper = u_lower * 1.1;
    }
    if (u - u_old - ratefact * net * dt < 0.0) {
        while (u_upper - u_old - ratefact * or_cool_rate_from_u(u_upper, nH, XH, J21, Z) * dt < 0.0)
            u_upper *= 1.1;
        u_lower = u_upper / 1.1;
    }
    int niter = 1;
    double du = u;
    while (fabs(du / u) > 1e-4 && niter < 100) {
        u = 0.5 * (u_lower + u_upper);
        net = or_cool_rate_from_u(u, nH, XH, J21, Z);
        if (u - u_old - ratefact * net * dt > 0.0) u_upper = u;
        else u_lower = u;
        du = fabs(u_upper - u_lower);
        niter++;
    }
    return u;
}

static double clampd(double v, double lo, double hi) { return fmin(fmax(v, lo), hi); }

/* Explicit per-step cooling update (CR/v4t.h:1768-1884 doCloudyHCooling, called after
 * the u update at CR/v3t.cu:1527-1536) with the exact rates of this file in place of the
 * CLOUDY table: rho_cgs = rho * units[0], nH = XH rho_cgs / m_H (units[3] = XH),
 * u_cgs = u * units[1]; nH, T(u), Z, J21 are clamped into the box 10^lo[a] .. 10^hi[a]
 * (a = nH, T, Z, J21) as doCloudyHCooling clamps to its table bounds; the net rate per
 * volume (Gamma - Lambda) nH^2 uses the unclamped nH; u_cgs' = max(u_cgs + rate / rho_cgs
 * * dt * units[2], units[4]) (1.8e11 erg/g in doCloudyHCooling), returned in code units.
 * floored[i] (may be NULL) = 1 where the floor was applied. */
void or_cool_update_n(int64_t n, const double *rho, const double *u, const double *Z, const double *J21, double dt,
                      const double *units, const double *lo, const double *hi, double *out, int32_t *floored)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; i++) {
        const double XH = units[3];
        const double rho_cgs = rho[i] * units[0];
        const double nH = XH * rho_cgs / CL_MH;
        const double u_cgs = u[i] * units[1];
        const double nHc = clampd(nH, pow(10.0, lo[0]), pow(10.0, hi[0]));
        const double zc = clampd(Z[i], pow(10.0, lo[2]), pow(10.0, hi[2]));
        const double jc = clampd(J21[i], pow(10.0, lo[3]), pow(10.0, hi[3]));
        const double Tc = clampd(or_cool_u_to_T(u_cgs, nHc, XH, jc), pow(10.0, lo[1]), pow(10.0, hi[1]));
        double hl[2];
        or_cool_heat_lambda(Tc, nHc, jc, zc, hl);
        double un = u_cgs + (hl[0] - hl[1]) * (nH * nH / rho_cgs) * (dt * units[2]);
        const int f = !(un >= units[4]);
        if (f) un = units[4];
        out[i] = un / units[1];
        if (floored) floored[i] = f;
    }
}
