/*
 * sph_nhcool.c -- CPU restatement of the reference's cooling update from a table over
 * (log n_H, log T, r, log N_H): doCloudyHCooling with find_nx0 and interpolate_4d_hypercube
 * (Particles_in_BH_Tree_III/CLOUDY_reloaded_Here/run_simulation/
 * hfvCLibs_v7_hfv_v8_gradRho_v4t.h:1450-1476, :1681-1765, :1767-1884), the table layout of
 * HCoolMu.bin as hfvGpuSphBHtree_gradRho_v3t.cu:45-122 reads it.
 *
 * TEST INFRASTRUCTURE ONLY (see sph_oracle.c).  FP32 arithmetic in the reference's order and
 * types, one rounding per operation (-ffp-contract=off).  Where the reference takes an
 * integer base, pow(10, float) resolves to the double pow, so 10^x terms are formed in
 * double and the sums that use them follow those types.  Its float log10 and sqrt are
 * taken here as the correctly rounded values: log10 and sqrt in double, rounded to float.
 * The interpolation keeps the reference's axis weights exactly: the second level of the
 * hypercube interpolates between corners that differ in r with the T fraction, the third
 * between corners that differ in T with the r fraction.
 */
#include <math.h>
#include <stdint.h>

#define NH_KB 1.3807e-16f
#define NH_MH 1.673534e-24f

static float lg10f(float v) { return (float)log10((double)v); }

/* find_nx0 (v4t.h:1450-1476): the grid point closest to x (first on ties), minus one when
 * x lies below it; never the last point, never below 0 */
int or_nh_nx0(const float *A, int N, float x)
{
    int c = 0;
    float best = fabsf(A[0] - x);
    for (int i = 1; i < N; i++) {
        float d = fabsf(A[i] - x);
        if (d < best) {
            best = d;
            c = i;
        }
    }
    int k = (x >= A[c]) ? c : c - 1;
    if (k == N - 1) k = c - 1;
    if (k == -1) k = 0;
    return k;
}

/* interpolate_4d_hypercube (v4t.h:1681-1765) on one table t[nH][T][r][NH] */
static float hyper(const float *t, const int *dims, int i0, int j0, int k0, int l0, float dx, float dy, float dz,
                   float dw)
{
    const int s1 = dims[1] * dims[2] * dims[3], s2 = dims[2] * dims[3], s3 = dims[3];
    float p[16];
    int q = 0;
    for (int i = i0; i <= i0 + 1; i++)
        for (int j = j0; j <= j0 + 1; j++)
            for (int k = k0; k <= k0 + 1; k++)
                for (int l = l0; l <= l0 + 1; l++) p[q++] = t[i * s1 + j * s2 + k * s3 + l];
    float c00 = p[0] * (1.0f - dx) + p[8] * dx;
    float c01 = p[1] * (1.0f - dx) + p[9] * dx;
    float c10 = p[2] * (1.0f - dx) + p[10] * dx;
    float c11 = p[3] * (1.0f - dx) + p[11] * dx;
    float c20 = p[4] * (1.0f - dx) + p[12] * dx;
    float c21 = p[5] * (1.0f - dx) + p[13] * dx;
    float c30 = p[6] * (1.0f - dx) + p[14] * dx;
    float c31 = p[7] * (1.0f - dx) + p[15] * dx;
    float c0 = c00 * (1.0f - dy) + c10 * dy;
    float c1 = c01 * (1.0f - dy) + c11 * dy;
    float c2 = c20 * (1.0f - dy) + c30 * dy;
    float c3 = c21 * (1.0f - dy) + c31 * dy;
    float c = c0 * (1.0f - dz) + c2 * dz;
    float d = c1 * (1.0f - dz) + c3 * dz;
    return c * (1.0f - dw) + d * dw;
}

static float clampb(float v, float lo, float hi)
{
    if (v < lo) v = (float)((double)lo + 0.0001);
    if (v > hi) v = (float)((double)hi - 0.0001);
    return v;
}

/* doCloudyHCooling (v4t.h:1767-1884) for n particles: grids g (n_nH, n_T, n_r, n_NH values in
 * that order), bounds b (nH, T, r, NH low/high pairs), tables heat, cool, mu, logu of
 * n_nH n_T n_r n_NH floats; pos (x, y, z per particle), rho, u, log_nh in code units as
 * floats; typ may be NULL (all gas).  f = (unit_density, unit_u, unit_time, XH, u_floor,
 * gamma, dt).  Returns the floored count. */
int64_t or_nh_cooling(int64_t n, const int *dims, const float *b, const float *g, const float *heat,
                      const float *cool, const float *mu, const float *logu, const float *pos, const float *rho,
                      const float *u, const float *log_nh, const int32_t *typ, const double *f, float *u_out)
{
    const float *gnH = g, *gT = g + dims[0], *gr = gT + dims[1], *gNH = gr + dims[2];
    const float unit_d = (float)f[0], unit_u = (float)f[1], unit_t = (float)f[2], XH = (float)f[3];
    const float gamma = (float)f[5], dt = (float)f[6];
    const double u_floor = f[4];
    const int ntot = dims[0] * dims[1] * dims[2] * dims[3];
    const int s1 = dims[1] * dims[2] * dims[3], s2 = dims[2] * dims[3], s3 = dims[3];
    int64_t nfloor = 0;
    for (int64_t i = 0; i < n; i++) {
        if (typ && typ[i] != 0) {
            u_out[i] = u[i];
            continue;
        }
        const float rhocgs = rho[i] * unit_d;
        float nH_p = lg10f(rhocgs * XH / NH_MH);
        const float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        float r_p = (float)sqrt((double)(x * x + y * y + z * z));
        float NH_p = log_nh[i];
        float u_p = u[i] * unit_u;
        nH_p = clampb(nH_p, b[0], b[1]);
        r_p = clampb(r_p, b[4], b[5]);
        NH_p = clampb(NH_p, b[6], b[7]);
        const int i0 = or_nh_nx0(gnH, dims[0], nH_p), k0 = or_nh_nx0(gr, dims[2], r_p),
                  l0 = or_nh_nx0(gNH, dims[3], NH_p);
        u_p = lg10f(u_p);
        int iu = -1;
        for (int j = 0; j < dims[1]; j++) {
            const int ndx = (i0 + 1) * s1 + j * s2 + (k0 + 1) * s3 + (l0 + 1);
            if (u_p <= logu[ndx]) {
                iu = ndx;
                break;
            }
        }
        if (iu == -1) iu = ntot - 1;
        const float mu_p = mu[iu];
        const float pre = (gamma - 1.0f) * mu_p * NH_MH / NH_KB;
        float T_p = (float)((double)pre * pow(10.0, (double)u_p));
        T_p = lg10f(T_p);
        T_p = clampb(T_p, b[2], b[3]);
        const int j0 = or_nh_nx0(gT, dims[1], T_p);
        const float dx = (nH_p - gnH[i0]) / (gnH[i0 + 1] - gnH[i0]);
        const float dy = (T_p - gT[j0]) / (gT[j0 + 1] - gT[j0]);
        const float dz = (r_p - gr[k0]) / (gr[k0 + 1] - gr[k0]);
        const float dw = (NH_p - gNH[l0]) / (gNH[l0 + 1] - gNH[l0]);
        const float GamH = hyper(heat, dims, i0, j0, k0, l0, dx, dy, dz, dw);
        const float LamC = hyper(cool, dims, i0, j0, k0, l0, dx, dy, dz, dw);
        const float gml = (float)(pow(10.0, (double)GamH) - pow(10.0, (double)LamC));
        const float dt_sec = dt * unit_t;
        float u_tmp = u[i] * unit_u + gml / rhocgs * dt_sec;
        if ((double)u_tmp < u_floor) {
            u_tmp = (float)u_floor;
            nfloor++;
        }
        u_out[i] = u_tmp / unit_u;
    }
    return nfloor;
}
