// sph_nhcool.hip -- cooling update from a table over (log n_H, log T, r, log N_H).
//
// Replaces doCloudyHCooling with find_nx0 and interpolate_4d_hypercube
// (Particles_in_BH_Tree_III/CLOUDY_reloaded_Here/run_simulation/hfvCLibs_v7_hfv_v8_gradRho_v4t.h:
// 1450-1476, 1681-1765, 1767-1884), called after the u update at
// hfvGpuSphBHtree_gradRho_v3t.cu:1527-1536 with the column density of the force evaluation.
// One thread per particle.  The arithmetic is FP32 in the reference's order and types
// (contraction off).  Where the reference writes pow(10, x) with an integer base, the double
// pow is taken and the expressions that use it follow those types.  Its float log10 and sqrt
// are taken as the correctly rounded values: log10 and sqrt in double, rounded to float.  The
// four grids are staged in LDS for the closest-point searches.  The hypercube interpolation
// keeps the reference's axis weights exactly: its second level interpolates between corners
// that differ in r with the T fraction, its third between corners that differ in T with the
// r fraction.
#include "sph_nhcool.h"

#pragma clang fp contract(off)

namespace {

constexpr float NH_KB = 1.3807e-16f;     // k_B as the reference driver holds it (float)
constexpr float NH_MH = 1.673534e-24f;   // m_H likewise

__device__ __forceinline__ float lg10f(float v) { return (float)log10((double)v); }

__device__ __forceinline__ float clampb(float v, float lo, float hi)
{
    if (v < lo) v = (float)((double)lo + 0.0001);
    if (v > hi) v = (float)((double)hi - 0.0001);
    return v;
}

// find_nx0: the grid point closest to x (the first on ties), minus one when x lies below it;
// never the last point, never below 0
__device__ __forceinline__ int nx0(const float *A, int N, float x)
{
    int c = 0;
    float best = fabsf(A[0] - x);
    for (int i = 1; i < N; i++) {
        const float d = fabsf(A[i] - x);
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

// interpolate_4d_hypercube on one table t[nH][T][r][NH]; corner q = 8 a + 4 b + 2 c + d for
// (i0 + a, j0 + b, k0 + c, l0 + d)
__device__ __forceinline__ float hyper(const float *__restrict__ t, int s1, int s2, int s3, int i0, int j0, int k0,
                                       int l0, float dx, float dy, float dz, float dw)
{
    float p[16];
#pragma unroll
    for (int q = 0; q < 16; q++)
        p[q] = t[(i0 + (q >> 3)) * s1 + (j0 + ((q >> 2) & 1)) * s2 + (k0 + ((q >> 1) & 1)) * s3 + l0 + (q & 1)];
    const float c00 = p[0] * (1.0f - dx) + p[8] * dx;
    const float c01 = p[1] * (1.0f - dx) + p[9] * dx;
    const float c10 = p[2] * (1.0f - dx) + p[10] * dx;
    const float c11 = p[3] * (1.0f - dx) + p[11] * dx;
    const float c20 = p[4] * (1.0f - dx) + p[12] * dx;
    const float c21 = p[5] * (1.0f - dx) + p[13] * dx;
    const float c30 = p[6] * (1.0f - dx) + p[14] * dx;
    const float c31 = p[7] * (1.0f - dx) + p[15] * dx;
    const float c0 = c00 * (1.0f - dy) + c10 * dy;
    const float c1 = c01 * (1.0f - dy) + c11 * dy;
    const float c2 = c20 * (1.0f - dy) + c30 * dy;
    const float c3 = c21 * (1.0f - dy) + c31 * dy;
    const float c = c0 * (1.0f - dz) + c2 * dz;
    const float d = c1 * (1.0f - dz) + c3 * dz;
    return c * (1.0f - dw) + d * dw;
}

template <typename RHO>
__global__ void __launch_bounds__(256) k_nh_cool(int64_t n, const double *__restrict__ x, const double *__restrict__ y,
                                                 const double *__restrict__ z, const RHO *__restrict__ rho,
                                                 const double *__restrict__ u, const double *__restrict__ log_nh,
                                                 const int32_t *__restrict__ typ, NHTable t, NHCoolArgs a,
                                                 double *__restrict__ u_out, unsigned long long *__restrict__ floors)
{
    __shared__ float G[SPH_NH_GRID_MAX];
    const int ng = t.dims[0] + t.dims[1] + t.dims[2] + t.dims[3];
    for (int q = threadIdx.x; q < ng; q += blockDim.x) G[q] = t.grids[q];
    __syncthreads();
    const float *gnH = G, *gT = G + t.dims[0], *gr = gT + t.dims[1], *gNH = gr + t.dims[2];
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool floored = false;
    if (i < n) {
        if (typ && typ[i] != 0) {
            u_out[i] = u[i];
        } else {
            const int s1 = t.dims[1] * t.dims[2] * t.dims[3], s2 = t.dims[2] * t.dims[3], s3 = t.dims[3];
            const float uf = (float)u[i];
            const float rhocgs = (float)rho[i] * a.unit_density;
            float nH_p = lg10f(rhocgs * a.XH / NH_MH);
            const float xf = (float)x[i], yf = (float)y[i], zf = (float)z[i];
            float r_p = (float)__builtin_sqrt((double)(xf * xf + yf * yf + zf * zf));
            float NH_p = (float)log_nh[i];
            float u_p = uf * a.unit_u;
            nH_p = clampb(nH_p, t.bounds[0], t.bounds[1]);
            r_p = clampb(r_p, t.bounds[4], t.bounds[5]);
            NH_p = clampb(NH_p, t.bounds[6], t.bounds[7]);
            const int i0 = nx0(gnH, t.dims[0], nH_p), k0 = nx0(gr, t.dims[2], r_p), l0 = nx0(gNH, t.dims[3], NH_p);
            u_p = lg10f(u_p);
            // mu of the first T row (at the upper nH, r, N_H corner) whose log u reaches u_p
            int iu = -1;
            for (int j = 0; j < t.dims[1]; j++) {
                const int ndx = (i0 + 1) * s1 + j * s2 + (k0 + 1) * s3 + (l0 + 1);
                if (u_p <= t.logu[ndx]) {
                    iu = ndx;
                    break;
                }
            }
            if (iu == -1) iu = t.dims[0] * s1 - 1;
            const float mu_p = t.mu[iu];
            const float pre = (a.gamma - 1.0f) * mu_p * NH_MH / NH_KB;
            float T_p = (float)((double)pre * pow(10.0, (double)u_p));
            T_p = lg10f(T_p);
            T_p = clampb(T_p, t.bounds[2], t.bounds[3]);
            const int j0 = nx0(gT, t.dims[1], T_p);
            const float dx = (nH_p - gnH[i0]) / (gnH[i0 + 1] - gnH[i0]);
            const float dy = (T_p - gT[j0]) / (gT[j0 + 1] - gT[j0]);
            const float dz = (r_p - gr[k0]) / (gr[k0 + 1] - gr[k0]);
            const float dw = (NH_p - gNH[l0]) / (gNH[l0 + 1] - gNH[l0]);
            const float GamH = hyper(t.heat, s1, s2, s3, i0, j0, k0, l0, dx, dy, dz, dw);
            const float LamC = hyper(t.cool, s1, s2, s3, i0, j0, k0, l0, dx, dy, dz, dw);
            const float gml = (float)(pow(10.0, (double)GamH) - pow(10.0, (double)LamC));
            const float dt_sec = a.dt * a.unit_time;
            float u_tmp = uf * a.unit_u + gml / rhocgs * dt_sec;
            if ((double)u_tmp < a.u_floor) {
                u_tmp = (float)a.u_floor;
                floored = true;
            }
            u_out[i] = (double)(u_tmp / a.unit_u);
        }
    }
    const uint64_t fm = __ballot(floored);
    if ((threadIdx.x & 63) == 0 && fm) spread_add(floors, 0, (unsigned long long)__popcll(fm));
}

}  // namespace

namespace sph {

template <typename RHO>
hipError_t nh_cool_update(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const RHO *rho,
                          const double *u, const double *log_nh, const int32_t *typ, const NHTable &t,
                          const NHCoolArgs &a, double *u_out, unsigned long long *floors)
{
    if (t.dims[0] + t.dims[1] + t.dims[2] + t.dims[3] > SPH_NH_GRID_MAX) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    k_nh_cool<RHO><<<sph_blocks(n, 256), 256, 0, st>>>(n, x, y, z, rho, u, log_nh, typ, t, a, u_out, floors);
    return hipGetLastError();
}

template hipError_t nh_cool_update<float>(hipStream_t, int64_t, const double *, const double *, const double *,
                                          const float *, const double *, const double *, const int32_t *,
                                          const NHTable &, const NHCoolArgs &, double *, unsigned long long *);
template hipError_t nh_cool_update<double>(hipStream_t, int64_t, const double *, const double *, const double *,
                                           const double *, const double *, const double *, const int32_t *,
                                           const NHTable &, const NHCoolArgs &, double *, unsigned long long *);

}  // namespace sph
