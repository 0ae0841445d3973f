// sph_agn.hip -- AGN source terms of the outflow driver: the isothermal halo's
// acceleration and the injection of outflow particle pairs.
//
// Replaces galaxy_isothermal_potential and outflow_injector2
// (Particles_in_BH_Tree_III/Evrard_Collapse_2024/hfvCLibs_v7_hfv_v8.1.h:1286-1428, called by
// Particles_in_BH_Tree_III/Main_Code/hfvGpuSphBHtree_newCool_v8.cu:1517 and :1636).  The
// reference injects in a single thread (a serial loop over the pairs, cuRAND states seeded by
// hash(seed, j)); here the host decides the number of pairs from the mass budget and one
// thread per pair draws its numbers from a counter-based generator, so the draws do not
// depend on the launch shape.
#include "sph_agn.h"

namespace {

__global__ void __launch_bounds__(256) k_halo(int64_t n, const double *__restrict__ x, const double *__restrict__ y,
                                              const double *__restrict__ z, double s2, double *__restrict__ ax,
                                              double *__restrict__ ay, double *__restrict__ az)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = x[i], yi = y[i], zi = z[i], r2 = xi * xi + yi * yi + zi * zi;
    if (!(r2 > 0.0)) return;
    const double f = -2.0 * s2 / r2;  // -G (2 sigma^2 r / G) / r^3 = -2 sigma^2 / r^2
    ax[i] += f * xi;
    ay[i] += f * yi;
    az[i] += f * zi;
}

__global__ void __launch_bounds__(256) k_h_of_caller(int64_t n, const int32_t *__restrict__ orig,
                                                     const double *__restrict__ h, int64_t caller,
                                                     double *__restrict__ out)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s < n && orig[s] == caller) *out = h[s];
}

__global__ void __launch_bounds__(256) k_inject(int64_t n, int64_t pairs, sph::InjectParams p,
                                                const double *__restrict__ hsrc, sph::InjectTarget t)
{
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= pairs) return;
    const double rt = inject_uniform(p.seed, p.step, (uint64_t)j, 0) * p.r_inject;
    const double theta = inject_uniform(p.seed, p.step, (uint64_t)j, 1) * (2.0 * SPH_PI);
    const double phi = acos(2.0 * inject_uniform(p.seed, p.step, (uint64_t)j, 2) - 1.0);
    const double ex = sin(phi) * cos(theta), ey = sin(phi) * sin(theta), ez = cos(phi);
    const double h0 = *hsrc;
    for (int k = 0; k < 2; k++) {
        const int64_t s = n + 2 * j + k;
        const double sg = k ? -1.0 : 1.0;
        t.x[s] = sg * rt * ex; t.y[s] = sg * rt * ey; t.z[s] = sg * rt * ez;
        t.vx[s] = sg * p.v_in * ex; t.vy[s] = sg * p.v_in * ey; t.vz[s] = sg * p.v_in * ez;
        t.m[s] = p.m_inject;
        t.h[s] = h0;
        t.u[s] = p.u_inject;
        t.duprev[s] = 0.0;
        t.ax[s] = 0.0; t.ay[s] = 0.0; t.az[s] = 0.0;
        t.orig[s] = (int32_t)s;
        if (t.zmet) t.zmet[s] = p.Z_inject;
        if (t.j21) t.j21[s] = p.J21_inject;
    }
}

}  // namespace

namespace sph {

hipError_t halo_accel(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, double sigma,
                      double *ax, double *ay, double *az)
{
    if (n <= 0 || sigma == 0.0) return hipSuccess;
    k_halo<<<sph_blocks(n, 256), 256, 0, st>>>(n, x, y, z, sigma * sigma, ax, ay, az);
    return hipGetLastError();
}

hipError_t h_of_caller(hipStream_t st, int64_t n, const int32_t *orig, const double *h, int64_t caller, double *hsrc)
{
    if (n <= 0) return hipErrorInvalidValue;
    k_h_of_caller<<<sph_blocks(n, 256), 256, 0, st>>>(n, orig, h, caller, hsrc);
    return hipGetLastError();
}

hipError_t inject_pairs(hipStream_t st, int64_t n, int64_t pairs, const InjectParams &p, const double *hsrc,
                        const InjectTarget &t)
{
    if (pairs <= 0) return hipSuccess;
    k_inject<<<sph_blocks(pairs, 256), 256, 0, st>>>(n, pairs, p, hsrc, t);
    return hipGetLastError();
}

}  // namespace sph
