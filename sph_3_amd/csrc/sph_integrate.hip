// sph_integrate.hip -- FP64 kick/drift, the trapezoidal u update, per-call
// preparation of P/rho^2 and the Balsara factor, and FP64 conservation
// reductions.
//
// Integrator follows sph_code/sph_serial_v2.py:148-203 (KDK leapfrog,
// u = u_prev + dt/2 (du + du_prev)); GPU twins EV/lib.h:1045-1073,1135-1148.
// Reductions restate getKE (sph_code/libsx.py:569-579) plus linear momentum.
#include "sph_integrate.h"

#pragma clang fp contract(off)

__global__ void __launch_bounds__(256) k_kick_drift(int64_t n, double *__restrict__ vx, double *__restrict__ vy,
                                                    double *__restrict__ vz, const double *__restrict__ ax,
                                                    const double *__restrict__ ay, const double *__restrict__ az,
                                                    double dtk, double dtd, double *__restrict__ x,
                                                    double *__restrict__ y, double *__restrict__ z, double L)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = vx[i] + ax[i] * dtk, b = vy[i] + ay[i] * dtk, c = vz[i] + az[i] * dtk;
    vx[i] = a; vy[i] = b; vz[i] = c;
    if (x) {
        double p[3] = {x[i] + a * dtd, y[i] + b * dtd, z[i] + c * dtd};
        if (L > 0.0) {  // back into [0, L) (one image; a step moves a particle less than L)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                if (p[d] >= L) p[d] = p[d] - L;
                else if (p[d] < 0.0) p[d] = p[d] + L;
            }
        }
        x[i] = p[0];
        y[i] = p[1];
        z[i] = p[2];
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_u_update(int64_t n, double *__restrict__ u, const R *__restrict__ du,
                                                  double *__restrict__ du_prev, double dt, int first)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = (double)du[i];
    if (first) u[i] = u[i] + d * dt;
    else u[i] = u[i] + 0.5 * dt * (d + du_prev[i]);
    du_prev[i] = d;
}

template <typename R>
__global__ void __launch_bounds__(256) k_prep(int64_t n, const R *__restrict__ rho, const R *__restrict__ P,
                                              const R *__restrict__ c, const R *__restrict__ divV,
                                              const R *__restrict__ curlV, const double *__restrict__ h,
                                              R *__restrict__ prho2, R *__restrict__ bal)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    R r = rho[i];
    prho2[i] = P[i] / (r * r);
    if (bal) {
        R d = divV[i];
        bal[i] = d / (d + curlV[i] + (R)0.0001 * c[i] / (R)h[i]);
    }
}

// [KE, px, py, pz, M] per block, then folded by one block (deterministic).  Velocity
// component k of particle i is vk[i * vs] (vs = 1: three arrays; vs = 3: one n x 3 array).
__global__ void __launch_bounds__(256) k_kin_partial(int64_t n, const double *__restrict__ vx,
                                                     const double *__restrict__ vy, const double *__restrict__ vz,
                                                     int64_t vs, const double *__restrict__ m,
                                                     double *__restrict__ part)
{
    __shared__ double s[5][256];
    double v[5] = {0, 0, 0, 0, 0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double a = vx[i * vs], b = vy[i * vs], c = vz[i * vs], mm = m[i];
        v[0] += 0.5 * mm * (a * a + b * b + c * c);
        v[1] += mm * a;
        v[2] += mm * b;
        v[3] += mm * c;
        v[4] += mm;
    }
    for (int d = 0; d < 5; d++) s[d][threadIdx.x] = v[d];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off)
            for (int d = 0; d < 5; d++) s[d][threadIdx.x] += s[d][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 5) part[blockIdx.x * 5 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) k_kin_final(int nb, const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double s[5][256];
    double v[5] = {0, 0, 0, 0, 0};
    for (int p = threadIdx.x; p < nb; p += blockDim.x)
        for (int d = 0; d < 5; d++) v[d] += part[p * 5 + d];
    for (int d = 0; d < 5; d++) s[d][threadIdx.x] = v[d];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off)
            for (int d = 0; d < 5; d++) s[d][threadIdx.x] += s[d][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 5) out[threadIdx.x] = s[threadIdx.x][0];
}

// Conserved totals of the state, FP64, one pass: [KE, U, PE, px, py, pz, Lx, Ly, Lz, M] per
// block (KE = sum m v^2 / 2, U = sum m u, PE = -1/2 sum m phi, p = sum m v, L = sum m x cross v);
// a null array contributes zeros.  Folded by one block in block order (deterministic).
#define SPH_NTOT 10
__global__ void __launch_bounds__(256) k_tot_partial(int64_t n, const double *__restrict__ x,
                                                     const double *__restrict__ y, const double *__restrict__ z,
                                                     const double *__restrict__ vx, const double *__restrict__ vy,
                                                     const double *__restrict__ vz, const double *__restrict__ m,
                                                     const double *__restrict__ u, const double *__restrict__ phi,
                                                     double *__restrict__ part)
{
    __shared__ double s[SPH_NTOT][256];
    double t[SPH_NTOT];
#pragma unroll
    for (int d = 0; d < SPH_NTOT; d++) t[d] = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double mm = m[i];
        const double a = vx ? vx[i] : 0.0, b = vy ? vy[i] : 0.0, c = vz ? vz[i] : 0.0;
        const double px = x ? x[i] : 0.0, py = y ? y[i] : 0.0, pz = z ? z[i] : 0.0;
        t[0] += 0.5 * mm * (a * a + b * b + c * c);
        if (u) t[1] += mm * u[i];
        if (phi) t[2] -= 0.5 * mm * phi[i];
        t[3] += mm * a;
        t[4] += mm * b;
        t[5] += mm * c;
        t[6] += mm * (py * c - pz * b);
        t[7] += mm * (pz * a - px * c);
        t[8] += mm * (px * b - py * a);
        t[9] += mm;
    }
#pragma unroll
    for (int d = 0; d < SPH_NTOT; d++) s[d][threadIdx.x] = t[d];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off)
            for (int d = 0; d < SPH_NTOT; d++) s[d][threadIdx.x] += s[d][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < SPH_NTOT) part[blockIdx.x * SPH_NTOT + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) k_tot_final(int nb, const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double s[SPH_NTOT][256];
    double t[SPH_NTOT];
#pragma unroll
    for (int d = 0; d < SPH_NTOT; d++) t[d] = 0.0;
    for (int p = threadIdx.x; p < nb; p += blockDim.x)
        for (int d = 0; d < SPH_NTOT; d++) t[d] += part[p * SPH_NTOT + d];
#pragma unroll
    for (int d = 0; d < SPH_NTOT; d++) s[d][threadIdx.x] = t[d];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off)
            for (int d = 0; d < SPH_NTOT; d++) s[d][threadIdx.x] += s[d][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < SPH_NTOT) out[threadIdx.x] = s[threadIdx.x][0];
}

namespace sph {

hipError_t totals(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const double *vx,
                  const double *vy, const double *vz, const double *m, const double *u, const double *phi,
                  double *part, double *out10)
{
    int nb = (int)((n + 255) / 256);
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    k_tot_partial<<<nb, 256, 0, st>>>(n, x, y, z, vx, vy, vz, m, u, phi, part);
    k_tot_final<<<1, 256, 0, st>>>(nb, part, out10);
    return hipGetLastError();
}

hipError_t kick_drift(hipStream_t st, int64_t n, double *vx, double *vy, double *vz, const double *ax,
                      const double *ay, const double *az, double dtk, double dtd, double *x, double *y, double *z,
                      double L)
{
    k_kick_drift<<<sph_blocks(n, 256), 256, 0, st>>>(n, vx, vy, vz, ax, ay, az, dtk, dtd, x, y, z, L);
    return hipGetLastError();
}

template <typename R>
hipError_t u_update(hipStream_t st, int64_t n, double *u, const R *du, double *du_prev, double dt, int first)
{
    k_u_update<R><<<sph_blocks(n, 256), 256, 0, st>>>(n, u, du, du_prev, dt, first);
    return hipGetLastError();
}

template <typename R>
hipError_t prep(hipStream_t st, int64_t n, const R *rho, const R *P, const R *c, const R *divV, const R *curlV,
                const double *h, R *prho2, R *bal)
{
    k_prep<R><<<sph_blocks(n, 256), 256, 0, st>>>(n, rho, P, c, divV, curlV, h, prho2, bal);
    return hipGetLastError();
}

hipError_t kinetics(hipStream_t st, int64_t n, const double *vx, const double *vy, const double *vz,
                    const double *m, double *part, double *out5, int64_t vstride)
{
    int nb = (int)((n + 255) / 256);
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    k_kin_partial<<<nb, 256, 0, st>>>(n, vx, vy, vz, vstride, m, part);
    k_kin_final<<<1, 256, 0, st>>>(nb, part, out5);
    return hipGetLastError();
}

template hipError_t u_update<float>(hipStream_t, int64_t, double *, const float *, double *, double, int);
template hipError_t u_update<double>(hipStream_t, int64_t, double *, const double *, double *, double, int);
template hipError_t prep<float>(hipStream_t, int64_t, const float *, const float *, const float *, const float *,
                                const float *, const double *, float *, float *);
template hipError_t prep<double>(hipStream_t, int64_t, const double *, const double *, const double *,
                                 const double *, const double *, const double *, double *, double *);

}  // namespace sph
