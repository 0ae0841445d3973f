// sph_grav.hip -- Barnes-Hut self-gravity: wavefront-cooperative stackless
// walk of the radix tree with a per-particle opening test.
//
// Replaces EV/bh.h:454-588 compute_forces_kernel (warp-shared stack, one
// opening decision per warp via __all_sync) and the CPU direct sums
// sph_code/libsx.py:623-662 / :908-947 (getAcc_g_smth[_mimj_mpi]).
//
// One lane per target particle, targets in Morton order.  The wave follows a
// single pre-order path; every active lane decides for itself whether the
// current node is far enough (FP64 test, identical to the CPU oracle's), so
// each particle's interaction set is exactly its own and does not depend on
// which particles share its wave.  A lane that accepts a node sleeps until
// the wave reaches that node's escape link.  Interactions are evaluated in
// the pass type R with Kahan-compensated sums; node centres of mass are FP64.
#include "sph_grav.h"

// Opening test (FP64, no contraction): use the node as a monopole iff
// size^2 < theta^2 |com - x|^2 and x lies outside the node's box.
__device__ __forceinline__ bool mac_accept(double x, double y, double z, const WNode *__restrict__ nd, double theta2)
{
#pragma clang fp contract(off)
    double dx = nd->cx - x, dy = nd->cy - y, dz = nd->cz - z;
    double d2 = dx * dx + dy * dy + dz * dz;
    bool outside = x < nd->lx || x > nd->hx || y < nd->ly || y > nd->hy || z < nd->lz || z > nd->hz;
    return (nd->size * nd->size < theta2 * d2) && outside;
}

template <typename R>
__device__ __forceinline__ void interact(double px, double py, double pz, double pm, double epsj, double xi,
                                         double yi, double zi, double epsi, R G, int soft_kind, R e2, Kahan<R> &sx,
                                         Kahan<R> &sy, Kahan<R> &sz)
{
    R dx = (R)(px - xi), dy = (R)(py - yi), dz = (R)(pz - zi);
    R r2 = dx * dx + dy * dy + dz * dz;
    R fk;
    if (soft_kind == 0) {
        fk = grav_fk_spline<R>(sph_sqrt<R>(r2), (R)(0.5 * (epsi + epsj)));
    } else {
        R inv = sph_rsqrt<R>(r2 + e2);
        fk = inv * inv * inv;
    }
    R w = G * fk * (R)pm;
    sx.add(w * dx);
    sy.add(w * dy);
    sz.add(w * dz);
}

template <typename R>
__global__ void __launch_bounds__(256) k_gravity(int64_t n, const double *__restrict__ x,
                                                 const double *__restrict__ y, const double *__restrict__ z,
                                                 const double *__restrict__ m, const double *__restrict__ eps,
                                                 const WNode *__restrict__ nodes, int bucket, double theta2,
                                                 double G, int soft_kind, double eps_plummer,
                                                 double *__restrict__ ax, double *__restrict__ ay,
                                                 double *__restrict__ az, unsigned long long *__restrict__ counters)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool valid = i < n;
    double xi = 0, yi = 0, zi = 0, epsi = 0;
    if (valid) {
        xi = x[i]; yi = y[i]; zi = z[i];
        epsi = eps ? eps[i] : 0.0;
    }
    const R Gr = (R)G, e2 = (R)(eps_plummer * eps_plummer);
    Kahan<R> sx, sy, sz;
    unsigned npp = 0, npn = 0;
    int node = 0;
    int resume = valid ? 0 : -2;
    while (node >= 0) {
        const WNode *nd = nodes + node;
        const int first = wave_uniform(nd->first), cnt = wave_uniform(nd->count);
        const int left = wave_uniform(nd->left), esc = wave_uniform(nd->escape);
        bool open = false;
        if (resume == node) {
            if (mac_accept(xi, yi, zi, nd, theta2)) {
                interact<R>(nd->cx, nd->cy, nd->cz, nd->mass, nd->epsmax, xi, yi, zi, epsi, Gr, soft_kind, e2, sx, sy,
                            sz);
                npn++;
                resume = esc;
            } else {
                open = true;
            }
        }
        const bool anyopen = wave_ballot(open) != 0;
        int next = esc;
        if (anyopen && cnt <= bucket) {
            for (int j = first; j < first + cnt; j++) {
                double px = x[j], py = y[j], pz = z[j], pm = m[j], pe = eps ? eps[j] : 0.0;
                if (open && j != i) {
                    interact<R>(px, py, pz, pm, pe, xi, yi, zi, epsi, Gr, soft_kind, e2, sx, sy, sz);
                    npp++;
                }
            }
            if (open) resume = esc;
        } else if (anyopen) {
            if (open) resume = left;
            next = left;
        }
        node = wave_uniform(next);
    }
    if (valid) {
        ax[i] = (double)sx.value();
        ay[i] = (double)sy.value();
        az[i] = (double)sz.value();
    }
    unsigned long long a = npp, b = npn;
    for (int off = 32; off >= 1; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    if ((threadIdx.x % SPH_WAVE) == 0) {
        atomicAdd(&counters[0], a);
        atomicAdd(&counters[1], b);
    }
}

namespace sph {

template <typename R>
hipError_t gravity(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const double *m,
                   const double *eps, const WNode *nodes, int bucket, double theta, double G, int soft_kind,
                   double eps_plummer, double *ax, double *ay, double *az, unsigned long long *counters)
{
    double theta2 = theta * theta;
    k_gravity<R><<<sph_blocks(n, 256), 256, 0, st>>>(n, x, y, z, m, eps, nodes, bucket, theta2, G, soft_kind,
                                                     eps_plummer, ax, ay, az, counters);
    return hipGetLastError();
}

template hipError_t gravity<float>(hipStream_t, int64_t, const double *, const double *, const double *,
                                   const double *, const double *, const WNode *, int, double, double, int, double,
                                   double *, double *, double *, unsigned long long *);
template hipError_t gravity<double>(hipStream_t, int64_t, const double *, const double *, const double *,
                                    const double *, const double *, const WNode *, int, double, double, int, double,
                                    double *, double *, double *, unsigned long long *);

}  // namespace sph
