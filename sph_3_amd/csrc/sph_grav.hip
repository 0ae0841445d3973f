// sph_grav.hip -- Barnes-Hut self-gravity with group interaction lists.
//
// Replaces EV/bh.h:454-588 compute_forces_kernel (one thread per particle,
// warp-shared stack, one opening decision per warp via __all_sync) and the CPU
// direct sums sph_code/libsx.py:623-662 / :908-947 (getAcc_g_smth[_mimj_mpi]).
//
// One wavefront per group of 64 consecutive Morton-ordered targets.  The walk
// is breadth-first and lane-parallel: each step pops up to 64 nodes from an
// LDS stack, every lane tests one node against the group's bounding box
// (FP64 opening test, identical to the CPU oracle's), and the wave appends in
// one ballot-compacted step: accepted nodes to the source buffer, opened
// buckets to a bucket list, children of opened internal nodes to the stack.
// Sources (node monopoles and bucket particles) are staged in LDS relative to
// the group centre; whenever 64 or more are staged all 64 lanes evaluate them
// against their own target (LDS broadcast reads, no divergence), with Kahan
// compensated sums in the pass type R.  Every particle of a group sees the
// same list, so the interaction count is n_nodes*g + n_parts*g - g per group.
#include "sph_grav.h"

#define GG 64        // targets per group (one wavefront)
#define GSTK 4096    // LDS stack capacity (node ids)
#define GSRC 128     // staged sources (< 64 before a flush + 64 appended)

// Group opening test (FP64, no contraction): monopole for the whole group iff
// size^2 < theta^2 d^2 (d: distance from the node's centre of mass to the
// group box) and the node's box is disjoint from the group box.
__device__ __forceinline__ bool mac_group(const double *gb, double cx, double cy, double cz, double lx, double ly,
                                          double lz, double hx, double hy, double hz, double size, double theta2)
{
#pragma clang fp contract(off)
    double ex = 0.0, ey = 0.0, ez = 0.0;
    if (gb[0] - cx > 0.0) ex = gb[0] - cx;
    if (cx - gb[3] > 0.0) ex = cx - gb[3];
    if (gb[1] - cy > 0.0) ey = gb[1] - cy;
    if (cy - gb[4] > 0.0) ey = cy - gb[4];
    if (gb[2] - cz > 0.0) ez = gb[2] - cz;
    if (cz - gb[5] > 0.0) ez = cz - gb[5];
    double d2 = ex * ex + ey * ey + ez * ez;
    bool disjoint = hx < gb[0] || lx > gb[3] || hy < gb[1] || ly > gb[4] || hz < gb[2] || lz > gb[5];
    return (size * size < theta2 * d2) && disjoint;
}

template <typename R, int SOFT>
__device__ __forceinline__ void eval_sources(int cnt, const R *__restrict__ sx, const R *__restrict__ sy,
                                             const R *__restrict__ sz, const R *__restrict__ sm,
                                             const R *__restrict__ se, R tx, R ty, R tz, R epsi, R e2,
                                             Kahan<R> &ax, Kahan<R> &ay, Kahan<R> &az)
{
    for (int k = 0; k < cnt; k++) {
        R dx = sx[k] - tx, dy = sy[k] - ty, dz = sz[k] - tz;
        R r2 = dx * dx + dy * dy + dz * dz;
        R w;
        if (SOFT == 1) {
            R inv = sph_rsqrt<R>(r2 + e2);
            w = sm[k] * (inv * inv * inv);
        } else {
            w = sm[k] * grav_fk_spline<R>(sph_sqrt<R>(r2), (R)0.5 * (epsi + se[k]));
        }
        w = r2 > (R)0 ? w : (R)0;  // the target itself (and exact coincidences) add nothing
        ax.add(w * dx);
        ay.add(w * dy);
        az.add(w * dz);
    }
}

template <typename R, int SOFT>
__global__ void __launch_bounds__(GG) k_gravity_group(int64_t n, const double *__restrict__ x,
                                                      const double *__restrict__ y, const double *__restrict__ z,
                                                      const double *__restrict__ m, const double *__restrict__ eps,
                                                      const WNode *__restrict__ nodes, int bucket, double theta2,
                                                      double G, double eps_plummer, double *__restrict__ ax,
                                                      double *__restrict__ ay, double *__restrict__ az,
                                                      unsigned long long *__restrict__ counters)
{
    __shared__ int stk[GSTK];
    __shared__ R sx[GSRC], sy[GSRC], sz[GSRC], sm[GSRC], se[GSRC];
    __shared__ int bfirst[GG], bcount[GG];
    const int lane = threadIdx.x;
    const int64_t i = blockIdx.x * (int64_t)GG + lane;
    const bool valid = i < n;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    double xi = 0, yi = 0, zi = 0, epsi = 0;
    if (valid) {
        xi = x[i]; yi = y[i]; zi = z[i];
        epsi = eps ? eps[i] : 0.0;
    }
    // group box (FP64 wave reductions)
    double gb[6];
    gb[0] = valid ? xi : INFINITY; gb[1] = valid ? yi : INFINITY; gb[2] = valid ? zi : INFINITY;
    gb[3] = valid ? xi : -INFINITY; gb[4] = valid ? yi : -INFINITY; gb[5] = valid ? zi : -INFINITY;
    for (int off = 32; off >= 1; off >>= 1)
        for (int d = 0; d < 3; d++) {
            gb[d] = fmin(gb[d], __shfl_xor(gb[d], off));
            gb[3 + d] = fmax(gb[3 + d], __shfl_xor(gb[3 + d], off));
        }
    const double gcx = 0.5 * (gb[0] + gb[3]), gcy = 0.5 * (gb[1] + gb[4]), gcz = 0.5 * (gb[2] + gb[5]);
    const R tx = (R)(xi - gcx), ty = (R)(yi - gcy), tz = (R)(zi - gcz), te = (R)epsi;
    const R e2 = (R)(eps_plummer * eps_plummer);
    const int ng = (int)__popcll(wave_ballot(valid));
    Kahan<R> kx, ky, kz;
    long long n_nodes = 0, n_parts = 0;
    int nsrc = 0;
    int sp = 1;
    if (lane == 0) stk[0] = 0;
    __syncthreads();
    while (sp > 0) {
        const int nt = (sp > GSTK - 2 * GG) ? 1 : (sp < GG ? sp : GG);
        const int base = sp - nt;
        const int node = lane < nt ? stk[base + lane] : -1;
        __syncthreads();
        sp = base;
        bool acc = false, bkt = false, opn = false;
        double ncx = 0, ncy = 0, ncz = 0, nmass = 0, neps = 0;
        int nfirst = 0, ncount = 0, nleft = 0, nright = 0;
        if (node >= 0) {
            const WNode *nd = nodes + node;
            ncx = nd->cx; ncy = nd->cy; ncz = nd->cz; nmass = nd->mass; neps = nd->epsmax;
            nfirst = nd->first; ncount = nd->count; nleft = nd->left; nright = nd->right;
            acc = mac_group(gb, ncx, ncy, ncz, nd->lx, nd->ly, nd->lz, nd->hx, nd->hy, nd->hz, nd->size, theta2);
            if (!acc) {
                if (ncount <= bucket) bkt = true;
                else opn = true;
            }
        }
        const uint64_t mo = wave_ballot(opn), ma = wave_ballot(acc), mb = wave_ballot(bkt);
        if (opn) {
            const int po = sp + 2 * (int)__popcll(mo & lt);
            if (po + 1 < GSTK) {
                stk[po] = nright;
                stk[po + 1] = nleft;
            } else {
                atomicAdd(&counters[2], 1ull);  // cannot happen for trees shallower than 2*GG levels
            }
        }
        sp += 2 * (int)__popcll(mo);
        if (sp > GSTK) sp = GSTK;
        if (acc) {
            const int k = nsrc + (int)__popcll(ma & lt);
            sx[k] = (R)(ncx - gcx); sy[k] = (R)(ncy - gcy); sz[k] = (R)(ncz - gcz);
            sm[k] = (R)nmass; se[k] = (R)neps;
        }
        nsrc += (int)__popcll(ma);
        n_nodes += (long long)__popcll(ma);
        if (bkt) {
            const int pb = (int)__popcll(mb & lt);
            bfirst[pb] = nfirst;
            bcount[pb] = ncount;
        }
        const int nb = (int)__popcll(mb);
        __syncthreads();
        if (nsrc >= GG) {
            eval_sources<R, SOFT>(nsrc, sx, sy, sz, sm, se, tx, ty, tz, te, e2, kx, ky, kz);
            nsrc = 0;
            __syncthreads();
        }
        for (int b = 0; b < nb; b++) {
            const int f = bfirst[b], c = bcount[b];
            if (lane < c) {
                const int j = f + lane, k = nsrc + lane;
                sx[k] = (R)(x[j] - gcx); sy[k] = (R)(y[j] - gcy); sz[k] = (R)(z[j] - gcz);
                sm[k] = (R)m[j]; se[k] = eps ? (R)eps[j] : (R)0;
            }
            nsrc += c;
            n_parts += c;
            __syncthreads();
            if (nsrc >= GG) {
                eval_sources<R, SOFT>(nsrc, sx, sy, sz, sm, se, tx, ty, tz, te, e2, kx, ky, kz);
                nsrc = 0;
                __syncthreads();
            }
        }
    }
    if (nsrc > 0) eval_sources<R, SOFT>(nsrc, sx, sy, sz, sm, se, tx, ty, tz, te, e2, kx, ky, kz);
    if (valid) {
        ax[i] = G * (double)kx.value();
        ay[i] = G * (double)ky.value();
        az[i] = G * (double)kz.value();
    }
    if (lane == 0) {
        atomicAdd(&counters[0], (unsigned long long)(n_parts * ng - ng));
        atomicAdd(&counters[1], (unsigned long long)(n_nodes * ng));
    }
}

namespace sph {

template <typename R>
hipError_t gravity(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const double *m,
                   const double *eps, const WNode *nodes, int bucket, double theta, double G, int soft_kind,
                   double eps_plummer, double *ax, double *ay, double *az, unsigned long long *counters)
{
    if (bucket > GG) return hipErrorInvalidValue;  // a bucket is expanded by one lane per particle
    double theta2 = theta * theta;
    unsigned nb = sph_blocks(n, GG);
    if (soft_kind == 0)
        k_gravity_group<R, 0><<<nb, GG, 0, st>>>(n, x, y, z, m, eps, nodes, bucket, theta2, G, eps_plummer, ax, ay,
                                                 az, counters);
    else
        k_gravity_group<R, 1><<<nb, GG, 0, st>>>(n, x, y, z, m, eps, nodes, bucket, theta2, G, eps_plummer, ax, ay,
                                                 az, counters);
    return hipGetLastError();
}

template hipError_t gravity<float>(hipStream_t, int64_t, const double *, const double *, const double *,
                                   const double *, const double *, const WNode *, int, double, double, int, double,
                                   double *, double *, double *, unsigned long long *);
template hipError_t gravity<double>(hipStream_t, int64_t, const double *, const double *, const double *,
                                    const double *, const double *, const WNode *, int, double, double, int, double,
                                    double *, double *, double *, unsigned long long *);

}  // namespace sph
