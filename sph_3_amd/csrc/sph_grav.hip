// sph_grav.hip -- Barnes-Hut self-gravity with group interaction lists.
//
// Replaces EV/bh.h:454-588 compute_forces_kernel (one thread per particle,
// warp-shared stack, one opening decision per warp via __all_sync) and the CPU
// direct sums sph_code/libsx.py:623-662 / :908-947 (getAcc_g_smth[_mimj_mpi]).
//
// Targets are processed in groups: the maximal tree nodes holding at most 64
// particles (a contiguous Morton range with a compact box), one wavefront per
// group.  The walk is breadth-first and lane-parallel: each step pops up to
// 64 nodes from an LDS stack, every lane tests one node against the group's
// box (FP64 opening test, identical to the CPU oracle's), and the wave
// appends in one ballot-compacted step: accepted nodes to the source buffer,
// opened buckets to a bucket list, children of opened internal nodes to the
// stack.  Sources (node monopoles and bucket particles) are staged in LDS
// relative to the group centre; whenever 64 or more are staged all lanes
// evaluate them against their own target (LDS broadcast reads, no
// divergence): plain FMA sums over each flush of <= 128 staged sources, added
// to Kahan-compensated totals, in the pass type R.  In the FP32
// form each relative coordinate is an unevaluated pair hi + lo of floats and
// a pair difference is (s_hi - t_hi) + (s_lo - t_lo), so near pairs keep full
// FP32 accuracy; sources whose node lies farther from the group box than a
// quarter of the group's extent skip the lo parts (their offsets dwarf the
// rounding).  With Plummer or no softening these far sources are staged as
// (x, y, z, |x|^2 + eps^2) and (m x, m y, m z, m): r^2 = |x_j|^2 + eps^2 +
// |x_i|^2 - 2 x_i.x_j is three FMAs and an add, and a_i = sum_j r^-3 (m_j x_j)
// - x_i sum_j r^-3 m_j needs no product by the mass, so a far interaction is
// ten VALU instructions and a reciprocal square root instead of twelve.
// Every particle of a group sees the same list: the counted
// interactions are n_nodes*g + n_parts*g - g per group of g particles.
#include "sph_grav.h"
#include "sph_walk.h"

#define GG 64        // maximum targets per group (one wavefront)
#define GSTK 512     // LDS stack capacity (node ids)
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

// A staged source: hi part of the offset + mass, lo part of the offset + softening.
template <typename R>
struct __align__(4 * sizeof(R)) Src4 {
    R x, y, z, w;
};

template <typename R>
__device__ __forceinline__ void split(double d, R &hi, R &lo)
{
    hi = (R)d;
    lo = (R)(d - (double)hi);
}

// One source against the lane's target.  POT: also the potential sum sp += m_j psi(r, e_ij)
// (Plummer: m_j / sqrt(r^2 + eps^2); spline: grav_pot_spline), coincident pairs excluded.
template <typename R, int SOFT, bool LO, bool POT>
__device__ __forceinline__ void interact(const Src4<R> &a, const Src4<R> &b, R tx, R ty, R tz, R tlx, R tly, R tlz,
                                         R epsi, R e2, R &sx, R &sy, R &sz, R &sp)
{
    constexpr bool SPLIT = LO && sizeof(R) == 4;
    R dx = a.x - tx, dy = a.y - ty, dz = a.z - tz;
    if (SPLIT) {
        dx += b.x - tlx;
        dy += b.y - tly;
        dz += b.z - tlz;
    }
    R r2, w;
    if (SOFT >= 1) {
        r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, e2)));  // r^2 + eps^2 (e2 = 0 for SOFT == 1)
        R inv = sph_rsqrt<R>(r2);
        w = a.w * (inv * inv * inv);
        if (POT) {
            const R d2 = dx * dx + dy * dy + dz * dz;
            sp = fma(a.w, d2 > (R)0 ? inv : (R)0, sp);
        }
    } else {
        r2 = dx * dx + dy * dy + dz * dz;
        const R rr = sph_sqrt<R>(r2), e = (R)0.5 * (epsi + b.w);
        w = a.w * grav_fk_spline<R>(rr, e);
        if (POT) sp = fma(a.w, grav_pot_spline<R>(rr, e), sp);
    }
    // the target itself (and exact coincidences) add nothing: with a positive Plummer
    // softening (SOFT == 2) w stays finite and w * 0 = 0 already
    if (SOFT != 2) w = r2 > (R)0 ? w : (R)0;
    sx = fma(w, dx, sx);
    sy = fma(w, dy, sy);
    sz = fma(w, dz, sz);
}

// Far source in the expanded form (SOFT >= 1): a = (x, y, z, |x|^2 + eps^2), b = m (x, y, z, 1),
// all relative to the group centre; the target enters as (-2 x_i, -2 y_i, -2 z_i) and |x_i|^2.
// With the D/4 separation of far sources the rounding of the expanded r^2 stays below about
// 80 ulp of r^2 (|x_j| <= |x_i| + r, |x_i| <= 0.87 D, r >= D/4).
template <typename R, bool POT>
__device__ __forceinline__ void interact_far(const Src4<R> &a, const Src4<R> &b, R m2x, R m2y, R m2z, R t2, R &fx,
                                             R &fy, R &fz, R &fw, R &sp)
{
    const R r2 = fma(m2x, a.x, fma(m2y, a.y, fma(m2z, a.z, a.w))) + t2;
    const R inv = sph_rsqrt<R>(r2);
    const R w = inv * inv * inv;
    fx = fma(w, b.x, fx);
    fy = fma(w, b.y, fy);
    fz = fma(w, b.z, fz);
    fw = fma(w, b.w, fw);
    if (POT) sp = fma(b.w, inv, sp);
}

// cnt far sources in the expanded form, two chains; each chain's flush partial
// sum r^-3 m (x_j - x_i) = sum r^-3 m x_j - x_i sum r^-3 m goes to the Kahan totals of the
// direct form (k[0..2], k[3..5]), and (POT) k[6]/k[7] += sum m / r per chain.
template <typename R, bool POT>
__device__ __forceinline__ void eval_far(int cnt, const Src4<R> *__restrict__ A, const Src4<R> *__restrict__ B, R tx,
                                         R ty, R tz, Kahan<R> *k)
{
    const R m2x = (R)-2 * tx, m2y = (R)-2 * ty, m2z = (R)-2 * tz, t2 = fma(tx, tx, fma(ty, ty, tz * tz));
    R s[10] = {(R)0, (R)0, (R)0, (R)0, (R)0, (R)0, (R)0, (R)0, (R)0, (R)0};
    int q = 0;
    for (; q + 1 < cnt; q += 2) {
        const Src4<R> a0 = A[q], b0 = B[q], a1 = A[q + 1], b1 = B[q + 1];
        interact_far<R, POT>(a0, b0, m2x, m2y, m2z, t2, s[0], s[1], s[2], s[3], s[8]);
        interact_far<R, POT>(a1, b1, m2x, m2y, m2z, t2, s[4], s[5], s[6], s[7], s[9]);
    }
    if (q < cnt) interact_far<R, POT>(A[q], B[q], m2x, m2y, m2z, t2, s[0], s[1], s[2], s[3], s[8]);
    k[0].add(fma(-tx, s[3], s[0]));
    k[1].add(fma(-ty, s[3], s[1]));
    k[2].add(fma(-tz, s[3], s[2]));
    k[3].add(fma(-tx, s[7], s[4]));
    k[4].add(fma(-ty, s[7], s[5]));
    k[5].add(fma(-tz, s[7], s[6]));
    if (POT) {
        k[6].add(s[8]);
        k[7].add(s[9]);
    }
}

// Evaluate cnt staged sources (LO: with the lo parts of the offsets).  Two
// independent chains of plain sums (ILP); each chain's partial over this flush
// (<= GSRC/2 terms) is then added to its Kahan accumulator, so the list total
// is compensated at the granularity of one flush.
template <typename R, int SOFT, bool LO, bool POT>
__device__ __forceinline__ void eval_run(int cnt, const Src4<R> *__restrict__ A, const Src4<R> *__restrict__ B,
                                         R tx, R ty, R tz, R tlx, R tly, R tlz, R epsi, R e2, Kahan<R> *k)
{
    constexpr int NK = POT ? 8 : 6;
    R s[8] = {(R)0, (R)0, (R)0, (R)0, (R)0, (R)0, (R)0, (R)0};
    int q = 0;
    for (; q + 1 < cnt; q += 2) {
        const Src4<R> a0 = A[q], b0 = B[q], a1 = A[q + 1], b1 = B[q + 1];
        interact<R, SOFT, LO, POT>(a0, b0, tx, ty, tz, tlx, tly, tlz, epsi, e2, s[0], s[1], s[2], s[6]);
        interact<R, SOFT, LO, POT>(a1, b1, tx, ty, tz, tlx, tly, tlz, epsi, e2, s[3], s[4], s[5], s[7]);
    }
    if (q < cnt) interact<R, SOFT, LO, POT>(A[q], B[q], tx, ty, tz, tlx, tly, tlz, epsi, e2, s[0], s[1], s[2], s[6]);
#pragma unroll
    for (int c = 0; c < NK; c++) k[c].add(s[c]);
}

// The staging buffer holds near sources at [0, nn) and far ones at
// [GSRC - nfar, GSRC): a source is near when its node's box lies closer to the
// group box than D/4 (D: the group's extent).  Group-relative FP32 coordinates
// round the source by <= 2^-24 (|d| + D/2) and the target by <= 2^-24 D/2 per
// axis, so a far offset (|d| >= D/4) keeps a relative error <= 5 * 2^-24
// without the lo parts; near ones need them.
template <typename R, int SOFT, bool POT>
__device__ __forceinline__ void eval_sources(int nn, int nfar, const Src4<R> *__restrict__ A,
                                             const Src4<R> *__restrict__ B, R tx, R ty, R tz, R tlx, R tly, R tlz,
                                             R epsi, R e2, Kahan<R> *k)
{
    if (nn > 0) eval_run<R, SOFT, true, POT>(nn, A, B, tx, ty, tz, tlx, tly, tlz, epsi, e2, k);
    if (nfar > 0) {
        if constexpr (SOFT >= 1)
            eval_far<R, POT>(nfar, A + (GSRC - nfar), B + (GSRC - nfar), tx, ty, tz, k);
        else
            eval_run<R, SOFT, false, POT>(nfar, A + (GSRC - nfar), B + (GSRC - nfar), tx, ty, tz, tlx, tly, tlz,
                                          epsi, e2, k);
    }
}

// Stage a source at slot q: near (or spline softening) as hi/lo offsets with the mass and the
// softening, far with Plummer or no softening in the expanded form of eval_far.
template <typename R, int SOFT>
__device__ __forceinline__ void stage_src(Src4<R> *A, Src4<R> *B, int q, bool near, double dx, double dy, double dz,
                                          double m, double eps, R e2)
{
    Src4<R> a, b;
    if (SOFT >= 1 && !near) {
        a.x = (R)dx;
        a.y = (R)dy;
        a.z = (R)dz;
        a.w = fma(a.x, a.x, fma(a.y, a.y, fma(a.z, a.z, e2)));
        b.w = (R)m;
        b.x = a.x * b.w;
        b.y = a.y * b.w;
        b.z = a.z * b.w;
    } else {
        split<R>(dx, a.x, b.x);
        split<R>(dy, a.y, b.y);
        split<R>(dz, a.z, b.z);
        a.w = (R)m;
        b.w = (R)eps;
    }
    A[q] = a;
    B[q] = b;
}

// POT: phi[i] = G sum_j m_j psi(r_ij) over the same interaction list as the force.
template <typename R, int SOFT, bool POT>
__global__ void __launch_bounds__(GG) k_gravity_group(const int32_t *__restrict__ groups, const double *__restrict__ x,
                                                      const double *__restrict__ y, const double *__restrict__ z,
                                                      const double *__restrict__ m, const double *__restrict__ eps,
                                                      const WNode *__restrict__ nodes, int bucket, double theta2,
                                                      double G, double eps_plummer, double *__restrict__ ax,
                                                      double *__restrict__ ay, double *__restrict__ az,
                                                      double *__restrict__ phi,
                                                      unsigned long long *__restrict__ counters)
{
    __shared__ int stk[GSTK];
    __shared__ Src4<R> A[GSRC], B[GSRC];
    __shared__ int bfirst[GG], bstart[GG + 1], bnear[GG];
    const int lane = threadIdx.x;
    const int gnode = groups[xcd_block(blockIdx.x, gridDim.x)];
    const int g0 = nodes[gnode].first, ng = nodes[gnode].count;
    const bool valid = lane < ng;
    const int64_t i = (int64_t)g0 + lane;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    double xi = 0, yi = 0, zi = 0, epsi = 0;
    if (valid) {
        xi = x[i]; yi = y[i]; zi = z[i];
        epsi = eps ? eps[i] : 0.0;
    }
    // the group box is the group node's box
    const WNode *gn = nodes + gnode;
    const double gb[6] = {gn->lx, gn->ly, gn->lz, gn->hx, gn->hy, gn->hz};
    const double gcx = 0.5 * (gb[0] + gb[3]), gcy = 0.5 * (gb[1] + gb[4]), gcz = 0.5 * (gb[2] + gb[5]);
    const double D2 = 0.0625 * gn->size * gn->size;  // near: node box closer than D/4
    R tx, ty, tz, tlx, tly, tlz;
    split<R>(xi - gcx, tx, tlx);
    split<R>(yi - gcy, ty, tly);
    split<R>(zi - gcz, tz, tlz);
    const R te = (R)epsi;
    const R e2 = (R)(eps_plummer * eps_plummer);
    Kahan<R> k[POT ? 8 : 6];
    long long n_nodes = 0, n_parts = 0;
    int nn = 0, nfar = 0;  // staged near sources [0, nn), far ones [GSRC - nfar, GSRC)
    int sp = 1;
    if (lane == 0) stk[0] = 0;
    __syncthreads();
    while (sp > 0) {
        // pop up to 64 nodes; one at a time near capacity (growth then <= 1 per level)
        const int nt = (sp > GSTK - 2 * GG) ? 1 : (sp < GG ? sp : GG);
        const int base = sp - nt;
        const int node = lane < nt ? stk[base + lane] : -1;
        __syncthreads();
        sp = base;
        bool acc = false, bkt = false, opn = false, near = false;
        double ncx = 0, ncy = 0, ncz = 0, nmass = 0, neps = 0;
        int nfirst = 0, ncount = 0, nleft = 0, nright = 0;
        if (node >= 0) {
            const WNode *nd = nodes + node;
            ncx = nd->cx; ncy = nd->cy; ncz = nd->cz; nmass = nd->mass; neps = nd->epsmax;
            nfirst = nd->first; ncount = nd->count; nleft = nd->left; nright = nd->right;
            acc = mac_group(gb, ncx, ncy, ncz, nd->lx, nd->ly, nd->lz, nd->hx, nd->hy, nd->hz, nd->size, theta2);
            near = box_box_dist2(gb, nd) < D2;
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
                spread_add(counters, 2, 1ull);  // cannot happen for trees shallower than 2*GG levels
            }
        }
        sp += 2 * (int)__popcll(mo);
        if (sp > GSTK) sp = GSTK;
        const uint64_t man = wave_ballot(acc && near), maf = ma & ~man;
        if (acc) {
            const int q = near ? nn + (int)__popcll(man & lt) : GSRC - 1 - nfar - (int)__popcll(maf & lt);
            stage_src<R, SOFT>(A, B, q, near, ncx - gcx, ncy - gcy, ncz - gcz, nmass, neps, e2);
        }
        nn += (int)__popcll(man);
        nfar += (int)__popcll(maf);
        n_nodes += (long long)__popcll(ma);
        // opened buckets: prefix sums of their counts (lane order = ballot order), then their
        // particles are staged 64 at a time, all loads of a chunk issued together
        // (the prefix by bit planes of the counts, counts <= bucket: one ballot and one masked
        // population count per bit, no lane exchanges)
        int excl = 0, total = 0;
        for (int b = 0; (1 << b) <= bucket; b++) {
            const uint64_t mbit = wave_ballot(bkt && ((ncount >> b) & 1));
            excl += (int)__popcll(mbit & lt) << b;
            total += (int)__popcll(mbit) << b;
        }
        const int incl = excl + (bkt ? ncount : 0);
        if (bkt) {
            const int pb = (int)__popcll(mb & lt);
            bfirst[pb] = nfirst;
            bstart[pb] = incl - ncount;
            bnear[pb] = near ? 1 : 0;
        }
        const int nb = (int)__popcll(mb);
        if (lane == 0) bstart[nb] = total;
        __syncthreads();
        if (nn + nfar >= GG) {
            eval_sources<R, SOFT, POT>(nn, nfar, A, B, tx, ty, tz, tlx, tly, tlz, te, e2, k);
            nn = nfar = 0;
            __syncthreads();
        }
        n_parts += total;
        for (int s0 = 0; s0 < total; s0 += GG) {
            // nn + nfar < GG here, so a chunk of <= GG fits the GSRC = 2 GG slots
            const int sidx = s0 + lane;
            const bool have = sidx < total;
            int bsel = 0;
            if (have) {
                int lo = 0, hi = nb - 1;  // last bucket b with bstart[b] <= sidx
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (bstart[mid] <= sidx) lo = mid;
                    else hi = mid - 1;
                }
                bsel = lo;
            }
            const bool pn = have && bnear[bsel], pf = have && !bnear[bsel];
            const uint64_t mn = wave_ballot(pn), mf = wave_ballot(pf);
            if (have) {
                const int j = bfirst[bsel] + (sidx - bstart[bsel]);
                const int q = pn ? nn + (int)__popcll(mn & lt) : GSRC - 1 - nfar - (int)__popcll(mf & lt);
                stage_src<R, SOFT>(A, B, q, pn, x[j] - gcx, y[j] - gcy, z[j] - gcz, m[j], eps ? eps[j] : 0.0, e2);
            }
            nn += (int)__popcll(mn);
            nfar += (int)__popcll(mf);
            __syncthreads();
            if (nn + nfar >= GG) {
                eval_sources<R, SOFT, POT>(nn, nfar, A, B, tx, ty, tz, tlx, tly, tlz, te, e2, k);
                nn = nfar = 0;
                __syncthreads();
            }
        }
    }
    if (nn + nfar > 0) eval_sources<R, SOFT, POT>(nn, nfar, A, B, tx, ty, tz, tlx, tly, tlz, te, e2, k);
    if (valid) {
        ax[i] = G * ((double)k[0].value() + (double)k[3].value());
        ay[i] = G * ((double)k[1].value() + (double)k[4].value());
        az[i] = G * ((double)k[2].value() + (double)k[5].value());
        if constexpr (POT) phi[i] = G * ((double)k[6].value() + (double)k[7].value());
    }
    if (lane == 0) {
        spread_add(counters, 0, (unsigned long long)(n_parts * ng - ng));
        spread_add(counters, 1, (unsigned long long)(n_nodes * ng));
    }
}

// Groups: node k with count <= gmax whose parent holds more (or the root).
// Compaction with one global atomic per 1024-thread block (wave offsets by
// ballot, block offsets in LDS): same-address atomics serialise.
__global__ void __launch_bounds__(1024) k_find_groups(int64_t nn, const int32_t *__restrict__ ncount,
                                                      const int32_t *__restrict__ parent, int gmax,
                                                      int32_t *__restrict__ groups, unsigned int *__restrict__ ngroups)
{
    __shared__ unsigned int wcount[16], base;
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bool g = false;
    if (k < nn && ncount[k] <= gmax) g = (k == 0) || ncount[parent[k]] > gmax;
    const uint64_t m = __ballot(g);
    if (lane == 0) wcount[wave] = (unsigned int)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
            const unsigned int c = wcount[w];
            wcount[w] = t;
            t += c;
        }
        base = t ? atomicAdd(ngroups, t) : 0u;
    }
    __syncthreads();
    if (g) {
        const uint64_t below = lane ? (m & (~0ull >> (64 - lane))) : 0ull;
        groups[base + wcount[wave] + (unsigned int)__popcll(below)] = (int32_t)k;
    }
}

namespace sph {

hipError_t gravity_groups(hipStream_t st, int64_t n, const int32_t *ncount, const int32_t *parent, int32_t *groups,
                          unsigned int *ngroups_dev, int64_t *ngroups)
{
    fill<unsigned int>(st, 1, ngroups_dev, 0u);
    k_find_groups<<<sph_blocks(2 * n - 1, 1024), 1024, 0, st>>>(2 * n - 1, ncount, parent, GG, groups, ngroups_dev);
    unsigned int h = 0;
    hipError_t e = hipMemcpyAsync(&h, ngroups_dev, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
    *ngroups = h;
    return e;
}

template <typename R, bool POT>
static void launch_gravity(hipStream_t st, int soft, int64_t ngroups, const int32_t *groups, const double *x,
                           const double *y, const double *z, const double *m, const double *eps, const WNode *nodes,
                           int bucket, double theta2, double G, double eps_plummer, double *ax, double *ay, double *az,
                           double *phi, unsigned long long *counters)
{
    if (soft == 2)
        k_gravity_group<R, 2, POT><<<(unsigned)ngroups, GG, 0, st>>>(groups, x, y, z, m, eps, nodes, bucket, theta2, G,
                                                                    eps_plummer, ax, ay, az, phi, counters);
    else if (soft == 0)
        k_gravity_group<R, 0, POT><<<(unsigned)ngroups, GG, 0, st>>>(groups, x, y, z, m, eps, nodes, bucket, theta2, G,
                                                                    eps_plummer, ax, ay, az, phi, counters);
    else
        k_gravity_group<R, 1, POT><<<(unsigned)ngroups, GG, 0, st>>>(groups, x, y, z, m, eps, nodes, bucket, theta2, G,
                                                                    eps_plummer, ax, ay, az, phi, counters);
}

template <typename R>
hipError_t gravity(hipStream_t st, int64_t ngroups, const int32_t *groups, const double *x, const double *y,
                   const double *z, const double *m, const double *eps, const WNode *nodes, int bucket, double theta,
                   double G, int soft_kind, double eps_plummer, double m_max, double *ax, double *ay, double *az,
                   unsigned long long *counters, double *phi)
{
    if (bucket > GG) return hipErrorInvalidValue;  // a bucket is expanded by one lane per particle
    if (ngroups <= 0) return hipSuccess;
    double theta2 = theta * theta;
    // Plummer softening without the self-pair select: a particle's term on itself is
    // w * 0 with w = m / eps^3, exactly zero while w is finite in the pass type R
    const double wlim = sizeof(R) == 4 ? 1e30 : 1e300;
    const double e3 = eps_plummer * eps_plummer * eps_plummer;
    const bool self_zero = eps_plummer > 0.0 && (R)(eps_plummer * eps_plummer) > (R)0 && m_max < wlim * e3;
    const int soft = (soft_kind == 1 && self_zero) ? 2 : soft_kind == 0 ? 0 : 1;
    if (phi)
        launch_gravity<R, true>(st, soft, ngroups, groups, x, y, z, m, eps, nodes, bucket, theta2, G, eps_plummer, ax,
                                ay, az, phi, counters);
    else
        launch_gravity<R, false>(st, soft, ngroups, groups, x, y, z, m, eps, nodes, bucket, theta2, G, eps_plummer,
                                 ax, ay, az, phi, counters);
    return hipGetLastError();
}

template hipError_t gravity<float>(hipStream_t, int64_t, const int32_t *, const double *, const double *,
                                   const double *, const double *, const double *, const WNode *, int, double, double,
                                   int, double, double, double *, double *, double *, unsigned long long *, double *);
template hipError_t gravity<double>(hipStream_t, int64_t, const int32_t *, const double *, const double *,
                                    const double *, const double *, const double *, const WNode *, int, double,
                                    double, int, double, double, double *, double *, double *, unsigned long long *,
                                    double *);

}  // namespace sph
