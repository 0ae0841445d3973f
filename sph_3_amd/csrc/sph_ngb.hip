// sph_ngb.hip -- tree neighbour search: smoothing-length solves and compact
// CSR neighbour lists.
//
// Replaces the reference's grid search with fixed 200-slot lists
// (EV/ngb.h:125-264 ngbDB_new_v3), its count-based h update
// (EV/lib.h:485-579 smoothing_h_ngb) and the CPU h searches
// (sph_code/libsx.py:752-900 do_smoothingX, h_smooth_fast).
//
// The CSR build and the Newton / k-th-distance h solves use the group walk of
// sph_walk.h: one wavefront per group of <= 64 targets, a conservative
// group-level box test per node (64 nodes tested per step), and the particles
// of every surviving bucket staged in LDS where each lane applies its own
// exact per-target test.  The reference's count stepping (k_hstep) keeps a
// per-lane cooperative stackless walk: one lane per target, the wave follows
// one pre-order path, each lane keeps a single "resume" node id.
//
// FP64 throughout, contraction off: membership r^2 < (h_i+h_j)^2 is decided
// exactly as the CPU oracle decides it.
#include "sph_ngb.h"
#include "sph_walk.h"

#include <algorithm>

#pragma clang fp contract(off)

__device__ __forceinline__ double box_dist2(double x, double y, double z, const WNode *__restrict__ nd)
{
    double ex = 0.0, ey = 0.0, ez = 0.0;
    if (nd->lx - x > 0.0) ex = nd->lx - x;
    if (x - nd->hx > 0.0) ex = x - nd->hx;
    if (nd->ly - y > 0.0) ey = nd->ly - y;
    if (y - nd->hy > 0.0) ey = y - nd->hy;
    if (nd->lz - z > 0.0) ez = nd->lz - z;
    if (z - nd->hz > 0.0) ez = z - nd->hz;
    return ex * ex + ey * ey + ez * ez;
}

// Generic cooperative radius walk.  prune(nd) -> true skips the node for this
// lane; leaf(j) is called for every particle j of an opened bucket.
template <class Prune, class Leaf>
__device__ __forceinline__ void coop_walk(const WNode *__restrict__ nodes, bool valid, int bucket, Prune prune,
                                          Leaf leaf)
{
    int node = 0;
    int resume = valid ? 0 : -2;
    while (node >= 0) {
        const WNode *nd = nodes + node;
        // link fields as wave-uniform scalars (scalar loads, scalar select)
        const int first = wave_uniform(nd->first), cnt = wave_uniform(nd->count);
        const int left = wave_uniform(nd->left), esc = wave_uniform(nd->escape);
        bool open = false;
        if (resume == node) {
            if (prune(nd)) resume = esc;
            else open = true;
        }
        const bool anyopen = wave_ballot(open) != 0;
        int next = esc;
        if (anyopen && cnt <= bucket) {
            for (int j = first; j < first + cnt; j++)
                if (open) leaf(j);
            if (open) resume = esc;
        } else if (anyopen) {
            if (open) resume = left;
            next = left;
        }
        node = wave_uniform(next);
    }
}

// ---------------------------------------------------------------- h solves
// mode 0: Newton-Raphson root of (32/3) sum_j w(r_j/h) - n_ngb (FP64),
// mode 1: the reference's +-0.3 % count stepping (k_hstep below),
// mode 2: h = 0.5 * (k+1)-th smallest distance (libsx.py:752-774, k = 50).
// Candidates (r^2 < R_t^2) of target t go to its row cidx[t*cap + c], c < cap, as source
// indices, with their r^2 beside them in cand[t*cap + c] when cand is given; without it the
// solve and the neighbour rows recompute r^2 from the positions with the walk's own
// operations (contraction off), so they see the walk's bits.  A target whose
// candidate set cannot decide its answer grows or shrinks R_t and stays
// pending (state[t] = 0); the host relaunches until every target is done.
struct HSolveArgs {
    int64_t m;                 // number of targets
    const double *tx, *ty, *tz;  // target positions
    const double *x, *y, *z;   // sorted source positions
    const WNode *nodes;
    int bucket;
    int mode;
    double n_ngb, tol;         // mode 0
    int n_low, n_up, max_steps;  // mode 1
    int kth;                   // mode 2
    int cap;
    int32_t *cidx;             // m rows of cap candidate indices (modes 0 and 2)
    double *cand;              // null, or m rows of cap squared distances beside them
    int32_t *ccnt;             // candidates per target (written with cidx)
    double *R;                 // per target search radius (in/out)
    double *h;                 // per target h (in: start, out: result)
    int32_t *steps;            // mode 1 step counter (persists over rounds)
    int32_t *state;            // 0 pending, 1 done, 2 done-unconverged
    unsigned long long *pending;  // spread counters: 0 pending, 1 unconverged, 2 walk overflows,
                                  // 3 distance tests, 4 Newton kernel terms
    const int32_t *groups;     // tree-node groups of the targets (null: blocks of 64 targets)
    const WNode *gnodes;       // node records the group ids refer to
    int32_t *big_list;         // mode 0 (may be null): targets whose solution needs more
    unsigned long long *big_count;  //   candidates than a row holds (state 3), for k_hsolve_big
    double per;                // side of a periodic box (0: open boundaries; modes 0 and 2)
    // rounds after the first: the groups (glist) and targets (tlist) that still hold pending
    // targets, with their device counts (null in the first round: every group and target)
    const int32_t *glist, *tlist;
    const unsigned long long *gcount, *tcount;
};

// Group range: tree node groups[blk] or the block [64 blk, 64 blk + 64) of targets.
__device__ __forceinline__ void group_range(const int32_t *groups, const WNode *gnodes, int64_t m, int64_t gi,
                                            int64_t &first, int &count)
{
    if (groups) {
        const WNode *g = gnodes + groups[gi];
        first = g->first;
        count = g->count;
    } else {
        first = gi * 64;
        count = (int)((m - first) < 64 ? (m - first) : 64);
    }
}

template <bool PER>
__global__ void __launch_bounds__(64) k_hwalk(HSolveArgs a)
{
    __shared__ WalkLDS L;
    __shared__ double sx[WSRC], sy[WSRC], sz[WSRC];
    __shared__ int sj[WSRC];
    __shared__ double ttx[64], tty[64], ttz[64], tR2[64];  // the group's targets (R^2 < 0: none)
    int64_t gi = xcd_block(blockIdx.x, gridDim.x);
    if (a.glist) {
        if (blockIdx.x >= *a.gcount) return;  // (uniform)
        gi = a.glist[blockIdx.x];
    }
    int64_t first;
    int ng;
    group_range(a.groups, a.gnodes, a.m, gi, first, ng);
    const int lane = threadIdx.x;
    const int64_t t = first + lane;
    const bool valid = lane < ng && a.state[t] == 0;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return;  // whole group decided (uniform)
    double xi = 0, yi = 0, zi = 0, R = 0;
    if (valid) {
        xi = a.tx[t]; yi = a.ty[t]; zi = a.tz[t];
        R = a.R[t];
    }
    double gb[6];
    group_box(valid, xi, yi, zi, gb);
    double Rg = R;
    for (int off = 32; off >= 1; off >>= 1) Rg = fmax(Rg, __shfl_xor(Rg, off));
    const double R2 = R * R, Rg2 = Rg * Rg;
    ttx[lane] = xi; tty[lane] = yi; ttz[lane] = zi;
    tR2[lane] = valid ? R2 : -1.0;
    const int cap = a.cap;
    int32_t *const ibase = a.cidx + first * (int64_t)cap;  // rows of the group's targets
    double *const cbase = a.cand ? a.cand + first * (int64_t)cap : nullptr;
    int c = 0;  // lane t: candidates of target t so far
    long long ntest = 0;  // algorithmic work: distance tests
    // Candidates: lanes take the staged sources, the group's targets in turn (their
    // coordinates and radius are wave-uniform LDS reads); the hits of a target are
    // compacted by a ballot into its row, so a row keeps the staging order.
    const bool ovf = group_walk(
        a.nodes, a.bucket, L, [&](const WNode *nd) { return box_box_dist2_p<PER>(gb, nd, a.per) >= Rg2; },
        [&](int q, int j) {
            sx[q] = a.x[j];
            sy[q] = a.y[j];
            sz[q] = a.z[j];
            sj[q] = j;
        },
        [&](int cnt) {
            for (int k0 = 0; k0 < cnt; k0 += 64) {
                const int k = k0 + lane;
                const bool src = k < cnt;
                const double px = src ? sx[k] : 0.0, py = src ? sy[k] : 0.0, pz = src ? sz[k] : 0.0;
                const int pj = src ? sj[k] : 0;
                const int nsrc = (cnt - k0) < 64 ? (cnt - k0) : 64;
                uint64_t need = vmask;
                if constexpr (!PER) {
                    // skip the targets whose distance to the chunk's bounding box is >= R:
                    // by monotone rounding no source of the chunk can then pass r^2 < R^2
                    double bl[3] = {src ? px : INFINITY, src ? py : INFINITY, src ? pz : INFINITY};
                    double bh[3] = {src ? px : -INFINITY, src ? py : -INFINITY, src ? pz : -INFINITY};
                    for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            bl[d] = fmin(bl[d], __shfl_xor(bl[d], off));
                            bh[d] = fmax(bh[d], __shfl_xor(bh[d], off));
                        }
                    const double tp[3] = {xi, yi, zi};
                    double e2 = 0.0;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        double g = 0.0;
                        if (bl[d] - tp[d] > g) g = bl[d] - tp[d];
                        if (tp[d] - bh[d] > g) g = tp[d] - bh[d];
                        e2 = e2 + g * g;
                    }
                    need = __ballot(valid && e2 < R2);
                }
                if (lane < 64 && ((need >> lane) & 1)) ntest += nsrc;  // distance tests performed
                // targets in turn; the next target's coordinates are read from LDS while the
                // current one is tested (two register sets used alternately)
                auto test = [&](int tt, double qx, double qy, double qz, double qr) {
                    const double dx = wrap_d<PER>(px - qx, a.per), dy = wrap_d<PER>(py - qy, a.per),
                                 dz = wrap_d<PER>(pz - qz, a.per);
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const bool hit = src && r2 < qr;
                    const uint64_t hm = __ballot(hit);
                    if (hm) {
                        const int ct = __builtin_amdgcn_readlane(c, tt);
                        if (hit) {
                            const uint32_t pos =
                                (uint32_t)ct + __builtin_amdgcn_mbcnt_hi((uint32_t)(hm >> 32),
                                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)hm, 0u));
                            if (pos < (uint32_t)cap) {
                                // the group's uniform row base + a 32-bit offset (scalar-base addressing)
                                const uint32_t e = (uint32_t)(tt * cap) + pos;
                                *(int32_t *)((char *)ibase + e * 4u) = pj;
                                if (cbase) *(double *)((char *)cbase + e * 8u) = r2;
                            }
                        }
                        c = (lane == tt) ? ct + (int)__popcll(hm) : c;
                    }
                };
                uint64_t todo = need;
                if (todo) {
                    int ta = __builtin_ctzll(todo);
                    double ax = ttx[ta], ay = tty[ta], az = ttz[ta], ar = tR2[ta];
                    for (;;) {
                        todo &= todo - 1;
                        const int tb = todo ? __builtin_ctzll(todo) : ta;
                        const double bx = ttx[tb], by = tty[tb], bz = ttz[tb], br = tR2[tb];
                        test(ta, ax, ay, az, ar);
                        if (!todo) break;
                        todo &= todo - 1;
                        ta = todo ? __builtin_ctzll(todo) : tb;
                        ax = ttx[ta]; ay = tty[ta]; az = ttz[ta]; ar = tR2[ta];
                        test(tb, bx, by, bz, br);
                        if (!todo) break;
                    }
                }
            }
        });
    if (ovf && lane == 0) spread_add(a.pending, 2, 1ull);
    if (valid) a.ccnt[t] = c;  // > cap: the row overflowed (k_hnewton shrinks R)
    for (int off = 32; off >= 1; off >>= 1) ntest += __shfl_xor(ntest, off);
    if (lane == 0) spread_add(a.pending, 3, (unsigned long long)ntest);
}

// Newton (mode 0) or k-th distance (mode 2) solve of the pending targets from their
// candidate rows, 8 targets per wavefront and 8 lanes per target: lane l of a sub-group
// holds the squared distances of candidates l, l+8, ... (up to 12 in registers; R2: read
// from the distance rows, else recomputed from the positions as the walk computed them;
// longer rows are read again in every evaluation) and the sub-group reduces with DPP lane
// exchanges, which give
// its 8 lanes bitwise the same sums, so each sub-group's iteration is uniform.  A target
// whose row overflowed gets R * 0.75, one whose solution lies beyond R/2 a larger R estimated from
// its weighted count at R/2, and stays pending for another round.
#define HN_LANES 8  // lanes per target of k_hnewton

// One DPP lane move of a double (two 32-bit moves); every source lane of the patterns used
// here lies in the reading lane's own row of 16.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over each aligned group of W = 4, 8 or 16 lanes with the same bits in every lane of
// the group: each step adds a partner lane's value and partners exchange (IEEE addition
// commutes): xor 1 and xor 2 (quad permutes), then (W >= 8) the mirror within 8 lanes and
// (W = 16) the mirror within 16 lanes.
template <int W>
__device__ __forceinline__ double lane_group_sum(double v)
{
    static_assert(W == 4 || W == 8 || W == 16, "groups of 4, 8 or 16 lanes");
    v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    if constexpr (W >= 8) v += dpp_f64<0x141>(v);  // row_half_mirror
    if constexpr (W == 16) v += dpp_f64<0x140>(v);  // row_mirror
    return v;
}

// sqrt(x) of a squared distance x > 0 for the Newton sums: the hardware reciprocal square
// root refined by two Newton-Raphson corrections (error about 1 ulp; the solve needs relative
// 1e-13), without the scaling and special-value handling of the correctly rounded library
// square root.  Valid for x >= 0x1p-900 (callers take the library path for a wave holding a
// smaller nonzero value); x = 0 (the target itself) gives 0.
__device__ __forceinline__ double sqrt_dist2(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y, hh = 0.5 * y;
    const double e = fma(-hh, s, 0.5);
    s = fma(s, e, s);
    hh = fma(hh, e, hh);
    double d = fma(-s, s, x);
    s = fma(d, hh, s);
    d = fma(-s, s, x);
    s = fma(d, hh, s);
    return x > 0.0 ? s : 0.0;
}

// 1/x for x > 0 in the normal range: the hardware reciprocal refined by two Newton-Raphson
// corrections (within 1 ulp), instead of the correctly rounded division sequence.
__device__ __forceinline__ double rcp_refined(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

#define HN_SLOTS 12  // candidates per lane held in registers (HN_LANES * HN_SLOTS per target)

template <bool PER, bool R2>
__global__ void __launch_bounds__(64) k_hnewton(HSolveArgs a)
{
    constexpr int HL = HN_LANES, HS = HN_SLOTS, HR = HL * HS;
    const int lane = threadIdx.x, lsub = lane % HL;
    int64_t j = (int64_t)blockIdx.x * (64 / HL) + lane / HL;
    bool inr = j < a.m;
    if (a.tlist) {
        inr = (unsigned long long)j < *a.tcount;
        j = inr ? a.tlist[j] : 0;
    }
    const bool valid = inr && a.state[j] == 0;
    if (__ballot(valid) == 0) return;
    const int cap = a.cap;
    const int cjs = valid ? a.ccnt[j] : 0;
    const bool over = valid && cjs > cap;
    const bool solve = valid && !over;
    const int cj = solve ? cjs : 0;
    const double Rj = valid ? a.R[j] : 0.0, hj0 = valid ? a.h[j] : 0.0;
    const int32_t *rowj = a.cidx + (solve ? j : 0) * (int64_t)cap;
    const double *rowr = R2 ? a.cand + (solve ? j : 0) * (int64_t)cap : nullptr;
    const double qx = (!R2 && solve) ? a.tx[j] : 0.0, qy = (!R2 && solve) ? a.ty[j] : 0.0,
                 qz = (!R2 && solve) ? a.tz[j] : 0.0;
    // r^2 of candidate k: from the row, or with the walk's test operands and order (source
    // minus target)
    auto dist2 = [&](int k) {
        if constexpr (R2) return rowr[k];
        const int s = rowj[k];
        const double dx = wrap_d<PER>(a.x[s] - qx, a.per), dy = wrap_d<PER>(a.y[s] - qy, a.per),
                     dz = wrap_d<PER>(a.z[s] - qz, a.per);
        return dx * dx + dy * dy + dz * dz;
    };
    // the most candidates of a target in the wave (uniform): register slots in use, and
    // whether some row continues past them (then read from the row in every evaluation)
    int cw = cj;
    for (int off = 32; off >= 1; off >>= 1) cw = max(cw, __shfl_xor(cw, off));
    cw = __builtin_amdgcn_readfirstlane(cw);
    const int nsl = min((cw + HL - 1) / HL, HS);
    const bool tail = cw > HR;
    long long nterm = 0;
    // slot u of lane lsub holds candidate lsub + HL u (static register indices); a slot past
    // the target's candidates holds INFINITY, later 4 R, whose kernel terms are exactly zero
    // for every h <= R/2
    double r[HS];
#pragma unroll
    for (int u = 0; u < HS; u++) {
        const int k = lsub + HL * u;
        r[u] = (u < nsl && k < cj) ? dist2(k) : INFINITY;
    }
    double hnew = hj0, Rnew = Rj;
    int st = 0;
    if (solve && a.mode == 0) {
        // distances; empty slots get 4 R, beyond the support of every h <= R/2
        bool tiny = false;
#pragma unroll
        for (int u = 0; u < HS; u++)
            if (u < nsl) tiny = tiny || (r[u] > 0.0 && r[u] < 0x1p-900);
        if (__ballot(tiny)) {
#pragma unroll
            for (int u = 0; u < HS; u++)
                if (u < nsl) r[u] = (r[u] < INFINITY) ? __builtin_sqrt(r[u]) : 4.0 * Rj;
        } else {
#pragma unroll
            for (int u = 0; u < HS; u++)
                if (u < nsl) r[u] = (r[u] < INFINITY) ? sqrt_dist2(r[u]) : 4.0 * Rj;
        }
        // f(h) = (32/3) sum_j w(r_j/h) - n_ngb and f'(h), with the branch-free
        // cubic spline w(q) = (2-q)_+^3/4 - (1-q)_+^3 and q w'(q) = q (3 (1-q)_+^2 - 3/4 (2-q)_+^2)
        auto feval = [&](double hh, double &f, double &df) {
            if (lsub == 0) nterm += cj;
            double sl = 0.0, sq = 0.0;
            const double ih = rcp_refined(hh);
            auto term = [&](double rr) {
                const double q = rr * ih;
                const double aa = fmax(2.0 - q, 0.0), bb = fmax(1.0 - q, 0.0);
                const double a2 = aa * aa, b2 = bb * bb;
                sl = fma(0.25 * a2, aa, fma(-b2, bb, sl));
                sq = fma(fma(3.0, b2, -0.75 * a2), q, sq);
            };
#pragma unroll
            for (int u = 0; u < HS; u++)
                if (u < nsl) term(r[u]);
            if (tail)
                for (int k = HR + lsub; k < cj; k += HL) term(__builtin_sqrt(dist2(k)));
            sl = lane_group_sum<HL>(sl);
            sq = lane_group_sum<HL>(sq);
            f = (32.0 / 3.0) * sl - a.n_ngb;
            df = -(32.0 / 3.0) * sq * ih;
        };
        // Safeguarded Newton from the previous h on the bracket (0, R/2]: the
        // candidates (r < R) are complete for every h <= R/2.  f(R/2) is only
        // evaluated when no iterate has shown f >= 0 and Newton leaves the
        // bracket upwards; f(R/2) < 0 means the root lies beyond: grow R.
        // Stop after a Newton step dh <= sqrt(tol/10) h: in the quadratic
        // regime the error after it is K dh^2 <= tol h for curvature
        // K = f''/(2 f') <= 10/h (K = 1/h for a uniform density), or after
        // any step dh <= tol h.
        const double hmaxv = 0.5 * Rj, qtol = __builtin_sqrt(0.1 * a.tol);
        double lo = 0.0, hi = hmaxv, hh = hj0, f, df;
        if (!(hh > 0.0 && hh < hi)) hh = 0.5 * hi;
        bool hik = false, grow = false;
        double fgrow = 0.0;  // f(R/2) when the root lies beyond R/2
        st = 2;
        for (int it = 0; it < 100; it++) {
            feval(hh, f, df);
            if (f < 0.0) lo = hh;
            else {
                hi = hh;
                hik = true;
            }
            double hn = (df > 0.0) ? hh - f * rcp_refined(df) : -1.0;
            bool newton = true;
            if (!(hn > lo && hn < hi)) {
                if (!hik) {
                    double fm, dfm;
                    feval(hmaxv, fm, dfm);
                    if (fm < 0.0) {
                        grow = true;
                        fgrow = fm;
                        break;
                    }
                    hik = true;
                    hn = (dfm > 0.0) ? hmaxv - fm * rcp_refined(dfm) : -1.0;
                    if (!(hn > lo && hn < hi)) hn = 0.5 * (lo + hi);
                } else {
                    hn = 0.5 * (lo + hi);
                }
                newton = false;
            }
            const double dh = fabs(hn - hh);
            hh = hn;
            if (dh <= a.tol * hh || (newton && dh <= qtol * hh)) {
                st = 1;
                break;
            }
        }
        if (grow) {
            // R for the next round from the weighted count n(R/2) = f(R/2) + N_ngb: at a locally
            // uniform density n grows as h^3, so the root lies near (R/2) (N_ngb / n)^(1/3); take
            // 1.1 times that, at least 1.25 R and at most 4 R (a target that still falls short
            // grows again in the next round)
            const double neff = fmax(fgrow + a.n_ngb, a.n_ngb / 64.0);
            Rnew = Rj * fmin(4.0, fmax(1.25, 1.1 * cbrt(a.n_ngb / neff)));
            st = 0;
        } else {
            hnew = hh;
        }
    } else if (solve) {
        // (kth+1)-th smallest squared distance: ascending distinct values with multiplicity
        const int need = a.kth + 1;
        if (cj < need) {
            Rnew = 1.5 * Rj;
        } else {
            double cur = -1.0, v = 0.0;
            int taken = 0;
            for (;;) {
                double mn = INFINITY;
#pragma unroll
                for (int u = 0; u < HS; u++)
                    if (u < nsl && r[u] > cur && r[u] < mn) mn = r[u];
                if (tail)
                    for (int k = HR + lsub; k < cj; k += HL) {
                        const double x = dist2(k);
                        if (x > cur && x < mn) mn = x;
                    }
                for (int off = HL / 2; off >= 1; off >>= 1) mn = fmin(mn, __shfl_xor(mn, off, HL));
                int mult = 0;
#pragma unroll
                for (int u = 0; u < HS; u++)
                    if (u < nsl) mult += (r[u] == mn);
                if (tail)
                    for (int k = HR + lsub; k < cj; k += HL) mult += (dist2(k) == mn);
                for (int off = HL / 2; off >= 1; off >>= 1) mult += __shfl_xor(mult, off, HL);
                if (taken + mult >= need) { v = mn; break; }
                taken += mult;
                cur = mn;
            }
            hnew = __builtin_sqrt(v) * 0.5;
            st = 1;
        }
    }
    if (over) Rnew = 0.75 * Rj;
    for (int off = 32; off >= 1; off >>= 1) nterm += __shfl_xor(nterm, off);
    if (lane == 0) spread_add(a.pending, 4, (unsigned long long)nterm);
    if (!valid || lsub != 0) return;
    if (a.mode == 0 && a.big_list) {
        // a target that must grow R after its candidates overflowed a row needs more
        // candidates than a row holds at its solution: k_hsolve_big solves it
        const int fl = a.steps[j];
        if (over) {
            a.steps[j] = fl | 1;
        } else if (st == 0 && (fl & 1)) {
            st = 3;
            a.big_list[atomicAdd(a.big_count, 1ull)] = (int32_t)j;
        }
    }
    a.h[j] = hnew;
    a.R[j] = Rnew;
    a.state[j] = st;
    if (st == 0) spread_add(a.pending, 0, 1ull);
    if (st == 2) spread_add(a.pending, 1, 1ull);
}

// Newton solve of targets whose solution needs more candidates than a row holds (a
// target beside a much denser structure): one wavefront per target; each evaluation
// of f(h) and f'(h) walks the tree over all sources within 2h (terms beyond 2h are
// zero), the 64 lanes splitting the staged sources, FP64 sums reduced by butterflies
// (uniform in the wave).  f is non-decreasing in h with f(0+) < 0, so Newton is
// safeguarded by the bracket [lo, hi] (hi doubles until f(hi) >= 0); the stopping
// rule is the row solver's.  The target's candidate count is set to cap + 1: its row
// of candidates is incomplete (neighbour rows built from candidates must not use it).
template <bool PER>
__global__ void __launch_bounds__(64) k_hsolve_big(HSolveArgs a, const int32_t *__restrict__ list)
{
    __shared__ WalkLDS L;
    __shared__ double sx[WSRC], sy[WSRC], sz[WSRC];
    const int64_t t = list[blockIdx.x];
    const int lane = threadIdx.x;
    const double xi = a.tx[t], yi = a.ty[t], zi = a.tz[t];
    const double gb[6] = {xi, yi, zi, xi, yi, zi};
    long long ntest = 0, nterm = 0;
    auto feval = [&](double hh, double &f, double &df) {
        const double lim = 2.0 * hh, L2 = lim * lim, ih = 1.0 / hh;
        double sl = 0.0, sq = 0.0;
        (void)group_walk(
            a.nodes, a.bucket, L, [&](const WNode *nd) { return box_box_dist2_p<PER>(gb, nd, a.per) >= L2; },
            [&](int q, int j) {
                sx[q] = a.x[j];
                sy[q] = a.y[j];
                sz[q] = a.z[j];
            },
            [&](int cnt) {
                if (lane == 0) ntest += cnt;
                for (int k = lane; k < cnt; k += 64) {
                    const double dx = wrap_d<PER>(sx[k] - xi, a.per), dy = wrap_d<PER>(sy[k] - yi, a.per),
                                 dz = wrap_d<PER>(sz[k] - zi, a.per);
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 < L2) {
                        nterm++;
                        const double q = __builtin_sqrt(r2) * ih;
                        const double aa = fmax(2.0 - q, 0.0), bb = fmax(1.0 - q, 0.0);
                        const double a2 = aa * aa, b2 = bb * bb;
                        sl = fma(0.25 * a2, aa, fma(-b2, bb, sl));
                        sq = fma(fma(3.0, b2, -0.75 * a2), q, sq);
                    }
                }
            });
        for (int off = 32; off >= 1; off >>= 1) {
            sl += __shfl_xor(sl, off);
            sq += __shfl_xor(sq, off);
        }
        f = (32.0 / 3.0) * sl - a.n_ngb;
        df = -(32.0 / 3.0) * sq * ih;
    };
    const double qtol = __builtin_sqrt(0.1 * a.tol);
    double lo = 0.0, hi = INFINITY, hh = a.h[t], f, df;
    if (!(hh > 0.0)) hh = 0.5 * a.R[t];
    int st = 2;
    for (int it = 0; it < 200; it++) {
        feval(hh, f, df);
        if (f < 0.0) lo = hh;
        else hi = hh;
        double hn = (df > 0.0) ? hh - f / df : -1.0;
        bool newton = true;
        if (!(hn > lo && hn < hi)) {
            hn = (hi < INFINITY) ? 0.5 * (lo + hi) : 2.0 * hh;
            newton = false;
        }
        const double dh = fabs(hn - hh);
        hh = hn;
        if (dh <= a.tol * hh || (newton && dh <= qtol * hh)) {
            st = 1;
            break;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) nterm += __shfl_xor(nterm, off);
    if (lane == 0) {
        spread_add(a.pending, 3, (unsigned long long)ntest);
        spread_add(a.pending, 4, (unsigned long long)nterm);
        a.h[t] = hh;
        a.R[t] = 2.0 * hh;
        a.state[t] = st;
        if (a.ccnt) a.ccnt[t] = a.cap + 1;
    }
}

// mode 1: the reference's count stepping, exactly (libsx.py:804-845): each
// step counts #(dist < 2h) over ALL sources with a cooperative count walk, so
// no candidate radius is involved.  dist = sqrt(r^2) correctly rounded, as in
// the reference; a node is skipped only when its box distance^2 exceeds
// (2h)^2, which cannot drop a particle with sqrt(r^2) < 2h.
__global__ void __launch_bounds__(256) k_hstep(HSolveArgs a)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool valid = t < a.m;
    double xi = 0, yi = 0, zi = 0, hh = 0;
    if (valid) { xi = a.tx[t]; yi = a.ty[t]; zi = a.tz[t]; hh = a.h[t]; }
    int steps = 0, st = valid ? 0 : 1;
    for (;;) {
        bool need = st == 0;
        if (wave_ballot(need) == 0) break;
        const double lim = 2.0 * hh, L2 = lim * lim;
        int n = 0;
        coop_walk(
            a.nodes, need, a.bucket, [&](const WNode *nd) { return box_dist2(xi, yi, zi, nd) > L2; },
            [&](int j) {
                double dx = a.x[j] - xi, dy = a.y[j] - yi, dz = a.z[j] - zi;
                n += __builtin_sqrt(dx * dx + dy * dy + dz * dz) < lim;
            });
        if (need) {
            if (n <= a.n_up && n >= a.n_low) st = 1;
            else if (steps >= a.max_steps) st = 2;
            else {
                if (n > a.n_up) hh -= 0.003 * hh;
                if (n < a.n_low) hh += 0.003 * hh;
                steps++;
            }
        }
    }
    if (!valid) return;
    a.h[t] = hh;
    a.steps[t] = steps;
    a.state[t] = st;
    if (st == 2) spread_add(a.pending, 1, 1ull);
}

__global__ void __launch_bounds__(256) k_hsolve_init(int64_t m, const double *__restrict__ h0, double fac,
                                                     double *__restrict__ R, double *__restrict__ h,
                                                     int32_t *__restrict__ steps, int32_t *__restrict__ state)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= m) return;
    double hv = h0[t];
    h[t] = hv;
    R[t] = 2.0 * hv * fac;
    steps[t] = 0;
    state[t] = 0;
}

// ---------------------------------------------------------------- CSR lists
// S_i = { j : r_ij^2 < (h_i + h_j)^2 }, self included.  A node is skipped for
// the whole group iff boxdist(group box, node box)^2 >= (hmax_group +
// hmax_node)^2, which by monotone rounding cannot drop a member.  Row order is
// the (deterministic) order in which the walk stages buckets.
// MODE 0: counts only; 1: fill CSR rows at off[i]; 2: fill padded rows
// idx[i*stride + c] (c < stride) and counts in one walk, with spread counters
// 0 += rows longer than stride, 1 += entries.
template <int MODE, bool PER>
__global__ void __launch_bounds__(64) k_ngb(const int32_t *__restrict__ groups, const double *__restrict__ x,
                                            const double *__restrict__ y, const double *__restrict__ z,
                                            const double *__restrict__ h, const WNode *__restrict__ nodes, int bucket,
                                            int32_t *__restrict__ cnt, const int64_t *__restrict__ off,
                                            int32_t *__restrict__ idx, int stride,
                                            unsigned long long *__restrict__ overflow,
                                            unsigned long long *__restrict__ spread,
                                            const double *__restrict__ hscale, double per)
{
    __shared__ WalkLDS L;
    __shared__ double sx[WSRC], sy[WSRC], sz[WSRC], sh[WSRC];
    __shared__ int sj[WSRC];
    const WNode *g = nodes + groups[xcd_block(blockIdx.x, gridDim.x)];
    const int64_t first = g->first;
    const int ng = g->count;
    const int lane = threadIdx.x;
    const int64_t i = first + lane;
    const bool valid = lane < ng;
    double xi = 0, yi = 0, zi = 0, hi = 0;
    if (valid) { xi = x[i]; yi = y[i]; zi = z[i]; hi = h[i]; }
    const double gb[6] = {g->lx, g->ly, g->lz, g->hx, g->hy, g->hz};
    // node h_max from an earlier h: *hscale >= max_i h_i / h_old_i keeps the prune conservative
    const double hs = hscale ? *hscale : 1.0;
    const double hg = g->hmax * hs;
    const int64_t o = !valid ? 0 : MODE == 1 ? off[i] : MODE == 2 ? i * (int64_t)stride : 0;
    int c = 0;
    const bool ovf = group_walk(
        nodes, bucket, L,
        [&](const WNode *nd) {
            double t = hg + nd->hmax * hs;
            return box_box_dist2_p<PER>(gb, nd, per) >= t * t;
        },
        [&](int q, int j) {
            sx[q] = x[j];
            sy[q] = y[j];
            sz[q] = z[j];
            sh[q] = h[j];
            sj[q] = j;
        },
        [&](int n) {
            if (valid)
                for (int k = 0; k < n; k++) {
                    double dx = wrap_d<PER>(xi - sx[k], per), dy = wrap_d<PER>(yi - sy[k], per), dz = wrap_d<PER>(zi - sz[k], per);
                    double r2 = dx * dx + dy * dy + dz * dz, t = hi + sh[k];
                    if (r2 < t * t) {
                        if (MODE == 1 || (MODE == 2 && c < stride)) idx[o + c] = sj[k];
                        c++;
                    }
                }
        });
    if (ovf && lane == 0) atomicAdd(overflow, 1ull);
    if (valid && MODE != 1) cnt[i] = c;
    if (MODE == 2) {
        unsigned long long nlong = (valid && c > stride) ? 1ull : 0ull, nent = valid ? (unsigned long long)c : 0ull;
        for (int sh2 = 32; sh2 >= 1; sh2 >>= 1) {
            nlong += __shfl_xor(nlong, sh2);
            nent += __shfl_xor(nent, sh2);
        }
        if (lane == 0) {
            if (nlong) spread_add(spread, 0, nlong);
            spread_add(spread, 1, nent);
        }
    }
}

// ---------------------------------------------------------------- rows from h-solve candidates
// After the h solve every target i holds its final candidate row
// C_i = { j : r_ij^2 < R_i^2 } (FP64 r^2 with the association of the neighbour
// test, recomputed here from x_j - x_i as the walk computed it; a - b = -(b - a)
// exactly) and 2 h_i <= R_i exactly (the Newton bracket stops at R_i/2).  Hence for S_i = { j : r_ij^2 < (h_i+h_j)^2 }:
//   h_j <= h_i:  (h_i+h_j)^2 <= (2 h_i)^2 <= R_i^2 by monotone rounding, so j in C_i;
//   h_j >  h_i:  likewise i in C_j.
// k_cand_rows keeps { j in C_i : r^2 < (h_i+h_j)^2 } in candidate order and
// records every pair with i in C_j, h_j > h_i, j not in C_i (r^2 >= R_i^2) as an
// extra (target i, member j) from j's row; k_ext_append appends the extras and
// k_row_tail_sort puts each row's appended part in ascending order, so a row's
// order (and every sum over it) does not depend on the order of the appends.
// Spread counters: 0 += rows longer than stride, 1 += entries.
// hr[j] = (h_j, R_j) (with distance rows) or rec[j] = (x_j, y_j, z_j, h_j, R_j) (without),
// R_j replaced by -1 for a row of k_hsolve_big's (incomplete: such a j's pairs are found by
// its own walk, k_big_rows), so one gather gives k_cand_rows both tests (and the distance).
__global__ void __launch_bounds__(256) k_pack_hr(int64_t n, int cap, const double *__restrict__ h,
                                                 const double *__restrict__ R, const int32_t *__restrict__ ccnt,
                                                 double2 *__restrict__ hr)
{
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j < n) hr[j] = make_double2(h[j], ccnt[j] <= cap ? R[j] : -1.0);
}

__global__ void __launch_bounds__(256) k_pack_rec(int64_t n, int cap, const double *__restrict__ x,
                                                  const double *__restrict__ y, const double *__restrict__ z,
                                                  const double *__restrict__ h, const double *__restrict__ R,
                                                  const int32_t *__restrict__ ccnt, sph::CandRec *__restrict__ rec)
{
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j < n) {
        sph::CandRec c;
        c.x = x[j]; c.y = y[j]; c.z = z[j]; c.h = h[j];
        c.R = ccnt[j] <= cap ? R[j] : -1.0;
        c.pad = 0.0;
        rec[j] = c;
    }
}

#define CR_UNROLL 5  // candidates per lane step of k_cand_rows
// R2: distances from the rows (cand) and (h_j, R_j) from hr; else recomputed from rec.
template <int G, bool PER, bool R2>
__global__ void __launch_bounds__(256) k_cand_rows(int64_t n, int cap, const double *__restrict__ cand,
                                                   const int32_t *__restrict__ cidx, const int32_t *__restrict__ ccnt,
                                                   const double2 *__restrict__ hr,
                                                   const sph::CandRec *__restrict__ rec, double per, int stride,
                                                   int32_t *__restrict__ cnt, int32_t *__restrict__ cnt0,
                                                   int32_t *__restrict__ idx, int2 *__restrict__ ext, int64_t ext_cap,
                                                   unsigned long long *__restrict__ next,
                                                   unsigned long long *__restrict__ spread)
{
    const int64_t gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t i = gid / G;
    const int l = (int)(gid % G), lane = threadIdx.x & 63;
    const bool valid = i < n;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint64_t gmask = (G == 64 ? ~0ull : ((1ull << G) - 1)) << (lane & ~(G - 1));
    double hi = 0.0, xi = 0.0, yi = 0.0, zi = 0.0;
    if constexpr (R2) {
        hi = hr[valid ? i : 0].x;
    } else {
        const sph::CandRec own = rec[valid ? i : 0];
        hi = own.h; xi = own.x; yi = own.y; zi = own.z;
    }
    const int craw = valid ? ccnt[i] : 0;
    const bool incomplete = craw > cap;  // solved without a complete candidate row (k_hsolve_big)
    const int c = incomplete ? 0 : craw;
    const int32_t *rj = cidx + (valid ? i : 0) * (int64_t)cap;
    const double *rr = R2 ? cand + (valid ? i : 0) * (int64_t)cap : nullptr;
    int32_t *row = idx + (valid ? i : 0) * (int64_t)stride;
    int w = 0;
    // one candidate of the row: member test, ballot-compacted append, extra pair
    auto visit = [&](bool have, int j, double r2, double hj, double Rj) {
        bool in = false, extra = false;
        if (have) {
            const double t = hi + hj, t2 = t * t;
            in = r2 < t2;
            // (Rj < 0: a row of k_hsolve_big's, whose pairs k_big_rows finds by a walk)
            extra = in && hi > hj && Rj > 0.0 && r2 >= Rj * Rj;
        }
        const uint64_t mb = __ballot(in) & gmask;
        if (in) {
            const int pos = w + (int)__popcll(mb & below);
            if (pos < stride) row[pos] = j;
        }
        w += (int)__popcll(mb);
        if (extra) {  // region of this workgroup (SPH_SPREAD regions of ext_cap / SPH_SPREAD pairs)
            const int64_t rcap = ext_cap / SPH_SPREAD, reg = blockIdx.x & (SPH_SPREAD - 1);
            const unsigned long long e = atomicAdd(next + reg * SPH_SPREAD_PAD, 1ull);
            if (e < (unsigned long long)rcap) ext[reg * rcap + (int64_t)e] = make_int2(j, (int)i);
        }
    };
    // candidates k0 + l + u G (u < CR_UNROLL) per step: all their indices (and distances),
    // then all their records, are loaded before any is visited (CR_UNROLL gathers in flight
    // per lane), visits in row order; without distance rows r^2 is recomputed with the walk's
    // operands and order (x_j - x_i)
    if constexpr (R2) {
        for (int k0 = 0; k0 < c; k0 += CR_UNROLL * G) {
            int jv[CR_UNROLL];
            double rv[CR_UNROLL];
            double2 hv[CR_UNROLL];
#pragma unroll
            for (int u = 0; u < CR_UNROLL; u++) {
                const int k = k0 + l + u * G;
                jv[u] = k < c ? rj[k] : 0;
                rv[u] = k < c ? rr[k] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < CR_UNROLL; u++)
                hv[u] = (k0 + l + u * G < c) ? hr[jv[u]] : make_double2(0.0, 0.0);
#pragma unroll
            for (int u = 0; u < CR_UNROLL; u++) visit(k0 + l + u * G < c, jv[u], rv[u], hv[u].x, hv[u].y);
        }
    } else {
        for (int k0 = 0; k0 < c; k0 += CR_UNROLL * G) {
            int jv[CR_UNROLL];
            sph::CandRec cv[CR_UNROLL];
#pragma unroll
            for (int u = 0; u < CR_UNROLL; u++) {
                const int k = k0 + l + u * G;
                jv[u] = k < c ? rj[k] : 0;
            }
#pragma unroll
            for (int u = 0; u < CR_UNROLL; u++) cv[u] = rec[jv[u]];
#pragma unroll
            for (int u = 0; u < CR_UNROLL; u++) {
                const double dx = wrap_d<PER>(cv[u].x - xi, per), dy = wrap_d<PER>(cv[u].y - yi, per),
                             dz = wrap_d<PER>(cv[u].z - zi, per);
                const double r2 = dx * dx + dy * dy + dz * dz;
                visit(k0 + l + u * G < c, jv[u], r2, cv[u].h, cv[u].R);
            }
        }
    }
    // the row and counts of an incomplete candidate row are written by k_big_rows
    const bool own = valid && l == 0 && !incomplete;
    if (own) {
        cnt[i] = w;
        cnt0[i] = w;
    }
    unsigned long long nlong = (own && w > stride) ? 1ull : 0ull;
    unsigned long long nent = own ? (unsigned long long)w : 0ull;
    for (int off = 32; off >= 1; off >>= 1) {
        nlong += __shfl_xor(nlong, off);
        nent += __shfl_xor(nent, off);
    }
    if (lane == 0) {
        if (nlong) spread_add(spread, 0, nlong);
        spread_add(spread, 1, nent);
    }
}

// Rows of the targets solved by k_hsolve_big (their candidate rows are incomplete):
// one wavefront per target j walks the tree for S_j (prune as k_ngb, node h_max
// scaled by *hscale when they belong to an earlier h) and writes row j in walk order
// (lanes compacted by ballots).  For a member i whose candidate row is complete, with
// h_i < h_j and r^2 >= R_i^2 (j is not among i's candidates, so k_cand_rows cannot
// add it), it records the extra (i, j).  r^2 uses x_j - x_i, the same bits as i's
// candidate test.  Spread counters: 0 += rows longer than stride, 1 += entries.
template <bool PER>
__global__ void __launch_bounds__(64) k_big_rows(const int32_t *__restrict__ list, const double *__restrict__ x,
                                                 const double *__restrict__ y, const double *__restrict__ z,
                                                 const double *__restrict__ h, const double *__restrict__ R,
                                                 const int32_t *__restrict__ ccnt, int cap,
                                                 const WNode *__restrict__ nodes, int bucket,
                                                 const double *__restrict__ hscale, int stride,
                                                 int32_t *__restrict__ cnt, int32_t *__restrict__ cnt0,
                                                 int32_t *__restrict__ idx, int2 *__restrict__ ext, int64_t ext_cap,
                                                 unsigned long long *__restrict__ next,
                                                 unsigned long long *__restrict__ spread, double per)
{
    __shared__ WalkLDS L;
    __shared__ double sx[WSRC], sy[WSRC], sz[WSRC], sh[WSRC];
    __shared__ int sj[WSRC];
    const int64_t jt = list[blockIdx.x];
    const int lane = threadIdx.x;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const double xj = x[jt], yj = y[jt], zj = z[jt], hj = h[jt];
    const double gb[6] = {xj, yj, zj, xj, yj, zj};
    const double hs = hscale ? *hscale : 1.0;
    int32_t *row = idx + jt * (int64_t)stride;
    const int64_t rcap = ext_cap / SPH_SPREAD, reg = blockIdx.x & (SPH_SPREAD - 1);
    int w = 0;
    (void)group_walk(
        nodes, bucket, L,
        [&](const WNode *nd) {
            const double t = hj + nd->hmax * hs;
            return box_box_dist2_p<PER>(gb, nd, per) >= t * t;
        },
        [&](int q, int i) {
            sx[q] = x[i];
            sy[q] = y[i];
            sz[q] = z[i];
            sh[q] = h[i];
            sj[q] = i;
        },
        [&](int ns) {
            for (int k0 = 0; k0 < ns; k0 += 64) {
                const int k = k0 + lane;
                bool in = false, extra = false;
                int i = 0;
                if (k < ns) {
                    const double dx = wrap_d<PER>(xj - sx[k], per), dy = wrap_d<PER>(yj - sy[k], per),
                                 dz = wrap_d<PER>(zj - sz[k], per);
                    const double r2 = dx * dx + dy * dy + dz * dz, t = hj + sh[k];
                    in = r2 < t * t;
                    i = sj[k];
                    if (in && sh[k] < hj && ccnt[i] <= cap) {
                        const double Ri = R[i];
                        extra = r2 >= Ri * Ri;
                    }
                }
                const uint64_t mb = __ballot(in);
                if (in) {
                    const int pos = w + (int)__popcll(mb & below);
                    if (pos < stride) row[pos] = i;
                }
                w += (int)__popcll(mb);
                if (extra) {
                    const unsigned long long e = atomicAdd(next + reg * SPH_SPREAD_PAD, 1ull);
                    if (e < (unsigned long long)rcap) ext[reg * rcap + (int64_t)e] = make_int2(i, (int)jt);
                }
            }
        });
    if (lane == 0) {
        cnt[jt] = w;
        cnt0[jt] = w;
        if (w > stride) spread_add(spread, 0, 1ull);
        spread_add(spread, 1, (unsigned long long)w);
    }
}

// The extras of region r (ext[r rcap + e], e < min(next[r], rcap)) appended to their rows:
// blocks r, r + SPH_SPREAD, ... stride over region r, so the work follows the pairs recorded
// rather than the regions' capacity.
__global__ void __launch_bounds__(256) k_ext_append(int64_t ext_cap, const int2 *__restrict__ ext,
                                                    const unsigned long long *__restrict__ next,
                                                    int32_t *__restrict__ cnt, int32_t *__restrict__ idx, int stride,
                                                    unsigned long long *__restrict__ spread)
{
    const int64_t rcap = ext_cap / SPH_SPREAD;
    const int reg = blockIdx.x % SPH_SPREAD;
    const int64_t step = (int64_t)(gridDim.x / SPH_SPREAD) * blockDim.x;
    const unsigned long long nr = next[reg * SPH_SPREAD_PAD];
    const int64_t ne = nr < (unsigned long long)rcap ? (int64_t)nr : rcap;
    unsigned long long nlong = 0, nadd = 0;
    for (int64_t e = (int64_t)(blockIdx.x / SPH_SPREAD) * blockDim.x + threadIdx.x; e < ne; e += step) {
        const int2 p = ext[reg * rcap + e];
        const int pos = atomicAdd(&cnt[p.x], 1);
        if (pos < stride) idx[(int64_t)p.x * stride + pos] = p.y;
        nlong += pos == stride ? 1ull : 0ull;  // the row just became too long (counted once)
        nadd += 1ull;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        nlong += __shfl_xor(nlong, off);
        nadd += __shfl_xor(nadd, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (nlong) spread_add(spread, 0, nlong);
        if (nadd) spread_add(spread, 1, nadd);
    }
}

// Ascending order of idx[i*stride + cnt0[i] .. i*stride + min(cnt[i], stride)):
// insertion sort per row (the appended parts hold about one entry on average).
__global__ void __launch_bounds__(256) k_row_tail_sort(int64_t n, const int32_t *__restrict__ cnt0,
                                                       const int32_t *__restrict__ cnt, int32_t *__restrict__ idx,
                                                       int stride)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = cnt0[i], e = min(cnt[i], stride);
    if (e - b < 2) return;
    int32_t *row = idx + i * (int64_t)stride;
    for (int k = b + 1; k < e; k++) {
        const int32_t v = row[k];
        int q = k - 1;
        while (q >= b && row[q] > v) {
            row[q + 1] = row[q];
            q--;
        }
        row[q + 1] = v;
    }
}

// *out = max_i h_new[i] / h_old[i] (1 + 8 eps): positive doubles order like
// their bit patterns, so one unsigned 64-bit atomicMax per workgroup.
__global__ void __launch_bounds__(256) k_hratio(int64_t n, const double *__restrict__ hn,
                                                const double *__restrict__ ho, unsigned long long *__restrict__ out)
{
    __shared__ double wmax[4];
    double v = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v = fmax(v, hn[i] / ho[i]);
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        v = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3])) * (1.0 + 8.0 * 2.220446049250313e-16);
        atomicMax(out, (unsigned long long)__double_as_longlong(v));
    }
}

namespace sph {

hipError_t neighbour_rows_from_candidates_begin(hipStream_t st, int64_t n, int cap, const double *x,
                                               const double *y, const double *z, double per, const double *cand,
                                               const int32_t *cidx, const int32_t *ccnt, const double *h,
                                               const double *R, int stride,
                                               int32_t *cnt, int32_t *cnt0, int32_t *idx, int2 *ext,
                                               int64_t ext_cap, unsigned long long *next,
                                               unsigned long long *spread, unsigned long long *totals,
                                               unsigned long long *host_words, hipEvent_t done,
                                               const BigRows &big, void *rec)
{
    if (stride < 1 || cap < 1 || ext_cap < SPH_SPREAD) return hipErrorInvalidValue;
    fill<unsigned long long>(st, SPH_SPREAD_WORDS, next, 0ull);
    fill<unsigned long long>(st, 2 * SPH_SPREAD_WORDS, spread, 0ull);
    const unsigned nbr = (unsigned)sph_blocks(n * 16, 256);
    double2 *hr = static_cast<double2 *>(rec);
    CandRec *cr = static_cast<CandRec *>(rec);
    if (cand) {
        k_pack_hr<<<sph_blocks(n, 256), 256, 0, st>>>(n, cap, h, R, ccnt, hr);
        k_cand_rows<16, false, true><<<nbr, 256, 0, st>>>(n, cap, cand, cidx, ccnt, hr, nullptr, per, stride, cnt,
                                                          cnt0, idx, ext, ext_cap, next, spread);
    } else {
        k_pack_rec<<<sph_blocks(n, 256), 256, 0, st>>>(n, cap, x, y, z, h, R, ccnt, cr);
        if (per > 0.0)
            k_cand_rows<16, true, false><<<nbr, 256, 0, st>>>(n, cap, nullptr, cidx, ccnt, nullptr, cr, per, stride,
                                                              cnt, cnt0, idx, ext, ext_cap, next, spread);
        else
            k_cand_rows<16, false, false><<<nbr, 256, 0, st>>>(n, cap, nullptr, cidx, ccnt, nullptr, cr, per, stride,
                                                               cnt, cnt0, idx, ext, ext_cap, next, spread);
    }
    if (big.count > 0) {
        const unsigned nb = (unsigned)big.count;
        if (big.per > 0.0)
            k_big_rows<true><<<nb, 64, 0, st>>>(big.list, big.x, big.y, big.z, h, R, ccnt, cap, big.nodes, big.bucket,
                                                big.hscale, stride, cnt, cnt0, idx, ext, ext_cap, next, spread,
                                                big.per);
        else
            k_big_rows<false><<<nb, 64, 0, st>>>(big.list, big.x, big.y, big.z, h, R, ccnt, cap, big.nodes, big.bucket,
                                                 big.hscale, stride, cnt, cnt0, idx, ext, ext_cap, next, spread,
                                                 big.per);
    }
    k_ext_append<<<SPH_SPREAD * 32, 256, 0, st>>>(ext_cap, ext, next, cnt, idx, stride, spread);
    k_row_tail_sort<<<sph_blocks(n, 256), 256, 0, st>>>(n, cnt0, cnt, idx, stride);
    spread_total(st, 2, spread, totals);
    hipError_t e = hipMemcpyAsync(host_words, totals, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipMemcpy2DAsync(host_words + 2, sizeof(unsigned long long), next, SPH_SPREAD_PAD * sizeof(unsigned long long),
                         sizeof(unsigned long long), SPH_SPREAD, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipEventRecord(done, st);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t neighbour_rows_from_candidates_end(hipEvent_t done, const unsigned long long *host_words, int64_t ext_cap,
                                             int64_t *n_long, int64_t *total, int64_t *n_extra)
{
    hipError_t e = hipEventSynchronize(done);
    if (e != hipSuccess) return e;
    int64_t nextra = 0, lost = 0;
    for (int r = 0; r < SPH_SPREAD; r++) {
        nextra += (int64_t)host_words[2 + r];
        if (host_words[2 + r] > (unsigned long long)(ext_cap / SPH_SPREAD)) lost = 1;  // a lost extra voids the rows
    }
    *n_long = (int64_t)host_words[0] + lost;
    *total = (int64_t)host_words[1];
    *n_extra = nextra;
    return hipSuccess;
}

hipError_t h_ratio_max(hipStream_t st, int64_t n, const double *h_new, const double *h_old, double *out)
{
    fill<unsigned long long>(st, 1, (unsigned long long *)out, 0ull);
    const int64_t nb = sph_blocks(n, 256) < 1024 ? sph_blocks(n, 256) : 1024;
    k_hratio<<<(unsigned)nb, 256, 0, st>>>(n, h_new, h_old, (unsigned long long *)out);
    return hipGetLastError();
}

// The groups that still hold pending targets (state 0) and those targets, for the rounds after
// the first: one wavefront per group, ballot-compacted appends (one atomic per list and group).
// List order follows the atomics; each group's and target's result does not depend on it.
__global__ void __launch_bounds__(64) k_pending_lists(const int32_t *__restrict__ groups, const WNode *__restrict__ gnodes,
                                                      int64_t m, const int32_t *__restrict__ state,
                                                      int32_t *__restrict__ glist, int32_t *__restrict__ tlist,
                                                      unsigned long long *__restrict__ counts)
{
    int64_t first;
    int ng;
    group_range(groups, gnodes, m, blockIdx.x, first, ng);
    const int lane = threadIdx.x;
    const bool pend = lane < ng && state[first + lane] == 0;
    const uint64_t pm = __ballot(pend);
    if (pm == 0) return;
    unsigned long long gb = 0, tb = 0;
    if (lane == 0) {
        gb = atomicAdd(counts, 1ull);
        tb = atomicAdd(counts + 1, (unsigned long long)__popcll(pm));
    }
    gb = __shfl(gb, 0);
    tb = __shfl(tb, 0);
    if (lane == 0) glist[gb] = (int32_t)blockIdx.x;
    if (pend)
        tlist[tb + __builtin_amdgcn_mbcnt_hi((uint32_t)(pm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u))] =
            (int32_t)(first + lane);
}

namespace {

HSolveArgs hsolve_args(const HSolveParams &p)
{
    HSolveArgs a;
    a.m = p.m; a.tx = p.tx; a.ty = p.ty; a.tz = p.tz;
    a.x = p.x; a.y = p.y; a.z = p.z; a.nodes = p.nodes; a.bucket = p.bucket;
    a.mode = p.mode; a.n_ngb = p.n_ngb; a.tol = p.tol;
    a.n_low = p.n_low; a.n_up = p.n_up; a.max_steps = p.max_steps; a.kth = p.kth;
    a.cap = p.cap; a.cand = p.cand; a.R = p.R; a.h = p.h_out; a.steps = p.steps; a.state = p.state;
    a.cidx = p.cidx; a.ccnt = p.ccnt;
    a.pending = p.counters;
    a.groups = p.groups;
    a.gnodes = p.nodes;
    a.big_list = p.mode == 0 ? p.big_list : nullptr;
    a.big_count = p.big_count;
    a.per = p.per;
    a.glist = a.tlist = nullptr;
    a.gcount = a.tcount = nullptr;
    return a;
}

// The targets k_hsolve left to the candidate-free solver (count read back by the caller).
hipError_t hsolve_big(hipStream_t st, const HSolveParams &p, const HSolveArgs &a, unsigned long long nbig)
{
    if (p.n_big) *p.n_big = (int64_t)nbig;
    if (nbig == 0 || !a.big_list) return hipSuccess;
    if (a.per > 0.0)
        k_hsolve_big<true><<<(unsigned)nbig, 64, 0, st>>>(a, a.big_list);
    else
        k_hsolve_big<false><<<(unsigned)nbig, 64, 0, st>>>(a, a.big_list);
    return hipGetLastError();
}

// One Newton / k-th round: counters cleared, the solve, the fold into p.totals.
// One round.  pending > 0 (a round after the first, with pending targets counted by the
// previous round): only the groups and targets still pending are launched, from lists built on
// the device (p.plist); the grids are sized by the host's pending count, which bounds both lists.
void hsolve_round(hipStream_t st, const HSolveParams &p, HSolveArgs a, int64_t pending = 0)
{
    int64_t nblk = p.groups ? p.ngroups : (p.m + 63) / 64;
    int64_t nt = p.m;
    if (pending > 0 && p.plist) {
        unsigned long long *cnt = p.plist_counts;
        int32_t *gl = p.plist, *tl = p.plist + nblk;
        fill<unsigned long long>(st, 2, cnt, 0ull);
        k_pending_lists<<<(unsigned)nblk, 64, 0, st>>>(p.groups, p.nodes, p.m, p.state, gl, tl, cnt);
        a.glist = gl;
        a.tlist = tl;
        a.gcount = cnt;
        a.tcount = cnt + 1;
        nblk = std::min<int64_t>(nblk, pending);
        nt = pending;
    }
    fill<unsigned long long>(st, 5 * SPH_SPREAD_WORDS, p.counters, 0ull);
    if (a.per > 0.0)
        k_hwalk<true><<<(unsigned)nblk, 64, 0, st>>>(a);
    else
        k_hwalk<false><<<(unsigned)nblk, 64, 0, st>>>(a);
    const unsigned nbn = (unsigned)((nt + 64 / HN_LANES - 1) / (64 / HN_LANES));
    if (a.cand)
        k_hnewton<false, true><<<nbn, 64, 0, st>>>(a);  // (distances read from the rows)
    else if (a.per > 0.0)
        k_hnewton<true, false><<<nbn, 64, 0, st>>>(a);
    else
        k_hnewton<false, false><<<nbn, 64, 0, st>>>(a);
    spread_total(st, 5, p.counters, p.totals);
}

// Book-keeping of one round's totals hc; true when every target is decided.
hipError_t hsolve_account(const HSolveParams &p, const unsigned long long *hc, int round, bool *done)
{
    if (p.work) {
        p.work[0] += (int64_t)hc[3];
        p.work[1] += (int64_t)hc[4];
    }
    if (hc[2]) return hipErrorIllegalState;  // walk stack overflow (cannot happen for valid trees)
    if (p.unconverged_total) *p.unconverged_total += (int64_t)hc[1];
    if (p.rounds) *p.rounds = round + 1;
    *done = hc[0] == 0;
    return hipSuccess;
}

}  // namespace

hipError_t hsolve_begin(hipStream_t st, const HSolveParams &p)
{
    if (p.mode == 1 || p.m <= 0 || !p.host_totals || !p.done) return hipErrorInvalidValue;
    if (p.bucket > 64 || p.cap > SPH_HCAP_SLOTS || p.cap < 1) return hipErrorInvalidValue;
    const HSolveArgs a = hsolve_args(p);
    k_hsolve_init<<<sph_blocks(p.m, 256), 256, 0, st>>>(p.m, p.h_in, p.r_fac, p.R, p.h_out, p.steps, p.state);
    if (a.big_list) fill<unsigned long long>(st, 1, p.big_count, 0ull);
    hsolve_round(st, p, a);
    hipError_t e = hipMemcpyAsync(p.host_totals, p.totals, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    if (a.big_list) {
        e = hipMemcpyAsync(p.host_totals + 5, p.big_count, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return e;
    }
    e = hipEventRecord(p.done, st);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t hsolve_end(hipStream_t st, const HSolveParams &p, int max_rounds, int64_t *n_unconverged,
                      bool *more_rounds)
{
    *more_rounds = false;
    hipError_t e = hipEventSynchronize(p.done);
    if (e != hipSuccess) return e;
    bool done = false;
    e = hsolve_account(p, p.host_totals, 0, &done);
    if (e != hipSuccess) return e;
    const HSolveArgs a = hsolve_args(p);
    if (done) {
        *n_unconverged = 0;
        return hsolve_big(st, p, a, a.big_list ? p.host_totals[5] : 0);
    }
    *more_rounds = true;
    unsigned long long hc[6] = {p.host_totals[0], 0, 0, 0, 0, 0};
    for (int r = 1; r < max_rounds; r++) {
        hsolve_round(st, p, a, (int64_t)hc[0]);
        e = hipMemcpyAsync(hc, p.totals, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return e;
        if (a.big_list) {
            e = hipMemcpyAsync(hc + 5, p.big_count, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
            if (e != hipSuccess) return e;
        }
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        e = hsolve_account(p, hc, r, &done);
        if (e != hipSuccess) return e;
        if (done) {
            *n_unconverged = 0;
            e = hsolve_big(st, p, a, hc[5]);
            if (e != hipSuccess) return e;
            return hipGetLastError();
        }
    }
    *n_unconverged = (int64_t)hc[0];
    return hipGetLastError();
}

hipError_t hsolve(hipStream_t st, const HSolveParams &p, int max_rounds, int64_t *n_unconverged)
{
    if (p.m <= 0) { *n_unconverged = 0; return hipSuccess; }
    if (p.bucket > 64 || p.cap > SPH_HCAP_SLOTS || p.cap < 1) return hipErrorInvalidValue;
    const HSolveArgs a = hsolve_args(p);
    k_hsolve_init<<<sph_blocks(p.m, 256), 256, 0, st>>>(p.m, p.h_in, p.r_fac, p.R, p.h_out, p.steps, p.state);
    if (a.big_list) fill<unsigned long long>(st, 1, p.big_count, 0ull);
    unsigned long long hc[6] = {0, 0, 0, 0, 0, 0};
    if (p.mode == 1) {
        if (p.per > 0.0) return hipErrorInvalidValue;  // the count walk has open boundaries only
        fill<unsigned long long>(st, 5 * SPH_SPREAD_WORDS, p.counters, 0ull);
        k_hstep<<<sph_blocks(p.m, 256), 256, 0, st>>>(a);
        spread_total(st, 5, p.counters, p.totals);
        hipError_t e = hipMemcpyAsync(hc, p.totals, sizeof(hc), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        if (p.unconverged_total) *p.unconverged_total += (int64_t)hc[1];
        *n_unconverged = 0;
        return hipGetLastError();
    }
    for (int r = 0; r < max_rounds; r++) {
        hsolve_round(st, p, a, r > 0 ? (int64_t)hc[0] : 0);
        hipError_t e = hipMemcpyAsync(hc, p.totals, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return e;
        if (a.big_list) {
            e = hipMemcpyAsync(hc + 5, p.big_count, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
            if (e != hipSuccess) return e;
        }
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        bool done = false;
        e = hsolve_account(p, hc, r, &done);
        if (e != hipSuccess) return e;
        if (done) {
            *n_unconverged = 0;
            e = hsolve_big(st, p, a, hc[5]);
            if (e != hipSuccess) return e;
            return hipGetLastError();
        }
    }
    *n_unconverged = (int64_t)hc[0];
    return hipGetLastError();
}

hipError_t neighbours(hipStream_t st, int64_t n, int64_t ngroups, const int32_t *groups, const double *x,
                      const double *y, const double *z, const double *h, const WNode *nodes, int bucket, int32_t *cnt,
                      int64_t *off, int64_t *scan_tmp, unsigned long long *overflow, int32_t *idx, int64_t idx_cap,
                      int64_t *total, const double *hscale, double per)
{
    if (bucket > 64) return hipErrorInvalidValue;
    fill<unsigned long long>(st, 1, overflow, 0ull);
    if (per > 0.0)
        k_ngb<0, true><<<(unsigned)ngroups, 64, 0, st>>>(groups, x, y, z, h, nodes, bucket, cnt, nullptr, nullptr, 0, overflow,
                                               nullptr, hscale, per);
    else
        k_ngb<0, false><<<(unsigned)ngroups, 64, 0, st>>>(groups, x, y, z, h, nodes, bucket, cnt, nullptr, nullptr, 0, overflow,
                                               nullptr, hscale, per);
    hipError_t e = scan_counts(st, n, cnt, off, scan_tmp);
    if (e != hipSuccess) return e;
    unsigned long long hov = 0;
    e = hipMemcpyAsync(total, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(&hov, overflow, sizeof(hov), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    if (hov) return hipErrorIllegalState;
    if (*total > idx_cap || idx == nullptr) return hipSuccess;  // caller grows idx and calls neighbours_fill
    if (per > 0.0)
        k_ngb<1, true><<<(unsigned)ngroups, 64, 0, st>>>(groups, x, y, z, h, nodes, bucket, nullptr, off, idx, 0, overflow,
                                               nullptr, hscale, per);
    else
        k_ngb<1, false><<<(unsigned)ngroups, 64, 0, st>>>(groups, x, y, z, h, nodes, bucket, nullptr, off, idx, 0, overflow,
                                               nullptr, hscale, per);
    return hipGetLastError();
}

hipError_t neighbours_fill(hipStream_t st, int64_t ngroups, const int32_t *groups, const double *x, const double *y,
                           const double *z, const double *h, const WNode *nodes, int bucket, const int64_t *off,
                           int32_t *idx, unsigned long long *overflow, const double *hscale, double per)
{
    if (per > 0.0)
        k_ngb<1, true><<<(unsigned)ngroups, 64, 0, st>>>(groups, x, y, z, h, nodes, bucket, nullptr, off, idx, 0, overflow,
                                               nullptr, hscale, per);
    else
        k_ngb<1, false><<<(unsigned)ngroups, 64, 0, st>>>(groups, x, y, z, h, nodes, bucket, nullptr, off, idx, 0, overflow,
                                               nullptr, hscale, per);
    return hipGetLastError();
}

hipError_t neighbour_rows(hipStream_t st, int64_t ngroups, const int32_t *groups, const double *x, const double *y,
                          const double *z, const double *h, const WNode *nodes, int bucket, int stride, int32_t *cnt,
                          int32_t *idx, unsigned long long *overflow, unsigned long long *spread,
                          unsigned long long *totals, int64_t *n_long, int64_t *total, const double *hscale,
                          double per)
{
    if (bucket > 64 || stride < 1) return hipErrorInvalidValue;
    fill<unsigned long long>(st, 1, overflow, 0ull);
    fill<unsigned long long>(st, 2 * SPH_SPREAD_WORDS, spread, 0ull);
    if (per > 0.0)
        k_ngb<2, true><<<(unsigned)ngroups, 64, 0, st>>>(groups, x, y, z, h, nodes, bucket, cnt, nullptr, idx, stride, overflow,
                                               spread, hscale, per);
    else
        k_ngb<2, false><<<(unsigned)ngroups, 64, 0, st>>>(groups, x, y, z, h, nodes, bucket, cnt, nullptr, idx, stride, overflow,
                                               spread, hscale, per);
    spread_total(st, 2, spread, totals);
    unsigned long long hc[3] = {0, 0, 0};
    hipError_t e = hipMemcpyAsync(hc, totals, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(hc + 2, overflow, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    if (hc[2]) return hipErrorIllegalState;
    *n_long = (int64_t)hc[0];
    *total = (int64_t)hc[1];
    return hipGetLastError();
}

}  // namespace sph
