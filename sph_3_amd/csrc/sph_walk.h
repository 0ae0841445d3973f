// sph_walk.h -- group walk shared by the neighbour and smoothing-length
// kernels (one wavefront per target group, 64 lanes).
//
// A group is a contiguous range [first, first+count) of targets (count <= 64):
// either a tree node holding at most 64 particles (targets = the particles)
// or a block of 64 consecutive entries of an arbitrary target list.  The walk
// is breadth-first and lane-parallel: each step pops up to 64 node ids from an
// LDS stack and every lane tests one node with the caller's prune predicate
// (a conservative group-level test).  Children of surviving internal nodes are
// pushed; the particles of surviving buckets (count <= bucket) are staged in
// LDS by the caller's `stage(slot, j)`, all buckets of a step at once (each lane
// finds its particle's bucket from the step's bucket prefix sums, so the loads
// of a step are issued together instead of one bucket after another), in
// bucket order; whenever the WSRC slots are full the caller's `eval(cnt)` runs
// on all lanes over the staged slots.  No lane diverges from the wave's
// control flow.
#pragma once
#include "sph_common.h"

#define WSTK 512   // LDS stack capacity (node ids)
#define WSRC 64    // staged sources before a flush (one chunk of the 64-lane evaluations)

// Box of the group's targets by wave reduction (lanes >= count pass +-inf).
__device__ __forceinline__ void group_box(bool valid, double x, double y, double z, double *gb)
{
    gb[0] = valid ? x : INFINITY; gb[1] = valid ? y : INFINITY; gb[2] = valid ? z : INFINITY;
    gb[3] = valid ? x : -INFINITY; gb[4] = valid ? y : -INFINITY; gb[5] = valid ? z : -INFINITY;
    for (int off = 32; off >= 1; off >>= 1)
        for (int d = 0; d < 3; d++) {
            gb[d] = fmin(gb[d], __shfl_xor(gb[d], off));
            gb[3 + d] = fmax(gb[3 + d], __shfl_xor(gb[3 + d], off));
        }
}

// Squared distance between two boxes (0 if they overlap), FP64, association
// ((ex*ex + ey*ey) + ez*ez).  For any points a in A and b in B the computed
// |a - b|^2 (same association, one rounding per operation) is >= this value,
// because rounding is monotone.
__device__ __forceinline__ double box_box_dist2(const double *gb, const WNode *__restrict__ nd)
{
#pragma clang fp contract(off)
    double ex = 0.0, ey = 0.0, ez = 0.0;
    if (nd->lx - gb[3] > 0.0) ex = nd->lx - gb[3];
    if (gb[0] - nd->hx > 0.0) ex = gb[0] - nd->hx;
    if (nd->ly - gb[4] > 0.0) ey = nd->ly - gb[4];
    if (gb[1] - nd->hy > 0.0) ey = gb[1] - nd->hy;
    if (nd->lz - gb[5] > 0.0) ez = nd->lz - gb[5];
    if (gb[2] - nd->hz > 0.0) ez = gb[2] - nd->hz;
    return ex * ex + ey * ey + ez * ez;
}

// The same for a periodic box of side L (PER = false: box_box_dist2): per axis the
// smallest gap over the node's images at -L, 0, +L, less 4 ulp of L (the wraps of the
// pair differences round), so no pair of points of the boxes has a computed
// minimum-image distance below it.
template <bool PER>
__device__ __forceinline__ double box_box_dist2_p(const double *gb, const WNode *__restrict__ nd, double L)
{
#pragma clang fp contract(off)
    if constexpr (!PER) return box_box_dist2(gb, nd);
    const double lo[3] = {nd->lx, nd->ly, nd->lz}, hi[3] = {nd->hx, nd->hy, nd->hz};
    const double slack = 8.9e-16 * L;
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double e = INFINITY;
#pragma unroll
        for (int k = -1; k <= 1; k++) {
            const double a = lo[d] + k * L, b = hi[d] + k * L;
            double g = 0.0;
            if (a - gb[3 + d] > g) g = a - gb[3 + d];
            if (gb[d] - b > g) g = gb[d] - b;
            e = fmin(e, g);
        }
        e = fmax(e - slack, 0.0);
        s = s + e * e;
    }
    return s;
}

struct WalkLDS {
    int stk[WSTK];
    int bfirst[64], bstart[65];
};

// Returns true if the stack ever overflowed (impossible for trees shallower
// than 128 levels, reported so that a caller can never use a truncated walk).
template <class Prune, class Stage, class Eval>
__device__ __forceinline__ bool group_walk(const WNode *__restrict__ nodes, int bucket, WalkLDS &L, Prune prune,
                                           Stage stage, Eval eval)
{
    bool overflow = false;
    const int lane = threadIdx.x;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int nsrc = 0, sp = 1;
    if (lane == 0) L.stk[0] = 0;
    __syncthreads();
    while (sp > 0) {
        // pop up to 64 nodes; one at a time near capacity (growth then <= 1 per level)
        const int nt = (sp > WSTK - 128) ? 1 : (sp < 64 ? sp : 64);
        const int base = sp - nt;
        const int node = lane < nt ? L.stk[base + lane] : -1;
        __syncthreads();
        sp = base;
        bool bkt = false, opn = false;
        int nfirst = 0, ncount = 0, nleft = 0, nright = 0;
        if (node >= 0) {
            const WNode *nd = nodes + node;
            if (!prune(nd)) {
                nfirst = nd->first; ncount = nd->count; nleft = nd->left; nright = nd->right;
                if (ncount <= bucket) bkt = true;
                else opn = true;
            }
        }
        const uint64_t mo = __ballot(opn), mb = __ballot(bkt);
        if (opn) {
            const int po = sp + 2 * (int)__popcll(mo & lt);
            if (po + 1 < WSTK) {
                L.stk[po] = nright;
                L.stk[po + 1] = nleft;
            }
        }
        if (sp + 2 * (int)__popcll(mo) > WSTK) overflow = true;
        sp += 2 * (int)__popcll(mo);
        if (sp > WSTK) sp = WSTK;
        // bucket prefix sums in ballot (= lane) order, by bit planes of the counts (counts <=
        // bucket): one ballot and one masked population count per bit, no lane exchanges
        int excl = 0, total = 0;
        for (int b = 0; (1 << b) <= bucket; b++) {
            const uint64_t mbit = wave_ballot(bkt && ((ncount >> b) & 1));
            excl += (int)__popcll(mbit & lt) << b;
            total += (int)__popcll(mbit) << b;
        }
        const int incl = excl + (bkt ? ncount : 0);
        if (bkt) {
            const int pb = (int)__popcll(mb & lt);
            L.bfirst[pb] = nfirst;
            L.bstart[pb] = incl - ncount;
        }
        const int nb = (int)__popcll(mb);
        if (lane == 0) L.bstart[nb] = total;
        __syncthreads();
        for (int s0 = 0; s0 < total;) {
            const int take = (WSRC - nsrc) < (total - s0) ? (WSRC - nsrc) : (total - s0);
            for (int k = lane; k < take; k += 64) {
                const int sidx = s0 + k;
                int lo = 0, hi = nb - 1;  // last bucket b with bstart[b] <= sidx
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (L.bstart[mid] <= sidx) lo = mid;
                    else hi = mid - 1;
                }
                stage(nsrc + k, L.bfirst[lo] + (sidx - L.bstart[lo]));
            }
            nsrc += take;
            s0 += take;
            __syncthreads();
            if (nsrc == WSRC) {
                eval(nsrc);
                nsrc = 0;
                __syncthreads();
            }
        }
    }
    if (nsrc > 0) eval(nsrc);
    __syncthreads();
    return overflow;
}
