// sph_walk.h -- group walk shared by the neighbour and smoothing-length
// kernels (one wavefront per target group, 64 lanes).
//
// A group is a contiguous range [first, first+count) of targets (count <= 64):
// either a tree node holding at most 64 particles (targets = the particles)
// or a block of 64 consecutive entries of an arbitrary target list.  The walk
// is breadth-first and lane-parallel: each step pops up to 64 node ids from an
// LDS stack and every lane tests one node with the caller's prune predicate
// (a conservative group-level test).  Children of surviving internal nodes are
// pushed; surviving buckets (count <= bucket) have their particles staged in
// LDS by the caller's `stage(slot, j)` and, once 64 or more are staged, the
// caller's `eval(cnt)` runs on all lanes over the staged slots.  No lane
// diverges from the wave's control flow.
#pragma once
#include "sph_common.h"

#define WSTK 2048  // LDS stack capacity (node ids)
#define WSRC 128   // staged sources before a flush (< 64 + one bucket of <= 64)

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

struct WalkLDS {
    int stk[WSTK];
    int bfirst[64], bcount[64];
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
        if (bkt) {
            const int pb = (int)__popcll(mb & lt);
            L.bfirst[pb] = nfirst;
            L.bcount[pb] = ncount;
        }
        const int nb = (int)__popcll(mb);
        __syncthreads();
        for (int b = 0; b < nb; b++) {
            const int f = L.bfirst[b], c = L.bcount[b];
            if (lane < c) stage(nsrc + lane, f + lane);
            nsrc += c;
            __syncthreads();
            if (nsrc >= 64) {
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
