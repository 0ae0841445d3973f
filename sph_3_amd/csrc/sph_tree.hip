// sph_tree.hip -- binary radix tree over Morton-sorted particles, built
// bottom-up without locks, with FP64 node moments and escape links for the
// stackless walks.
//
// Replaces the reference's lock-based top-down octree insertion
// (EV/bh.h:149-367 build_tree_kernel, CAS on child slots, one block),
// centre_of_mass_kernel (EV/bh.h:374-389) and the DFS sort_kernel
// (EV/bh.h:394-450).  Construction follows Karras (HPG 2012): every internal
// node is found independently from the sorted keys; moments are combined by
// the second child to arrive (one atomic per internal node, workgroup scope).
#include "sph_tree.h"

#pragma clang fp contract(off)

__device__ __forceinline__ int clz64d(uint64_t x) { return x ? __clzll(x) : 64; }
__device__ __forceinline__ int clz32d(uint32_t x) { return x ? __clz(x) : 32; }

// Common-prefix length of sorted keys i and j; equal keys fall back to the
// index bits so that every split is unique.  -1 outside [0, n).
__device__ __forceinline__ int delta(const uint64_t *__restrict__ k, int64_t n, int64_t i, int64_t j)
{
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a != b) return clz64d(a ^ b);
    return 64 + clz32d((uint32_t)i ^ (uint32_t)j);
}

__global__ void __launch_bounds__(256) k_topology(int64_t n, const uint64_t *__restrict__ k,
                                                  WNode *__restrict__ nodes,
                                                  int32_t *__restrict__ right,
                                                  int32_t *__restrict__ parent,
                                                  int32_t *__restrict__ ncount)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t nl = n - 1;
    if (i < n) {  // leaf of sorted particle i
        WNode *L = &nodes[nl + i];
        L->first = (int)i;
        L->count = 1;
        L->left = -1;
        L->right = -1;
        ncount[nl + i] = 1;
    }
    if (i >= n - 1) return;
    int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) > 0 ? 1 : -1;
    int dmin = delta(k, n, i, i - d);
    int64_t lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = delta(k, n, i, j);
    int64_t s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    int32_t lc = (int32_t)(lo == gamma ? nl + gamma : gamma);
    int32_t rc = (int32_t)(hi == gamma + 1 ? nl + gamma + 1 : gamma + 1);
    WNode *P = &nodes[i];
    P->first = (int)lo;
    P->count = (int)(hi - lo + 1);
    ncount[i] = (int)(hi - lo + 1);
    P->left = lc;
    P->right = rc;
    right[i] = rc;
    parent[lc] = (int32_t)i;
    parent[rc] = (int32_t)i;
    if (i == 0) parent[0] = -1;
}

// Escape link: the topmost node whose range starts right after this node's
// range (internal b+1 if its range starts at b+1, else leaf b+1); -1 past the end.
__global__ void __launch_bounds__(256) k_escape(int64_t n, WNode *__restrict__ nodes)
{
    int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= 2 * n - 1) return;
    int64_t nl = n - 1;
    int64_t b = (int64_t)nodes[id].first + nodes[id].count - 1;
    int e;
    if (b == n - 1) e = -1;
    else if (b + 1 < n - 1 && nodes[b + 1].first == b + 1) e = (int)(b + 1);
    else e = (int)(nl + b + 1);
    nodes[id].escape = e;
}

__device__ __forceinline__ void combine(WNode *__restrict__ nodes, int64_t p, int64_t a, int64_t b)
{
    // children were written by other threads: read through L2 (the acquire
    // fence before this call invalidates this CU's vector L1).
    const WNode A = nodes[a];
    const WNode B = nodes[b];
    WNode &P = nodes[p];
    double lx = A.lx < B.lx ? A.lx : B.lx, ly = A.ly < B.ly ? A.ly : B.ly, lz = A.lz < B.lz ? A.lz : B.lz;
    double hx = A.hx > B.hx ? A.hx : B.hx, hy = A.hy > B.hy ? A.hy : B.hy, hz = A.hz > B.hz ? A.hz : B.hz;
    double mt = A.mass + B.mass;
    double cx, cy, cz;
    if (mt > 0.0) {
        cx = (A.mass * A.cx + B.mass * B.cx) / mt;
        cy = (A.mass * A.cy + B.mass * B.cy) / mt;
        cz = (A.mass * A.cz + B.mass * B.cz) / mt;
    } else {
        cx = 0.5 * (lx + hx);
        cy = 0.5 * (ly + hy);
        cz = 0.5 * (lz + hz);
    }
    double s = hx - lx;
    if (hy - ly > s) s = hy - ly;
    if (hz - lz > s) s = hz - lz;
    P.lx = lx; P.ly = ly; P.lz = lz;
    P.hx = hx; P.hy = hy; P.hz = hz;
    P.cx = cx; P.cy = cy; P.cz = cz;
    P.mass = mt;
    P.size = s;
    P.hmax = A.hmax > B.hmax ? A.hmax : B.hmax;
    P.epsmax = A.epsmax > B.epsmax ? A.epsmax : B.epsmax;
}

// Bottom-up moments without device-scope fences.  Device-wide release
// fences write back the L2 of every XCD and cost microseconds each, and
// same-address device atomics serialise, so the climb runs in levels whose
// workgroups each own a contiguous range of node ids and particles:
//  (1) k_moments_local: one thread per internal node of the workgroup's
//      MOM_BLOCK ids.  Only the buckets (nodes of at most `bucket` particles
//      whose parent holds more, or the root) and the nodes above them get
//      records -- every walk stops at a bucket and stages its particles, so
//      nothing reads a node below one.  A bucket's record is summed directly
//      from its particles (in index order) by the thread of its id, or, for a
//      single-particle bucket (a leaf child of a larger node), by the thread of
//      its parent; that thread then climbs while the parent's particle range
//      lies inside this workgroup's MOM_BLOCK particles (internal node k has k
//      as an end of its range, so both children of such a parent belong to
//      this workgroup); the second child to arrive combines (workgroup-scope
//      atomics on the arrival counters).  A finished node whose parent leaves
//      the range joins the workgroup's frontier, kept in LDS and written to
//      the block's own region of MOM_FMAX slots.
//  (2) k_moments_up: one workgroup per MOM_FAN level-1 blocks climbs from
//      their frontier nodes the same way inside MOM_BLOCK*MOM_FAN particles,
//      and a last single workgroup climbs from the remaining frontier to the
//      root.
// A block's frontier nodes are the maximal subtrees inside its range (at most
// two per tree level; the radix tree has at most 97 levels: 64 key plus 32
// index bits of common prefix), the buckets that cross the range's ends (at
// most two), and leaf buckets outside the range, each the child of a node of
// this block that crosses one of the range's ends (nested, so at most 97 per
// end): 2*97 + 2 + 2*97 = 390 < MOM_FMAX slots.
#define MOM_BLOCK 1024
#define MOM_FAN 64
#define MOM_FMAX 512

__device__ __forceinline__ void front_push(int32_t *flist, unsigned int *fn, int32_t node)
{
    const unsigned int k = atomicAdd(fn, 1u);
    if (k < MOM_FMAX) flist[k] = node;
}

__device__ __forceinline__ void front_flush(int32_t *flist, unsigned int fn, int32_t *reg, int32_t *cnt, int64_t blk)
{
    const int c = fn < MOM_FMAX ? (int)fn : MOM_FMAX;
    for (int t = threadIdx.x; t < c; t += blockDim.x) reg[blk * MOM_FMAX + t] = flist[t];
    if (threadIdx.x == 0) cnt[blk] = c;
}

// Record of a bucket: box, mass, centre of mass sum_j m_j x_j / sum_j m_j (sums in index
// order; the box centre when the mass is zero), size = largest box extent, hmax, epsmax.
__device__ __forceinline__ void bucket_record(WNode *__restrict__ nodes, int64_t id, int first, int cnt,
                                              const double *__restrict__ x, const double *__restrict__ y,
                                              const double *__restrict__ z, const double *__restrict__ m,
                                              const double *__restrict__ h, const double *__restrict__ eps)
{
    double lx = INFINITY, ly = INFINITY, lz = INFINITY, hx = -INFINITY, hy = -INFINITY, hz = -INFINITY;
    double mt = 0.0, sx = 0.0, sy = 0.0, sz = 0.0, hm = 0.0, em = 0.0;
    for (int j = first; j < first + cnt; j++) {
        const double px = x[j], py = y[j], pz = z[j], mj = m ? m[j] : 0.0;
        lx = px < lx ? px : lx; ly = py < ly ? py : ly; lz = pz < lz ? pz : lz;
        hx = px > hx ? px : hx; hy = py > hy ? py : hy; hz = pz > hz ? pz : hz;
        mt = mt + mj;
        sx = sx + mj * px; sy = sy + mj * py; sz = sz + mj * pz;
        if (h && h[j] > hm) hm = h[j];
        if (eps && eps[j] > em) em = eps[j];
    }
    WNode &P = nodes[id];
    P.lx = lx; P.ly = ly; P.lz = lz;
    P.hx = hx; P.hy = hy; P.hz = hz;
    if (mt > 0.0) {
        P.cx = sx / mt; P.cy = sy / mt; P.cz = sz / mt;
    } else {
        P.cx = 0.5 * (lx + hx); P.cy = 0.5 * (ly + hy); P.cz = 0.5 * (lz + hz);
    }
    P.mass = mt;
    double sz3 = hx - lx;
    if (hy - ly > sz3) sz3 = hy - ly;
    if (hz - lz > sz3) sz3 = hz - lz;
    P.size = sz3;
    P.hmax = hm;
    P.epsmax = em;
}

// Climb from the finished node id: combine each parent whose range lies inside this
// workgroup's particle block when the second child arrives; hand the node to the frontier
// when the parent leaves the block.  Such a parent's id lies in the block too (an internal
// node's id is an end of its range), so its arrival counter is lflag[p - base] in LDS.
__device__ __forceinline__ void climb_local(WNode *__restrict__ nodes, const int32_t *__restrict__ right,
                                            const int32_t *__restrict__ parent, uint32_t *lflag, int64_t base,
                                            int32_t *flist, unsigned int *fn, int64_t id)
{
    int64_t p = parent[id];
    while (p >= 0) {
        const int64_t f = nodes[p].first, last = f + nodes[p].count - 1;
        if (f / MOM_BLOCK != last / MOM_BLOCK) {  // parent leaves the block: hand over
            front_push(flist, fn, (int32_t)id);
            return;
        }
        const uint32_t old =
            __hip_atomic_fetch_add(&lflag[p - base], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old == 0) return;  // first arrival: the sibling finishes this node
        combine(nodes, p, nodes[p].left, right[p]);
        id = p;
        p = parent[p];
    }
}

__global__ void __launch_bounds__(MOM_BLOCK) k_moments_local(int64_t n, int bucket, const double *__restrict__ x,
                                                             const double *__restrict__ y,
                                                             const double *__restrict__ z,
                                                             const double *__restrict__ m,
                                                             const double *__restrict__ h,
                                                             const double *__restrict__ eps,
                                                             WNode *__restrict__ nodes,
                                                             const int32_t *__restrict__ right,
                                                             const int32_t *__restrict__ parent,
                                                             const int32_t *__restrict__ ncount,
                                                             uint32_t *__restrict__ flags,
                                                             int32_t *__restrict__ reg, int32_t *__restrict__ cnt)
{
    __shared__ int32_t flist[MOM_FMAX];
    __shared__ uint32_t lflag[MOM_BLOCK];  // arrival counters of this block's internal nodes
    __shared__ unsigned int fn;
    if (threadIdx.x == 0) fn = 0;
    lflag[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t nl = n - 1;
    const int64_t base = blockIdx.x * (int64_t)MOM_BLOCK;
    const int64_t i = base + threadIdx.x;
    if (n == 1) {
        if (i == 0) bucket_record(nodes, 0, 0, 1, x, y, z, m, h, eps);  // the root is the single leaf
    } else if (i < nl) {
        const int ci = ncount[i];
        if (ci <= bucket) {
            const int64_t pi = parent[i];
            if (pi < 0 || ncount[pi] > bucket) {
                bucket_record(nodes, i, nodes[i].first, ci, x, y, z, m, h, eps);
                climb_local(nodes, right, parent, lflag, base, flist, &fn, i);
            }
        } else {
            // a leaf child of a node above the buckets is a bucket of one particle
            const int64_t c[2] = {nodes[i].left, right[i]};
            for (int k = 0; k < 2; k++)
                if (c[k] >= nl) {
                    bucket_record(nodes, c[k], (int)(c[k] - nl), 1, x, y, z, m, h, eps);
                    climb_local(nodes, right, parent, lflag, base, flist, &fn, c[k]);
                }
        }
    }
    __syncthreads();
    front_flush(flist, fn, reg, cnt, blockIdx.x);
}

// Climb from the frontier regions of `fan` consecutive blocks of the level
// below (region b: reg_in[b*MOM_FMAX + i], i < cnt_in[b]).  span > 0: stay
// inside this workgroup's span particles and hand the rest to its region of
// reg_out/cnt_out; span <= 0: climb to the root (single workgroup).
__global__ void __launch_bounds__(1024) k_moments_up(int64_t nblk_in, int fan, int64_t span,
                                                     WNode *__restrict__ nodes, const int32_t *__restrict__ right,
                                                     const int32_t *__restrict__ parent,
                                                     uint32_t *__restrict__ flags, const int32_t *__restrict__ reg_in,
                                                     const int32_t *__restrict__ cnt_in, int32_t *__restrict__ reg_out,
                                                     int32_t *__restrict__ cnt_out)
{
    __shared__ int32_t pre[1025];
    __shared__ int32_t flist[MOM_FMAX];
    __shared__ unsigned int fn;
    const int64_t b0 = blockIdx.x * (int64_t)fan;
    const int nb = (int)((nblk_in - b0) < fan ? (nblk_in - b0) : fan);
    if (threadIdx.x == 0) {
        fn = 0;
        pre[0] = 0;
    }
    if ((int)threadIdx.x < nb) pre[threadIdx.x + 1] = cnt_in[b0 + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int j = 0; j < nb; j++) pre[j + 1] += pre[j];
    __syncthreads();
    const int T = pre[nb];
    for (int e = threadIdx.x; e < T; e += blockDim.x) {
        int lo = 0, hi = nb;  // pre[lo] <= e < pre[lo + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (pre[mid] <= e) lo = mid;
            else hi = mid;
        }
        int64_t id = reg_in[(b0 + lo) * MOM_FMAX + (e - pre[lo])];
        int64_t p = parent[id];
        while (p >= 0) {
            if (span > 0) {
                const int64_t f = nodes[p].first, last = f + nodes[p].count - 1;
                if (f / span != last / span) {
                    front_push(flist, &fn, (int32_t)id);
                    break;
                }
            }
            const uint32_t old =
                __hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0) break;
            combine(nodes, p, nodes[p].left, right[p]);
            id = p;
            p = parent[p];
        }
    }
    if (span > 0) {
        __syncthreads();
        front_flush(flist, fn, reg_out, cnt_out, blockIdx.x);
    }
}

namespace sph {

hipError_t tree_topology(hipStream_t st, int64_t n, const uint64_t *skeys, WNode *nodes, int32_t *right,
                         int32_t *parent, int32_t *ncount)
{
    k_topology<<<sph_blocks(n, 256), 256, 0, st>>>(n, skeys, nodes, right, parent, ncount);
    return hipGetLastError();
}

hipError_t tree_escape(hipStream_t st, int64_t n, WNode *nodes)
{
    k_escape<<<sph_blocks(2 * n - 1, 256), 256, 0, st>>>(n, nodes);
    return hipGetLastError();
}

int64_t moments_scratch_entries(int64_t n)
{
    const int64_t b1 = (n + MOM_BLOCK - 1) / MOM_BLOCK, b2 = (b1 + MOM_FAN - 1) / MOM_FAN;
    return (b1 + b2) * (MOM_FMAX + 1);
}

hipError_t tree_moments(hipStream_t st, int64_t n, int bucket, const double *x, const double *y, const double *z,
                        const double *m, const double *h, const double *eps, WNode *nodes,
                        const int32_t *right, const int32_t *parent, const int32_t *ncount, uint32_t *flags,
                        int32_t *frontier)
{
    if (bucket < 1) return hipErrorInvalidValue;
    // flags[0..n-2]: arrival counters of the internal nodes
    const int64_t b1 = (n + MOM_BLOCK - 1) / MOM_BLOCK, b2 = (b1 + MOM_FAN - 1) / MOM_FAN;
    if (b2 > 1024) return hipErrorInvalidValue;  // the last level's prefix table holds 1024 regions
    int32_t *reg1 = frontier, *reg2 = reg1 + b1 * MOM_FMAX, *cnt1 = reg2 + b2 * MOM_FMAX, *cnt2 = cnt1 + b1;
    if (n > 1) fill<uint32_t>(st, n - 1, flags, 0u);
    k_moments_local<<<(unsigned)b1, MOM_BLOCK, 0, st>>>(n, bucket, x, y, z, m, h, eps, nodes, right, parent, ncount,
                                                        flags, reg1, cnt1);
    if (b1 > 1) {
        k_moments_up<<<(unsigned)b2, 1024, 0, st>>>(b1, MOM_FAN, (int64_t)MOM_BLOCK * MOM_FAN, nodes, right, parent,
                                                    flags, reg1, cnt1, reg2, cnt2);
        if (b2 > 1)
            k_moments_up<<<1, 1024, 0, st>>>(b2, (int)b2, 0, nodes, right, parent, flags, reg2, cnt2, nullptr,
                                             nullptr);
    }
    return hipGetLastError();
}

}  // namespace sph
