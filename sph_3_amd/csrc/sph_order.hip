// sph_order.hip -- spatial ordering: bounding box, Morton keys, LSD radix
// sort (hand-written, LDS-staged, stable), prefix scans and the gathers that
// move particle data into Morton order.
//
// Replaces the reference's single-block grid counting sort
// (EV/ngb.h:70-118 CountBodies/ComputeOffset/GroupBodies,
// EV/bh.h:76-145 compute_bounding_box_kernel) with one Morton order that
// serves both the neighbour search and the gravity tree.
#include "sph_order.h"

#pragma clang fp contract(off)

// ---------------------------------------------------------------- bbox
// Stage 1: per-block min/max of x, y, z (FP64, exact regardless of order).
__global__ void __launch_bounds__(256) k_bbox_partial(int64_t n, const double *__restrict__ x,
                                                      const double *__restrict__ y,
                                                      const double *__restrict__ z,
                                                      double *__restrict__ part)
{
    __shared__ double s[6][256];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double a = x[i], b = y[i], c = z[i];
        v[0] = fmin(v[0], a); v[1] = fmin(v[1], b); v[2] = fmin(v[2], c);
        v[3] = fmax(v[3], a); v[4] = fmax(v[4], b); v[5] = fmax(v[5], c);
    }
    for (int d = 0; d < 6; d++) s[d][threadIdx.x] = v[d];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            for (int d = 0; d < 3; d++) s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + off]);
            for (int d = 3; d < 6; d++) s[d][threadIdx.x] = fmax(s[d][threadIdx.x], s[d][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// Stage 2: one block folds the partials; writes box[0..5] and box[6] = scale
// = 65536 / (max extent), matching the oracle's key definition.
__global__ void __launch_bounds__(256) k_bbox_final(int nparts, const double *__restrict__ part,
                                                    double *__restrict__ box)
{
    __shared__ double s[6][256];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) {
        for (int d = 0; d < 3; d++) v[d] = fmin(v[d], part[p * 6 + d]);
        for (int d = 3; d < 6; d++) v[d] = fmax(v[d], part[p * 6 + d]);
    }
    for (int d = 0; d < 6; d++) s[d][threadIdx.x] = v[d];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            for (int d = 0; d < 3; d++) s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + off]);
            for (int d = 3; d < 6; d++) s[d][threadIdx.x] = fmax(s[d][threadIdx.x], s[d][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int d = 0; d < 6; d++) box[d] = s[d][0];
        double L = box[3] - box[0];
        if (box[4] - box[1] > L) L = box[4] - box[1];
        if (box[5] - box[2] > L) L = box[5] - box[2];
        box[6] = (L > 0.0) ? 65536.0 / L : 0.0;
    }
}

// ---------------------------------------------------------------- keys
__device__ __forceinline__ uint64_t spread16(uint64_t v)
{
    v &= 0xFFFFull;
    v = (v | (v << 16)) & 0x0000FF0000FFull;
    v = (v | (v << 8)) & 0x00F00F00F00Full;
    v = (v | (v << 4)) & 0x0C30C30C30C3ull;
    v = (v | (v << 2)) & 0x249249249249ull;
    return v;
}

__device__ __forceinline__ uint32_t cell16(double x, double lo, double scale)
{
    double t = (x - lo) * scale;
    if (!(t > 0.0)) return 0;
    if (t >= 65535.0) return 65535;
    return (uint32_t)t;
}

__global__ void __launch_bounds__(256) k_morton(int64_t n, const double *__restrict__ x,
                                                const double *__restrict__ y,
                                                const double *__restrict__ z,
                                                const double *__restrict__ box,
                                                uint64_t *__restrict__ keys,
                                                int32_t *__restrict__ vals)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sc = box[6];
    uint64_t cx = cell16(x[i], box[0], sc), cy = cell16(y[i], box[1], sc), cz = cell16(z[i], box[2], sc);
    keys[i] = (spread16(cx) << 2) | (spread16(cy) << 1) | spread16(cz);
    vals[i] = (int32_t)i;
}

// ---------------------------------------------------------------- radix sort
// 8-bit digits, tiles of RS_TILE keys per 256-thread block.  Stable: within a
// tile, keys are ranked chunk by chunk (256 keys per chunk, in order) with
// 64-lane ballots to find equal-digit peers.
#define RS_BLOCK 256
#define RS_CHUNKS 16
#define RS_TILE (RS_BLOCK * RS_CHUNKS)

__global__ void __launch_bounds__(RS_BLOCK) k_rs_hist(int64_t n, const uint64_t *__restrict__ keys,
                                                      int shift, int nblk,
                                                      uint32_t *__restrict__ hist)
{
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int c = 0; c < RS_CHUNKS; c++) {
        int64_t i = base + c * RS_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

__global__ void __launch_bounds__(RS_BLOCK) k_rs_scatter(int64_t n, const uint64_t *__restrict__ kin,
                                                         const int32_t *__restrict__ vin,
                                                         uint64_t *__restrict__ kout,
                                                         int32_t *__restrict__ vout, int shift,
                                                         int nblk,
                                                         const uint32_t *__restrict__ hscan)
{
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[RS_BLOCK / SPH_WAVE][256];
    const int t = threadIdx.x, w = t / SPH_WAVE, lane = t % SPH_WAVE;
    run[t] = hscan[(int64_t)t * nblk + blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int c = 0; c < RS_CHUNKS; c++) {
        for (int k = 0; k < RS_BLOCK / SPH_WAVE; k++) wcnt[k][t] = 0;
        __syncthreads();
        int64_t i = base + c * RS_BLOCK + t;
        bool valid = i < n;
        uint64_t key = valid ? kin[i] : 0;
        int32_t val = valid ? vin[i] : 0;
        uint32_t d = (uint32_t)((key >> shift) & 0xFF);
        uint64_t peers = wave_ballot(valid);
        for (int b = 0; b < 8; b++) {
            uint64_t m = wave_ballot(((d >> b) & 1u) != 0);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        // highest lane of each peer group publishes the group's size
        if (valid && (peers >> lane) == 1ull) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = run[d] + rank;
            for (int k = 0; k < w; k++) off += wcnt[k][d];
            kout[off] = key;
            vout[off] = val;
        }
        __syncthreads();
        uint32_t tot = 0;
        for (int k = 0; k < RS_BLOCK / SPH_WAVE; k++) tot += wcnt[k][t];
        run[t] += tot;
        __syncthreads();
    }
}

// Exclusive scan of the digit-major histogram hist[d * nblk + b]: one
// workgroup per digit scans its nblk counts and publishes the digit total;
// a second launch adds the prefix of the digit totals.
__global__ void __launch_bounds__(256) k_rs_scan_digit(int nblk, uint32_t *__restrict__ hist,
                                                       uint32_t *__restrict__ dtot)
{
    __shared__ uint32_t part[256];
    uint32_t *a = hist + (int64_t)blockIdx.x * nblk;
    int per = (nblk + 255) / 256;
    int b = threadIdx.x * per, e = b + per < nblk ? b + per : nblk;
    uint32_t s = 0;
    for (int i = b; i < e; i++) s += a[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (int i = b; i < e; i++) {
        uint32_t v = a[i];
        a[i] = run;
        run += v;
    }
    if (threadIdx.x == 255) dtot[blockIdx.x] = part[255];
}

__global__ void __launch_bounds__(256) k_rs_scan_base(int nblk, uint32_t *__restrict__ hist,
                                                      const uint32_t *__restrict__ dtot)
{
    // base = sum of the totals of the digits below this one (one load per thread, wave
    // sums, then the four wave sums)
    __shared__ uint32_t wsum[256 / SPH_WAVE];
    uint32_t v = threadIdx.x < blockIdx.x ? dtot[threadIdx.x] : 0u;
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & (SPH_WAVE - 1)) == 0) wsum[threadIdx.x / SPH_WAVE] = v;
    __syncthreads();
    uint32_t base = 0;
    for (int w = 0; w < 256 / SPH_WAVE; w++) base += wsum[w];
    uint32_t *a = hist + (int64_t)blockIdx.x * nblk;
    for (int i = threadIdx.x; i < nblk; i += 256) a[i] += base;
}

// ---------------------------------------------------------------- scans
// Exclusive scan of uint32 -> uint32 (radix histograms) in one block.
__global__ void __launch_bounds__(1024) k_scan_u32_small(int64_t n, uint32_t *__restrict__ a)
{
    __shared__ uint32_t part[1024];
    int64_t per = (n + 1023) / 1024;
    int64_t b = threadIdx.x * per, e = b + per < n ? b + per : n;
    uint32_t s = 0;
    for (int64_t i = b; i < e; i++) s += a[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (int64_t i = b; i < e; i++) {
        uint32_t v = a[i];
        a[i] = run;
        run += v;
    }
}

// Exclusive scan of int32 counts into int64 offsets (n+1 entries), three
// launches: block sums, scan of block sums, block-local scan + add.
#define SCAN_BS 256
#define SCAN_PER 8
#define SCAN_TILE (SCAN_BS * SCAN_PER)

__global__ void __launch_bounds__(SCAN_BS) k_scan_blocksum(int64_t n, const int32_t *__restrict__ in,
                                                           int64_t *__restrict__ bsum)
{
    __shared__ int64_t s[SCAN_BS];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t v = 0;
    for (int k = 0; k < SCAN_PER; k++) {
        int64_t i = base + k * SCAN_BS + threadIdx.x;
        if (i < n) v += in[i];
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = SCAN_BS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(1024) k_scan_i64_small(int64_t n, int64_t *__restrict__ a)
{
    __shared__ int64_t part[1024];
    int64_t per = (n + 1023) / 1024;
    int64_t b = threadIdx.x * per, e = b + per < n ? b + per : n;
    int64_t s = 0;
    for (int64_t i = b; i < e; i++) s += a[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int64_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (int64_t i = b; i < e; i++) {
        int64_t v = a[i];
        a[i] = run;
        run += v;
    }
}

__global__ void __launch_bounds__(SCAN_BS) k_scan_apply(int64_t n, const int32_t *__restrict__ in,
                                                        const int64_t *__restrict__ bscan,
                                                        int64_t *__restrict__ out)
{
    __shared__ int64_t s[SCAN_TILE];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    for (int k = 0; k < SCAN_PER; k++) {
        int64_t i = base + k * SCAN_BS + threadIdx.x;
        s[k * SCAN_BS + threadIdx.x] = (i < n) ? in[i] : 0;
    }
    __syncthreads();
    // each thread scans SCAN_PER consecutive entries, then a block scan of totals
    __shared__ int64_t tot[SCAN_BS];
    int64_t run = 0;
    for (int k = 0; k < SCAN_PER; k++) {
        int64_t v = s[threadIdx.x * SCAN_PER + k];
        s[threadIdx.x * SCAN_PER + k] = run;
        run += v;
    }
    tot[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < SCAN_BS; off <<= 1) {
        int64_t v = threadIdx.x >= (unsigned)off ? tot[threadIdx.x - off] : 0;
        __syncthreads();
        tot[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t pre = bscan[blockIdx.x] + (threadIdx.x ? tot[threadIdx.x - 1] : 0);
    for (int k = 0; k < SCAN_PER; k++) s[threadIdx.x * SCAN_PER + k] += pre;
    __syncthreads();
    for (int k = 0; k < SCAN_PER; k++) {
        int64_t i = base + k * SCAN_BS + threadIdx.x;
        if (i < n) out[i] = s[k * SCAN_BS + threadIdx.x];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        int64_t last = n - 1;
        out[n] = (n > 0) ? s[last - base] + in[last] : 0;
    }
}

// ---------------------------------------------------------------- gathers
template <typename T, typename U>
__global__ void __launch_bounds__(256) k_gather(int64_t n, const int32_t *__restrict__ perm,
                                                const T *__restrict__ src, int stride, int comp,
                                                U *__restrict__ dst)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (U)src[(int64_t)perm[i] * stride + comp];
}

// Up to SPH_GATHER_MAX FP64 arrays and one optional int32 array in one pass:
// dst_k[i] = src_k[perm[i]], idst[i] = isrc[perm[i]].
struct GatherSet {
    const double *src[SPH_GATHER_MAX];
    double *dst[SPH_GATHER_MAX];
    const int32_t *isrc;
    int32_t *idst;
    int k;
};

__global__ void __launch_bounds__(256) k_gather_set(int64_t n, const int32_t *__restrict__ perm, GatherSet g)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t s = perm[i];
    double v[SPH_GATHER_MAX];
#pragma unroll
    for (int k = 0; k < SPH_GATHER_MAX; k++)
        if (k < g.k) v[k] = g.src[k][s];  // all loads issued before the stores
#pragma unroll
    for (int k = 0; k < SPH_GATHER_MAX; k++)
        if (k < g.k) g.dst[k][i] = v[k];
    if (g.idst) g.idst[i] = g.isrc[s];
}

template <typename T>
__global__ void __launch_bounds__(256) k_scatter(int64_t n, const int32_t *__restrict__ perm,
                                                 const T *__restrict__ src, int stride, int comp,
                                                 double *__restrict__ dst)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[(int64_t)perm[i] * stride + comp] = (double)src[i];
}

__global__ void __launch_bounds__(256) k_compose(int64_t n, const int32_t *__restrict__ a,
                                                 const int32_t *__restrict__ b,
                                                 int32_t *__restrict__ out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[b[i]];
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill(int64_t n, T *__restrict__ a, T v)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ================================================================ host side
namespace sph {

hipError_t bbox(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                double *part, double *box)
{
    int nb = (int)((n + 255) / 256);
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    k_bbox_partial<<<nb, 256, 0, st>>>(n, x, y, z, part);
    k_bbox_final<<<1, 256, 0, st>>>(nb, part, box);
    return hipGetLastError();
}

hipError_t morton(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                  const double *box, uint64_t *keys, int32_t *vals)
{
    k_morton<<<sph_blocks(n, 256), 256, 0, st>>>(n, x, y, z, box, keys, vals);
    return hipGetLastError();
}

int64_t radix_hist_entries(int64_t n) { return 256 * ((n + RS_TILE - 1) / RS_TILE) + 256; }

hipError_t radix_sort(hipStream_t st, int64_t n, int nbits, uint64_t *k0, int32_t *v0,
                      uint64_t *k1, int32_t *v1, uint32_t *hist)
{
    int nblk = (int)((n + RS_TILE - 1) / RS_TILE);
    if (nblk < 1) return hipSuccess;
    for (int shift = 0; shift < nbits; shift += 8) {
        k_rs_hist<<<nblk, RS_BLOCK, 0, st>>>(n, k0, shift, nblk, hist);
        k_rs_scan_digit<<<256, 256, 0, st>>>(nblk, hist, hist + (int64_t)256 * nblk);
        k_rs_scan_base<<<256, 256, 0, st>>>(nblk, hist, hist + (int64_t)256 * nblk);
        k_rs_scatter<<<nblk, RS_BLOCK, 0, st>>>(n, k0, v0, k1, v1, shift, nblk, hist);
        uint64_t *tk = k0; k0 = k1; k1 = tk;
        int32_t *tv = v0; v0 = v1; v1 = tv;
    }
    return hipGetLastError();
}

int64_t scan_tmp_entries(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

hipError_t scan_counts(hipStream_t st, int64_t n, const int32_t *cnt, int64_t *off, int64_t *tmp)
{
    int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb < 1) {
        k_fill<int64_t><<<1, 1, 0, st>>>(1, off, 0);
        return hipGetLastError();
    }
    k_scan_blocksum<<<(unsigned)nb, SCAN_BS, 0, st>>>(n, cnt, tmp);
    k_scan_i64_small<<<1, 1024, 0, st>>>(nb, tmp);
    k_scan_apply<<<(unsigned)nb, SCAN_BS, 0, st>>>(n, cnt, tmp, off);
    return hipGetLastError();
}

template <typename T, typename U>
hipError_t gather(hipStream_t st, int64_t n, const int32_t *perm, const T *src, int stride, int comp, U *dst)
{
    k_gather<T, U><<<sph_blocks(n, 256), 256, 0, st>>>(n, perm, src, stride, comp, dst);
    return hipGetLastError();
}

template <typename T>
hipError_t scatter(hipStream_t st, int64_t n, const int32_t *perm, const T *src, int stride, int comp, double *dst)
{
    k_scatter<T><<<sph_blocks(n, 256), 256, 0, st>>>(n, perm, src, stride, comp, dst);
    return hipGetLastError();
}

hipError_t gather_set(hipStream_t st, int64_t n, const int32_t *perm, int k, const double *const *src,
                      double *const *dst, const int32_t *isrc, int32_t *idst)
{
    if (k < 0 || k > SPH_GATHER_MAX) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    GatherSet g{};
    for (int j = 0; j < k; j++) {
        g.src[j] = src[j];
        g.dst[j] = dst[j];
    }
    g.k = k;
    g.isrc = isrc;
    g.idst = idst;
    k_gather_set<<<sph_blocks(n, 256), 256, 0, st>>>(n, perm, g);
    return hipGetLastError();
}

hipError_t compose(hipStream_t st, int64_t n, const int32_t *a, const int32_t *b, int32_t *out)
{
    k_compose<<<sph_blocks(n, 256), 256, 0, st>>>(n, a, b, out);
    return hipGetLastError();
}

// One wavefront per counter: lane l sums slot l (SPH_SPREAD == 64).
__global__ void __launch_bounds__(64) k_spread_total(const unsigned long long *__restrict__ c,
                                                     unsigned long long *__restrict__ out)
{
    static_assert(SPH_SPREAD == 64, "one lane per slot");
    unsigned long long v = c[(size_t)blockIdx.x * SPH_SPREAD_WORDS + threadIdx.x * SPH_SPREAD_PAD];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

hipError_t spread_total(hipStream_t st, int ncnt, const unsigned long long *c, unsigned long long *out)
{
    if (ncnt <= 0) return hipSuccess;
    k_spread_total<<<ncnt, 64, 0, st>>>(c, out);
    return hipGetLastError();
}

template <typename T>
hipError_t fill(hipStream_t st, int64_t n, T *a, T v)
{
    if (n <= 0) return hipSuccess;
    k_fill<T><<<sph_blocks(n, 256), 256, 0, st>>>(n, a, v);
    return hipGetLastError();
}

template hipError_t gather<double, double>(hipStream_t, int64_t, const int32_t *, const double *, int, int, double *);
template hipError_t gather<double, float>(hipStream_t, int64_t, const int32_t *, const double *, int, int, float *);
template hipError_t gather<float, float>(hipStream_t, int64_t, const int32_t *, const float *, int, int, float *);
template hipError_t gather<int32_t, int32_t>(hipStream_t, int64_t, const int32_t *, const int32_t *, int, int, int32_t *);
template hipError_t scatter<double>(hipStream_t, int64_t, const int32_t *, const double *, int, int, double *);
template hipError_t scatter<float>(hipStream_t, int64_t, const int32_t *, const float *, int, int, double *);
template hipError_t fill<int32_t>(hipStream_t, int64_t, int32_t *, int32_t);
template hipError_t fill<int64_t>(hipStream_t, int64_t, int64_t *, int64_t);
template hipError_t fill<uint32_t>(hipStream_t, int64_t, uint32_t *, uint32_t);
template hipError_t fill<double>(hipStream_t, int64_t, double *, double);
template hipError_t fill<float>(hipStream_t, int64_t, float *, float);
template hipError_t fill<unsigned long long>(hipStream_t, int64_t, unsigned long long *, unsigned long long);

}  // namespace sph
