// sph_common.h -- shared device definitions for the gfx950 SPH step library.
//
// Layout conventions (all device arrays, one GPU):
//   * particles live in Morton order ("sorted index" s); orig[s] is the
//     caller's index of sorted particle s;
//   * positions, velocities, masses, smoothing lengths, softenings and
//     accelerations are FP64 structure-of-arrays;
//   * hydro scalars (u, rho, P, c, P/rho^2, divV, curlV, Balsara f, du/dt) are
//     stored in the arithmetic type R of the pass: float for the mixed form,
//     double for the all-FP64 form (each buffer is allocated 8 bytes/particle);
//   * the tree has 2N-1 node ids: internal k -> k (0 <= k < N-1), leaf of
//     sorted particle s -> N-1+s; the root is id 0.  Walks follow escape
//     links (pre-order successor after a subtree), so they need no stack.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SPH_WAVE 64
#define SPH_PI 3.14159265358979323846
// hydrogen atom mass [g] of the column-density formula (the value of
// Particles_in_BH_Tree_III/CLOUDY_reloaded_Here/run_simulation/hfvGpuSphBHtree_gradRho_v3t.cu:42)
#define SPH_M_H 1.673534e-24

// Packed per-node record read by the walks with wave-uniform (scalar) loads.
struct __align__(16) WNode {
    double cx, cy, cz, mass;   // centre of mass (FP64) and mass
    double lx, ly, lz, size;   // box low corner, max box extent
    double hx, hy, hz, hmax;   // box high corner, max smoothing length
    double epsmax;             // max softening length
    int first, count;          // sorted-particle range [first, first+count)
    int left, escape;          // left child id (internal), pre-order successor
    int right, pad;            // right child id (internal)
};
static_assert(sizeof(WNode) == 128, "WNode must be 128 bytes");

__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

// Error-free transformation a + b = s + e (Knuth TwoSum).
template <typename R>
__device__ __forceinline__ void two_sum(R a, R b, R &s, R &e)
{
    s = a + b;
    R bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Kahan-compensated accumulator: value = s - c.
template <typename R>
struct Kahan {
    R s, c;
    __device__ __forceinline__ Kahan() : s(0), c(0) {}
    __device__ __forceinline__ void add(R x)
    {
        R y = x - c;
        R t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ R value() const { return s - c; }
};

// Reduce a compensated value over groups of G consecutive lanes (G | 64) with
// TwoSum at every level; every lane of the group gets the result.
template <typename R, int G>
__device__ __forceinline__ R group_sum(R v, R err)
{
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) {
        R ov = __shfl_xor(v, off, G);
        R oe = __shfl_xor(err, off, G);
        R s, e;
        two_sum(v, ov, s, e);
        v = s;
        err = err + oe + e;
    }
    return v + err;
}

// Cubic spline, sph_code/libsx.py:29-39 (W) and :130-144 (grad W factor).
template <typename R>
__device__ __forceinline__ R spline_W(R rr, R hij)
{
    const R sig = (R)(1.0 / SPH_PI);
    R q = rr / hij, h3 = hij * hij * hij, W = 0;
    if (q <= (R)1) W = sig / h3 * ((R)1 - (R)1.5 * q * q + (R)0.75 * q * q * q);
    else if (q <= (R)2) { R t = (R)2 - q; W = sig / h3 * (R)0.25 * t * t * t; }
    return W;
}

template <typename R>
__device__ __forceinline__ R spline_nW(R rr, R hij)
{
    const R sig = (R)(1.0 / SPH_PI);
    R q = rr / hij, h2 = hij * hij, h5 = h2 * h2 * hij, nW = 0;
    if (q <= (R)1) nW = sig / h5 * ((R)-3 + (R)2.25 * q);
    else if (q <= (R)2) { R t = (R)2 - q; nW = (R)-0.75 * sig / h5 * t * t / (q + (R)1e-20); }
    return nW;
}

// FP32 forms of the two kernels for the mixed-precision passes: one hardware reciprocal of
// h_ij (v_rcp_f32, 1 ulp) instead of the correctly rounded divisions, and both pieces
// evaluated and selected instead of branching (the pieces vanish past q = 2 by t = 0).
// Within a few ulp of the generic forms, which the FP64 passes keep.
template <>
__device__ __forceinline__ float spline_W<float>(float rr, float hij)
{
    const float sig = (float)(1.0 / SPH_PI);
    const float ih = __builtin_amdgcn_rcpf(hij), q = rr * ih;
    const float w1 = 1.0f - 1.5f * q * q + 0.75f * q * q * q;
    const float t = fmaxf(2.0f - q, 0.0f);
    const float w2 = 0.25f * t * t * t;
    return sig * (ih * ih * ih) * (q <= 1.0f ? w1 : w2);
}

template <>
__device__ __forceinline__ float spline_nW<float>(float rr, float hij)
{
    const float sig = (float)(1.0 / SPH_PI);
    const float ih = __builtin_amdgcn_rcpf(hij), q = rr * ih, ih2 = ih * ih;
    const float n1 = -3.0f + 2.25f * q;
    const float t = fmaxf(2.0f - q, 0.0f);
    const float n2 = -0.75f * t * t * __builtin_amdgcn_rcpf(q + 1e-20f);
    return sig * (ih2 * ih2 * ih) * (q <= 1.0f ? n1 : n2);
}

// Spline-softened gravity factor, sph_code/libsx.py:640-652.
template <typename R>
__device__ __forceinline__ R grav_fk_spline(R rr, R e)
{
    R q = rr / e;
    if (q <= (R)1) return ((R)1 / (e * e * e)) * ((R)(4.0 / 3.0) - (R)1.2 * q * q + (R)0.5 * q * q * q);
    R inv_r3 = (R)1 / (rr * rr * rr);
    if (q <= (R)2) {
        R q2 = q * q, q3 = q2 * q;
        return inv_r3 * ((R)(-1.0 / 15.0) + (R)(8.0 / 3.0) * q3 - (R)3 * q3 * q + (R)1.2 * q3 * q2 -
                         (R)(1.0 / 6.0) * q3 * q3);
    }
    return inv_r3;
}

// Spline-softened potential per unit source mass, psi(r, e) with phi_i = G sum_j m_j psi(r_ij, e_ij)
// and PE = -1/2 sum_i m_i phi_i: sph_code/libsx.py:597-613 (getPE's fk / m_j).  Coincident pairs
// (q = 0) add nothing there, and here.
template <typename R>
__device__ __forceinline__ R grav_pot_spline(R rr, R e)
{
    R q = rr / e;
    if (!(rr > (R)0)) return (R)0;
    if (q <= (R)1) {
        R q2 = q * q;
        return ((R)-2 / e) * ((R)(1.0 / 3.0) * q2 - (R)0.15 * q2 * q2 + (R)0.05 * q2 * q2 * q) + (R)1.4 / e;
    }
    if (q <= (R)2) {
        R q2 = q * q, q3 = q2 * q;
        return (R)(-1.0 / 15.0) / rr -
               ((R)1 / e) * ((R)(4.0 / 3.0) * q2 - q3 + (R)0.3 * q3 * q - (R)(1.0 / 30.0) * q3 * q2) + (R)1.6 / e;
    }
    return (R)1 / rr;
}

template <typename R> __device__ __forceinline__ R sph_sqrt(R x);
// FP32: the hardware square root (v_sqrt_f32, 1 ulp) without the correctly rounded expansion
template <> __device__ __forceinline__ float sph_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <> __device__ __forceinline__ double sph_sqrt<double>(double x) { return __builtin_sqrt(x); }

template <typename R> __device__ __forceinline__ R sph_rsqrt(R x);
template <> __device__ __forceinline__ float sph_rsqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <> __device__ __forceinline__ double sph_rsqrt<double>(double x) { return 1.0 / __builtin_sqrt(x); }

// Workgroups are dispatched round-robin over the eight XCDs (XCD = blockIdx % 8),
// each with its own L2.  xcd_block(b, nb) is a bijection of [0, nb) that gives
// each XCD one contiguous range of work items, so spatially adjacent (Morton
// consecutive) groups share an L2 instead of being spread over all eight.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb)
{
    const int64_t q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// Minimum-image difference of two coordinates in a periodic box of side L
// (L = 0: open boundaries): the wrap is decided on the FP64 difference, one
// rounding per operation, identically in every kernel and in the oracle.
// PER = false compiles to nothing (kernels are instantiated for both).
template <bool PER>
__device__ __forceinline__ double wrap_d(double d, double L)
{
    if constexpr (PER) {
        if (d > 0.5 * L) d = d - L;
        else if (d < -0.5 * L) d = d + L;
    }
    return d;
}

// Host-side launch helper.
static inline unsigned sph_blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Event counters bumped by many wavefronts.  Same-address device atomics
// serialise (about 10 ns each), so each logical counter k owns SPH_SPREAD
// slots 64 bytes apart and a workgroup adds into slot blockIdx.x % SPH_SPREAD;
// sph::spread_total folds the slots.  A spread array of c counters holds
// c * SPH_SPREAD_WORDS 64-bit words and must be cleared before use.
#define SPH_SPREAD 64
#define SPH_SPREAD_PAD 8
#define SPH_SPREAD_WORDS (SPH_SPREAD * SPH_SPREAD_PAD)

__device__ __forceinline__ void spread_add(unsigned long long *c, int k, unsigned long long v)
{
    atomicAdd(c + (size_t)k * SPH_SPREAD_WORDS + (blockIdx.x & (SPH_SPREAD - 1)) * SPH_SPREAD_PAD, v);
}

namespace sph {
// out[k] = sum of the slots of counter k, k < ncnt (device pointers, stream-ordered).
hipError_t spread_total(hipStream_t st, int ncnt, const unsigned long long *c, unsigned long long *out);
}  // namespace sph
