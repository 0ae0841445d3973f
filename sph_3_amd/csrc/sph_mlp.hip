// sph_mlp.hip -- learned surrogate of the per-particle cooling rate: an MLP
// 4 -> 128 -> 128 -> 128 -> 1 with ReLU activations mapping the normalised
// (log10 n_H, log10 T, log10 Z, log10 J21) to the normalised log10 Lambda.
//
// Design point: the Keras surrogate of
// Particles_in_BH_Tree_III/CLOUDY_replace_CHIMES/HCooling_Here/ML/prog2.py:24-52
// (4 -> 128 -> 64 -> 1 with ReLU activations, Adam on the mean squared error), which the
// reference trains and runs offline only.  ReLU costs one maximum per hidden unit, so the
// two MFMA layers are not held back by activation arithmetic on the VALU.  Here training and inference are device kernels:
//   * weights and activations BF16, accumulation FP32; the two 128 x 128
//     layers run on v_mfma_f32_32x32x8_bf16, a workgroup of four wavefronts per
//     tile of 32 samples (each wavefront owns 32 of the 128 features of a
//     layer), activations staged in LDS;
//   * backward on MFMA as well (dh = dz W, dW = dz^T h); weight gradients are
//     summed over tiles with FP32 atomics; Adam keeps FP32 master weights and
//     refreshes the BF16 copies (row-major and transposed);
//   * inference gates each particle: outside the training box, or in a cell
//     whose measured deviation from the table exceeds the tolerance, the
//     result is the table's 4D multilinear interpolation instead.
//
// Numerics of one sample (what the CPU oracle emulates):
//   z1 = W1 x + b1 (FP32),           h1 = bf16(relu(z1))
//   z2 = bf16(W2) h1 + b2 (FP32 acc), h2 = bf16(relu(z2))   (same for layer 3)
//   y  = W4 . h3 + b4 (FP32)
//   backward: dz3 = bf16(dy W4 relu'(z3)); dz2 = bf16((dz3 bf16(W3)) relu'(z2));
//   dz1 = bf16((dz2 bf16(W2)) relu'(z1)); z_l are kept in BF16 for relu' (its sign).
//
// MFMA operand layout (32x32x8, wave64): lane l supplies A[l%32][4(l/32)+i]
// and B[4(l/32)+i][l%32], i = 0..3; accumulator r holds
// D[8(r/4) + 4(l/32) + r%4][l%32], r = 0..15.
#include "sph_mlp.h"
#include "sph_order.h"

#include <algorithm>

#include <hip/amd_detail/amd_hip_unsafe_atomics.h>

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

#define TS 32            // samples per wave tile
#define HW MLP_HIDDEN    // hidden width
#define LDP 136          // LDS row pitch in bf16 elements (272 B: rows 68 dwords apart)
#define TILE (TS * LDP)  // one [32][LDP] bf16 buffer

// flat parameter layout
#define O_W1 0
#define O_B1 512
#define O_W2 640
#define O_B2 (O_W2 + HW * HW)
#define O_W3 (O_B2 + HW)
#define O_B3 (O_W3 + HW * HW)
#define O_W4 (O_B3 + HW)
#define O_B4 (O_W4 + HW)
static_assert(O_B4 + 1 == MLP_NPARAM, "parameter layout");

__device__ __forceinline__ unsigned short f2bf(float f)
{
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest, ties to even (finite inputs)
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }
// ReLU and its derivative (taken as 0 at z = 0).
__device__ __forceinline__ float relu(float z) { return fmaxf(z, 0.0f); }
__device__ __forceinline__ float drelu(float z) { return z > 0.0f ? 1.0f : 0.0f; }

__device__ __forceinline__ bf16x4 as_bf16x4(uint2 v)
{
    bf16x4 r;
    r[0] = (short)(v.x & 0xFFFF);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xFFFF);
    r[3] = (short)(v.y >> 16);
    return r;
}

__device__ __forceinline__ int drow(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// 32 x 32 tile of A (32 x 128, LDS rows of pitch LDP) times B (128 x 32) with
// B(k, n) = Wrow[n*128 + k] (a row-major BF16 matrix whose rows n are the
// tile's output columns).
__device__ __forceinline__ floatx16 tile_gemm(const unsigned short *A, const unsigned short *__restrict__ Wrow)
{
    const int lane = threadIdx.x & 63, r32 = lane & 31, hk = (lane >> 5) * 4;
    floatx16 acc = {};
#pragma unroll 4
    for (int kk = 0; kk < HW; kk += 8) {
        const bf16x4 a = as_bf16x4(*reinterpret_cast<const uint2 *>(A + r32 * LDP + kk + hk));
        const bf16x4 b = as_bf16x4(*reinterpret_cast<const uint2 *>(Wrow + (size_t)r32 * HW + kk + hk));
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// ---------------------------------------------------------------- inference
// Inference works on 64 samples per wavefront and computes each hidden layer
// transposed, Z^T = W H^T: the MFMA A operand is a 32 x 16 slice of the
// row-major BF16 weight (16-byte global loads), the B operand 16 features of
// 32 samples from the LDS activation rows, and each lane's accumulators hold
// 4 consecutive features of one sample, so the BF16 activations go back to
// LDS as 8-byte stores (in place: a layer's reads all precede its writes) and
// the output layer is a register dot product in the layer-3 epilogue.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

#define IS 64                 // samples per wave
#define ITILE (IS * LDP)      // activation buffer [64][LDP]

__device__ __forceinline__ unsigned int pack2(float a, float b)
{
    bf16x2 v;
    v[0] = (__bf16)a;  // v_cvt_pk_bf16_f32: round to nearest even
    v[1] = (__bf16)b;
    return __builtin_bit_cast(unsigned int, v);
}

__device__ __forceinline__ float bf16_round(float a) { return (float)(__bf16)a; }

// Normalised output of sample (blockIdx.x*64 + lane) for this lane's input xin.
__device__ __forceinline__ float mlp_forward64(const float xin[4], const float *__restrict__ P,
                                               const unsigned short *__restrict__ W2b,
                                               const unsigned short *__restrict__ W3b, unsigned short *act)
{
    const int lane = threadIdx.x & 63, l32 = lane & 31, half = lane >> 5;
    // layer 1 on VALU: this lane's sample, all 128 features (weights are wave-uniform)
    for (int f0 = 0; f0 < HW; f0 += 8) {
        float h[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int f = f0 + j;
            float z = P[O_B1 + f];
#pragma unroll
            for (int k = 0; k < 4; k++) z = fmaf(P[O_W1 + f * 4 + k], xin[k], z);
            h[j] = relu(z);
        }
        uint4 w;
        w.x = pack2(h[0], h[1]);
        w.y = pack2(h[2], h[3]);
        w.z = pack2(h[4], h[5]);
        w.w = pack2(h[6], h[7]);
        *reinterpret_cast<uint4 *>(act + lane * LDP + f0) = w;
    }
    __syncthreads();
    float y[2] = {0.0f, 0.0f};
#pragma unroll
    for (int L = 0; L < 2; L++) {
        const unsigned short *W = L == 0 ? W2b : W3b;
        floatx16 acc[4][2];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) acc[mt][0] = acc[mt][1] = floatx16{};
#pragma unroll 1
        for (int kk = 0; kk < HW; kk += 16) {
            bf16x8 bop[2];
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
                bop[nt] = *reinterpret_cast<const bf16x8 *>(act + (nt * 32 + l32) * LDP + kk + 8 * half);
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                const bf16x8 aop =
                    *reinterpret_cast<const bf16x8 *>(W + (size_t)(mt * 32 + l32) * HW + kk + 8 * half);
#pragma unroll
                for (int nt = 0; nt < 2; nt++)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aop, bop[nt], acc[mt][nt], 0, 0, 0);
            }
        }
        __syncthreads();
        const int ob = L == 0 ? O_B2 : O_B3;
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                // accumulators 4q..4q+3: features f..f+3 of samples nt*32 + l32
                const int f = mt * 32 + 8 * q + 4 * half;
                const float4 bias = *reinterpret_cast<const float4 *>(P + ob + f);
                if (L == 0) {
#pragma unroll
                    for (int nt = 0; nt < 2; nt++) {
                        const float h0 = relu(acc[mt][nt][4 * q] + bias.x);
                        const float h1 = relu(acc[mt][nt][4 * q + 1] + bias.y);
                        const float h2 = relu(acc[mt][nt][4 * q + 2] + bias.z);
                        const float h3 = relu(acc[mt][nt][4 * q + 3] + bias.w);
                        uint2 w;
                        w.x = pack2(h0, h1);
                        w.y = pack2(h2, h3);
                        *reinterpret_cast<uint2 *>(act + (nt * 32 + l32) * LDP + f) = w;
                    }
                } else {
                    const float4 w4 = *reinterpret_cast<const float4 *>(P + O_W4 + f);
#pragma unroll
                    for (int nt = 0; nt < 2; nt++) {
                        y[nt] = fmaf(w4.x, bf16_round(relu(acc[mt][nt][4 * q] + bias.x)), y[nt]);
                        y[nt] = fmaf(w4.y, bf16_round(relu(acc[mt][nt][4 * q + 1] + bias.y)), y[nt]);
                        y[nt] = fmaf(w4.z, bf16_round(relu(acc[mt][nt][4 * q + 2] + bias.z)), y[nt]);
                        y[nt] = fmaf(w4.w, bf16_round(relu(acc[mt][nt][4 * q + 3] + bias.w)), y[nt]);
                    }
                }
            }
        __syncthreads();
    }
    // lanes l and l^32 hold the two feature halves of samples l32 and 32 + l32
    y[0] += __shfl_xor(y[0], 32);
    y[1] += __shfl_xor(y[1], 32);
    return (half ? y[1] : y[0]) + P[O_B4];
}

// Network output for rows of normalised inputs x[n][4].
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_mlp_forward(int64_t n, const float *__restrict__ x, const float *__restrict__ P,
              const unsigned short *__restrict__ W2b, const unsigned short *__restrict__ W3b, float *__restrict__ y)
{
    __shared__ __align__(16) unsigned short act[ITILE];
    const int64_t s = blockIdx.x * (int64_t)IS + threadIdx.x;
    float xin[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (s < n) {
        const float4 v = *reinterpret_cast<const float4 *>(x + 4 * s);
        xin[0] = v.x; xin[1] = v.y; xin[2] = v.z; xin[3] = v.w;
    }
    const float out = mlp_forward64(xin, P, W2b, W3b, act);
    if (s < n) y[s] = out;
}

__device__ __forceinline__ int mlp_cell(const float v[4], const MLPNorm &nm, const int nc[4], bool &inside)
{
    int id = 0;
    inside = true;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const float t = (v[a] - nm.lo[a]) / (nm.hi[a] - nm.lo[a]);
        if (!(t >= 0.0f && t <= 1.0f)) inside = false;
        const int c = (int)(fminf(fmaxf(t, 0.0f), 0.999999f) * nc[a]);
        id = id * nc[a] + c;
    }
    return id;
}

// Gated inference of samples blockIdx.x*64 + lane from log10 columns x[a*n + i]: the network's
// value, and the gate (outside the training box, or a cell whose map value exceeds tol: the
// table instead) evaluated in the same wavefront; the samples that take the table are appended
// to list (ballot-compacted, one atomic per wavefront) for k_mlp_fallback, so the table's
// gathers run on full wavefronts of fallback samples only.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_mlp_infer(int64_t n, const float *__restrict__ x, MLPNorm nm, const float *__restrict__ P,
            const unsigned short *__restrict__ W2b, const unsigned short *__restrict__ W3b,
            const float *__restrict__ emap, int ec0, int ec1, int ec2, int ec3, float tol, float *__restrict__ yn,
            float *__restrict__ outL, int *__restrict__ flag, int32_t *__restrict__ list,
            unsigned long long *__restrict__ nlist)
{
    __shared__ __align__(16) unsigned short act[ITILE];
    const int64_t s = blockIdx.x * (int64_t)IS + threadIdx.x;
    const bool valid = s < n;
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f}, xin[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (valid)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = x[k * n + s];
            xin[k] = 2.0f * (v[k] - nm.lo[k]) / (nm.hi[k] - nm.lo[k]) - 1.0f;
        }
    const float out = mlp_forward64(xin, P, W2b, W3b, act);
    bool fb = false;
    if (valid) {
        const int nc[4] = {ec0, ec1, ec2, ec3};
        bool inside;
        const int cell = mlp_cell(v, nm, nc, inside);
        fb = !inside || !(emap[cell] <= tol);
        yn[s] = out;
        outL[s] = fmaf(out, nm.sigma, nm.mu);
        if (flag) flag[s] = fb ? 1 : 0;
    }
    const uint64_t m = __ballot(fb);
    if (m) {
        unsigned long long base = 0;
        if (threadIdx.x == 0) base = atomicAdd(nlist, (unsigned long long)__popcll(m));
        base = __shfl(base, 0);
        if (fb)
            list[base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] =
                (int32_t)s;
    }
}

// Table interpolation of the listed samples (grid-stride over the device count); the
// count is then added to *nfall.
__global__ void __launch_bounds__(256) k_mlp_fallback(int64_t n, const float *__restrict__ x, CoolAxes ax,
                                                      const float *__restrict__ tabL, const int32_t *__restrict__ list,
                                                      const unsigned long long *__restrict__ nlist,
                                                      float *__restrict__ outL, unsigned long long *__restrict__ nfall)
{
    const int64_t cnt = (int64_t)*nlist;
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (int64_t k = g; k < cnt; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = list[k];
        const float v[4] = {x[s], x[n + s], x[2 * n + s], x[3 * n + s]};
        float res = 0.0f, gg;
        cool_table_interp(ax, tabL, nullptr, v, res, gg);
        outL[s] = res;
    }
    if (g == 0 && cnt) atomicAdd(nfall, (unsigned long long)cnt);
}

// ---------------------------------------------------------------- training
// Forward and backward of one tile of 32 batch samples by a workgroup of four wavefronts;
// batch sample j uses training sample (mult*j + off) mod n.  Every phase splits the 128
// features (or the 32 x 32 MFMA tiles) of a layer among the four wavefronts, the activations
// and pre-activations stay in LDS, and the phases are separated by workgroup barriers.
// Gradients of the batch-mean squared error are added to G, the batch-mean loss to *loss.
#define TW 4              // wavefronts per training tile
__global__ void __launch_bounds__(64 * TW) k_mlp_train(int64_t n, int64_t first, int64_t bsz, int64_t mult,
                                                       int64_t off, const float *__restrict__ x,
                                                       const float *__restrict__ t, const float *__restrict__ P,
                                                       const unsigned short *__restrict__ W2b,
                                                       const unsigned short *__restrict__ W3b,
                                                       const unsigned short *__restrict__ W2tb,
                                                       const unsigned short *__restrict__ W3tb, float *__restrict__ G,
                                                       float *__restrict__ loss)
{
    __shared__ __align__(16) unsigned short buf[6 * TILE];
    __shared__ float xs[TS][4], dys[TS], ypart[TW][TS];
    unsigned short *h1 = buf, *h2 = buf + TILE, *h3 = buf + 2 * TILE;
    unsigned short *z1 = buf + 3 * TILE, *z2 = buf + 4 * TILE, *z3 = buf + 5 * TILE;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, sl = lane & 31, half = lane >> 5;
    const int fb = 32 * w;  // this wavefront's 32 features
    const int64_t j = blockIdx.x * (int64_t)TS + sl;
    const bool valid = j < bsz;
    const int64_t src = valid ? (int64_t)(((unsigned long long)mult * (unsigned long long)(first + j) +
                                           (unsigned long long)off) % (unsigned long long)n)
                              : 0;
    float xin[4];
#pragma unroll
    for (int k = 0; k < 4; k++) xin[k] = valid ? x[src * 4 + k] : 0.0f;
    if (w == 0 && lane < 32)
#pragma unroll
        for (int k = 0; k < 4; k++) xs[sl][k] = xin[k];
    // layer 1 (FP32 on the VALU): sample sl, features fb + 16 half + q
    for (int q = 0; q < 16; q++) {
        const int f = fb + 16 * half + q;
        float zz = P[O_B1 + f];
#pragma unroll
        for (int k = 0; k < 4; k++) zz = fmaf(P[O_W1 + f * 4 + k], xin[k], zz);
        h1[sl * LDP + f] = f2bf(relu(zz));
        z1[sl * LDP + f] = f2bf(zz);
    }
    __syncthreads();
    // layers 2 and 3: wavefront w computes the output tile of features fb..fb+31
#pragma unroll
    for (int L = 0; L < 2; L++) {
        const unsigned short *W = L == 0 ? W2b : W3b;
        const int ob = L == 0 ? O_B2 : O_B3;
        const unsigned short *hin = L == 0 ? h1 : h2;
        unsigned short *hout = L == 0 ? h2 : h3;
        unsigned short *zout = L == 0 ? z2 : z3;
        const floatx16 acc = tile_gemm(hin, W + (size_t)fb * HW);
        const int f = fb + sl;
        const float bias = P[ob + f];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = drow(r, lane);
            const float zz = acc[r] + bias;
            hout[row * LDP + f] = f2bf(relu(zz));
            zout[row * LDP + f] = f2bf(zz);
        }
        __syncthreads();
    }
    // output layer: partial dot products over this wavefront's features, then their sum
    {
        float y = 0.0f;
        for (int q = 0; q < 16; q++) {
            const int f = fb + 16 * half + q;
            y = fmaf(P[O_W4 + f], bf2f(h3[sl * LDP + f]), y);
        }
        y += __shfl_xor(y, 32);
        if (lane < 32) ypart[w][sl] = y;
    }
    __syncthreads();
    const float inv_b = 1.0f / (float)bsz;
    if (w == 0) {
        float y = P[O_B4];
#pragma unroll
        for (int k = 0; k < TW; k++) y += ypart[k][sl];
        const float err = valid ? (y - t[src]) : 0.0f;
        if (lane < 32) dys[sl] = 2.0f * err * inv_b;
        float l2 = lane < 32 ? err * err : 0.0f;
        for (int o = 16; o >= 1; o >>= 1) l2 += __shfl_xor(l2, o);
        if (lane == 0) unsafeAtomicAdd(loss, l2 * inv_b);
    }
    __syncthreads();

    // output layer: gW4 = sum_s dy_s h3[s], gb4 = sum_s dy_s
    if (lane < 32) {
        const int f = fb + lane;
        float gw = 0.0f;
        for (int s = 0; s < TS; s++) gw = fmaf(dys[s], bf2f(h3[s * LDP + f]), gw);
        unsafeAtomicAdd(&G[O_W4 + f], gw);
    }
    if (w == 0 && lane == 0) {
        float gb = 0.0f;
        for (int s = 0; s < TS; s++) gb += dys[s];
        unsafeAtomicAdd(&G[O_B4], gb);
    }
    __syncthreads();  // h3 is read above; z3 becomes dz3 below and h3 dz2 later
    // dz3 = dy W4 relu'(z3), in place over z3
    for (int q = 0; q < 16; q++) {
        const int f = fb + 16 * half + q;
        const float zz = bf2f(z3[sl * LDP + f]);
        z3[sl * LDP + f] = f2bf(dys[sl] * P[O_W4 + f] * drelu(zz));
    }
    __syncthreads();

    // hidden layers 3 then 2: dz (32 x 128) in LDS
    // dz3 lives over z3, dz2 over h3, dz1 over z3 again (dz3 is dead by then)
    for (int L = 0; L < 2; L++) {
        const unsigned short *dz = L == 0 ? z3 : h3;
        const unsigned short *hp = L == 0 ? h2 : h1;
        const unsigned short *zp = L == 0 ? z2 : z1;
        const unsigned short *Wt = L == 0 ? W3tb : W2tb;
        const int ow = L == 0 ? O_W3 : O_W2, obb = L == 0 ? O_B3 : O_B2;
        // dW[o][i] = sum_s dz[s][o] hp[s][i]: A(o, s) = dz[s][o], B(s, i) = hp[s][i]; wavefront w
        // takes the output rows o in fb..fb+31
        {
            const int mt = w;
            for (int nt = 0; nt < 4; nt++) {
                floatx16 acc = {};
#pragma unroll
                for (int kk = 0; kk < TS; kk += 8) {
                    const int k0 = kk + 4 * half;
                    bf16x4 a, b;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        a[i] = (short)dz[(k0 + i) * LDP + mt * 32 + sl];
                        b[i] = (short)hp[(k0 + i) * LDP + nt * 32 + sl];
                    }
                    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 16; r++)
                    unsafeAtomicAdd(&G[ow + (mt * 32 + drow(r, lane)) * HW + nt * 32 + sl], acc[r]);
            }
        }
        if (lane < 32) {
            const int f = fb + lane;
            float gb = 0.0f;
            for (int s = 0; s < TS; s++) gb += bf2f(dz[s * LDP + f]);
            unsafeAtomicAdd(&G[obb + f], gb);
        }
        // dh_prev = dz W (B(o, i) = W[o][i] = Wt[i][o]); dz_prev = dh_prev relu'(z_prev); wavefront
        // w takes the input features i in fb..fb+31
        unsigned short *nx = L == 0 ? h3 : z3;
        {
            const floatx16 acc = tile_gemm(dz, Wt + (size_t)fb * HW);
            const int i = fb + sl;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int s = drow(r, lane);
                nx[s * LDP + i] = f2bf(acc[r] * drelu(bf2f(zp[s * LDP + i])));
            }
        }
        __syncthreads();
    }
    // input layer: gW1[o][k] = sum_s dz1[s][o] x[s][k], gb1[o] = sum_s dz1[s][o]
    const unsigned short *dz1 = z3;
    if (lane < 32) {
        const int o = fb + lane;
        float g[4] = {0.0f, 0.0f, 0.0f, 0.0f}, gb = 0.0f;
        for (int s = 0; s < TS; s++) {
            const float d = bf2f(dz1[s * LDP + o]);
            gb += d;
#pragma unroll
            for (int k = 0; k < 4; k++) g[k] = fmaf(d, xs[s][k], g[k]);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) unsafeAtomicAdd(&G[O_W1 + o * 4 + k], g[k]);
        unsafeAtomicAdd(&G[O_B1 + o], gb);
    }
}

__device__ __forceinline__ void store_bf16_copies(int i, float w, unsigned short *__restrict__ W2b,
                                                  unsigned short *__restrict__ W3b, unsigned short *__restrict__ W2tb,
                                                  unsigned short *__restrict__ W3tb)
{
    if (i >= O_W2 && i < O_W2 + HW * HW) {
        const int k = i - O_W2, o = k / HW, in = k % HW;
        W2b[k] = f2bf(w);
        W2tb[in * HW + o] = f2bf(w);
    } else if (i >= O_W3 && i < O_W3 + HW * HW) {
        const int k = i - O_W3, o = k / HW, in = k % HW;
        W3b[k] = f2bf(w);
        W3tb[in * HW + o] = f2bf(w);
    }
}

// Adam (Kingma & Ba 2015) with bias correction c1 = 1 - b1^t, c2 = 1 - b2^t.
__global__ void __launch_bounds__(256) k_adam(int np, float *__restrict__ P, float *__restrict__ G,
                                              float *__restrict__ M, float *__restrict__ V, float lr, float b1,
                                              float b2, float eps, float c1, float c2,
                                              unsigned short *__restrict__ W2b, unsigned short *__restrict__ W3b,
                                              unsigned short *__restrict__ W2tb, unsigned short *__restrict__ W3tb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float g = G[i];
    const float m = b1 * M[i] + (1.0f - b1) * g;
    const float v = b2 * V[i] + (1.0f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    const float w = P[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[i] = w;
    G[i] = 0.0f;
    store_bf16_copies(i, w, W2b, W3b, W2tb, W3tb);
}

__global__ void __launch_bounds__(256) k_refresh_bf16(int np, const float *__restrict__ P,
                                                      unsigned short *__restrict__ W2b,
                                                      unsigned short *__restrict__ W3b,
                                                      unsigned short *__restrict__ W2tb,
                                                      unsigned short *__restrict__ W3tb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    store_bf16_copies(i, P[i], W2b, W3b, W2tb, W3tb);
}

// ---------------------------------------------------------------- samples
__device__ __forceinline__ uint64_t mix64(uint64_t v)
{
    // SplitMix64 finaliser (Steele, Lea & Flood 2014)
    v += 0x9E3779B97F4A7C15ull;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}
__device__ __forceinline__ double unit_draw(uint64_t seed, uint64_t i, int k)
{
    return (double)(mix64(seed ^ mix64(i * 4 + (uint64_t)k)) >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void __launch_bounds__(256) k_mlp_samples(int64_t n, uint64_t seed, MLPNorm nm, int nc0, int nc1, int nc2,
                                                     int nc3, int per, float *__restrict__ xraw,
                                                     float *__restrict__ xn)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[4];
    if (per > 0) {
        int64_t c = i / per;
        const int nc[4] = {nc0, nc1, nc2, nc3};
        int ci[4];
        for (int a = 3; a >= 0; a--) {
            ci[a] = (int)(c % nc[a]);
            c /= nc[a];
        }
        for (int a = 0; a < 4; a++) u[a] = (ci[a] + unit_draw(seed, i, a)) / nc[a];
    } else {
        for (int a = 0; a < 4; a++) u[a] = unit_draw(seed, i, a);
    }
    for (int a = 0; a < 4; a++) {
        const float v = (float)(nm.lo[a] + (double)(nm.hi[a] - nm.lo[a]) * u[a]);
        xraw[i * 4 + a] = v;
        xn[i * 4 + a] = 2.0f * (v - nm.lo[a]) / (nm.hi[a] - nm.lo[a]) - 1.0f;
    }
}

// Per-cell maxima over groups of `per` consecutive samples: |network - table|,
// |network - exact| and |table - exact|.  Maps hold float bits compared as
// unsigned ints (valid for non-negative values) and must start at zero.
__global__ void __launch_bounds__(256) k_emap(int64_t n, int per, const float *__restrict__ yn,
                                              const float *__restrict__ xraw, const float *__restrict__ exact,
                                              MLPNorm nm, CoolAxes ax, const float *__restrict__ tabL,
                                              unsigned int *__restrict__ e_nt, unsigned int *__restrict__ e_ne,
                                              unsigned int *__restrict__ e_te)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float pred = fmaf(yn[i], nm.sigma, nm.mu);
    const float v[4] = {xraw[4 * i], xraw[4 * i + 1], xraw[4 * i + 2], xraw[4 * i + 3]};
    float tl, g;
    cool_table_interp(ax, tabL, nullptr, v, tl, g);
    const int64_t c = i / per;
    atomicMax(&e_nt[c], __float_as_uint(fabsf(pred - tl)));
    atomicMax(&e_ne[c], __float_as_uint(fabsf(pred - exact[i])));
    atomicMax(&e_te[c], __float_as_uint(fabsf(tl - exact[i])));
}

// Gate maps: exact mode keeps the network's error only where the table measured
// better in the same cell (elsewhere the network is never worse than its fallback).
__global__ void __launch_bounds__(256) k_gate_map(int64_t ncell, const float *__restrict__ e_ne,
                                                  const float *__restrict__ e_te, float *__restrict__ gate)
{
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    gate[c] = e_ne[c] > e_te[c] ? e_ne[c] : 0.0f;
}

namespace sph {

hipError_t mlp_forward_raw(hipStream_t st, int64_t n, const float *x, const MLPDev &d, float *y)
{
    k_mlp_forward<<<(unsigned)((n + IS - 1) / IS), 64, 0, st>>>(n, x, d.P, d.W2b, d.W3b, y);
    return hipGetLastError();
}

hipError_t mlp_infer(hipStream_t st, int64_t n, const float *x, const MLPNorm &nm, const MLPDev &d,
                     const float *emap, const int *ec, float tol, const CoolAxes &ax, const float *tabL, float *yn,
                     float *outL, int *flag, unsigned long long *nfall, int32_t *scratch)
{
    if (n < 1) return hipSuccess;
    unsigned long long *nlist = reinterpret_cast<unsigned long long *>(scratch);  // 64-int header: the count
    int32_t *list = scratch + 64;
    fill<unsigned long long>(st, 1, nlist, 0ull);
    k_mlp_infer<<<(unsigned)((n + IS - 1) / IS), 64, 0, st>>>(n, x, nm, d.P, d.W2b, d.W3b, emap, ec[0], ec[1], ec[2],
                                                              ec[3], tol, yn, outL, flag, list, nlist);
    const int64_t nb = std::min<int64_t>(sph_blocks(n, 256), 2048);
    k_mlp_fallback<<<(unsigned)nb, 256, 0, st>>>(n, x, ax, tabL, list, nlist, outL, nfall);
    return hipGetLastError();
}

hipError_t mlp_train_batch(hipStream_t st, int64_t n, int64_t first, int64_t bsz, int64_t mult, int64_t off,
                           const float *x, const float *t, const MLPDev &d, float *loss)
{
    k_mlp_train<<<(unsigned)((bsz + TS - 1) / TS), 64 * TW, 0, st>>>(n, first, bsz, mult, off, x, t, d.P, d.W2b, d.W3b,
                                                               d.W2tb, d.W3tb, d.G, loss);
    return hipGetLastError();
}

hipError_t mlp_adam(hipStream_t st, const MLPDev &d, float lr, float b1, float b2, float eps, int step)
{
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    k_adam<<<sph_blocks(MLP_NPARAM, 256), 256, 0, st>>>(MLP_NPARAM, d.P, d.G, d.M, d.V, lr, b1, b2, eps, c1, c2, d.W2b,
                                                        d.W3b, d.W2tb, d.W3tb);
    return hipGetLastError();
}

hipError_t mlp_refresh(hipStream_t st, const MLPDev &d)
{
    k_refresh_bf16<<<sph_blocks(MLP_NPARAM, 256), 256, 0, st>>>(MLP_NPARAM, d.P, d.W2b, d.W3b, d.W2tb, d.W3tb);
    return hipGetLastError();
}

hipError_t mlp_samples(hipStream_t st, int64_t n, uint64_t seed, const MLPNorm &nm, const int *nc, int per,
                       float *xraw, float *xn)
{
    k_mlp_samples<<<sph_blocks(n, 256), 256, 0, st>>>(n, seed, nm, nc ? nc[0] : 1, nc ? nc[1] : 1, nc ? nc[2] : 1,
                                                      nc ? nc[3] : 1, per, xraw, xn);
    return hipGetLastError();
}

hipError_t mlp_emap(hipStream_t st, int64_t n, int per, const float *yn, const float *xraw, const float *exact,
                    const MLPNorm &nm, const CoolAxes &ax, const float *tabL, float *e_nt, float *e_ne, float *e_te)
{
    k_emap<<<sph_blocks(n, 256), 256, 0, st>>>(n, per, yn, xraw, exact, nm, ax, tabL, (unsigned int *)e_nt,
                                               (unsigned int *)e_ne, (unsigned int *)e_te);
    return hipGetLastError();
}

hipError_t mlp_gate_map(hipStream_t st, int64_t ncell, const float *e_ne, const float *e_te, float *gate)
{
    k_gate_map<<<sph_blocks(ncell, 256), 256, 0, st>>>(ncell, e_ne, e_te, gate);
    return hipGetLastError();
}

}  // namespace sph
