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
// Inference keeps every activation in registers.  A wavefront takes 64 samples as two
// 32-sample tiles (nt = 0, 1) and computes each layer transposed, Z^T = W H^T, on
// v_mfma_f32_32x32x16_bf16: the A operand is a 32 x 16 block of the layer's weights, the B
// operand 16 inputs of the tile's 32 samples.  Accumulator r of a lane in output block mt
// holds unit 32 mt + 8 (r/4) + 4 half + r%4 of sample l%32.  After the bias (the first MFMA's C
// input), ReLU and BF16 rounding, accumulators 8p..8p+7 of block mt ARE the lane's B operand
// for the next layer's input block kk = 2 mt + p, provided that layer's weight columns are
// stored in the matching order: B slot i of lane half h in block kk carries hidden unit
// hperm(16 kk + 8 h + i) = 32 (kk/2) + 16 (kk%2) + 8 (i/4) + 4 h + i%4.  No activation goes
// through LDS.  The first layer (4 inputs) and the output layer run on the same MFMA with split
// operands: x = xh + xm + xl and W1 = Wh + Wm + Wl (BF16 each, exact FP32 remainders) give
// W1 x from the six products whose order is at least 2^-16 (two 16-deep MFMA blocks), within
// about 2^-24 relative like the FP32 sum, and y = W4h.h3 + W4m.h3 + W4l.h3 (rows 0..2 of the
// output tile) carries the FP32 W4.  The weights, in operand order (the image, built by
// k_mlp_image), are copied to LDS once per workgroup; one workgroup of eight wavefronts per CU
// loops over the tiles.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));

#define IS 64                        // samples per wavefront tile pair
#define IWAVES 8                     // wavefronts per inference workgroup
#define IMG_W1 0                     // [mt 4][kb 2][lane 64]
#define IMG_W2 512                   // [mt 4][kk 8][lane 64]
#define IMG_W3 (IMG_W2 + 2048)       // [mt 4][kk 8][lane 64]
#define IMG_W4 (IMG_W3 + 2048)       // [kk 8][lane 64]: rows 0..2 (W4h, W4m, W4l)
#define IMG_B (IMG_W4 + 512)         // [layer 3][mt 4][half 2][4]: biases in accumulator order
#define IMG_B4 (IMG_B + 96)          // [half 2][4]: b4 at accumulator 0 of half 0
#define IMG_U4 (IMG_B4 + 8)          // records of 16 bytes
static_assert(IMG_U4 * 16 == MLP_IMG_BYTES, "inference image size");

__device__ __forceinline__ int hperm(int k)
{
    const int kk = k >> 4, h = (k >> 3) & 1, i = k & 7;
    return 32 * (kk >> 1) + 16 * (kk & 1) + 8 * (i >> 2) + 4 * h + (i & 3);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ unsigned int pack2(float a, float b)
{
    // one v_cvt_pk_bf16_f32 (round to nearest even) for the pair
    return __builtin_bit_cast(unsigned int, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

__device__ __forceinline__ uint4 pack8(const unsigned short v[8])
{
    return make_uint4(v[0] | ((unsigned)v[1] << 16), v[2] | ((unsigned)v[3] << 16), v[4] | ((unsigned)v[5] << 16),
                      v[6] | ((unsigned)v[7] << 16));
}

// BF16 parts of an FP32 value: w = hi + mid + lo to 24 bits (each remainder exact in FP32).
__device__ __forceinline__ void split3(float w, unsigned short &hi, unsigned short &mid, unsigned short &lo)
{
    hi = f2bf(w);
    const float r = w - bf2f(hi);
    mid = f2bf(r);
    lo = f2bf(r - bf2f(mid));
}

// One record of the image per thread (the W4 rows from 3 on and the unused slots are zero).
// Layer 1, output unit f = 32 mt + l%32, against the B slots of mlp_tiles:
//   block 0, half 0: [Wh, Wh] x [xh, xm];  block 0, half 1: [Wm, Wh] x [xh, xl];
//   block 1, half 0: [Wm, Wl] x [xm, xh];  block 1, half 1: zero.
__global__ void __launch_bounds__(256) k_mlp_image(const float *__restrict__ P, uint4 *__restrict__ img)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= IMG_U4) return;
    if (e < IMG_B) {
        unsigned short v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (e < IMG_W2) {
            const int mt = e >> 7, kb = (e >> 6) & 1, lane = e & 63, f = 32 * mt + (lane & 31);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                unsigned short wh, wm, wl;
                split3(P[O_W1 + 4 * f + i], wh, wm, wl);
                if (kb == 0 && lane < 32) {
                    v[i] = wh;
                    v[4 + i] = wh;
                } else if (kb == 0) {
                    v[i] = wm;
                    v[4 + i] = wh;
                } else if (lane < 32) {
                    v[i] = wm;
                    v[4 + i] = wl;
                }
            }
        } else if (e < IMG_W4) {  // hidden layers: row 32 mt + l%32, columns in B-slot order
            const bool l3 = e >= IMG_W3;
            const int q = e - (l3 ? IMG_W3 : IMG_W2), mt = q >> 9, kk = (q >> 6) & 7, lane = q & 63;
            const float *W = P + (l3 ? O_W3 : O_W2) + (size_t)(32 * mt + (lane & 31)) * HW;
#pragma unroll
            for (int i = 0; i < 8; i++) v[i] = f2bf(W[hperm(16 * kk + 8 * (lane >> 5) + i)]);
        } else {  // output layer: rows 0, 1, 2 the BF16 parts of W4
            const int q = e - IMG_W4, kk = q >> 6, lane = q & 63, row = lane & 31;
            if (row < 3)
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    unsigned short wh, wm, wl;
                    split3(P[O_W4 + hperm(16 * kk + 8 * (lane >> 5) + i)], wh, wm, wl);
                    v[i] = row == 0 ? wh : (row == 1 ? wm : wl);
                }
        }
        img[e] = pack8(v);
    } else {
        float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (e < IMG_B4) {  // accumulator r = 4 j + c of half h in block mt: unit 32 mt + 8 j + 4 h + c
            const int q = e - IMG_B, L = q >> 5, mt = (q >> 3) & 3, h = (q >> 2) & 1, j = q & 3;
            const float *b = P + (L == 0 ? O_B1 : (L == 1 ? O_B2 : O_B3));
#pragma unroll
            for (int c = 0; c < 4; c++) f[c] = b[32 * mt + 8 * j + 4 * h + c];
        } else if (e == IMG_B4) {
            f[0] = P[O_B4];
        }
        img[e] = make_uint4(__float_as_uint(f[0]), __float_as_uint(f[1]), __float_as_uint(f[2]),
                            __float_as_uint(f[3]));
    }
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// Four consecutive records as one accumulator set (the C input of a block's first MFMA).
__device__ __forceinline__ floatx16 lds_acc(const uint4 *p)
{
    const uint4 a = p[0], b = p[1], c = p[2], d = p[3];
    floatx16 r;
    r[0] = __uint_as_float(a.x); r[1] = __uint_as_float(a.y); r[2] = __uint_as_float(a.z); r[3] = __uint_as_float(a.w);
    r[4] = __uint_as_float(b.x); r[5] = __uint_as_float(b.y); r[6] = __uint_as_float(b.z); r[7] = __uint_as_float(b.w);
    r[8] = __uint_as_float(c.x); r[9] = __uint_as_float(c.y); r[10] = __uint_as_float(c.z); r[11] = __uint_as_float(c.w);
    r[12] = __uint_as_float(d.x); r[13] = __uint_as_float(d.y); r[14] = __uint_as_float(d.z); r[15] = __uint_as_float(d.w);
    return r;
}

// Accumulators 8p..8p+7 -> ReLU -> BF16 (round to nearest even), as a B operand.  ReLU acts on
// the packed BF16 pairs: read as signed 16-bit integers the negative BF16 values are exactly the
// negative integers, so a packed integer maximum with 0 is relu(bf16(z)) = bf16(relu(z)).
__device__ __forceinline__ bf16x8 relu_bf16(const floatx16 &a, int p)
{
    uint4 u;
    u.x = pack2(a[8 * p + 0], a[8 * p + 1]);
    u.y = pack2(a[8 * p + 2], a[8 * p + 3]);
    u.z = pack2(a[8 * p + 4], a[8 * p + 5]);
    u.w = pack2(a[8 * p + 6], a[8 * p + 7]);
    const s16x2 z = {0, 0};
    u.x = __builtin_bit_cast(unsigned int, __builtin_elementwise_max(__builtin_bit_cast(s16x2, u.x), z));
    u.y = __builtin_bit_cast(unsigned int, __builtin_elementwise_max(__builtin_bit_cast(s16x2, u.y), z));
    u.z = __builtin_bit_cast(unsigned int, __builtin_elementwise_max(__builtin_bit_cast(s16x2, u.z), z));
    u.w = __builtin_bit_cast(unsigned int, __builtin_elementwise_max(__builtin_bit_cast(s16x2, u.w), z));
    return as_bf16x8(u);
}

// Hidden layer: out[kk][nt] (B operands of the next layer) from in[kk][nt].
__device__ __forceinline__ void hidden_layer(const uint4 *__restrict__ W, const uint4 *__restrict__ bias,
                                             const bf16x8 (&in)[8][2], bf16x8 (&out)[8][2], int lane)
{
    const int h = lane >> 5;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const floatx16 c = lds_acc(bias + mt * 8 + h * 4);
        floatx16 acc[2];
        const bf16x8 a0 = as_bf16x8(W[(mt * 8) * 64 + lane]);
        acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, in[0][0], c, 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, in[0][1], c, 0, 0, 0);
#pragma unroll
        for (int kk = 1; kk < 8; kk++) {
            const bf16x8 a = as_bf16x8(W[(mt * 8 + kk) * 64 + lane]);
            acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, in[kk][0], acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, in[kk][1], acc[1], 0, 0, 0);
        }
#pragma unroll
        for (int nt = 0; nt < 2; nt++) {
            out[2 * mt][nt] = relu_bf16(acc[nt], 0);
            out[2 * mt + 1][nt] = relu_bf16(acc[nt], 1);
        }
        // keep the scheduler from hoisting the next blocks' weight reads (registers)
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Network outputs of the wavefront's 64 samples from the normalised inputs xin[nt] of samples
// 32 nt + l%32: returns, in every lane l, the output of sample l.
__device__ __forceinline__ float mlp_tiles(const uint4 *__restrict__ img, const float (&xin)[2][4], int lane)
{
    const int h = lane >> 5;
    bf16x8 b1[2][2];  // [block][tile]
#pragma unroll
    for (int nt = 0; nt < 2; nt++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float x = xin[nt][i];
            const __bf16 xh = (__bf16)x;
            const float r = x - (float)xh;
            const __bf16 xm = (__bf16)r;
            const __bf16 xl = (__bf16)(r - (float)xm);
            b1[0][nt][i] = xh;
            b1[0][nt][4 + i] = h ? xl : xm;
            b1[1][nt][i] = h ? (__bf16)0.0f : xm;
            b1[1][nt][4 + i] = h ? (__bf16)0.0f : xh;
        }
    bf16x8 ha[8][2], hb[8][2];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const floatx16 c = lds_acc(img + IMG_B + mt * 8 + h * 4);
        const bf16x8 a0 = as_bf16x8(img[IMG_W1 + mt * 128 + lane]);
        const bf16x8 a1 = as_bf16x8(img[IMG_W1 + mt * 128 + 64 + lane]);
#pragma unroll
        for (int nt = 0; nt < 2; nt++) {
            floatx16 acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1[0][nt], c, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1[1][nt], acc, 0, 0, 0);
            ha[2 * mt][nt] = relu_bf16(acc, 0);
            ha[2 * mt + 1][nt] = relu_bf16(acc, 1);
        }
    }
    hidden_layer(img + IMG_W2, img + IMG_B + 32, ha, hb, lane);
    hidden_layer(img + IMG_W3, img + IMG_B + 64, hb, ha, lane);
    const floatx16 c = lds_acc(img + IMG_B4 + h * 4);
    floatx16 acc[2];
    const bf16x8 a0 = as_bf16x8(img[IMG_W4 + lane]);
    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, ha[0][0], c, 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, ha[0][1], c, 0, 0, 0);
#pragma unroll
    for (int kk = 1; kk < 8; kk++) {
        const bf16x8 a = as_bf16x8(img[IMG_W4 + kk * 64 + lane]);
        acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, ha[kk][0], acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, ha[kk][1], acc[1], 0, 0, 0);
    }
    // rows 0..2 (accumulators 0..2 of the lanes l < 32) of tile nt: sample 32 nt + l
    const float y0 = acc[0][0] + (acc[0][1] + acc[0][2]), y1 = acc[1][0] + (acc[1][1] + acc[1][2]);
    const float y1s = __shfl(y1, lane & 31);
    return h ? y1s : y0;
}

// Copy the image to LDS (whole workgroup).
__device__ __forceinline__ void load_image(uint4 *img, const uint4 *__restrict__ gimg)
{
    for (int k = threadIdx.x; k < IMG_U4; k += blockDim.x) img[k] = gimg[k];
    __syncthreads();
}

// Network output for rows of normalised inputs x[n][4].
__global__ void __launch_bounds__(64 * IWAVES) k_mlp_forward(int64_t n, const float *__restrict__ x,
                                                              const uint4 *__restrict__ gimg, float *__restrict__ y)
{
    __shared__ uint4 img[IMG_U4];
    load_image(img, gimg);
    const int lane = threadIdx.x & 63;
    const int64_t ntile = (n + IS - 1) / IS;
    for (int64_t tile = (int64_t)blockIdx.x * IWAVES + (threadIdx.x >> 6); tile < ntile;
         tile += (int64_t)gridDim.x * IWAVES) {
        const int64_t base = tile * IS;
        float xin[2][4];
#pragma unroll
        for (int nt = 0; nt < 2; nt++) {
            const int64_t j = base + 32 * nt + (lane & 31);
            const float4 v = j < n ? *reinterpret_cast<const float4 *>(x + 4 * j) : make_float4(0, 0, 0, 0);
            xin[nt][0] = v.x; xin[nt][1] = v.y; xin[nt][2] = v.z; xin[nt][3] = v.w;
        }
        const float out = mlp_tiles(img, xin, lane);
        if (base + lane < n) y[base + lane] = out;
    }
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

// Gated inference from log10 columns x[a*n + i]: the network's value, and the gate (outside the
// training box, or a cell whose map value exceeds tol: the table instead) evaluated in the same
// wavefront; the samples that take the table are appended to list (ballot-compacted, one atomic
// per wavefront tile) for k_mlp_fallback, so the table's gathers run on full wavefronts of
// fallback samples only.
__global__ void __launch_bounds__(64 * IWAVES)
k_mlp_infer(int64_t n, const float *__restrict__ x, MLPNorm nm, const uint4 *__restrict__ gimg,
            const float *__restrict__ emap, int ec0, int ec1, int ec2, int ec3, float tol, float *__restrict__ yn,
            float *__restrict__ outL, int *__restrict__ flag, int32_t *__restrict__ list,
            unsigned long long *__restrict__ nlist)
{
    __shared__ uint4 img[IMG_U4];
    load_image(img, gimg);
    const int lane = threadIdx.x & 63, h = lane >> 5;
    const int64_t ntile = (n + IS - 1) / IS;
    for (int64_t tile = (int64_t)blockIdx.x * IWAVES + (threadIdx.x >> 6); tile < ntile;
         tile += (int64_t)gridDim.x * IWAVES) {
        const int64_t base = tile * IS;
        float v[2][4], xin[2][4];
#pragma unroll
        for (int nt = 0; nt < 2; nt++) {
            const int64_t j = base + 32 * nt + (lane & 31);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                v[nt][k] = j < n ? x[k * n + j] : 0.0f;
                // the normalisation of the training samples (k_mlp_samples), operation for operation
                xin[nt][k] = 2.0f * (v[nt][k] - nm.lo[k]) / (nm.hi[k] - nm.lo[k]) - 1.0f;
            }
        }
        const float out = mlp_tiles(img, xin, lane);
        const int64_t s = base + lane;  // this lane's sample: tile h, row l%32
        bool fb = false;
        if (s < n) {
            const float vs[4] = {h ? v[1][0] : v[0][0], h ? v[1][1] : v[0][1], h ? v[1][2] : v[0][2],
                                 h ? v[1][3] : v[0][3]};
            const int nc[4] = {ec0, ec1, ec2, ec3};
            bool inside;
            const int cell = mlp_cell(vs, nm, nc, inside);
            fb = !inside || !(emap[cell] <= tol);
            yn[s] = out;
            outL[s] = fmaf(out, nm.sigma, nm.mu);
            if (flag) flag[s] = fb ? 1 : 0;
        }
        const uint64_t m = __ballot(fb);
        if (m) {
            unsigned long long lb = 0;
            if (lane == 0) lb = atomicAdd(nlist, (unsigned long long)__popcll(m));
            lb = __shfl(lb, 0);
            if (fb)
                list[lb + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] =
                    (int32_t)s;
        }
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

// Persistent inference grid: one workgroup per CU (the image takes half of a CU's LDS and the
// registers allow two wavefronts per SIMD), no more than the tiles need.
static unsigned infer_grid(int64_t n)
{
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (cus[dev] <= 0) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
        cus[dev] = c;
    }
    const int64_t wgs = ((n + IS - 1) / IS + IWAVES - 1) / IWAVES;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(wgs, cus[dev]));
}

hipError_t mlp_forward_raw(hipStream_t st, int64_t n, const float *x, const MLPDev &d, float *y)
{
    if (n < 1) return hipSuccess;
    k_mlp_forward<<<infer_grid(n), 64 * IWAVES, 0, st>>>(n, x, d.img, y);
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
    k_mlp_infer<<<infer_grid(n), 64 * IWAVES, 0, st>>>(n, x, nm, d.img, emap, ec[0], ec[1], ec[2], ec[3], tol, yn,
                                                       outL, flag, list, nlist);
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
    const hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : mlp_image(st, d);
}

hipError_t mlp_image(hipStream_t st, const MLPDev &d)
{
    k_mlp_image<<<sph_blocks(IMG_U4, 256), 256, 0, st>>>(d.P, d.img);
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
