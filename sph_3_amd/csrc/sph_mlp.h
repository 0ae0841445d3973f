// sph_mlp.h -- host entry points of the cooling-surrogate MLP kernels (sph_mlp.hip).
#pragma once
#include "sph_common.h"
#include "sph_cool.h"

#define MLP_HIDDEN 128
#define MLP_NPARAM 33793  // 4*128+128 + 2*(128*128+128) + 128+1

// Input box (log n_H, log T, log Z, log J21) mapped to [-1, 1]^4 and the
// normalisation of the target log10 Lambda: y_norm = (log10 Lambda - mu) / sigma.
struct MLPNorm {
    float lo[4], hi[4];
    float mu, sigma;
};

// Inference weights in MFMA operand order (sph_mlp.hip, k_mlp_image): 5,224 records of 16 bytes.
#define MLP_IMG_BYTES 83584

// Device parameters: FP32 masters, gradient and Adam moments (MLP_NPARAM each,
// flat layout W1[128][4], b1, W2[128][128], b2, W3[128][128], b3, W4[128], b4),
// BF16 copies of W2, W3 (row-major [out][in]) and their transposes for training,
// and the inference image (MLP_IMG_BYTES) built from the masters.
struct MLPDev {
    float *P, *G, *M, *V;
    unsigned short *W2b, *W3b, *W2tb, *W3tb;
    uint4 *img;
};

namespace sph {

// Network output (normalised) for n normalised inputs x[n][4].
hipError_t mlp_forward_raw(hipStream_t st, int64_t n, const float *x, const MLPDev &d, float *y);
// Gated inference at log10 columns x[a*n + i] (a = nH, T, Z, J21): the network inside the box
// where the cell map emap <= tol, the table elsewhere; yn[n] receives the network output,
// flag[n] (may be null) 1 where the table was used, *nfall += the number of those.
// scratch: mlp_infer_scratch_ints(n) int32 (the compacted list of table samples).
hipError_t mlp_infer(hipStream_t st, int64_t n, const float *x, const MLPNorm &nm, const MLPDev &d,
                     const float *emap, const int *ec, float tol, const CoolAxes &ax, const float *tabL, float *yn,
                     float *outL, int *flag, unsigned long long *nfall, int32_t *scratch);
inline int64_t mlp_infer_scratch_ints(int64_t n) { return n + 64; }
// Gradient of the batch MSE (samples src = (mult*j + off) mod n, j = first..first+bsz-1) added to d.G;
// loss (batch mean) added to *loss.
hipError_t mlp_train_batch(hipStream_t st, int64_t n, int64_t first, int64_t bsz, int64_t mult, int64_t off,
                           const float *x, const float *t, const MLPDev &d, float *loss);
// Adam step (bias-corrected, step >= 1) on the masters; refreshes the BF16 training copies and
// clears d.G (the inference image is rebuilt by mlp_image once the training calls are done).
hipError_t mlp_adam(hipStream_t st, const MLPDev &d, float lr, float b1, float b2, float eps, int step);
// BF16 training copies and the inference image from the masters.
hipError_t mlp_refresh(hipStream_t st, const MLPDev &d);
// Inference image from the masters (after the weights changed).
hipError_t mlp_image(hipStream_t st, const MLPDev &d);
// Sample coordinates: per > 0 stratified (per points in each cell of the grid nc), else uniform.
// xraw[n][4] log10 values, xn[n][4] normalised.
hipError_t mlp_samples(hipStream_t st, int64_t n, uint64_t seed, const MLPNorm &nm, const int *nc, int per,
                       float *xraw, float *xn);
// Per-cell maxima over consecutive groups of `per` samples of |network - table| (e_nt),
// |network - exact| (e_ne) and |table - exact| (e_te), in dex; the maps must start at zero.
hipError_t mlp_emap(hipStream_t st, int64_t n, int per, const float *yn, const float *xraw, const float *exact,
                    const MLPNorm &nm, const CoolAxes &ax, const float *tabL, float *e_nt, float *e_ne, float *e_te);
// gate[c] = e_ne[c] where e_ne[c] > e_te[c], else 0.
hipError_t mlp_gate_map(hipStream_t st, int64_t ncell, const float *e_ne, const float *e_te, float *gate);

}  // namespace sph
