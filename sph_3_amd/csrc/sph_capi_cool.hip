// sph_capi_cool.hip -- cooling entry points: the exact rates, the table, the surrogate
// (training, calibration, gated inference), the stateless cooling update and the step's
// cooling state.
#include "sph_ctx.h"

using namespace sphc;

// =========================================================== cooling
namespace sphc {

int cool_prep(sph_ctx *ctx, int64_t n)
{
    if (!ctx) return SPH_ERR_ARG;
    if (n < 1) return fail(ctx, SPH_ERR_ARG, "n must be >= 1");
    HIPCK(hipSetDevice(ctx->device));
    if (!ctx->rf1_ok) {
        CK(ensure(ctx, ctx->rf1, 8 * 8));
        HIPCK(sph::cool_radfield(ctx->st, D(rf1)));
        ctx->rf1_ok = true;
    }
    for (int k = 0; k < 6; k++) CK(ensure(ctx, ctx->cin[k], (size_t)n * 8));
    for (int k = 0; k < 2; k++) CK(ensure(ctx, ctx->cout[k], (size_t)n * 8));
    return SPH_OK;
}

// upload an optional array into cin[k]; returns the device pointer or null
int up(sph_ctx *ctx, int k, const double *src, int64_t n, const double **dptr)
{
    *dptr = nullptr;
    if (!src) return SPH_OK;
    HIPCK(hipMemcpyAsync(ctx->cin[k].p, src, sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
    *dptr = (const double *)ctx->cin[k].p;
    return SPH_OK;
}

int down(sph_ctx *ctx, int k, double *dst, int64_t n)
{
    HIPCK(hipMemcpyAsync(dst, ctx->cout[k].p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
    return SPH_OK;
}

}  // namespace sphc

extern "C" int sph_cooling_rates(sph_ctx *ctx, int64_t n, const double *T, const double *nH, const double *J21,
                                 double J21c, const double *Z, double Zc, double *heat_out, double *lambda_out)
{
    CK(cool_prep(ctx, n));
    if (!T || !nH || !heat_out || !lambda_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    const double *dT, *dn, *dJ, *dZ;
    CK(up(ctx, 0, T, n, &dT));
    CK(up(ctx, 1, nH, n, &dn));
    CK(up(ctx, 2, J21, n, &dJ));
    CK(up(ctx, 3, Z, n, &dZ));
    HIPCK(sph::cool_rates_exact(ctx->st, n, dT, dn, dJ, J21c, dZ, Zc, D(rf1), D(cout[0]), D(cout[1])));
    CK(down(ctx, 0, heat_out, n));
    CK(down(ctx, 1, lambda_out, n));
    HIPCK(hipStreamSynchronize(ctx->st));
    return SPH_OK;
}

extern "C" int sph_cooling_from_u(sph_ctx *ctx, int64_t n, int what, const double *u, const double *nH, double XH,
                                  const double *J21, double J21c, const double *Z, double Zc, double *out)
{
    CK(cool_prep(ctx, n));
    if (!u || !nH || !out) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (what != 0 && what != 1) return fail(ctx, SPH_ERR_ARG, "what must be 0 (T) or 1 (net rate)");
    const double *du, *dn, *dJ, *dZ;
    CK(up(ctx, 0, u, n, &du));
    CK(up(ctx, 1, nH, n, &dn));
    CK(up(ctx, 2, J21, n, &dJ));
    CK(up(ctx, 3, Z, n, &dZ));
    HIPCK(sph::cool_u_exact(ctx->st, n, what, du, dn, XH, dJ, J21c, dZ, Zc, D(rf1), D(cout[0])));
    CK(down(ctx, 0, out, n));
    HIPCK(hipStreamSynchronize(ctx->st));
    return SPH_OK;
}

extern "C" int sph_do_cooling(sph_ctx *ctx, int64_t n, const double *rho, const double *u_old, const double *dt,
                              double dtc, double XH, const double *J21, double J21c, const double *Z, double Zc,
                              double u_floor, double *u_out)
{
    CK(cool_prep(ctx, n));
    if (!rho || !u_old || !u_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    const double *dr, *du, *dd, *dJ, *dZ;
    CK(up(ctx, 0, rho, n, &dr));
    CK(up(ctx, 1, u_old, n, &du));
    CK(up(ctx, 2, dt, n, &dd));
    CK(up(ctx, 3, J21, n, &dJ));
    CK(up(ctx, 4, Z, n, &dZ));
    HIPCK(sph::do_cooling_exact(ctx->st, n, dr, du, dd, dtc, XH, dJ, J21c, dZ, Zc, u_floor, D(rf1), D(cout[0])));
    CK(down(ctx, 0, u_out, n));
    HIPCK(hipStreamSynchronize(ctx->st));
    return SPH_OK;
}

extern "C" int sph_cooling_table_build(sph_ctx *ctx, const double *lo, const double *hi, const int *npts)
{
    CK(cool_prep(ctx, 1));
    if (!lo || !hi || !npts) return fail(ctx, SPH_ERR_ARG, "null argument");
    CoolAxes ax{};
    int64_t total = 1;
    for (int a = 0; a < 4; a++) {
        if (npts[a] < 2 || !(hi[a] > lo[a])) return fail(ctx, SPH_ERR_ARG, "each axis needs npts >= 2 and hi > lo");
        ax.lo[a] = lo[a];
        ax.n[a] = npts[a];
        ax.d[a] = (hi[a] - lo[a]) / (npts[a] - 1);
        total *= npts[a];
    }
    if (total > ((int64_t)1 << 28)) return fail(ctx, SPH_ERR_ARG, "cooling table too large");
    CK(ensure(ctx, ctx->ctabL, (size_t)total * 4));
    CK(ensure(ctx, ctx->ctabG, (size_t)total * 4));
    HIPCK(sph::cool_table_build(ctx->st, ax, D(rf1), P_<float>(ctx->ctabL), P_<float>(ctx->ctabG)));
    HIPCK(hipStreamSynchronize(ctx->st));
    ctx->cax = ax;
    ctx->ctab_ok = true;
    ctx->ttab_XH = -1.0;  // the temperature table spans this table's box
    return SPH_OK;
}

namespace sphc {
__global__ void k_d2f(int64_t n, const double *__restrict__ a, float *__restrict__ b)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void k_f2d(int64_t n, const float *__restrict__ a, double *__restrict__ b)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}
}  // namespace sphc

extern "C" int sph_cooling_table_lookup(sph_ctx *ctx, int64_t n, const double *lognH, const double *logT,
                                        const double *logZ, const double *logJ21, double *logL_out, double *logG_out)
{
    CK(cool_prep(ctx, n));
    if (!ctx->ctab_ok) return fail(ctx, SPH_ERR_STATE, "no cooling table (call sph_cooling_table_build)");
    if (!lognH || !logT || !logZ || !logJ21 || !logL_out || !logG_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    const double *src[4] = {lognH, logT, logZ, logJ21};
    for (int k = 0; k < 4; k++) {
        HIPCK(hipMemcpyAsync(ctx->cin[k].p, src[k], sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
        // float copy in the upper half of the same buffer
        k_d2f<<<sph_blocks(n, 256), 256, 0, ctx->st>>>(n, D(cin[k]), (float *)ctx->cin[4 + (k >> 1)].p + (k & 1) * n);
    }
    float *lf = (float *)ctx->cin[4].p, *lg = (float *)ctx->cin[5].p;
    float *oL = (float *)ctx->cout[0].p, *oG = oL + n;
    (void)lg;
    HIPCK(sph::cool_table_lookup(ctx->st, n, ctx->cax, P_<float>(ctx->ctabL), P_<float>(ctx->ctabG), lf, lf + n,
                                 (float *)ctx->cin[5].p, (float *)ctx->cin[5].p + n, oL, oG));
    k_f2d<<<sph_blocks(n, 256), 256, 0, ctx->st>>>(n, oL, D(cin[0]));
    k_f2d<<<sph_blocks(n, 256), 256, 0, ctx->st>>>(n, oG, D(cin[1]));
    HIPCK(hipMemcpyAsync(logL_out, ctx->cin[0].p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipMemcpyAsync(logG_out, ctx->cin[1].p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    return SPH_OK;
}

// =========================================================== cooling surrogate
namespace sphc {

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double unit(uint64_t &s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

MLPDev mlp_dev(sph_ctx *ctx)
{
    MLPDev d;
    d.P = P_<float>(ctx->mP);
    d.G = P_<float>(ctx->mG);
    d.M = P_<float>(ctx->mM);
    d.V = P_<float>(ctx->mV);
    d.W2b = P_<unsigned short>(ctx->mW[0]);
    d.W3b = P_<unsigned short>(ctx->mW[1]);
    d.W2tb = P_<unsigned short>(ctx->mW[2]);
    d.W3tb = P_<unsigned short>(ctx->mW[3]);
    d.img = P_<uint4>(ctx->mimg);
    return d;
}

int mlp_alloc(sph_ctx *ctx)
{
    if (!ctx) return SPH_ERR_ARG;
    HIPCK(hipSetDevice(ctx->device));
    Buf *ps[] = {&ctx->mP, &ctx->mG, &ctx->mM, &ctx->mV};
    for (Buf *b : ps) CK(ensure(ctx, *b, (size_t)MLP_NPARAM * 4));
    for (int k = 0; k < 4; k++) CK(ensure(ctx, ctx->mW[k], (size_t)MLP_HIDDEN * MLP_HIDDEN * 2));
    CK(ensure(ctx, ctx->mimg, (size_t)MLP_IMG_BYTES));
    CK(ensure(ctx, ctx->mcnt, 64));
    return SPH_OK;
}

// Upload parameters, refresh the BF16 copies, clear gradient and Adam moments.
int mlp_load(sph_ctx *ctx, const float *params)
{
    CK(mlp_alloc(ctx));
    HIPCK(hipMemcpyAsync(ctx->mP.p, params, (size_t)MLP_NPARAM * 4, hipMemcpyHostToDevice, ctx->st));
    HIPCK(hipMemsetAsync(ctx->mG.p, 0, (size_t)MLP_NPARAM * 4, ctx->st));
    HIPCK(hipMemsetAsync(ctx->mM.p, 0, (size_t)MLP_NPARAM * 4, ctx->st));
    HIPCK(hipMemsetAsync(ctx->mV.p, 0, (size_t)MLP_NPARAM * 4, ctx->st));
    HIPCK(sph::mlp_refresh(ctx->st, mlp_dev(ctx)));
    HIPCK(hipStreamSynchronize(ctx->st));
    ctx->mlp_step = 0;
    ctx->mlp_ok = true;
    ctx->mlp_cal = false;
    return SPH_OK;
}

int mlp_need(sph_ctx *ctx, bool table, bool cal)
{
    if (!ctx) return SPH_ERR_ARG;
    HIPCK(hipSetDevice(ctx->device));
    if (!ctx->mlp_ok) return fail(ctx, SPH_ERR_STATE, "no surrogate parameters (call sph_mlp_init or sph_mlp_set_params)");
    if (table && !ctx->ctab_ok) return fail(ctx, SPH_ERR_STATE, "no cooling table (call sph_cooling_table_build)");
    if (cal && !ctx->mlp_cal) return fail(ctx, SPH_ERR_STATE, "no error map (call sph_mlp_calibrate)");
    return SPH_OK;
}

__global__ void k_affine(int64_t n, float *__restrict__ a, float mu, float inv_sigma)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) a[i] = (a[i] - mu) * inv_sigma;
}

__global__ void k_rows_to_cols(int64_t n, const float *__restrict__ rows, float *__restrict__ cols)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n)
        for (int a = 0; a < 4; a++) cols[a * n + i] = rows[4 * i + a];
}

// Training samples: n points uniform (per = 0) or stratified over cells, log10
// coordinates in mx, normalised in mxn, exact log10 Lambda (normalised if
// `normalise`) in mt.
int mlp_make_samples(sph_ctx *ctx, int64_t n, uint64_t seed, const int *cells, int per, bool normalise)
{
    CK(ensure(ctx, ctx->mx, (size_t)n * 16));
    CK(ensure(ctx, ctx->mxn, (size_t)n * 16));
    CK(ensure(ctx, ctx->mt, (size_t)n * 4));
    HIPCK(sph::mlp_samples(ctx->st, n, seed, ctx->mnorm, cells, per, P_<float>(ctx->mx), P_<float>(ctx->mxn)));
    HIPCK(sph::cool_log_lambda_points(ctx->st, n, P_<float>(ctx->mx), D(rf1), P_<float>(ctx->mt)));
    if (normalise)
        k_affine<<<sph_blocks(n, 256), 256, 0, ctx->st>>>(n, P_<float>(ctx->mt), ctx->mnorm.mu,
                                                          1.0f / ctx->mnorm.sigma);
    return SPH_OK;
}

}  // namespace sphc

extern "C" int sph_mlp_init(sph_ctx *ctx, uint64_t seed)
{
    CK(cool_prep(ctx, 1));
    if (!ctx->ctab_ok) return fail(ctx, SPH_ERR_STATE, "no cooling table (call sph_cooling_table_build)");
    const CoolAxes &ax = ctx->cax;
    MLPNorm nm{};
    for (int a = 0; a < 4; a++) {
        nm.lo[a] = (float)ax.lo[a];
        nm.hi[a] = (float)(ax.lo[a] + (ax.n[a] - 1) * ax.d[a]);
    }
    // target normalisation from the table's log10 Lambda values
    const int64_t total = (int64_t)ax.n[0] * ax.n[1] * ax.n[2] * ax.n[3];
    std::vector<float> tab(total);
    HIPCK(hipMemcpyAsync(tab.data(), ctx->ctabL.p, (size_t)total * 4, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    double s1 = 0.0, s2 = 0.0;
    for (float v : tab) {
        s1 += v;
        s2 += (double)v * v;
    }
    const double mu = s1 / total, var = std::max(s2 / total - mu * mu, 1e-12);
    nm.mu = (float)mu;
    nm.sigma = (float)std::sqrt(var);
    // He-uniform hidden layers, Glorot-uniform output, zero biases
    std::vector<float> P(MLP_NPARAM, 0.0f);
    uint64_t s = seed;
    auto fill = [&](int off, int count, int fan_in, double gain) {
        const double a = std::sqrt(gain / fan_in);
        for (int i = 0; i < count; i++) P[off + i] = (float)((2.0 * unit(s) - 1.0) * a);
    };
    const int H = MLP_HIDDEN;
    fill(0, H * 4, 4, 6.0);
    fill(H * 4 + H, H * H, H, 6.0);
    fill(H * 4 + H + H * H + H, H * H, H, 6.0);
    fill(H * 4 + H + 2 * (H * H + H), H, H, 3.0);
    ctx->mnorm = nm;
    return mlp_load(ctx, P.data());
}

extern "C" int sph_mlp_set_params(sph_ctx *ctx, const float *params, const double *norm)
{
    if (!ctx) return SPH_ERR_ARG;
    if (!params || !norm) return fail(ctx, SPH_ERR_ARG, "null argument");
    MLPNorm nm{};
    for (int a = 0; a < 4; a++) {
        nm.lo[a] = (float)norm[a];
        nm.hi[a] = (float)norm[4 + a];
        if (!(nm.hi[a] > nm.lo[a])) return fail(ctx, SPH_ERR_ARG, "norm: hi must exceed lo on every axis");
    }
    nm.mu = (float)norm[8];
    nm.sigma = (float)norm[9];
    if (!(nm.sigma > 0.0f)) return fail(ctx, SPH_ERR_ARG, "norm: sigma must be positive");
    ctx->mnorm = nm;
    return mlp_load(ctx, params);
}

extern "C" int sph_mlp_get_params(sph_ctx *ctx, float *params, double *norm)
{
    CK(mlp_need(ctx, false, false));
    if (params) {
        HIPCK(hipMemcpyAsync(params, ctx->mP.p, (size_t)MLP_NPARAM * 4, hipMemcpyDeviceToHost, ctx->st));
        HIPCK(hipStreamSynchronize(ctx->st));
    }
    if (norm) {
        for (int a = 0; a < 4; a++) {
            norm[a] = ctx->mnorm.lo[a];
            norm[4 + a] = ctx->mnorm.hi[a];
        }
        norm[8] = ctx->mnorm.mu;
        norm[9] = ctx->mnorm.sigma;
    }
    return SPH_OK;
}

extern "C" int sph_mlp_forward(sph_ctx *ctx, int64_t n, const float *x, float *y)
{
    CK(mlp_need(ctx, false, false));
    if (n < 1 || !x || !y) return fail(ctx, SPH_ERR_ARG, "need n >= 1 and non-null x, y");
    CK(ensure(ctx, ctx->mxn, (size_t)n * 16));
    CK(ensure(ctx, ctx->my, (size_t)n * 4));
    HIPCK(hipMemcpyAsync(ctx->mxn.p, x, (size_t)n * 16, hipMemcpyHostToDevice, ctx->st));
    HIPCK(sph::mlp_forward_raw(ctx->st, n, P_<float>(ctx->mxn), mlp_dev(ctx), P_<float>(ctx->my)));
    HIPCK(hipMemcpyAsync(y, ctx->my.p, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    return SPH_OK;
}

namespace sphc {
// Gradient of the mean squared error over the n given samples into the device
// gradient buffer (cleared first); *loss gets the mean.
int mlp_grad_on(sph_ctx *ctx, int64_t n, const float *x, const float *t, double *loss)
{
    CK(ensure(ctx, ctx->mxn, (size_t)n * 16));
    CK(ensure(ctx, ctx->mt, (size_t)n * 4));
    CK(ensure(ctx, ctx->mloss, 64));
    HIPCK(hipMemcpyAsync(ctx->mxn.p, x, (size_t)n * 16, hipMemcpyHostToDevice, ctx->st));
    HIPCK(hipMemcpyAsync(ctx->mt.p, t, (size_t)n * 4, hipMemcpyHostToDevice, ctx->st));
    HIPCK(hipMemsetAsync(ctx->mG.p, 0, (size_t)MLP_NPARAM * 4, ctx->st));
    HIPCK(hipMemsetAsync(ctx->mloss.p, 0, 4, ctx->st));
    HIPCK(sph::mlp_train_batch(ctx->st, n, 0, n, 1, 0, P_<float>(ctx->mxn), P_<float>(ctx->mt), mlp_dev(ctx),
                               P_<float>(ctx->mloss)));
    if (loss) {
        float l = 0.0f;
        HIPCK(hipMemcpyAsync(&l, ctx->mloss.p, 4, hipMemcpyDeviceToHost, ctx->st));
        HIPCK(hipStreamSynchronize(ctx->st));
        *loss = l;
    }
    return SPH_OK;
}
}  // namespace sphc

extern "C" int sph_mlp_gradient(sph_ctx *ctx, int64_t n, const float *x, const float *t, float *grad, double *loss)
{
    CK(mlp_need(ctx, false, false));
    if (n < 1 || !x || !t || !grad) return fail(ctx, SPH_ERR_ARG, "need n >= 1 and non-null x, t, grad");
    CK(mlp_grad_on(ctx, n, x, t, loss));
    HIPCK(hipMemcpyAsync(grad, ctx->mG.p, (size_t)MLP_NPARAM * 4, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipMemsetAsync(ctx->mG.p, 0, (size_t)MLP_NPARAM * 4, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    return SPH_OK;
}

extern "C" int sph_mlp_train_batch(sph_ctx *ctx, int64_t n, const float *x, const float *t, double lr, double *loss)
{
    CK(mlp_need(ctx, false, false));
    if (n < 1 || !x || !t || !(lr > 0.0)) return fail(ctx, SPH_ERR_ARG, "need n >= 1, non-null x, t and lr > 0");
    CK(mlp_grad_on(ctx, n, x, t, loss));
    HIPCK(sph::mlp_adam(ctx->st, mlp_dev(ctx), (float)lr, 0.9f, 0.999f, 1e-8f, ++ctx->mlp_step));
    HIPCK(sph::mlp_image(ctx->st, mlp_dev(ctx)));
    HIPCK(hipStreamSynchronize(ctx->st));
    ctx->mlp_cal = false;
    return SPH_OK;
}

extern "C" int sph_mlp_train(sph_ctx *ctx, int64_t n_samples, int epochs, int batch, double lr, uint64_t seed,
                             double *loss_out)
{
    CK(cool_prep(ctx, 1));
    CK(mlp_need(ctx, true, false));
    if (n_samples < 1 || n_samples > ((int64_t)1 << 27) || epochs < 1 || batch < 32 || !(lr > 0.0))
        return fail(ctx, SPH_ERR_ARG, "need 1 <= n_samples <= 2^27, epochs >= 1, batch >= 32, lr > 0");
    CK(mlp_make_samples(ctx, n_samples, seed, nullptr, 0, true));
    CK(ensure(ctx, ctx->mloss, (size_t)epochs * 4));
    HIPCK(hipMemsetAsync(ctx->mloss.p, 0, (size_t)epochs * 4, ctx->st));
    const int64_t nb = (n_samples + batch - 1) / batch;
    const double total = (double)nb * epochs, lr_min = lr * 0.01;
    uint64_t rs = seed ^ 0x5DEECE66Dull;
    const MLPDev d = mlp_dev(ctx);
    for (int e = 0; e < epochs; e++) {
        // a fresh permutation j -> (mult*j + off) mod n with gcd(mult, n) = 1
        int64_t mult = 1, off = (int64_t)(splitmix(rs) % (uint64_t)n_samples);
        for (int tries = 0; tries < 64; tries++) {
            const int64_t cand = 1 + (int64_t)(splitmix(rs) % (uint64_t)n_samples);
            int64_t a = cand, b = n_samples;
            while (b) {
                const int64_t r = a % b;
                a = b;
                b = r;
            }
            if (a == 1) {
                mult = cand % n_samples == 0 ? 1 : cand % n_samples;
                break;
            }
        }
        for (int64_t b = 0; b < nb; b++) {
            const int64_t first = b * batch, bsz = std::min<int64_t>(batch, n_samples - first);
            const double tt = (double)(e * nb + b) / total;
            const double lr_t = lr_min + 0.5 * (lr - lr_min) * (1.0 + std::cos(SPH_PI * tt));
            HIPCK(sph::mlp_train_batch(ctx->st, n_samples, first, bsz, mult, off, P_<float>(ctx->mxn),
                                       P_<float>(ctx->mt), d, P_<float>(ctx->mloss) + e));
            HIPCK(sph::mlp_adam(ctx->st, d, (float)lr_t, 0.9f, 0.999f, 1e-8f, ++ctx->mlp_step));
        }
    }
    HIPCK(sph::mlp_image(ctx->st, d));
    std::vector<float> l(epochs);
    HIPCK(hipMemcpyAsync(l.data(), ctx->mloss.p, (size_t)epochs * 4, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    if (loss_out)
        for (int e = 0; e < epochs; e++) loss_out[e] = (double)l[e] / (double)nb;
    ctx->mlp_cal = false;
    return SPH_OK;
}

extern "C" int sph_mlp_calibrate(sph_ctx *ctx, const int *cells, int per, uint64_t seed, float *emap_table,
                                 float *emap_exact, float *emap_table_exact)
{
    CK(cool_prep(ctx, 1));
    CK(mlp_need(ctx, true, false));
    if (!cells || per < 1) return fail(ctx, SPH_ERR_ARG, "need cells[4] and per >= 1");
    int64_t ncell = 1;
    for (int a = 0; a < 4; a++) {
        if (cells[a] < 1 || cells[a] > 4096) return fail(ctx, SPH_ERR_ARG, "cells per axis must be in [1, 4096]");
        ncell *= cells[a];
    }
    const int64_t nval = ncell * per;
    if (ncell > ((int64_t)1 << 24) || nval > ((int64_t)1 << 27))
        return fail(ctx, SPH_ERR_ARG, "error map too large (cells <= 2^24, cells*per <= 2^27)");
    CK(mlp_make_samples(ctx, nval, seed, cells, per, false));
    CK(ensure(ctx, ctx->my, (size_t)nval * 4));
    CK(ensure(ctx, ctx->memap, (size_t)ncell * 4));
    CK(ensure(ctx, ctx->memap2, (size_t)ncell * 4));
    CK(ensure(ctx, ctx->memap3, (size_t)ncell * 4));
    CK(ensure(ctx, ctx->mgate, (size_t)ncell * 4));
    HIPCK(hipMemsetAsync(ctx->memap.p, 0, (size_t)ncell * 4, ctx->st));
    HIPCK(hipMemsetAsync(ctx->memap2.p, 0, (size_t)ncell * 4, ctx->st));
    HIPCK(hipMemsetAsync(ctx->memap3.p, 0, (size_t)ncell * 4, ctx->st));
    HIPCK(sph::mlp_forward_raw(ctx->st, nval, P_<float>(ctx->mxn), mlp_dev(ctx), P_<float>(ctx->my)));
    HIPCK(sph::mlp_emap(ctx->st, nval, per, P_<float>(ctx->my), P_<float>(ctx->mx), P_<float>(ctx->mt), ctx->mnorm,
                        ctx->cax, P_<float>(ctx->ctabL), P_<float>(ctx->memap), P_<float>(ctx->memap2),
                        P_<float>(ctx->memap3)));
    HIPCK(sph::mlp_gate_map(ctx->st, ncell, P_<float>(ctx->memap2), P_<float>(ctx->memap3), P_<float>(ctx->mgate)));
    if (emap_table)
        HIPCK(hipMemcpyAsync(emap_table, ctx->memap.p, (size_t)ncell * 4, hipMemcpyDeviceToHost, ctx->st));
    if (emap_exact)
        HIPCK(hipMemcpyAsync(emap_exact, ctx->memap2.p, (size_t)ncell * 4, hipMemcpyDeviceToHost, ctx->st));
    if (emap_table_exact)
        HIPCK(hipMemcpyAsync(emap_table_exact, ctx->memap3.p, (size_t)ncell * 4, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    for (int a = 0; a < 4; a++) ctx->mcells[a] = cells[a];
    ctx->mlp_cal = true;
    return SPH_OK;
}

extern "C" int sph_mlp_cooling(sph_ctx *ctx, int64_t n, const double *lognH, const double *logT, const double *logZ,
                               const double *logJ21, double tol, int ref, double *logL_out, int32_t *fallback_out,
                               int64_t *n_fallback)
{
    CK(cool_prep(ctx, n));
    CK(mlp_need(ctx, true, true));
    if (!lognH || !logT || !logZ || !logJ21 || !logL_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (ref != SPH_MLP_REF_EXACT && ref != SPH_MLP_REF_TABLE) return fail(ctx, SPH_ERR_ARG, "ref must be 0 or 1");
    const float *emap = ref == SPH_MLP_REF_EXACT ? P_<float>(ctx->mgate) : P_<float>(ctx->memap);
    CK(ensure(ctx, ctx->mx, (size_t)n * 16));
    CK(ensure(ctx, ctx->my, (size_t)n * 4));
    float *cols = P_<float>(ctx->mx);  // log10 inputs as four FP32 columns
    const double *src[4] = {lognH, logT, logZ, logJ21};
    for (int k = 0; k < 4; k++) {
        HIPCK(hipMemcpyAsync(ctx->cin[k].p, src[k], sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
        k_d2f<<<sph_blocks(n, 256), 256, 0, ctx->st>>>(n, D(cin[k]), cols + k * n);
    }
    float *oL = (float *)ctx->cout[0].p;
    int *flag = (int *)ctx->cout[1].p;
    HIPCK(hipMemsetAsync(ctx->mcnt.p, 0, 8, ctx->st));
    CK(ensure(ctx, ctx->mfl, (size_t)sph::mlp_infer_scratch_ints(n) * 4));
    HIPCK(sph::mlp_infer(ctx->st, n, cols, ctx->mnorm, mlp_dev(ctx), emap, ctx->mcells, (float)tol, ctx->cax,
                         P_<float>(ctx->ctabL), P_<float>(ctx->my), oL, flag, P_<unsigned long long>(ctx->mcnt),
                         I32(mfl)));
    k_f2d<<<sph_blocks(n, 256), 256, 0, ctx->st>>>(n, oL, D(cin[0]));
    HIPCK(hipMemcpyAsync(logL_out, ctx->cin[0].p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
    if (fallback_out)
        HIPCK(hipMemcpyAsync(fallback_out, flag, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->st));
    unsigned long long nf = 0;
    HIPCK(hipMemcpyAsync(&nf, ctx->mcnt.p, 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    if (n_fallback) *n_fallback = (int64_t)nf;
    return SPH_OK;
}

extern "C" int sph_mlp_timing(sph_ctx *ctx, int64_t n, int reps, uint64_t seed, double tol, double *out)
{
    CK(cool_prep(ctx, 1));
    CK(mlp_need(ctx, true, true));
    if (n < 1 || n > ((int64_t)1 << 26) || reps < 1 || !out)
        return fail(ctx, SPH_ERR_ARG, "need 1 <= n <= 2^26, reps >= 1, out[4]");
    CK(mlp_make_samples(ctx, n, seed, nullptr, 0, false));
    CK(ensure(ctx, ctx->my, (size_t)n * 16));
    CK(ensure(ctx, ctx->cout[0], (size_t)n * 8));
    CK(ensure(ctx, ctx->cout[1], (size_t)n * 8));
    float *cols = P_<float>(ctx->my);  // log10 inputs as four columns
    k_rows_to_cols<<<sph_blocks(n, 256), 256, 0, ctx->st>>>(n, P_<float>(ctx->mx), cols);
    float *oL = (float *)ctx->cout[0].p;
    int *flag = (int *)ctx->cout[1].p;
    const MLPDev d = mlp_dev(ctx);
    CK(ensure(ctx, ctx->mfl, (size_t)sph::mlp_infer_scratch_ints(n) * 4));
    hipEvent_t e0, e1;
    HIPCK(hipEventCreate(&e0));
    HIPCK(hipEventCreate(&e1));
    float ms[3] = {0, 0, 0};
    for (int which = 0; which < 3; which++) {
        for (int r = -1; r < reps; r++) {  // r = -1: warm-up launch
            if (r == 0) HIPCK(hipEventRecord(e0, ctx->st));
            if (which == 0) {
                HIPCK(hipMemsetAsync(ctx->mcnt.p, 0, 8, ctx->st));
                HIPCK(sph::mlp_infer(ctx->st, n, cols, ctx->mnorm, d, P_<float>(ctx->mgate), ctx->mcells, (float)tol,
                                     ctx->cax, P_<float>(ctx->ctabL), P_<float>(ctx->mt), oL, flag,
                                     P_<unsigned long long>(ctx->mcnt), I32(mfl)));
            } else if (which == 1) {
                HIPCK(sph::cool_table_lookup(ctx->st, n, ctx->cax, P_<float>(ctx->ctabL), nullptr, cols, cols + n,
                                             cols + 2 * n, cols + 3 * n, oL, nullptr));
            } else {
                HIPCK(sph::mlp_forward_raw(ctx->st, n, P_<float>(ctx->mxn), d, oL));
            }
        }
        HIPCK(hipEventRecord(e1, ctx->st));
        HIPCK(hipEventSynchronize(e1));
        HIPCK(hipEventElapsedTime(&ms[which], e0, e1));
    }
    unsigned long long nf = 0;
    HIPCK(hipMemcpy(&nf, ctx->mcnt.p, 8, hipMemcpyDeviceToHost));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out[0] = ms[0] / reps;
    out[1] = ms[1] / reps;
    out[2] = ms[2] / reps;
    out[3] = (double)nf / (double)n;
    return SPH_OK;
}

// =========================================================== cooling update
namespace sphc {

int cool_check(sph_ctx *ctx, const sph_cool_params *cp)
{
    if (!cp) return fail(ctx, SPH_ERR_ARG, "null cooling parameters");
    if (cp->mode != SPH_COOL_EXACT && cp->mode != SPH_COOL_TABLE && cp->mode != SPH_COOL_MLP &&
        cp->mode != SPH_COOL_NH)
        return fail(ctx, SPH_ERR_ARG, "mode must be SPH_COOL_EXACT, SPH_COOL_TABLE, SPH_COOL_MLP or SPH_COOL_NH");
    if (!(cp->unit_density_cgs > 0.0) || !(cp->unit_u_cgs > 0.0) || !(cp->unit_time_s > 0.0) ||
        !(cp->XH > 0.0 && cp->XH <= 1.0) || !(cp->u_floor_cgs >= 0.0))
        return fail(ctx, SPH_ERR_ARG, "cooling units must be > 0, 0 < XH <= 1 and u_floor_cgs >= 0");
    if (cp->mode == SPH_COOL_NH) {
        if (!ctx->nh_ok) return fail(ctx, SPH_ERR_STATE, "no N_H cooling table (call sph_nh_table_set)");
        return SPH_OK;
    }
    if (!ctx->ctab_ok) return fail(ctx, SPH_ERR_STATE, "no cooling table (call sph_cooling_table_build)");
    if (cp->mode == SPH_COOL_MLP) {
        CK(mlp_need(ctx, true, true));
        if (cp->ref != SPH_MLP_REF_EXACT && cp->ref != SPH_MLP_REF_TABLE)
            return fail(ctx, SPH_ERR_ARG, "ref must be SPH_MLP_REF_EXACT or SPH_MLP_REF_TABLE");
        if (!(cp->tol > 0.0)) return fail(ctx, SPH_ERR_ARG, "tol must be > 0");
    }
    return SPH_OK;
}

double tab_hi(const CoolAxes &ax, int a) { return ax.lo[a] + (ax.n[a] - 1) * ax.d[a]; }

// Temperature table over the 4D table's n_H and J21 ranges (0.1 and 0.125 dex) and the
// log u range that maps its T range for mean molecular weights 1.3 (neutral) to 0.55
// (fully ionised) at gamma = 5/3, 0.02 dex.
int ensure_ttable(sph_ctx *ctx, double XH)
{
    if (ctx->ttab_XH == XH) return SPH_OK;
    const CoolAxes &ax = ctx->cax;
    const double kb = 1.3807e-16, mh = 1.6726e-24;
    double lo[3] = {ax.lo[0], std::log10(kb * std::pow(10.0, ax.lo[1]) / (2.0 / 3.0 * 1.3 * mh)) - 0.1, ax.lo[3]};
    double hi[3] = {tab_hi(ax, 0), std::log10(kb * std::pow(10.0, tab_hi(ax, 1)) / (2.0 / 3.0 * 0.55 * mh)) + 0.1,
                    tab_hi(ax, 3)};
    const double step[3] = {0.1, 0.02, 0.125};
    CoolTAxes t{};
    int64_t total = 1;
    for (int a = 0; a < 3; a++) {
        t.lo[a] = lo[a];
        t.n[a] = std::max(2, (int)std::ceil((hi[a] - lo[a]) / step[a] - 1e-9) + 1);
        t.d[a] = (hi[a] - lo[a]) / (t.n[a] - 1);
        total *= t.n[a];
    }
    if (total > ((int64_t)1 << 26)) return fail(ctx, SPH_ERR_ARG, "temperature table too large");
    CK(ensure(ctx, ctx->ttab, (size_t)total * 4));
    HIPCK(sph::cool_ttable_build(ctx->st, t, XH, D(rf1), P_<float>(ctx->ttab)));
    HIPCK(hipStreamSynchronize(ctx->st));
    ctx->tax = t;
    ctx->ttab_XH = XH;
    return SPH_OK;
}

CoolUpdate make_update(const sph_ctx *ctx, const sph_cool_params &cp, double dt)
{
    CoolUpdate cu{};
    cu.rho_cgs = cp.unit_density_cgs;
    cu.u_cgs = cp.unit_u_cgs;
    cu.dt_s = dt * cp.unit_time_s;
    cu.XH = cp.XH;
    cu.u_floor = cp.u_floor_cgs;
    for (int a = 0; a < 4; a++) {
        cu.lo[a] = ctx->cax.lo[a];
        cu.hi[a] = tab_hi(ctx->cax, a);
        cu.lin_lo[a] = std::pow(10.0, cu.lo[a]);
        cu.lin_hi[a] = std::pow(10.0, cu.hi[a]);
    }
    return cu;
}

template <typename RHO>
int cool_run(sph_ctx *ctx, int64_t n, const RHO *rho, const double *u, const double *Z, const double *J21,
             double dt, const sph_cool_params &cp, double *u_out, unsigned long long *counts)
{
    hipStream_t st = ctx->st;
    const CoolUpdate cu = make_update(ctx, cp, dt);
    unsigned long long *spr = P_<unsigned long long>(ctx->spread) + SPR_C * SPH_SPREAD_WORDS;
    HIPCK(sph::fill<unsigned long long>(st, SPH_SPREAD_WORDS, spr, 0ull));
    HIPCK(sph::fill<unsigned long long>(st, 1, counts + 1, 0ull));
    if (cp.mode != SPH_COOL_EXACT) CK(ensure_ttable(ctx, cp.XH));
    const float *tabT = P_<float>(ctx->ttab);
    if (cp.mode != SPH_COOL_MLP) {
        const int kind = cp.mode == SPH_COOL_EXACT ? SPH_COOLK_EXACT : SPH_COOLK_TABLE;
        HIPCK(sph::cool_update<RHO>(st, n, kind, rho, u, Z, J21, cu, ctx->cax, P_<float>(ctx->ctabL),
                                    P_<float>(ctx->ctabG), ctx->tax, tabT, D(rf1), u_out, nullptr, nullptr, spr));
    } else {
        CK(ensure(ctx, ctx->ucols, (size_t)n * 16));
        CK(ensure(ctx, ctx->uyn, (size_t)n * 4));
        CK(ensure(ctx, ctx->uL, (size_t)n * 4));
        CK(ensure(ctx, ctx->uG, (size_t)n * 4));
        CK(ensure(ctx, ctx->uflag, (size_t)n * 4));
        float *cols = P_<float>(ctx->ucols);
        HIPCK(sph::cool_update<RHO>(st, n, SPH_COOLK_PREP, rho, u, Z, J21, cu, ctx->cax, P_<float>(ctx->ctabL),
                                    P_<float>(ctx->ctabG), ctx->tax, tabT, D(rf1), nullptr, cols,
                                    P_<float>(ctx->uG), nullptr));
        const float *emap = cp.ref == SPH_MLP_REF_EXACT ? P_<float>(ctx->mgate) : P_<float>(ctx->memap);
        CK(ensure(ctx, ctx->mfl, (size_t)sph::mlp_infer_scratch_ints(n) * 4));
        HIPCK(sph::mlp_infer(st, n, cols, ctx->mnorm, mlp_dev(ctx), emap, ctx->mcells, (float)cp.tol, ctx->cax,
                             P_<float>(ctx->ctabL), P_<float>(ctx->uyn), P_<float>(ctx->uL), P_<int>(ctx->uflag),
                             counts + 1, I32(mfl)));
        HIPCK(sph::cool_apply<RHO>(st, n, rho, u, P_<float>(ctx->uL), P_<float>(ctx->uG), cu, u_out, spr));
    }
    HIPCK(sph::spread_total(st, 1, spr, counts));
    return SPH_OK;
}

NHTable nh_table(const sph_ctx *ctx)
{
    NHTable t;
    int ng = 0;
    for (int k = 0; k < 4; k++) {
        t.dims[k] = ctx->nh_dims[k];
        ng += ctx->nh_dims[k];
    }
    for (int k = 0; k < 8; k++) t.bounds[k] = ctx->nh_bounds[k];
    const size_t cells = (size_t)t.dims[0] * t.dims[1] * t.dims[2] * t.dims[3];
    const float *base = static_cast<const float *>(ctx->nhtab.p);
    const size_t off = ((size_t)ng + 63) / 64 * 64;
    t.grids = base;
    t.heat = base + off;
    t.cool = t.heat + cells;
    t.mu = t.cool + cells;
    t.logu = t.mu + cells;
    return t;
}

template <typename RHO>
int nh_run(sph_ctx *ctx, int64_t n, const double *x, const double *y, const double *z, const RHO *rho,
           const double *u, const double *log_nh, const int32_t *typ, double gamma, double dt,
           const sph_cool_params &cp, double *u_out, unsigned long long *counts)
{
    if (!ctx->nh_ok) return fail(ctx, SPH_ERR_STATE, "no N_H cooling table (call sph_nh_table_set)");
    hipStream_t st = ctx->st;
    unsigned long long *spr = P_<unsigned long long>(ctx->spread) + SPR_C * SPH_SPREAD_WORDS;
    HIPCK(sph::fill<unsigned long long>(st, SPH_SPREAD_WORDS, spr, 0ull));
    HIPCK(sph::fill<unsigned long long>(st, 1, counts + 1, 0ull));
    const NHCoolArgs a{(float)cp.unit_density_cgs, (float)cp.unit_u_cgs, (float)cp.unit_time_s, (float)cp.XH,
                            (float)gamma, (float)dt, cp.u_floor_cgs};
    HIPCK(sph::nh_cool_update<RHO>(st, n, x, y, z, rho, u, log_nh, typ, nh_table(ctx), a, u_out, spr));
    HIPCK(sph::spread_total(st, 1, spr, counts));
    return SPH_OK;
}

template int nh_run<float>(sph_ctx *, int64_t, const double *, const double *, const double *, const float *,
                           const double *, const double *, const int32_t *, double, double, const sph_cool_params &,
                           double *, unsigned long long *);
template int nh_run<double>(sph_ctx *, int64_t, const double *, const double *, const double *, const double *,
                            const double *, const double *, const int32_t *, double, double, const sph_cool_params &,
                            double *, unsigned long long *);

template int cool_run<float>(sph_ctx *, int64_t, const float *, const double *, const double *, const double *, double,
                             const sph_cool_params &, double *, unsigned long long *);
template int cool_run<double>(sph_ctx *, int64_t, const double *, const double *, const double *, const double *,
                              double, const sph_cool_params &, double *, unsigned long long *);

}  // namespace sphc

extern "C" int sph_nh_table_set(sph_ctx *ctx, const int32_t *dims, const float *bounds, const float *nH_grid,
                                const float *T_grid, const float *r_grid, const float *NH_grid, const float *heat,
                                const float *cool, const float *mu, const float *logu)
{
    if (!ctx) return SPH_ERR_ARG;
    if (!dims || !bounds || !nH_grid || !T_grid || !r_grid || !NH_grid || !heat || !cool || !mu || !logu)
        return fail(ctx, SPH_ERR_ARG, "null argument");
    int ng = 0;
    int64_t cells = 1;
    for (int k = 0; k < 4; k++) {
        if (dims[k] < 2) return fail(ctx, SPH_ERR_ARG, "every N_H table axis needs at least 2 points");
        ng += dims[k];
        cells *= dims[k];
    }
    if (ng > SPH_NH_GRID_MAX) return fail(ctx, SPH_ERR_ARG, "N_H table grids exceed 2048 points together");
    if (cells > ((int64_t)1 << 28)) return fail(ctx, SPH_ERR_ARG, "N_H table too large");
    const float *grids[4] = {nH_grid, T_grid, r_grid, NH_grid};
    for (int k = 0; k < 4; k++) {
        for (int q = 0; q < dims[k]; q++)
            if (!std::isfinite(grids[k][q]) || (q > 0 && !(grids[k][q] > grids[k][q - 1])))
                return fail(ctx, SPH_ERR_ARG, "N_H table grids must be finite and strictly ascending");
        if (!std::isfinite(bounds[2 * k]) || !std::isfinite(bounds[2 * k + 1]))
            return fail(ctx, SPH_ERR_ARG, "N_H table bounds must be finite");
    }
    HIPCK(hipSetDevice(ctx->device));
    const size_t off = ((size_t)ng + 63) / 64 * 64;
    CK(ensure(ctx, ctx->nhtab, (off + 4 * (size_t)cells) * sizeof(float)));
    std::vector<float> g;
    g.reserve(off);
    for (int k = 0; k < 4; k++) g.insert(g.end(), grids[k], grids[k] + dims[k]);
    g.resize(off, 0.0f);
    float *base = P_<float>(ctx->nhtab);
    HIPCK(hipMemcpyAsync(base, g.data(), off * sizeof(float), hipMemcpyHostToDevice, ctx->st));
    const float *tabs[4] = {heat, cool, mu, logu};
    for (int k = 0; k < 4; k++)
        HIPCK(hipMemcpyAsync(base + off + k * (size_t)cells, tabs[k], (size_t)cells * sizeof(float),
                             hipMemcpyHostToDevice, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    for (int k = 0; k < 4; k++) ctx->nh_dims[k] = dims[k];
    for (int k = 0; k < 8; k++) ctx->nh_bounds[k] = bounds[k];
    ctx->nh_ok = true;
    return SPH_OK;
}

extern "C" int sph_nh_cooling(sph_ctx *ctx, int64_t n, const double *pos, const double *rho, const double *u,
                              const double *log_nh, const int32_t *typ, double gamma, double dt,
                              const sph_cool_params *cp, double *u_out, int64_t *n_floor)
{
    CK(cool_prep(ctx, n));
    if (!pos || !rho || !u || !log_nh || !u_out || !cp) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (!(dt >= 0.0) || !(gamma > 1.0)) return fail(ctx, SPH_ERR_ARG, "dt must be >= 0 and gamma > 1");
    sph_cool_params c = *cp;
    c.mode = SPH_COOL_NH;
    CK(cool_check(ctx, &c));
    CK(ensure(ctx, ctx->spread, (size_t)SPR_TOTAL * SPH_SPREAD_WORDS * 8));
    CK(ensure(ctx, ctx->ucnt, 2 * 8));
    if (typ) CK(ensure(ctx, ctx->uflag, (size_t)n * 4));
    std::vector<double> xyz((size_t)n * 3);
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) xyz[(size_t)d * n + i] = pos[3 * i + d];
    for (int d = 0; d < 3; d++)
        HIPCK(hipMemcpyAsync(ctx->cin[d].p, xyz.data() + (size_t)d * n, sizeof(double) * n, hipMemcpyHostToDevice,
                             ctx->st));
    const double *src[3] = {rho, u, log_nh};
    for (int k = 0; k < 3; k++)
        HIPCK(hipMemcpyAsync(ctx->cin[3 + k].p, src[k], sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
    if (typ) HIPCK(hipMemcpyAsync(ctx->uflag.p, typ, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->st));
    CK(nh_run<double>(ctx, n, D(cin[0]), D(cin[1]), D(cin[2]), D(cin[3]), D(cin[4]), D(cin[5]),
                      typ ? I32(uflag) : nullptr, gamma, dt, c, D(cout[0]), P_<unsigned long long>(ctx->ucnt)));
    HIPCK(hipMemcpyAsync(u_out, ctx->cout[0].p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
    unsigned long long cc[2] = {0, 0};
    HIPCK(hipMemcpyAsync(cc, ctx->ucnt.p, sizeof(cc), hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    if (n_floor) *n_floor = (int64_t)cc[0];
    return SPH_OK;
}

extern "C" int sph_cooling_update(sph_ctx *ctx, int64_t n, const double *rho, const double *u, const double *Z,
                                  const double *J21, double dt, const sph_cool_params *cp, double *u_out,
                                  int32_t *fallback_out, int64_t *n_floor, int64_t *n_fallback)
{
    CK(cool_prep(ctx, n));
    if (!rho || !u || !Z || !J21 || !u_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (!(dt >= 0.0)) return fail(ctx, SPH_ERR_ARG, "dt must be >= 0");
    CK(cool_check(ctx, cp));
    CK(ensure(ctx, ctx->spread, (size_t)SPR_TOTAL * SPH_SPREAD_WORDS * 8));
    CK(ensure(ctx, ctx->ucnt, 2 * 8));
    const double *src[4] = {rho, u, Z, J21};
    for (int k = 0; k < 4; k++)
        HIPCK(hipMemcpyAsync(ctx->cin[k].p, src[k], sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
    CK(cool_run<double>(ctx, n, D(cin[0]), D(cin[1]), D(cin[2]), D(cin[3]), dt, *cp, D(cout[0]),
                        P_<unsigned long long>(ctx->ucnt)));
    HIPCK(hipMemcpyAsync(u_out, ctx->cout[0].p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
    if (fallback_out && cp->mode == SPH_COOL_MLP)
        HIPCK(hipMemcpyAsync(fallback_out, ctx->uflag.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->st));
    unsigned long long cc[2] = {0, 0};
    HIPCK(hipMemcpyAsync(cc, ctx->ucnt.p, sizeof(cc), hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    if (fallback_out && cp->mode != SPH_COOL_MLP) std::memset(fallback_out, 0, sizeof(int32_t) * n);
    if (n_floor) *n_floor = (int64_t)cc[0];
    if (n_fallback) *n_fallback = (int64_t)cc[1];
    return SPH_OK;
}

extern "C" int sph_cooling_temperature(sph_ctx *ctx, int64_t n, const double *u, const double *nH, const double *J21,
                                       double XH, double *T_out)
{
    CK(cool_prep(ctx, n));
    if (!u || !nH || !J21 || !T_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (!(XH > 0.0 && XH <= 1.0)) return fail(ctx, SPH_ERR_ARG, "need 0 < XH <= 1");
    if (!ctx->ctab_ok) return fail(ctx, SPH_ERR_STATE, "no cooling table (call sph_cooling_table_build)");
    CK(ensure_ttable(ctx, XH));
    const double *src[3] = {u, nH, J21};
    for (int k = 0; k < 3; k++)
        HIPCK(hipMemcpyAsync(ctx->cin[k].p, src[k], sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
    HIPCK(sph::cool_ttable_lookup(ctx->st, n, ctx->tax, P_<float>(ctx->ttab), D(cin[0]), D(cin[1]), D(cin[2]), 0.0,
                                  D(cout[0])));
    HIPCK(hipMemcpyAsync(T_out, ctx->cout[0].p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    return SPH_OK;
}


extern "C" int sph_state_set_cooling(sph_ctx *ctx, const double *Z, const double *J21, const sph_cool_params *cp)
{
    if (!ctx) return SPH_ERR_ARG;
    if (!ctx->resident) return fail(ctx, SPH_ERR_STATE, "no resident state (call sph_state_set first)");
    HIPCK(hipSetDevice(ctx->device));
    if (!cp) {
        ctx->cool_on = false;
        return SPH_OK;
    }
    if (cp->mode == SPH_COOL_NH) {  // Z and J21 are not used; the step's column density must be on
        CK(cool_check(ctx, cp));
        ctx->cool_p = *cp;
        ctx->cool_on = true;
        return SPH_OK;
    }
    if (!Z || !J21) return fail(ctx, SPH_ERR_ARG, "null argument");
    CK(cool_prep(ctx, 1));
    CK(cool_check(ctx, cp));
    CK(ensure(ctx, ctx->zmet, (size_t)std::max(ctx->n, ctx->n_reserve) * 8));
    CK(ensure(ctx, ctx->j21, (size_t)std::max(ctx->n, ctx->n_reserve) * 8));
    CK(put_scalar(ctx, Z, ctx->zmet, nullptr, nullptr, nullptr));
    CK(put_scalar(ctx, J21, ctx->j21, nullptr, nullptr, nullptr));
    if (cp->mode != SPH_COOL_EXACT) CK(ensure_ttable(ctx, cp->XH));
    HIPCK(hipStreamSynchronize(ctx->st));
    ctx->cool_p = *cp;
    ctx->cool_on = true;
    return SPH_OK;
}

