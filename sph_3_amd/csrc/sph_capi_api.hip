// sph_capi_api.hip -- the per-function entry points that mirror the reference's Python call
// surface (smoothing lengths, neighbours, density, div/curl, acceleration, du/dt, gravity,
// potential energy, column density) on the context's cached state.
#include "sph_ctx.h"

using namespace sphc;

// =========================================================== h solves
static_assert(SPH_PH_COUNT <= 20, "sph_stats holds 20 phases");

namespace sphc {

// Parameters of an h solve over m targets (scratch in the context); host counters go to
// unc (unconverged), rounds_out and work (distance tests, kernel terms).
int hsolve_params(sph_ctx *ctx, int64_t m, const double *tx, const double *ty, const double *tz, int mode,
                  double param, int n_low, int n_up, int max_steps, double tol, int64_t *unc, int64_t *rounds_out,
                  int64_t *work, sph::HSolveParams &p)
{
    if (mode != SPH_H_REF_COUNT) CK(ensure(ctx, ctx->ccidx, (size_t)m * SPH_HCAP * 4));
    // FP64 distances beside the indices while they fit the context's limit; beyond it the
    // solve recomputes them (the same bits) and the buffer is given back
    const bool rows_r2 = mode != SPH_H_REF_COUNT && (uint64_t)m * SPH_HCAP * 8 <= ctx->r2_limit;
    if (rows_r2) {
        CK(ensure(ctx, ctx->cand, (size_t)m * SPH_HCAP * 8));
    } else if (ctx->cand.p) {
        HIPCK(hipFree(ctx->cand.p));
        ctx->cand.p = nullptr;
        ctx->cand.bytes = 0;
    }
    CK(ensure(ctx, ctx->cccnt, (size_t)m * 4));
    p.m = m;
    p.tx = tx; p.ty = ty; p.tz = tz;
    p.x = D(x); p.y = D(y); p.z = D(z);
    p.nodes = P_<WNode>(ctx->nodes);
    p.bucket = SPH_BUCKET;
    p.mode = mode;
    p.n_ngb = param;
    p.tol = tol;
    p.n_low = n_low; p.n_up = n_up; p.max_steps = max_steps;
    p.kth = (int)param;
    p.cap = SPH_HCAP;
    p.r_fac = 1.06;  // candidate radius 2.12 h_prev: one round while h grows < 6 % per call
    p.h_in = D(hin);
    p.h_out = D(hout);
    p.R = D(hR);
    p.steps = I32(hsteps);
    p.state = I32(hstate);
    p.counters = P_<unsigned long long>(ctx->spread);
    p.totals = P_<unsigned long long>(ctx->hcount);
    p.unconverged_total = unc;
    p.groups = (tx == D(x)) ? I32(groups) : nullptr;  // targets are the sorted particles: tree-node groups
    p.ngroups = ctx->ngroups;
    p.rounds = rounds_out;
    p.work = work;
    p.cidx = mode != SPH_H_REF_COUNT ? I32(ccidx) : nullptr;
    p.cand = rows_r2 ? D(cand) : nullptr;
    p.ccnt = I32(cccnt);
    p.host_totals = ctx->hpin;
    p.done = ctx->ev_h;
    if (mode != SPH_H_REF_COUNT) {
        const int64_t ng = p.groups ? ctx->ngroups : 0;
        CK(ensure(ctx, ctx->hplist, (size_t)(sph::hsolve_plist_ints(m, ng) + 64) * 4));
        p.plist_counts = P_<unsigned long long>(ctx->hplist);
        p.plist = I32(hplist) + 64;
    }
    if (mode == SPH_H_NEWTON) {
        CK(ensure(ctx, ctx->hbig, (size_t)m * 4));
        p.big_list = I32(hbig);
        p.big_count = P_<unsigned long long>(ctx->hcount) + 14;
        p.n_big = &ctx->stats.h_big;
    }
    return SPH_OK;
}

int run_hsolve(sph_ctx *ctx, int64_t m, const double *tx, const double *ty, const double *tz, int mode,
               double param, int n_low, int n_up, int max_steps, double tol, int64_t *n_unconv, int64_t *rounds_out,
               int64_t *work = nullptr)
{
    sph::HSolveParams p;
    int64_t unc = 0;
    CK(hsolve_params(ctx, m, tx, ty, tz, mode, param, n_low, n_up, max_steps, tol, &unc, rounds_out, work, p));
    int64_t pending = 0;
    if (mode == SPH_H_REF_COUNT) HIPCK(sph::tree_escape(ctx->st, ctx->n, P_<WNode>(ctx->nodes)));  // stackless walk
    HIPCK(sph::hsolve(ctx->st, p, SPH_HROUNDS, &pending));
    if (pending) return fail(ctx, SPH_ERR_NOCONV, "smoothing-length solve: " + std::to_string(pending) +
                                                      " targets undecided after the maximum number of radius rounds");
    *n_unconv = unc;
    return SPH_OK;
}

}  // namespace sphc

extern "C" int sph_smoothing_length(sph_ctx *ctx, int64_t n, const double *pos, int64_t m, const double *targets,
                                    const double *h_in, int mode, double param, int n_low, int n_up, int max_steps,
                                    double tol, double *h_out, int64_t *n_unconverged)
{
    CK(check_common(ctx, n, SPH_PREC_FP64));
    if (!pos || !h_out || (mode != SPH_H_KTH && !h_in))
        return fail(ctx, SPH_ERR_ARG, "pos, h_out (and h_in unless mode = SPH_H_KTH) are required");
    if (mode < 0 || mode > 2) return fail(ctx, SPH_ERR_ARG, "unknown h mode");
    if (mode == SPH_H_KTH && (param < 0 || param >= (double)n))
        return fail(ctx, SPH_ERR_ARG, "kth must satisfy 0 <= kth < n");
    if (mode == SPH_H_KTH && (int)param + 1 > SPH_HCAP)
        return fail(ctx, SPH_ERR_ARG, "kth + 1 exceeds the candidate capacity (256)");
    if (targets && m < 1) return fail(ctx, SPH_ERR_ARG, "m must be >= 1");
    if (targets) {
        std::string why;
        for (int64_t i = 0; i < 3 * m; i++)
            if (!std::isfinite(targets[i])) return fail(ctx, SPH_ERR_ARG, "a target coordinate is not finite");
        if (h_in && !particles_ok(m, nullptr, nullptr, nullptr, nullptr, h_in, why)) return fail(ctx, SPH_ERR_ARG, why);
    } else if (h_in) {
        CK(check_particles(ctx, n, nullptr, nullptr, nullptr, nullptr, h_in));
    }
    CK(set_pos(ctx, n, pos));
    CK(ensure_moments(ctx));
    hipStream_t st = ctx->st;
    int64_t mm = targets ? m : n;
    if (targets) {
        if (mm > n) {
            CK(ensure(ctx, ctx->tx, (size_t)mm * 8));
            CK(ensure(ctx, ctx->ty, (size_t)mm * 8));
            CK(ensure(ctx, ctx->tz, (size_t)mm * 8));
            CK(ensure(ctx, ctx->hin, (size_t)mm * 8));
            CK(ensure(ctx, ctx->hout, (size_t)mm * 8));
            CK(ensure(ctx, ctx->hR, (size_t)mm * 8));
            CK(ensure(ctx, ctx->hsteps, (size_t)mm * 4));
            CK(ensure(ctx, ctx->hstate, (size_t)mm * 4));
            CK(ensure(ctx, ctx->stage, (size_t)mm * 24));
        }
        HIPCK(hipMemcpyAsync(ctx->stage.p, targets, sizeof(double) * 3 * mm, hipMemcpyHostToDevice, st));
        k_deinterleave<<<sph_blocks(mm, 256), 256, 0, st>>>(mm, D(stage), D(tx), D(ty), D(tz));
    }
    const double *tx = targets ? D(tx) : D(x), *ty = targets ? D(ty) : D(y), *tz = targets ? D(tz) : D(z);
    if (h_in) {
        if (targets) HIPCK(hipMemcpyAsync(ctx->hin.p, h_in, sizeof(double) * mm, hipMemcpyHostToDevice, st));
        else CK(put_scalar(ctx, h_in, ctx->hin, nullptr, nullptr, nullptr));
    } else {
        // k-th distance guess from the mean density of the bounding box
        double box[8];
        HIPCK(hipMemcpyAsync(box, ctx->box.p, sizeof(double) * 7, hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
        double L = box[6] > 0 ? 65536.0 / box[6] : 1.0;
        double g = 0.5 * L * std::cbrt((param + 1.0) / (double)n);
        HIPCK(sph::fill<double>(st, mm, D(hin), g));
    }
    int64_t unc = 0;
    CK(run_hsolve(ctx, mm, tx, ty, tz, mode, param, n_low, n_up, max_steps, tol, &unc, nullptr));
    if (n_unconverged) *n_unconverged = unc;
    if (targets) {
        HIPCK(hipMemcpyAsync(h_out, ctx->hout.p, sizeof(double) * mm, hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
    } else {
        CK(get_scalar<double>(ctx, ctx->hout, h_out));
    }
    return SPH_OK;
}

// =========================================================== neighbours
extern "C" int sph_neighbours(sph_ctx *ctx, int64_t n, const double *pos, const double *h, int64_t *offsets,
                              int32_t *idx, int64_t idx_cap)
{
    CK(check_common(ctx, n, SPH_PREC_FP64));
    if (!pos || !h || !offsets) return fail(ctx, SPH_ERR_ARG, "pos, h and offsets are required");
    CK(geom(ctx, n, pos, h, nullptr));
    CK(ensure_csr(ctx, true));  // offsets need the exact CSR layout
    int64_t total = ctx->csr_total;
    std::vector<int64_t> soff(n + 1);
    std::vector<int32_t> orig(n);
    HIPCK(hipMemcpyAsync(soff.data(), ctx->off.p, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipMemcpyAsync(orig.data(), ctx->orig.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    // offsets in caller order
    std::vector<int64_t> cnt(n);
    for (int64_t s = 0; s < n; s++) cnt[orig[s]] = soff[s + 1] - soff[s];
    offsets[0] = 0;
    for (int64_t i = 0; i < n; i++) offsets[i + 1] = offsets[i] + cnt[i];
    if (!idx || idx_cap < total)
        return idx ? fail(ctx, SPH_ERR_CAPACITY, "idx capacity " + std::to_string(idx_cap) + " < " +
                                                     std::to_string(total))
                   : SPH_OK;
    std::vector<int32_t> sidx(total);
    HIPCK(hipMemcpyAsync(sidx.data(), ctx->idx.p, sizeof(int32_t) * total, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    for (int64_t s = 0; s < n; s++) {
        int64_t r = orig[s], o = offsets[r];
        for (int64_t k = soff[s]; k < soff[s + 1]; k++) idx[o++] = orig[sidx[k]];
        std::sort(idx + offsets[r], idx + offsets[r + 1]);
    }
    return SPH_OK;
}

// Neighbour sets of the resident state (the last sph_prime / sph_step), in the caller's
// numbering: the same output as sph_neighbours at the state's positions and h.  Particles
// injected after that evaluation (sorted slots >= n_eval) have no set yet: empty rows.
extern "C" int sph_state_neighbours(sph_ctx *ctx, int64_t *offsets, int32_t *idx, int64_t idx_cap)
{
    if (!ctx || !offsets) return SPH_ERR_ARG;
    if (!ctx->resident || !ctx->primed || ctx->n_eval < 1)
        return fail(ctx, SPH_ERR_STATE, "no forces evaluated yet (sph_prime)");
    HIPCK(hipSetDevice(ctx->device));
    const int64_t n = ctx->n, ne = ctx->n_eval, K = ctx->ngb_stride;
    std::vector<int32_t> orig(n), scnt(n, 0);
    std::vector<int64_t> soff(n + 1, 0);
    HIPCK(hipMemcpyAsync(orig.data(), ctx->orig.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->st));
    if (K > 0) {
        HIPCK(hipMemcpyAsync(scnt.data(), ctx->cnt.p, sizeof(int32_t) * ne, hipMemcpyDeviceToHost, ctx->st));
        HIPCK(hipStreamSynchronize(ctx->st));
        for (int64_t s = 0; s < ne; s++) soff[s] = s * K;
    } else {
        HIPCK(hipMemcpyAsync(soff.data(), ctx->off.p, sizeof(int64_t) * (ne + 1), hipMemcpyDeviceToHost, ctx->st));
        HIPCK(hipStreamSynchronize(ctx->st));
        for (int64_t s = 0; s < ne; s++) scnt[s] = (int32_t)(soff[s + 1] - soff[s]);
    }
    std::vector<int64_t> cnt(n);
    for (int64_t s = 0; s < n; s++) cnt[orig[s]] = scnt[s];
    offsets[0] = 0;
    for (int64_t i = 0; i < n; i++) offsets[i + 1] = offsets[i] + cnt[i];
    const int64_t total = offsets[n];
    if (!idx || idx_cap < total)
        return idx ? fail(ctx, SPH_ERR_CAPACITY, "idx capacity " + std::to_string(idx_cap) + " < " +
                                                     std::to_string(total))
                   : SPH_OK;
    const int64_t span = K > 0 ? ne * K : total;
    std::vector<int32_t> sidx(span);
    HIPCK(hipMemcpyAsync(sidx.data(), ctx->idx.p, sizeof(int32_t) * span, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    for (int64_t s = 0; s < ne; s++) {
        int64_t r = orig[s], o = offsets[r];
        for (int64_t k = soff[s]; k < soff[s] + scnt[s]; k++) idx[o++] = orig[sidx[k]];
        std::sort(idx + offsets[r], idx + offsets[r + 1]);
    }
    return SPH_OK;
}

// =========================================================== hydro
extern "C" int sph_density(sph_ctx *ctx, int64_t n, const double *pos, const double *m, const double *h,
                           int precision, double *rho_out)
{
    CK(check_common(ctx, n, precision));
    if (!pos || !m || !h || !rho_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    CK(geom(ctx, n, pos, h, m));
    NgbRows rows;
    CK(ngb_rows(ctx, rows));
    if (precision == SPH_PREC_MIXED) {
        HIPCK(sph::density<float>(ctx->st, n, nullptr, 0.0, rows, P_<float>(ctx->rho), nullptr, nullptr, nullptr));
        return get_scalar<float>(ctx, ctx->rho, rho_out);
    }
    HIPCK(sph::density<double>(ctx->st, n, nullptr, 0.0, rows, D(rho), nullptr, nullptr, nullptr));
    return get_scalar<double>(ctx, ctx->rho, rho_out);
}

namespace sphc {
template <typename R>
int divcurl_impl(sph_ctx *ctx, const double *vel, const double *rho, double *dv, double *cv)
{
    int64_t n = ctx->n;
    CK(put_vec3(ctx, vel, ctx->vx, ctx->vy, ctx->vz));
    CK(put_real<R>(ctx, rho, ctx->rho));
    NgbRows rows;
    CK(ngb_rows(ctx, rows));
    HIPCK(sph::divcurl<R>(ctx->st, n, P_<R>(ctx->rho), nullptr, rows, P_<R>(ctx->divV), P_<R>(ctx->curlV), nullptr));
    CK(get_scalar<R>(ctx, ctx->divV, dv));
    return get_scalar<R>(ctx, ctx->curlV, cv);
}

template <typename R>
int hydro_inputs(sph_ctx *ctx, const double *vel, const double *rho, const double *P, const double *c,
                 const double *divV, const double *curlV)
{
    int64_t n = ctx->n;
    CK(put_vec3(ctx, vel, ctx->vx, ctx->vy, ctx->vz));
    CK(put_real<R>(ctx, rho, ctx->rho));
    CK(put_real<R>(ctx, P, ctx->P));
    CK(put_real<R>(ctx, c, ctx->c));
    bool bal = divV && curlV;
    if (bal) {
        CK(put_real<R>(ctx, divV, ctx->divV));
        CK(put_real<R>(ctx, curlV, ctx->curlV));
    }
    HIPCK(sph::prep<R>(ctx->st, n, P_<R>(ctx->rho), P_<R>(ctx->P), P_<R>(ctx->c), P_<R>(ctx->divV),
                       P_<R>(ctx->curlV), D(h), P_<R>(ctx->prho2), bal ? P_<R>(ctx->bal) : nullptr));
    return SPH_OK;
}

template <typename R>
int acc_impl(sph_ctx *ctx, const double *vel, const double *rho, const double *P, const double *c,
             const double *divV, const double *curlV, double alpha, double *acc)
{
    CK(hydro_inputs<R>(ctx, vel, rho, P, c, divV, curlV));
    bool bal = divV && curlV;
    NgbRows rows;
    CK(ngb_rows(ctx, rows));
    CK(ensure(ctx, ctx->s4, (size_t)ctx->n * 4 * sizeof(R)));
    HIPCK(sph::pack_s4<R>(ctx->st, ctx->n, P_<R>(ctx->rho), P_<R>(ctx->prho2), P_<R>(ctx->c),
                          bal ? P_<R>(ctx->bal) : nullptr, P_<R4<R>>(ctx->s4)));
    HIPCK(sph::accsph<R>(ctx->st, ctx->n, P_<R4<R>>(ctx->s4), bal, alpha, rows, nullptr, nullptr, nullptr, D(ax),
                         D(ay), D(az)));
    return get_vec3(ctx, ctx->ax, ctx->ay, ctx->az, acc);
}

template <typename R>
int dudt_impl(sph_ctx *ctx, const double *vel, const double *rho, const double *P, const double *c,
              const double *divV, const double *curlV, double alpha, double *out)
{
    CK(hydro_inputs<R>(ctx, vel, rho, P, c, divV, curlV));
    bool bal = divV && curlV;
    NgbRows rows;
    CK(ngb_rows(ctx, rows));
    CK(ensure(ctx, ctx->s4, (size_t)ctx->n * 4 * sizeof(R)));
    HIPCK(sph::pack_s4<R>(ctx->st, ctx->n, P_<R>(ctx->rho), P_<R>(ctx->prho2), P_<R>(ctx->c),
                          bal ? P_<R>(ctx->bal) : nullptr, P_<R4<R>>(ctx->s4)));
    HIPCK(sph::dudt<R>(ctx->st, ctx->n, P_<R4<R>>(ctx->s4), bal, alpha, rows, P_<R>(ctx->du)));
    return get_scalar<R>(ctx, ctx->du, out);
}
}  // namespace sphc

extern "C" int sph_div_curl(sph_ctx *ctx, int64_t n, const double *pos, const double *vel, const double *rho,
                            const double *m, const double *h, int precision, double *divV_out, double *curlV_out)
{
    CK(check_common(ctx, n, precision));
    if (!pos || !vel || !rho || !m || !h || !divV_out || !curlV_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    CK(check_particles(ctx, n, nullptr, vel, nullptr, nullptr, nullptr));
    CK(geom(ctx, n, pos, h, m));
    return precision == SPH_PREC_MIXED ? divcurl_impl<float>(ctx, vel, rho, divV_out, curlV_out)
                                       : divcurl_impl<double>(ctx, vel, rho, divV_out, curlV_out);
}

extern "C" int sph_acc_hydro(sph_ctx *ctx, int64_t n, const double *pos, const double *vel, const double *rho,
                             const double *P, const double *c, const double *h, const double *m, const double *divV,
                             const double *curlV, double alpha, int precision, double *acc_out)
{
    CK(check_common(ctx, n, precision));
    if (!pos || !vel || !rho || !P || !c || !h || !m || !acc_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    if ((divV == nullptr) != (curlV == nullptr)) return fail(ctx, SPH_ERR_ARG, "divV and curlV go together");
    CK(check_particles(ctx, n, nullptr, vel, nullptr, nullptr, nullptr));
    CK(geom(ctx, n, pos, h, m));
    return precision == SPH_PREC_MIXED ? acc_impl<float>(ctx, vel, rho, P, c, divV, curlV, alpha, acc_out)
                                       : acc_impl<double>(ctx, vel, rho, P, c, divV, curlV, alpha, acc_out);
}

extern "C" int sph_dudt(sph_ctx *ctx, int64_t n, const double *pos, const double *vel, const double *rho,
                        const double *P, const double *c, const double *h, const double *m, const double *divV,
                        const double *curlV, double alpha, int precision, double *dudt_out)
{
    CK(check_common(ctx, n, precision));
    if (!pos || !vel || !rho || !P || !c || !h || !m || !dudt_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    if ((divV == nullptr) != (curlV == nullptr)) return fail(ctx, SPH_ERR_ARG, "divV and curlV go together");
    CK(check_particles(ctx, n, nullptr, vel, nullptr, nullptr, nullptr));
    CK(geom(ctx, n, pos, h, m));
    return precision == SPH_PREC_MIXED ? dudt_impl<float>(ctx, vel, rho, P, c, divV, curlV, alpha, dudt_out)
                                       : dudt_impl<double>(ctx, vel, rho, P, c, divV, curlV, alpha, dudt_out);
}

// =========================================================== gravity
namespace sphc {
// The tree walk of sph_gravity / sph_potential_energy up to its results in the sorted
// arrays (gx, gy, gz; phi when pot).
int gravity_pass(sph_ctx *ctx, int64_t n, const double *pos, const double *m, const double *eps, double G,
                 double theta, int soft_kind, double eps_plummer, int precision, bool pot, int64_t *n_interactions)
{
    CK(check_common(ctx, n, precision));
    if (!pos || !m) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (soft_kind != SPH_SOFT_SPLINE && soft_kind != SPH_SOFT_PLUMMER) return fail(ctx, SPH_ERR_ARG, "bad soft_kind");
    if (!(theta >= 0.0)) return fail(ctx, SPH_ERR_ARG, "theta must be >= 0");
    if (!std::isfinite(G) || !(eps_plummer >= 0.0) || !std::isfinite(eps_plummer))
        return fail(ctx, SPH_ERR_ARG, "G and eps_plummer must be finite, eps_plummer >= 0");
    CK(check_particles(ctx, n, nullptr, nullptr, m, nullptr, nullptr));
    if (eps)
        for (int64_t i = 0; i < n; i++)
            if (!(eps[i] >= 0.0) || !std::isfinite(eps[i]))
                return fail(ctx, SPH_ERR_ARG, "softening lengths must be finite and >= 0");
    CK(set_pos(ctx, n, pos));
    bool cm = false, ce = false;
    CK(put_scalar(ctx, m, ctx->m, &ctx->cm, &ctx->m_ok, &cm));
    if (eps) CK(put_scalar(ctx, eps, ctx->eps, &ctx->ceps, &ctx->eps_ok, &ce));
    else if (ctx->eps_ok) {
        ctx->eps_ok = false;
        ctx->ceps.clear();
        ce = true;
    }
    if (cm || ce) ctx->moments_ok = false;
    CK(ensure_moments(ctx));
    if (pot) CK(ensure(ctx, ctx->phi, (size_t)n * 8));
    unsigned long long *cnt = P_<unsigned long long>(ctx->spread) + SPR_G * SPH_SPREAD_WORDS;
    HIPCK(sph::fill<unsigned long long>(ctx->st, 3 * SPH_SPREAD_WORDS, cnt, 0ull));
    const double *e = ctx->eps_ok ? D(eps) : nullptr;
    const int64_t ng = ctx->ngroups;
    double *phi = pot ? D(phi) : nullptr;
    if (precision == SPH_PREC_MIXED)
        HIPCK(sph::gravity<float>(ctx->st, ng, I32(groups), D(x), D(y), D(z), D(m), e, P_<WNode>(ctx->nodes),
                                  SPH_BUCKET, theta, G, soft_kind, eps_plummer, ctx->m_max, D(gx), D(gy), D(gz),
                                  cnt, phi));
    else
        HIPCK(sph::gravity<double>(ctx->st, ng, I32(groups), D(x), D(y), D(z), D(m), e, P_<WNode>(ctx->nodes),
                                   SPH_BUCKET, theta, G, soft_kind, eps_plummer, ctx->m_max, D(gx), D(gy), D(gz),
                                   cnt, phi));
    unsigned long long hc[3];
    HIPCK(sph::spread_total(ctx->st, 3, cnt, P_<unsigned long long>(ctx->hcount) + 5));
    HIPCK(hipMemcpyAsync(hc, P_<unsigned long long>(ctx->hcount) + 5, sizeof(hc), hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    if (hc[2]) return fail(ctx, SPH_ERR_CAPACITY, "gravity walk stack overflow");
    if (n_interactions) {
        n_interactions[0] = (int64_t)hc[0];
        n_interactions[1] = (int64_t)hc[1];
    }
    return SPH_OK;
}
}  // namespace sphc

extern "C" int sph_gravity(sph_ctx *ctx, int64_t n, const double *pos, const double *m, const double *eps, double G,
                           double theta, int soft_kind, double eps_plummer, int precision, double *acc_out,
                           int64_t *n_interactions)
{
    if (ctx && !acc_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    CK(gravity_pass(ctx, n, pos, m, eps, G, theta, soft_kind, eps_plummer, precision, false, n_interactions));
    return get_vec3(ctx, ctx->gx, ctx->gy, ctx->gz, acc_out);
}

extern "C" int sph_potential_energy(sph_ctx *ctx, int64_t n, const double *pos, const double *m, const double *eps,
                                    double G, double theta, int soft_kind, double eps_plummer, int precision,
                                    double *pe_out, double *phi_out)
{
    if (ctx && !pe_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    CK(gravity_pass(ctx, n, pos, m, eps, G, theta, soft_kind, eps_plummer, precision, true, nullptr));
    CK(ensure(ctx, ctx->tpart, 10 * 1024 * 8));
    CK(ensure(ctx, ctx->tout, 16 * 8));
    HIPCK(sph::totals(ctx->st, n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, D(m), nullptr, D(phi), D(tpart),
                      D(tout)));
    double t[10];
    HIPCK(hipMemcpyAsync(t, ctx->tout.p, sizeof(t), hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    *pe_out = t[2];
    if (phi_out) CK(get_scalar<double>(ctx, ctx->phi, phi_out));
    return SPH_OK;
}

extern "C" int sph_validate_particles(int64_t n, const double *pos, const double *vel, const double *m,
                                      const double *u, const double *h, char *msg, int64_t msg_cap)
{
    std::string why;
    const bool ok = particles_ok(n, pos, vel, m, u, h, why);
    if (msg && msg_cap > 0) {
        const size_t k = std::min((size_t)(msg_cap - 1), why.size());
        std::memcpy(msg, why.data(), k);
        msg[k] = 0;
    }
    return ok ? SPH_OK : SPH_ERR_ARG;
}

extern "C" int sph_column_density(sph_ctx *ctx, int64_t n, const double *pos, const double *m, const double *h,
                                  const double *rho, const sph_column_params *cp, int precision, double *grad_out,
                                  double *log_nh_out)
{
    CK(check_common(ctx, n, precision));
    if (!pos || !m || !h || !rho || !cp || !log_nh_out) return fail(ctx, SPH_ERR_ARG, "null argument");
    CK(column_check(ctx, cp));
    CK(geom(ctx, n, pos, h, m));
    CK(ensure(ctx, ctx->colg, (size_t)n * 8));
    CK(ensure(ctx, ctx->coln, (size_t)n * 8));
    NgbRows rows;
    CK(ngb_rows(ctx, rows));
    const sph::ColumnParams p = column_params(cp);
    if (precision == SPH_PREC_MIXED) {
        CK(put_real<float>(ctx, rho, ctx->rho));
        HIPCK(sph::column_density<float>(ctx->st, n, P_<float>(ctx->rho), rows, p, D(colg), D(coln)));
    } else {
        CK(put_real<double>(ctx, rho, ctx->rho));
        HIPCK(sph::column_density<double>(ctx->st, n, D(rho), rows, p, D(colg), D(coln)));
    }
    if (grad_out) CK(get_scalar<double>(ctx, ctx->colg, grad_out));
    return get_scalar<double>(ctx, ctx->coln, log_nh_out);
}

