// sph_capi_step.hip -- the device-resident state and the KDK step: set/get, the force
// evaluation, the step, statistics, restart, outflow injection, the column-density and
// time-step read-backs.
#include "sph_ctx.h"
#include <type_traits>

using namespace sphc;

// =========================================================== resident step
extern "C" int sph_state_set(sph_ctx *ctx, int64_t n, const double *pos, const double *vel, const double *m,
                             const double *u, const double *h)
{
    CK(check_common(ctx, n, SPH_PREC_FP64));
    if (!pos || !vel || !m || !u || !h) return fail(ctx, SPH_ERR_ARG, "null argument");
    CK(check_particles(ctx, n, pos, vel, m, u, h));
    ctx->resident = false;
    ctx->cool_on = false;  // Z and J21 belong to the previous state
    ctx->col_on = ctx->col_ok = false;
    ctx->ts_ok = false;
    ctx->tot_ran = false;
    ctx->n_eval = 0;
    ctx->agn_on = false;
    ctx->agn_leftover = 0.0;
    ctx->agn_injected = 0;
    invalidate_all(ctx);
    CK(set_pos(ctx, n, pos));
    ctx->n_placed = std::max(n, ctx->n_reserve);  // set_pos placed every per-particle buffer for this many
    CK(put_vec3(ctx, vel, ctx->vx, ctx->vy, ctx->vz));
    CK(put_scalar(ctx, m, ctx->m, nullptr, nullptr, nullptr));
    CK(put_scalar(ctx, u, ctx->u, nullptr, nullptr, nullptr));
    CK(put_scalar(ctx, h, ctx->h, nullptr, nullptr, nullptr));
    HIPCK(sph::fill<double>(ctx->st, n, D(duprev), 0.0));
    HIPCK(sph::fill<double>(ctx->st, n, D(ax), 0.0));
    HIPCK(sph::fill<double>(ctx->st, n, D(ay), 0.0));
    HIPCK(sph::fill<double>(ctx->st, n, D(az), 0.0));
    HIPCK(hipStreamSynchronize(ctx->st));
    invalidate_all(ctx);  // stateless caches no longer describe the device arrays
    ctx->resident = true;
    ctx->primed = false;
    std::memset(&ctx->stats, 0, sizeof(ctx->stats));  // counters describe this state only
    ctx->stats.n = n;
    return SPH_OK;
}

extern "C" int sph_state_get(sph_ctx *ctx, double *pos, double *vel, double *u, double *h, double *rho, double *acc)
{
    if (!ctx) return SPH_ERR_ARG;
    if (!ctx->resident) return fail(ctx, SPH_ERR_STATE, "no resident state (call sph_state_set first)");
    HIPCK(hipSetDevice(ctx->device));
    if (pos) CK(get_vec3(ctx, ctx->x, ctx->y, ctx->z, pos));
    if (vel) CK(get_vec3(ctx, ctx->vx, ctx->vy, ctx->vz, vel));
    if (u) CK(get_scalar<double>(ctx, ctx->u, u));
    if (h) CK(get_scalar<double>(ctx, ctx->h, h));
    if (acc) CK(get_vec3(ctx, ctx->ax, ctx->ay, ctx->az, acc));
    if (rho) {
        if (!ctx->primed || ctx->n_eval < 1)
            return fail(ctx, SPH_ERR_STATE, "rho is available after sph_prime/sph_step");
        const int64_t ne = ctx->n_eval, extra = ctx->n - ne;  // injected after the evaluation: NaN
        if (ctx->rho_prec == SPH_PREC_FP64) {
            if (extra > 0) HIPCK(sph::fill<double>(ctx->st, extra, D(rho) + ne, NAN));
            CK(get_scalar<double>(ctx, ctx->rho, rho));
        } else {
            if (extra > 0) HIPCK(sph::fill<float>(ctx->st, extra, P_<float>(ctx->rho) + ne, NAN));
            CK(get_scalar<float>(ctx, ctx->rho, rho));
        }
    }
    return SPH_OK;
}

namespace sphc {

// Re-sort the resident state by the current positions.
int reorder_state(sph_ctx *ctx)
{
    int64_t n = ctx->n;
    hipStream_t st = ctx->st;
    CK(sort_current(ctx, n));
    const int32_t *perm = I32(vals0);
    // every per-particle state array and the caller-index map in one gather pass into the
    // shadow buffers, which then take the arrays' places
    Buf *arr[SPH_GATHER_MAX] = {&ctx->x, &ctx->y, &ctx->z, &ctx->vx, &ctx->vy, &ctx->vz, &ctx->m, &ctx->h,
                                &ctx->u, &ctx->duprev, &ctx->ax, &ctx->ay, &ctx->az};
    int k = 13;
    if (cool_columns(ctx)) {
        arr[k++] = &ctx->zmet;
        arr[k++] = &ctx->j21;
    }
    const double *src[SPH_GATHER_MAX];
    double *dst[SPH_GATHER_MAX];
    for (int j = 0; j < k; j++) {
        CK(ensure(ctx, ctx->shadow[j], sizeof(double) * std::max(n, ctx->n_reserve)));
        src[j] = (const double *)arr[j]->p;
        dst[j] = (double *)ctx->shadow[j].p;
    }
    HIPCK(sph::gather_set(st, n, perm, k, src, dst, I32(orig), I32(tmpi)));
    for (int j = 0; j < k; j++) std::swap(*arr[j], ctx->shadow[j]);
    std::swap(ctx->orig, ctx->tmpi);
    return SPH_OK;
}

// Cooling update of n particles (sph_cooling_update's semantics); counts[0] = floored
// particles, counts[1] = surrogate fallbacks (device words).
template <typename RHO>
int cool_run(sph_ctx *ctx, int64_t n, const RHO *rho, const double *u, const double *Z, const double *J21,
             double dt, const sph_cool_params &cp, double *u_out, unsigned long long *counts);

template <typename R>
int forces(sph_ctx *ctx, const sph_step_params *p, int first)
{
    int64_t n = ctx->n;
    hipStream_t st = ctx->st;
    // --- order + tree topology
    ph_begin(ctx, SPH_PH_ORDER);
    CK(reorder_state(ctx));
    ph_end(ctx, SPH_PH_ORDER);
    ph_begin(ctx, SPH_PH_TOPOLOGY);
    CK(topology(ctx));
    ph_end(ctx, SPH_PH_TOPOLOGY);
    // boxes are needed by the h walk; hmax is refreshed after the solve
    ph_begin(ctx, SPH_PH_HMOMENTS);
    HIPCK(sph::tree_moments(st, n, SPH_BUCKET, D(x), D(y), D(z), D(m), D(h), nullptr, P_<WNode>(ctx->nodes), I32(right),
                            I32(parent), I32(ncount), P_<uint32_t>(ctx->flags), I32(front)));
    ph_end(ctx, SPH_PH_HMOMENTS);
    ph_begin(ctx, SPH_PH_HSOLVE);
    HIPCK(hipMemcpyAsync(ctx->hin.p, ctx->h.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    ctx->hs_unc = 0;
    ctx->hs_work[0] = ctx->hs_work[1] = 0;
    sph::HSolveParams hp;
    CK(hsolve_params(ctx, n, D(x), D(y), D(z), SPH_H_NEWTON, p->n_ngb, 0, 0, 0, p->h_tol, &ctx->hs_unc,
                     &ctx->stats.h_rounds, ctx->hs_work, hp));
    hp.per = p->period;
    HIPCK(sph::hsolve_begin(st, hp));
    ph_end(ctx, SPH_PH_HSOLVE);
    std::swap(ctx->h, ctx->hout);  // the solve writes hp.h_out: the new h
    // The solve's first round is enqueued; its read-back (pending targets) is waited for
    // only after the gravity walk is enqueued behind it when gravity does not depend on h
    // (Plummer softening or no gravity), so the host wait does not leave the GPU idle.
    auto finish_h = [&]() -> int {
        int64_t pending = 0;
        bool more = false;
        HIPCK(sph::hsolve_end(st, hp, SPH_HROUNDS, &pending, &more));
        if (pending)
            return fail(ctx, SPH_ERR_NOCONV, "smoothing-length solve: " + std::to_string(pending) +
                                                 " targets undecided after the maximum number of radius rounds");
        ctx->stats.h_tests = ctx->hs_work[0];
        ctx->stats.h_terms = ctx->hs_work[1];
        ctx->stats.h_unconverged += ctx->hs_unc;
        return SPH_OK;
    };
    // --- node h_max (and eps_max = h_max for the spline softening) of the new h.  Boxes, masses
    // and centres of mass do not depend on h: with Plummer softening (or no gravity) the
    // neighbour walk prunes with the h_max of the boxes' pass scaled by max_i h_i/h_old_i
    // instead of a second bottom-up pass; the spline softening needs the exact eps_max.
    const bool spline = p->soft_kind == SPH_SOFT_SPLINE && p->G != 0.0;
    const double *eps = spline ? D(h) : nullptr;
    if (spline) CK(finish_h());
    ph_begin(ctx, SPH_PH_MOMENTS);
    if (spline)
        HIPCK(sph::tree_moments(st, n, SPH_BUCKET, D(x), D(y), D(z), D(m), D(h), eps, P_<WNode>(ctx->nodes), I32(right),
                                I32(parent), I32(ncount), P_<uint32_t>(ctx->flags), I32(front)));
    ph_end(ctx, SPH_PH_MOMENTS);
    ph_begin(ctx, SPH_PH_GRAVITY);
    unsigned long long *cnt = P_<unsigned long long>(ctx->spread) + SPR_G * SPH_SPREAD_WORDS;
    HIPCK(sph::fill<unsigned long long>(st, 3 * SPH_SPREAD_WORDS, cnt, 0ull));
    const bool pot = p->energy && p->G != 0.0;  // the energy option: potential from the same walk
    if (pot) CK(ensure(ctx, ctx->phi, (size_t)std::max(n, ctx->n_reserve) * 8));
    if (p->G != 0.0) {
        HIPCK(sph::gravity<R>(st, ctx->ngroups, I32(groups), D(x), D(y), D(z), D(m), eps, P_<WNode>(ctx->nodes), SPH_BUCKET,
                              p->theta, p->G, p->soft_kind, p->eps_plummer, ctx->m_max, D(gx), D(gy), D(gz), cnt,
                              pot ? D(phi) : nullptr));
    }
    HIPCK(sph::spread_total(st, 3, cnt, P_<unsigned long long>(ctx->hcount) + 5));
    ph_end(ctx, SPH_PH_GRAVITY);
    if (!spline) CK(finish_h());
    // --- neighbour rows from the h solve's candidates.  Their overflow check (a row longer
    // than the stride or a full extras region, then a tree walk) is read back after the
    // density and acceleration passes are enqueued behind them; on overflow those two passes
    // are run again on the rebuilt sets.
    CK(ensure(ctx, ctx->idx, (size_t)n * SPH_NGB_STRIDE * 4));
    const int64_t ext_cap = std::max<int64_t>(4 * n, SPH_SPREAD * 64);
    CK(ensure(ctx, ctx->cext, (size_t)ext_cap * 8));
    CK(ensure(ctx, ctx->cnext, (size_t)SPH_SPREAD_WORDS * 8));
    CK(ensure(ctx, ctx->chr, (size_t)n * (hp.cand ? sizeof(double2) : sizeof(sph::CandRec))));
    ph_begin(ctx, SPH_PH_NGB);
    sph::BigRows big;  // targets of the candidate-free h solver: rows by a walk each
    big.list = I32(hbig);
    big.count = ctx->stats.h_big;
    big.x = D(x); big.y = D(y); big.z = D(z);
    big.nodes = P_<WNode>(ctx->nodes);
    big.bucket = SPH_BUCKET;
    big.per = p->period;
    if (big.count > 0 && !spline) {  // node h_max are those of h_old: scale them
        HIPCK(sph::h_ratio_max(st, n, D(h), D(hin), D(kout) + 7));  // slot 7 (kinetics uses 0..4)
        big.hscale = D(kout) + 7;
    }
    HIPCK(sph::neighbour_rows_from_candidates_begin(
        st, n, SPH_HCAP, D(x), D(y), D(z), p->period, hp.cand, I32(ccidx), I32(cccnt), D(h), D(hR), SPH_NGB_STRIDE,
        I32(cnt), I32(rcnt0), I32(idx), P_<int2>(ctx->cext), ext_cap, P_<unsigned long long>(ctx->cnext),
        P_<unsigned long long>(ctx->spread) + SPR_N * SPH_SPREAD_WORDS, P_<unsigned long long>(ctx->hcount) + 10,
        ctx->hpin + 8, ctx->ev_n, big, ctx->chr.p));
    ph_end(ctx, SPH_PH_NGB);
    ctx->ngb_stride = SPH_NGB_STRIDE;
    ctx->ngb_per = p->period;
    R *bal = p->balsara ? P_<R>(ctx->bal) : nullptr;
    const bool g = p->G != 0.0;
    CK(ensure(ctx, ctx->s4, (size_t)n * 4 * sizeof(R)));
    // The mixed form reads the compact FP32 records in the density, acceleration and du/dt
    // passes; the FP64 records are packed only for the passes that read them (div v without
    // the Balsara switch, column density, time-step criteria).
    const bool cmp = std::is_same<R, float>::value;
    const bool full = !cmp || ctx->col_on || p->C_CFL > 0.0;
    // density + EOS, fused with div/curl v and the Balsara factor when the switch is on; then
    // the acceleration (a = a_grav + a_sph)
    auto hydro_forces = [&]() -> int {
        NgbRows rows;
        CK(ngb_rows(ctx, rows, false, cmp, full));
        ph_begin(ctx, SPH_PH_DENSITY);
        if (p->balsara)
            HIPCK(sph::density_divcurl<R>(st, n, D(u), p->gamma, rows, P_<R>(ctx->rho), P_<R>(ctx->P), P_<R>(ctx->c),
                                          P_<R>(ctx->prho2), P_<R>(ctx->divV), P_<R>(ctx->curlV), bal,
                                          P_<R4<R>>(ctx->s4)));
        else
            HIPCK(sph::density<R>(st, n, D(u), p->gamma, rows, P_<R>(ctx->rho), P_<R>(ctx->P), P_<R>(ctx->c),
                                  P_<R>(ctx->prho2), P_<R4<R>>(ctx->s4)));
        ph_end(ctx, SPH_PH_DENSITY);
        if (!p->balsara && p->C_CFL > 0.0) {
            // div v of the time-step criteria from the velocities of this force evaluation (before the
            // second kick), as the Balsara form's density pass computes it and as the reference's
            // div_curlVel_ngb runs in the force pass (MC/drv.cu:1271) before dt_array_indiv_dt (:1610)
            ph_begin(ctx, SPH_PH_DIVCURL);
            HIPCK(sph::divcurl<R>(st, n, P_<R>(ctx->rho), P_<R>(ctx->c), rows, P_<R>(ctx->divV), P_<R>(ctx->curlV),
                                  nullptr));
            ph_end(ctx, SPH_PH_DIVCURL);
        }
        if (ctx->col_on) {  // column density from this density pass (gradient over the same rows)
            CK(ensure(ctx, ctx->colg, (size_t)std::max(n, ctx->n_reserve) * 8));
            CK(ensure(ctx, ctx->coln, (size_t)std::max(n, ctx->n_reserve) * 8));
            ph_begin(ctx, SPH_PH_COLUMN);
            HIPCK(sph::column_density<R>(st, n, P_<R>(ctx->rho), rows, ctx->col_p, D(colg), D(coln)));
            ph_end(ctx, SPH_PH_COLUMN);
            ctx->col_ok = true;
        }
        ph_begin(ctx, SPH_PH_ACCSPH);
        HIPCK(sph::accsph<R>(st, n, P_<R4<R>>(ctx->s4), bal != nullptr, p->alpha, rows, g ? D(gx) : nullptr,
                             g ? D(gy) : nullptr, g ? D(gz) : nullptr, D(ax), D(ay), D(az)));
        ph_end(ctx, SPH_PH_ACCSPH);
        return SPH_OK;
    };
    CK(hydro_forces());
    {
        int64_t nlong = 0, total = 0, nextra = 0;
        HIPCK(sph::neighbour_rows_from_candidates_end(ctx->ev_n, ctx->hpin + 8, ext_cap, &nlong, &total, &nextra));
        ctx->stats.ngb_extra = nextra;
        if (nlong == 0) {
            ctx->csr_total = total;
            ctx->stats.ngb_total = total;
        } else {
            const double *hscale = nullptr;
            if (!spline) {  // node h_max are those of h_old: scale them
                HIPCK(sph::h_ratio_max(st, n, D(h), D(hin), D(kout) + 7));  // slot 7 (kinetics uses 0..4)
                hscale = D(kout) + 7;
            }
            CK(build_ngb(ctx, false, hscale, p->period));
            CK(hydro_forces());
        }
    }
    if (ctx->agn_on && ctx->agn.sigma != 0.0)  // isothermal halo, added to the total acceleration
        HIPCK(sph::halo_accel(st, n, D(x), D(y), D(z), ctx->agn.sigma, D(ax), D(ay), D(az)));
    if (!first) {
        ph_begin(ctx, SPH_PH_KICK);
        HIPCK(sph::kick_drift(st, n, D(vx), D(vy), D(vz), D(ax), D(ay), D(az), 0.5 * p->dt, 0.0, nullptr, nullptr,
                              nullptr));
        ph_end(ctx, SPH_PH_KICK);
    }
    ph_begin(ctx, SPH_PH_DUDT);
    NgbRows rows;
    CK(ngb_rows(ctx, rows, true, cmp, full));  // packs the kicked velocities (positions and h as in the force passes)
    HIPCK(sph::dudt<R>(st, n, P_<R4<R>>(ctx->s4), bal != nullptr, p->alpha, rows, P_<R>(ctx->du)));
    ph_end(ctx, SPH_PH_DUDT);
    ph_begin(ctx, SPH_PH_UUPDATE);
    HIPCK(sph::u_update<R>(st, n, D(u), P_<R>(ctx->du), D(duprev), p->dt, first));
    ph_end(ctx, SPH_PH_UUPDATE);
    // cooling after the u update of each step (CR/v3t.cu:1527-1536; none after the first update)
    ctx->cool_ran = ctx->cool_on && !first;
    if (ctx->cool_ran) {
        ph_begin(ctx, SPH_PH_COOLING);
        if (ctx->cool_p.mode == SPH_COOL_NH) {  // log N_H of this evaluation's column pass, r = |x|
            if (!ctx->col_on)
                return fail(ctx, SPH_ERR_STATE, "SPH_COOL_NH needs the column density on (sph_state_set_column)");
            CK(nh_run<R>(ctx, n, D(x), D(y), D(z), P_<R>(ctx->rho), D(u), D(coln), nullptr, p->gamma, p->dt,
                         ctx->cool_p, D(u), P_<unsigned long long>(ctx->hcount) + 12));
        } else {
            CK(cool_run<R>(ctx, n, P_<R>(ctx->rho), D(u), D(zmet), D(j21), p->dt, ctx->cool_p, D(u),
                           P_<unsigned long long>(ctx->hcount) + 12));
        }
        ph_end(ctx, SPH_PH_COOLING);
    }
    // per-particle time-step criteria of this evaluation (velocities after the kick)
    ctx->ts_ok = false;
    if (p->C_CFL > 0.0) {
        const size_t cap = (size_t)std::max(n, ctx->n_reserve) * 8;
        CK(ensure(ctx, ctx->tsdt, cap));
        CK(ensure(ctx, ctx->tsdh, cap));
        unsigned long long *dmin = P_<unsigned long long>(ctx->hcount) + 15;
        HIPCK(sph::fill<unsigned long long>(st, 1, dmin, ~0ull));
        // softening of the gravity criteria: the Plummer length, else each particle's h (the
        // reference's eps = h of its initial conditions)
        const bool eps_h = p->soft_kind == SPH_SOFT_SPLINE || !(p->eps_plummer > 0.0);
        const sph::StepCriteria sc{p->C_CFL, p->alpha, p->eps_plummer};
        const bool g = p->G != 0.0;
        ph_begin(ctx, SPH_PH_TIMESTEP);
        HIPCK(sph::timestep<R>(st, n, P_<R4<R>>(ctx->s4), P_<R>(ctx->divV), rows, D(ax), D(ay), D(az),
                               g ? D(gx) : nullptr, g ? D(gy) : nullptr, g ? D(gz) : nullptr, eps_h ? D(h) : nullptr,
                               sc, D(tsdt), D(tsdh), dmin));
        ph_end(ctx, SPH_PH_TIMESTEP);
        ctx->ts_ok = true;
        ctx->ts_n = n;
    }
    // FP64 totals of the state at the end of the evaluation (energy option)
    ctx->tot_ran = p->energy != 0;
    if (ctx->tot_ran) {
        CK(ensure(ctx, ctx->tpart, 10 * 1024 * 8));
        CK(ensure(ctx, ctx->tout, 16 * 8));
        HIPCK(sph::totals(st, n, D(x), D(y), D(z), D(vx), D(vy), D(vz), D(m), D(u), pot ? D(phi) : nullptr, D(tpart),
                          D(tout)));
    }
    ctx->n_eval = n;
    return SPH_OK;
}

int check_params(sph_ctx *ctx, const sph_step_params *p)
{
    if (!ctx || !p) return SPH_ERR_ARG;
    if (!ctx->resident) return fail(ctx, SPH_ERR_STATE, "no resident state (call sph_state_set first)");
    if (p->precision != SPH_PREC_MIXED && p->precision != SPH_PREC_FP64) return fail(ctx, SPH_ERR_ARG, "bad precision");
    if (p->soft_kind != SPH_SOFT_SPLINE && p->soft_kind != SPH_SOFT_PLUMMER) return fail(ctx, SPH_ERR_ARG, "bad soft_kind");
    if (!(p->dt > 0.0) || !(p->n_ngb > 0.0) || !(p->h_tol > 0.0) || !(p->theta >= 0.0))
        return fail(ctx, SPH_ERR_ARG, "dt, n_ngb, h_tol must be > 0 and theta >= 0");
    if (!(p->period >= 0.0)) return fail(ctx, SPH_ERR_ARG, "period must be >= 0");
    if (!(p->C_CFL >= 0.0)) return fail(ctx, SPH_ERR_ARG, "C_CFL must be >= 0");
    if (p->period > 0.0 && p->G != 0.0)
        return fail(ctx, SPH_ERR_ARG, "a periodic box needs G = 0 (periodic gravity is not provided)");
    // (4 pi/3)(2h)^3 sum_j W(r_ij, h) tends to (32/3) n from below as h grows: no h reaches more
    if (!(p->n_ngb < (32.0 / 3.0) * (double)ctx->n))
        return fail(ctx, SPH_ERR_ARG, "n_ngb = " + std::to_string(p->n_ngb) + " cannot be reached with " +
                                          std::to_string(ctx->n) + " particles (needs n_ngb < 32 n / 3)");
    if (p->energy != 0 && p->energy != 1) return fail(ctx, SPH_ERR_ARG, "energy must be 0 or 1");
    HIPCK(hipSetDevice(ctx->device));
    return SPH_OK;
}

// Outflow pairs after a step (the driver's outflow_injector2 after t += dt): the host spends
// M_dot dt plus the leftover on whole pairs, the device appends them at the end of the
// sorted arrays (the next step's ordering sorts them in).
int inject_outflow(sph_ctx *ctx, double dt)
{
    const sph_agn_params &a = ctx->agn;
    ctx->agn_step += 1;
    if (!(a.M_dot > 0.0)) return SPH_OK;
    double mass = a.M_dot * dt + ctx->agn_leftover;
    const int64_t pairs = (int64_t)std::floor(mass / (2.0 * a.m_inject));
    ctx->agn_leftover = mass - (double)pairs * 2.0 * a.m_inject;
    if (pairs <= 0) return SPH_OK;
    const int64_t n = ctx->n;
    if (n + 2 * pairs > ctx->n_placed || n + 2 * pairs > (int64_t)INT32_MAX)
        return fail(ctx, SPH_ERR_CAPACITY, "outflow injection needs " + std::to_string(n + 2 * pairs) +
                                               " particles; reserve more with sph_state_reserve");
    hipStream_t st = ctx->st;
    // h of the last injected particle (caller index n - 1), before the first injection of h_ref_index
    const int64_t src = ctx->agn_injected > 0 ? n - 1 : a.h_ref_index;
    double *hsrc = D(kout) + 6;  // slot 6 (kinetics uses 0..4, the h ratio 7)
    HIPCK(sph::h_of_caller(st, n, I32(orig), D(h), src, hsrc));
    sph::InjectParams ip{a.r_inject, a.v_in, a.m_inject, a.u_inject, a.Z_inject, a.J21_inject, a.seed, ctx->agn_step};
    sph::InjectTarget t{D(x), D(y), D(z), D(vx), D(vy), D(vz), D(m), D(h), D(u), D(duprev), D(ax), D(ay), D(az),
                        I32(orig), cool_columns(ctx) ? D(zmet) : nullptr, cool_columns(ctx) ? D(j21) : nullptr};
    HIPCK(sph::inject_pairs(st, n, pairs, ip, hsrc, t));
    HIPCK(hipStreamSynchronize(st));
    ctx->n = n + 2 * pairs;
    ctx->stats.n = ctx->n;
    ctx->agn_injected += 2 * pairs;
    ctx->m_max = std::max(ctx->m_max, std::fabs(a.m_inject));
    return SPH_OK;
}

int finish_step(sph_ctx *ctx, const sph_step_params *p)
{
    HIPCK(hipStreamSynchronize(ctx->st));
    unsigned long long hc[3];
    HIPCK(hipMemcpy(hc, P_<unsigned long long>(ctx->hcount) + 5, sizeof(hc), hipMemcpyDeviceToHost));
    if (hc[2]) return fail(ctx, SPH_ERR_CAPACITY, "gravity walk stack overflow");
    ctx->stats.grav_pp = (int64_t)hc[0];
    ctx->stats.grav_pn = (int64_t)hc[1];
    if (ctx->cool_ran) {
        unsigned long long cc[2];
        HIPCK(hipMemcpy(cc, P_<unsigned long long>(ctx->hcount) + 12, sizeof(cc), hipMemcpyDeviceToHost));
        ctx->stats.cool_floor = (int64_t)cc[0];
        ctx->stats.cool_fallback = (int64_t)cc[1];
        ctx->cool_ran = false;
    }
    if (ctx->timing) {
        // only the phases this step recorded (an unrecorded event pair would fail and
        // leave a sticky error behind)
        for (int ph = 0; ph < SPH_PH_COUNT; ph++) {
            if (!(ctx->ph_rec & (1u << ph))) continue;
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ctx->ev[2 * ph], ctx->ev[2 * ph + 1]) == hipSuccess) {
                ctx->stats.phase_ms[ph] += ms;
                ctx->stats.phase_calls[ph] += 1;
            }
        }
    }
    ctx->ph_rec = 0;
    ctx->rho_prec = p->precision;
    if (ctx->tot_ran) {
        double t[10];
        HIPCK(hipMemcpy(t, ctx->tout.p, sizeof(t), hipMemcpyDeviceToHost));
        for (int k = 0; k < 3; k++) {
            ctx->stats.energy[k] = t[k];
            ctx->stats.momentum[k] = t[3 + k];
            ctx->stats.ang_momentum[k] = t[6 + k];
        }
    }
    ctx->stats.dt_min = 0.0;
    if (ctx->ts_ok) {
        unsigned long long bits = 0;
        HIPCK(hipMemcpy(&bits, P_<unsigned long long>(ctx->hcount) + 15, sizeof(bits), hipMemcpyDeviceToHost));
        double v;
        std::memcpy(&v, &bits, sizeof(v));
        ctx->stats.dt_min = bits == ~0ull ? INFINITY : v;
    }
    return SPH_OK;
}

}  // namespace sphc

extern "C" int sph_prime(sph_ctx *ctx, const sph_step_params *p)
{
    CK(check_params(ctx, p));
    ctx->timing = false;
    int r = p->precision == SPH_PREC_MIXED ? forces<float>(ctx, p, 1) : forces<double>(ctx, p, 1);
    if (r != SPH_OK) return r;
    CK(finish_step(ctx, p));
    ctx->primed = true;
    return SPH_OK;
}

extern "C" int sph_step(sph_ctx *ctx, const sph_step_params *p)
{
    CK(check_params(ctx, p));
    if (!ctx->primed) return fail(ctx, SPH_ERR_STATE, "call sph_prime before sph_step");
    ctx->timing = p->timing != 0;
    ctx->ph_rec = 0;
    int64_t n = ctx->n;
    ph_begin(ctx, SPH_PH_KICKDRIFT);
    HIPCK(sph::kick_drift(ctx->st, n, D(vx), D(vy), D(vz), D(ax), D(ay), D(az), 0.5 * p->dt, p->dt, D(x), D(y), D(z),
                          p->period));
    ph_end(ctx, SPH_PH_KICKDRIFT);
    int r = p->precision == SPH_PREC_MIXED ? forces<float>(ctx, p, 0) : forces<double>(ctx, p, 0);
    if (r != SPH_OK) return r;
    CK(finish_step(ctx, p));
    ctx->stats.steps += 1;
    ctx->timing = false;
    if (ctx->agn_on) CK(inject_outflow(ctx, p->dt));
    return SPH_OK;
}

extern "C" int sph_get_stats(sph_ctx *ctx, sph_stats *out)
{
    if (!ctx || !out) return SPH_ERR_ARG;
    *out = ctx->stats;
    return SPH_OK;
}

extern "C" int sph_reset_stats(sph_ctx *ctx)
{
    if (!ctx) return SPH_ERR_ARG;
    int64_t n = ctx->stats.n;
    std::memset(&ctx->stats, 0, sizeof(ctx->stats));
    ctx->stats.n = n;
    return SPH_OK;
}

extern "C" int sph_state_get_timestep(sph_ctx *ctx, double *dt_out, double *dhdt_out)
{
    if (!ctx) return SPH_ERR_ARG;
    if (!ctx->resident) return fail(ctx, SPH_ERR_STATE, "no resident state (call sph_state_set first)");
    if (!ctx->ts_ok) return fail(ctx, SPH_ERR_STATE, "no time-step criteria (set C_CFL > 0 in sph_step_params)");
    HIPCK(hipSetDevice(ctx->device));
    if (ctx->n > ctx->ts_n) {  // particles injected after the evaluation
        HIPCK(sph::fill<double>(ctx->st, ctx->n - ctx->ts_n, D(tsdt) + ctx->ts_n, NAN));
        HIPCK(sph::fill<double>(ctx->st, ctx->n - ctx->ts_n, D(tsdh) + ctx->ts_n, NAN));
    }
    if (dt_out) CK(get_scalar<double>(ctx, ctx->tsdt, dt_out));
    if (dhdt_out) CK(get_scalar<double>(ctx, ctx->tsdh, dhdt_out));
    return SPH_OK;
}

extern "C" int sph_state_get_duprev(sph_ctx *ctx, double *duprev_out)
{
    if (!ctx || !duprev_out) return SPH_ERR_ARG;
    if (!ctx->resident) return fail(ctx, SPH_ERR_STATE, "no resident state (call sph_state_set first)");
    HIPCK(hipSetDevice(ctx->device));
    return get_scalar<double>(ctx, ctx->duprev, duprev_out);
}

extern "C" int sph_state_set_restart(sph_ctx *ctx, const double *acc, const double *duprev, double agn_leftover,
                                     uint64_t agn_step, int64_t agn_injected)
{
    if (!ctx) return SPH_ERR_ARG;
    if (!acc || !duprev) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (!ctx->resident) return fail(ctx, SPH_ERR_STATE, "no resident state (call sph_state_set first)");
    if (!(agn_leftover >= 0.0) || agn_injected < 0 || agn_injected > ctx->n)
        return fail(ctx, SPH_ERR_ARG, "bad injection bookkeeping");
    HIPCK(hipSetDevice(ctx->device));
    CK(put_vec3(ctx, acc, ctx->ax, ctx->ay, ctx->az));
    CK(put_scalar(ctx, duprev, ctx->duprev, nullptr, nullptr, nullptr));
    HIPCK(hipStreamSynchronize(ctx->st));
    ctx->agn_leftover = agn_leftover;
    ctx->agn_step = agn_step;
    ctx->agn_injected = agn_injected;
    ctx->primed = true;
    return SPH_OK;
}

extern "C" int sph_state_reserve(sph_ctx *ctx, int64_t n_max)
{
    if (!ctx) return SPH_ERR_ARG;
    if (n_max < 0 || n_max > (int64_t)INT32_MAX) return fail(ctx, SPH_ERR_ARG, "n_max out of range");
    ctx->n_reserve = std::max(ctx->n_reserve, n_max);  // placed by the next sph_state_set
    return SPH_OK;
}

extern "C" int sph_state_set_agn(sph_ctx *ctx, const sph_agn_params *p)
{
    if (!ctx) return SPH_ERR_ARG;
    if (!ctx->resident) return fail(ctx, SPH_ERR_STATE, "no resident state (call sph_state_set first)");
    if (!p) {
        ctx->agn_on = false;
        return SPH_OK;
    }
    if (!(p->sigma >= 0.0) || !(p->M_dot >= 0.0) || !std::isfinite(p->v_in) || !std::isfinite(p->u_inject) ||
        !(p->r_inject >= 0.0))
        return fail(ctx, SPH_ERR_ARG, "AGN source terms: sigma, M_dot, r_inject must be >= 0, v_in and u finite");
    if (p->M_dot > 0.0 && !(p->m_inject > 0.0))
        return fail(ctx, SPH_ERR_ARG, "AGN injection needs m_inject > 0");
    if (p->M_dot > 0.0 && (p->h_ref_index < 0 || p->h_ref_index >= ctx->n))
        return fail(ctx, SPH_ERR_ARG, "h_ref_index is not a particle of the resident state");
    ctx->agn = *p;
    ctx->agn_on = true;
    ctx->agn_step = 0;
    return SPH_OK;
}

extern "C" int sph_state_count(sph_ctx *ctx, int64_t *n, double *leftover_mass, uint64_t *agn_step,
                               int64_t *agn_injected)
{
    if (!ctx) return SPH_ERR_ARG;
    if (!ctx->resident) return fail(ctx, SPH_ERR_STATE, "no resident state (call sph_state_set first)");
    if (n) *n = ctx->n;
    if (leftover_mass) *leftover_mass = ctx->agn_leftover;
    if (agn_step) *agn_step = ctx->agn_step;
    if (agn_injected) *agn_injected = ctx->agn_injected;
    return SPH_OK;
}

extern "C" int sph_state_set_column(sph_ctx *ctx, const sph_column_params *cp)
{
    if (!ctx) return SPH_ERR_ARG;
    if (!ctx->resident) return fail(ctx, SPH_ERR_STATE, "no resident state (call sph_state_set first)");
    HIPCK(hipSetDevice(ctx->device));
    if (!cp) {
        ctx->col_on = false;
        return SPH_OK;
    }
    CK(column_check(ctx, cp));
    CK(ensure(ctx, ctx->colg, (size_t)std::max(ctx->n, ctx->n_reserve) * 8));
    CK(ensure(ctx, ctx->coln, (size_t)std::max(ctx->n, ctx->n_reserve) * 8));
    ctx->col_p = column_params(cp);
    ctx->col_on = true;
    return SPH_OK;
}

extern "C" int sph_state_get_column(sph_ctx *ctx, double *grad_out, double *log_nh_out)
{
    if (!ctx) return SPH_ERR_ARG;
    if (!ctx->resident) return fail(ctx, SPH_ERR_STATE, "no resident state (call sph_state_set first)");
    if (!ctx->col_ok)
        return fail(ctx, SPH_ERR_STATE, "no column density yet (sph_state_set_column, then sph_prime or sph_step)");
    HIPCK(hipSetDevice(ctx->device));
    if (ctx->n > ctx->n_eval) {  // particles injected after the evaluation
        HIPCK(sph::fill<double>(ctx->st, ctx->n - ctx->n_eval, D(colg) + ctx->n_eval, NAN));
        HIPCK(sph::fill<double>(ctx->st, ctx->n - ctx->n_eval, D(coln) + ctx->n_eval, NAN));
    }
    if (grad_out) CK(get_scalar<double>(ctx, ctx->colg, grad_out));
    if (log_nh_out) CK(get_scalar<double>(ctx, ctx->coln, log_nh_out));
    return SPH_OK;
}

