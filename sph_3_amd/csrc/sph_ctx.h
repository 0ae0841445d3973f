// sph_ctx.h -- internal to the C ABI sources (sph_capi*.hip): the context that holds the
// device state, its buffers, and the helpers the entry points share.  Not part of the ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sph_amd.h"
#include "sph_common.h"
#include "sph_agn.h"
#include "sph_cool.h"
#include "sph_grav.h"
#include "sph_hydro.h"
#include "sph_integrate.h"
#include "sph_mlp.h"
#include "sph_nhcool.h"
#include "sph_ngb.h"
#include "sph_order.h"
#include "sph_tree.h"

#define SPH_VERSION 1
#define SPH_BUCKET 16      // particles per tree bucket (walk leaves)
#define SPH_HCAP SPH_HCAP_SLOTS  // candidate slots per target in the h solves
#define SPH_HROUNDS 64     // maximum radius-adjustment rounds of an h solve
#define SPH_KEYBITS 48     // Morton key bits (16 per axis)
#define SPR_H 0            // spread counters of the h solve (5)
#define SPR_G 5            // spread counters of the gravity walk (3)
#define SPR_N 8            // spread counters of the padded neighbour rows (2)
#define SPR_C 10           // spread counter of the cooling update (1)
#define SPR_TOTAL 11
#define SPH_NGB_STRIDE 128 // entries per padded neighbour row


struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct sph_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;
    int64_t n = 0;

    // sorted particle state
    Buf x, y, z, vx, vy, vz, m, h, eps, u, duprev, ax, ay, az, gx, gy, gz, tmpd;
    Buf shadow[SPH_GATHER_MAX];  // targets of the reorder gathers (swapped with the state)
    double m_max = INFINITY;     // largest particle mass of the current masses
    int64_t n_reserve = 0;       // per-particle buffers hold at least this many (injection room)
    uint64_t r2_limit = 0;       // bytes the h solve's candidate distance rows may take (else indices only)
    int64_t n_placed = 0;        // room of the resident state's buffers (the reserve when it was set)
    bool agn_on = false;
    sph_agn_params agn{};
    double agn_leftover = 0.0;   // injection mass carried to the next step
    int64_t agn_injected = 0;    // particles injected since the state was set
    uint64_t agn_step = 0;       // steps since the source terms were set (the generator's step)
    Buf orig, tmpi;
    // ordering
    Buf keys0, keys1, vals0, vals1, hist, bpart, box;
    // tree
    Buf nodes, right, parent, ncount, flags, front, groups;
    // neighbour lists
    Buf cnt, rcnt0, off, scantmp, idx;  // rcnt0: row lengths before the appended pairs
    Buf p4, v4, s4;                      // packed (x, y, z, h), (vx, vy, vz, m), (rho, P/rho^2, c, f) records
    Buf pf, vf;                          // compact records of the mixed form (NgbRows::pf, vf)
    // hydro scalars (8 bytes per particle, used as float or double)
    Buf rho, P, c, prho2, divV, curlV, bal, du;
    // h solve scratch
    Buf cand, hR, hsteps, hstate, hcount, hout, hin, tx, ty, tz;
    Buf hplist;  // pending groups and targets of the h-solve rounds after the first (+ two counters)
    Buf ccidx, cccnt, cext, cnext;  // candidate indices, counts, extra pairs and their region counters
    Buf chr;                        // (x_j, h_j, R_j) packed for the rows from candidates
    Buf hbig;                       // targets of the candidate-free h solver
    // staging and reductions
    Buf stage, kpart, kout;
    // spread event counters (sph_common.h): SPR_H.. h solve (5), SPR_G.. gravity (3);
    // their totals land in hcount: [0..4] h solve, [5..7] gravity, [8] walk overflow, [9] group count,
    // [10..11] padded neighbour rows, [12] cooling floor, [13] surrogate fallbacks, [14] targets of
    // the candidate-free h solver
    Buf spread;
    // cooling: field integrals, table, elementwise scratch
    Buf rf1, ctabL, ctabG, cin[6], cout[2];
    bool rf1_ok = false;
    CoolAxes cax{};
    bool ctab_ok = false;
    Buf nhtab;                   // N_H cooling table: grids, heat, cool, mu, log u (floats)
    int nh_dims[4] = {0, 0, 0, 0};
    float nh_bounds[8] = {};
    bool nh_ok = false;
    // cooling update: temperature table (built for ttab_XH), per-particle Z and J21 of the
    // resident state (sorted), its parameters, and the surrogate path's columns
    Buf ttab, zmet, j21, ucols, uyn, uL, uG, uflag, ucnt;
    CoolTAxes tax{};
    double ttab_XH = -1.0;
    bool cool_on = false, cool_ran = false;
    bool col_on = false, col_ok = false;  // column density in the step; computed for the current state
    sph::ColumnParams col_p{};
    Buf colg, coln;  // |grad rho|, log10 N_H (sorted order, FP64)
    int64_t n_eval = 0;     // particles of the last force evaluation (later injections are not in it)
    Buf phi, tpart, tout;   // potential of the energy option, FP64 totals (partials, result)
    bool tot_ran = false;   // the last force evaluation computed the totals
    bool ts_ok = false;     // time-step criteria of the last force evaluation
    int64_t ts_n = 0;       // particles they cover (later injections are not covered)
    Buf tsdt, tsdh;         // dt_i, dh/dt_i (sorted order, FP64)
    sph_cool_params cool_p{};
    // cooling surrogate: parameters, Adam state, BF16 weight copies, samples, error map
    Buf mP, mG, mM, mV, mW[4], mimg, mx, mxn, mt, my, mloss, memap, memap2, memap3, mgate, mcnt, mfl;
    MLPNorm mnorm{};
    bool mlp_ok = false, mlp_cal = false;
    int mlp_step = 0;
    int mcells[4] = {0, 0, 0, 0};

    // caches of the stateless entry points (host copies of the last inputs)
    std::vector<double> cpos, ch, cm, ceps;
    bool order_ok = false, h_ok = false, m_ok = false, eps_ok = false, moments_ok = false, csr_ok = false;
    bool resident = false, primed = false;
    int64_t csr_total = 0;
    int ngb_stride = 0;  // layout of the neighbour sets: 0 CSR, else padded rows of this width
    double ngb_per = 0.0;  // side of the periodic box the sets were built in (0: open boundaries)
    int64_t ngroups = 0;  // tree-node target groups of the current tree

    // statistics
    sph_stats stats;
    hipEvent_t ev[2 * SPH_PH_COUNT];
    bool timing = false;
    uint32_t ph_rec = 0;  // phases recorded since the last finish_step
    // deferred read-backs of the step: pinned host words ([0..4] h-solve totals of the first
    // round, [8..9] neighbour-row totals, [10..10+SPH_SPREAD) extras region counters) and the
    // events recorded after their copies
    unsigned long long *hpin = nullptr;
    hipEvent_t ev_h = nullptr, ev_n = nullptr;
    int64_t hs_unc = 0, hs_work[2] = {0, 0};
    int rho_prec = SPH_PREC_MIXED;  // arithmetic type of the resident rho buffer
    double *scratch = nullptr;      // this device's address of the static scratch array
};

namespace {
// (each C ABI source gets its own copy)
__global__ void k_deinterleave(int64_t n, const double *__restrict__ aos, double *__restrict__ x,
                               double *__restrict__ y, double *__restrict__ z)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) {
        x[i] = aos[3 * i];
        y[i] = aos[3 * i + 1];
        z[i] = aos[3 * i + 2];
    }
}
}  // namespace

namespace sphc {

int fail(sph_ctx *c, int code, const std::string &msg);

#define HIPCK(...)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (__VA_ARGS__);                                                               \
        if (e_ != hipSuccess)                                                                        \
            return fail(ctx, SPH_ERR_HIP, std::string(#__VA_ARGS__) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CK(...)                                                                                      \
    do {                                                                                             \
        int r_ = (__VA_ARGS__);                                                                      \
        if (r_ != SPH_OK) return r_;                                                                 \
    } while (0)

template <typename T> T *P_(Buf &b) { return (T *)b.p; }
#define D(b) P_<double>(ctx->b)
#define I32(b) P_<int32_t>(ctx->b)

bool particles_ok(int64_t n, const double *pos, const double *vel, const double *m, const double *u,
                  const double *h, std::string &why);
int check_particles(sph_ctx *ctx, int64_t n, const double *pos, const double *vel, const double *m, const double *u,
                    const double *h);
int ensure(sph_ctx *ctx, Buf &b, size_t bytes);
int capacity(sph_ctx *ctx, int64_t n);
void invalidate_all(sph_ctx *ctx);
int sort_current(sph_ctx *ctx, int64_t n);
int topology(sph_ctx *ctx);
int set_pos(sph_ctx *ctx, int64_t n, const double *pos);
int put_scalar(sph_ctx *ctx, const double *src, Buf &dst, std::vector<double> *cache, bool *ok, bool *changed);
int put_vec3(sph_ctx *ctx, const double *src, Buf &a, Buf &b, Buf &c);
int get_vec3(sph_ctx *ctx, Buf &a, Buf &b, Buf &c, double *out);
int ensure_moments(sph_ctx *ctx);
int build_ngb(sph_ctx *ctx, bool csr, const double *hscale = nullptr, double per = 0.0);
int ngb_rows(sph_ctx *ctx, NgbRows &r, bool vel_only = false, bool compact = false, bool full = true);
int ensure_csr(sph_ctx *ctx, bool csr = false);
int geom(sph_ctx *ctx, int64_t n, const double *pos, const double *h, const double *m);
int check_common(sph_ctx *ctx, int64_t n, int precision);
void ph_begin(sph_ctx *ctx, int ph);
void ph_end(sph_ctx *ctx, int ph);

template <typename R>
int put_real(sph_ctx *ctx, const double *src, Buf &dst)
{
    int64_t n = ctx->n;
    HIPCK(hipMemcpyAsync(ctx->stage.p, src, sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
    HIPCK(sph::gather<double, R>(ctx->st, n, I32(orig), D(stage), 1, 0, (R *)dst.p));
    return SPH_OK;
}

template <typename T>
int get_scalar(sph_ctx *ctx, Buf &src, double *out)
{
    int64_t n = ctx->n;
    HIPCK(sph::scatter<T>(ctx->st, n, I32(orig), (const T *)src.p, 1, 0, D(stage)));
    HIPCK(hipMemcpyAsync(out, ctx->stage.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    return SPH_OK;
}

// column-density parameters: checked, and in the kernels' form
inline int column_check(sph_ctx *ctx, const sph_column_params *cp)
{
    if (!(cp->unit_density_cgs > 0.0) || !(cp->unit_length_cm > 0.0) || !(cp->XH > 0.0) || !(cp->n_ngb > 0.0))
        return fail(ctx, SPH_ERR_ARG, "column density: unit_density_cgs, unit_length_cm, XH, n_ngb must be > 0");
    return SPH_OK;
}
inline sph::ColumnParams column_params(const sph_column_params *cp)
{
    return sph::ColumnParams{cp->unit_density_cgs, cp->unit_length_cm, cp->XH, cp->n_ngb};
}

// sph_capi_api.hip
int hsolve_params(sph_ctx *ctx, int64_t m, const double *tx, const double *ty, const double *tz, int mode,
                  double param, int n_low, int n_up, int max_steps, double tol, int64_t *unc, int64_t *rounds_out,
                  int64_t *work, sph::HSolveParams &p);
// sph_capi_cool.hip: the cooling update of n particles (counts[0] floored, counts[1] surrogate fallbacks)
template <typename RHO>
int cool_run(sph_ctx *ctx, int64_t n, const RHO *rho, const double *u, const double *Z, const double *J21,
             double dt, const sph_cool_params &cp, double *u_out, unsigned long long *counts);
// Per-particle Z and J21 travel with the resident state (the cooling modes that read them; the
// N_H-table mode reads neither and they are not allocated for it).
inline bool cool_columns(const sph_ctx *ctx) { return ctx->cool_on && ctx->cool_p.mode != SPH_COOL_NH; }
// sph_capi_cool.hip: the N_H-table cooling update (doCloudyHCooling) of n particles; counts[0]
// floored, counts[1] = 0
template <typename RHO>
int nh_run(sph_ctx *ctx, int64_t n, const double *x, const double *y, const double *z, const RHO *rho,
           const double *u, const double *log_nh, const int32_t *typ, double gamma, double dt,
           const sph_cool_params &cp, double *u_out, unsigned long long *counts);

}  // namespace sphc
