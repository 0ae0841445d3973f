// sph_capi.hip -- the C ABI (include/sph_amd.h): context, device caches of
// the particle state in Morton order, the per-function entry points that
// mirror the reference's Python call surface, and the device-resident step.
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

namespace {

struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
};

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

__global__ void k_iota(int64_t n, int32_t *__restrict__ a)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) a[i] = (int32_t)i;
}

__global__ void k_copy_d(int64_t n, const double *__restrict__ s, double *__restrict__ d)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

}  // namespace

// ---- scratch of the allocation-free entry point (sph_kinetics_noalloc): one static device
// array per GPU, held from module load until the process ends.  Layout (doubles): the
// caller's n x 3 velocities at [0, 3n), masses at [3n, 4n), the reduction's 1,024 block
// partials of 5 values at [4 SPH_NOALLOC_MAX_N, + 5120), the 5 results after them.  In host
// code `scratch` names a host-side stand-in: kernels and copies use the address
// hipGetSymbolAddress returned for the device (g_scratch[device]).
#define SPH_NOALLOC_MAX_N (1 << 20)
#define SPH_NOALLOC_PART (4 * (size_t)SPH_NOALLOC_MAX_N)
#define SPH_NOALLOC_WORDS (SPH_NOALLOC_PART + 5 * 1024 + 8)
#define SPH_MAX_DEVICES 64
static __device__ __align__(256) double scratch[SPH_NOALLOC_WORDS];

namespace {
std::mutex g_scratch_mu;
double *g_scratch[SPH_MAX_DEVICES];           // device address of scratch per device (null: unresolved)
hipError_t g_scratch_err[SPH_MAX_DEVICES];    // why it is unresolved

// With `device` current: resolve that device's address of scratch once.
double *resolve_scratch(int device)
{
    if (device < 0 || device >= SPH_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess) g_scratch[device] = (double *)p;
        else g_scratch_err[device] = e;
    }
    return g_scratch[device];
}
}  // namespace

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
    // hydro scalars (8 bytes per particle, used as float or double)
    Buf rho, P, c, prho2, divV, curlV, bal, du;
    // h solve scratch
    Buf cand, hR, hsteps, hstate, hcount, hout, hin, tx, ty, tz;
    Buf hplist;  // pending groups and targets of the h-solve rounds after the first (+ two counters)
    Buf ccidx, cccnt, cext, cnext;  // candidate indices, counts, extra pairs and their region counters
    Buf chr;                        // (h_j, R_j) packed for the rows from candidates
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
    Buf mP, mG, mM, mV, mW[4], mx, mxn, mt, my, mloss, memap, memap2, memap3, mgate, mcnt, mfl;
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

int fail(sph_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

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

// Host checks of a particle set (sph_validate_particles); false with the reason in why.
bool particles_ok(int64_t n, const double *pos, const double *vel, const double *m, const double *u,
                  const double *h, std::string &why)
{
    auto bad = [&](const char *what, int64_t i, double v) {
        char b[160];
        std::snprintf(b, sizeof(b), "%s of particle %lld is %.17g", what, (long long)i, v);
        why = b;
        return false;
    };
    if (n < 1) {
        why = "n must be >= 1";
        return false;
    }
    if (pos) {
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) {
                const double v = pos[3 * i + d];
                if (!std::isfinite(v)) return bad("a position coordinate", i, v);
                lo[d] = std::min(lo[d], v);
                hi[d] = std::max(hi[d], v);
            }
        if (n > 1 && !(hi[0] > lo[0] || hi[1] > lo[1] || hi[2] > lo[2])) {
            why = "all particles coincide: the bounding box has zero extent, so there is no spatial order";
            return false;
        }
    }
    if (vel)
        for (int64_t i = 0; i < 3 * n; i++)
            if (!std::isfinite(vel[i])) return bad("a velocity component", i / 3, vel[i]);
    if (m)
        for (int64_t i = 0; i < n; i++)
            if (!(m[i] >= 0.0) || !std::isfinite(m[i])) return bad("the mass (must be finite and >= 0)", i, m[i]);
    if (u)
        for (int64_t i = 0; i < n; i++)
            if (!(u[i] >= 0.0) || !std::isfinite(u[i])) return bad("u (must be finite and >= 0)", i, u[i]);
    if (h)
        for (int64_t i = 0; i < n; i++)
            if (!(h[i] > 0.0) || !std::isfinite(h[i])) return bad("h (must be finite and > 0)", i, h[i]);
    return true;
}

int check_particles(sph_ctx *ctx, int64_t n, const double *pos, const double *vel, const double *m, const double *u,
                    const double *h)
{
    std::string why;
    if (!particles_ok(n, pos, vel, m, u, h, why)) return fail(ctx, SPH_ERR_ARG, why);
    return SPH_OK;
}

int ensure(sph_ctx *ctx, Buf &b, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (b.bytes >= bytes) return SPH_OK;
    if (b.p) HIPCK(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
    size_t nb = bytes + bytes / 8;
    HIPCK(hipMalloc(&b.p, nb));
    b.bytes = nb;
    return SPH_OK;
}

template <typename T> T *P_(Buf &b) { return (T *)b.p; }
#define D(b) P_<double>(ctx->b)
#define I32(b) P_<int32_t>(ctx->b)

int capacity(sph_ctx *ctx, int64_t n)
{
    n = std::max(n, ctx->n_reserve);  // room for injected particles: no state buffer ever moves
    size_t nd = (size_t)n * 8, ni = (size_t)n * 4;
    Buf *dbl[] = {&ctx->x, &ctx->y, &ctx->z, &ctx->vx, &ctx->vy, &ctx->vz, &ctx->m, &ctx->h, &ctx->eps,
                  &ctx->u, &ctx->duprev, &ctx->ax, &ctx->ay, &ctx->az, &ctx->gx, &ctx->gy, &ctx->gz,
                  &ctx->tmpd, &ctx->rho, &ctx->P, &ctx->c, &ctx->prho2, &ctx->divV, &ctx->curlV,
                  &ctx->bal, &ctx->du, &ctx->hR, &ctx->hout, &ctx->hin, &ctx->tx, &ctx->ty, &ctx->tz};
    for (Buf *b : dbl) CK(ensure(ctx, *b, nd));
    Buf *i32[] = {&ctx->orig, &ctx->tmpi, &ctx->vals0, &ctx->vals1, &ctx->cnt, &ctx->rcnt0, &ctx->hsteps, &ctx->hstate};
    for (Buf *b : i32) CK(ensure(ctx, *b, ni));
    CK(ensure(ctx, ctx->keys0, nd));
    CK(ensure(ctx, ctx->keys1, nd));
    CK(ensure(ctx, ctx->hist, (size_t)sph::radix_hist_entries(n) * 4));
    CK(ensure(ctx, ctx->bpart, 6 * 1024 * 8));
    CK(ensure(ctx, ctx->box, 8 * 8));
    CK(ensure(ctx, ctx->nodes, (size_t)(2 * n) * sizeof(WNode)));
    CK(ensure(ctx, ctx->right, ni));
    CK(ensure(ctx, ctx->parent, 2 * ni));
    CK(ensure(ctx, ctx->ncount, 2 * ni));
    CK(ensure(ctx, ctx->flags, ni));
    CK(ensure(ctx, ctx->front, std::max(2 * ni, (size_t)sph::moments_scratch_entries(n) * 4)));
    CK(ensure(ctx, ctx->groups, 2 * ni));
    CK(ensure(ctx, ctx->off, (size_t)(n + 1) * 8));
    CK(ensure(ctx, ctx->scantmp, (size_t)sph::scan_tmp_entries(n) * 8));
    CK(ensure(ctx, ctx->hcount, 16 * 8));
    CK(ensure(ctx, ctx->spread, (size_t)SPR_TOTAL * SPH_SPREAD_WORDS * 8));
    CK(ensure(ctx, ctx->stage, 3 * nd));
    CK(ensure(ctx, ctx->kpart, 5 * 1024 * 8));
    CK(ensure(ctx, ctx->kout, 8 * 8));
    return SPH_OK;
}

void invalidate_all(sph_ctx *ctx)
{
    ctx->order_ok = ctx->h_ok = ctx->m_ok = ctx->eps_ok = ctx->moments_ok = ctx->csr_ok = false;
    ctx->cpos.clear();
    ctx->ch.clear();
    ctx->cm.clear();
    ctx->ceps.clear();
}

// Sort the particles currently in (x, y, z) by Morton key; the permutation
// (sorted -> current index) lands in vals0; keys0 holds the sorted keys.
int sort_current(sph_ctx *ctx, int64_t n)
{
    hipStream_t st = ctx->st;
    HIPCK(sph::bbox(st, n, D(x), D(y), D(z), D(bpart), D(box)));
    HIPCK(sph::morton(st, n, D(x), D(y), D(z), D(box), P_<uint64_t>(ctx->keys0), I32(vals0)));
    HIPCK(sph::radix_sort(st, n, SPH_KEYBITS, P_<uint64_t>(ctx->keys0), I32(vals0), P_<uint64_t>(ctx->keys1),
                          I32(vals1), P_<uint32_t>(ctx->hist)));
    if ((SPH_KEYBITS / 8) % 2) {  // result in the alternate buffers
        std::swap(ctx->keys0, ctx->keys1);
        std::swap(ctx->vals0, ctx->vals1);
    }
    return SPH_OK;
}

// Tree topology of the sorted keys and the target groups (maximal nodes of
// at most 64 particles) used by every walk.
int topology(sph_ctx *ctx)
{
    HIPCK(sph::tree_topology(ctx->st, ctx->n, P_<uint64_t>(ctx->keys0), P_<WNode>(ctx->nodes), I32(right),
                             I32(parent), I32(ncount)));
    HIPCK(sph::gravity_groups(ctx->st, ctx->n, I32(ncount), I32(parent), I32(groups),
                              (unsigned int *)(P_<unsigned long long>(ctx->hcount) + 9), &ctx->ngroups));
    return SPH_OK;
}

// Stateless entry: establish the Morton order of the caller's positions.
int set_pos(sph_ctx *ctx, int64_t n, const double *pos)
{
    if (ctx->resident) {
        ctx->resident = false;
        ctx->primed = false;
        invalidate_all(ctx);
    }
    if (ctx->order_ok && n == ctx->n && (int64_t)ctx->cpos.size() == 3 * n &&
        std::memcmp(ctx->cpos.data(), pos, sizeof(double) * 3 * n) == 0)
        return SPH_OK;
    CK(check_particles(ctx, n, pos, nullptr, nullptr, nullptr, nullptr));
    invalidate_all(ctx);
    CK(capacity(ctx, n));
    ctx->n = n;
    hipStream_t st = ctx->st;
    HIPCK(hipMemcpyAsync(ctx->stage.p, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
    k_deinterleave<<<sph_blocks(n, 256), 256, 0, st>>>(n, D(stage), D(x), D(y), D(z));
    CK(sort_current(ctx, n));
    HIPCK(hipMemcpyAsync(ctx->orig.p, ctx->vals0.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
    for (int d = 0; d < 3; d++) {
        double *dst = d == 0 ? D(x) : d == 1 ? D(y) : D(z);
        HIPCK(sph::gather<double, double>(st, n, I32(orig), D(stage), 3, d, dst));
    }
    CK(topology(ctx));
    ctx->cpos.assign(pos, pos + 3 * n);
    ctx->order_ok = true;
    return SPH_OK;
}

// Upload a caller-order scalar array into sorted buffer dst (double).  With a
// cache: returns 1 in *changed when the data differ from the last upload.
int put_scalar(sph_ctx *ctx, const double *src, Buf &dst, std::vector<double> *cache, bool *ok, bool *changed)
{
    int64_t n = ctx->n;
    if (changed) *changed = false;
    if (cache && *ok && (int64_t)cache->size() == n && std::memcmp(cache->data(), src, sizeof(double) * n) == 0)
        return SPH_OK;
    HIPCK(hipMemcpyAsync(ctx->stage.p, src, sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
    HIPCK(sph::gather<double, double>(ctx->st, n, I32(orig), D(stage), 1, 0, (double *)dst.p));
    if (cache) {
        cache->assign(src, src + n);
        *ok = true;
    }
    if (changed) *changed = true;
    if (&dst == &ctx->m) {
        double mx = 0.0;
        for (int64_t i = 0; i < n; i++) mx = std::max(mx, std::fabs(src[i]));
        ctx->m_max = mx;
    }
    return SPH_OK;
}

template <typename R>
int put_real(sph_ctx *ctx, const double *src, Buf &dst)
{
    int64_t n = ctx->n;
    HIPCK(hipMemcpyAsync(ctx->stage.p, src, sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
    HIPCK(sph::gather<double, R>(ctx->st, n, I32(orig), D(stage), 1, 0, (R *)dst.p));
    return SPH_OK;
}

int put_vec3(sph_ctx *ctx, const double *src, Buf &a, Buf &b, Buf &c)
{
    int64_t n = ctx->n;
    HIPCK(hipMemcpyAsync(ctx->stage.p, src, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->st));
    Buf *t[3] = {&a, &b, &c};
    for (int d = 0; d < 3; d++) HIPCK(sph::gather<double, double>(ctx->st, n, I32(orig), D(stage), 3, d, (double *)t[d]->p));
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

int get_vec3(sph_ctx *ctx, Buf &a, Buf &b, Buf &c, double *out)
{
    int64_t n = ctx->n;
    Buf *t[3] = {&a, &b, &c};
    for (int d = 0; d < 3; d++)
        HIPCK(sph::scatter<double>(ctx->st, n, I32(orig), (const double *)t[d]->p, 3, d, D(stage)));
    HIPCK(hipMemcpyAsync(out, ctx->stage.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->st));
    HIPCK(hipStreamSynchronize(ctx->st));
    return SPH_OK;
}

int ensure_moments(sph_ctx *ctx)
{
    if (ctx->moments_ok) return SPH_OK;
    HIPCK(sph::tree_moments(ctx->st, ctx->n, SPH_BUCKET, D(x), D(y), D(z), ctx->m_ok ? D(m) : nullptr,
                            ctx->h_ok ? D(h) : nullptr, ctx->eps_ok ? D(eps) : nullptr, P_<WNode>(ctx->nodes),
                            I32(right), I32(parent), I32(ncount), P_<uint32_t>(ctx->flags), I32(front)));
    ctx->moments_ok = true;
    return SPH_OK;
}

// Neighbour sets of the current positions and h: padded rows of
// SPH_NGB_STRIDE entries built in one walk, or (csr, or a row longer than the
// stride) the exact two-pass CSR.
int build_ngb(sph_ctx *ctx, bool csr, const double *hscale = nullptr, double per = 0.0)
{
    const int64_t n = ctx->n;
    ctx->ngb_per = per;
    unsigned long long *ovf = P_<unsigned long long>(ctx->hcount) + 8;
    if (!csr) {
        CK(ensure(ctx, ctx->idx, (size_t)n * SPH_NGB_STRIDE * 4));
        int64_t nlong = 0, total = 0;
        HIPCK(sph::neighbour_rows(ctx->st, ctx->ngroups, I32(groups), D(x), D(y), D(z), D(h), P_<WNode>(ctx->nodes),
                                  SPH_BUCKET, SPH_NGB_STRIDE, I32(cnt), I32(idx), ovf,
                                  P_<unsigned long long>(ctx->spread) + SPR_N * SPH_SPREAD_WORDS,
                                  P_<unsigned long long>(ctx->hcount) + 10, &nlong, &total, hscale, per));
        if (nlong == 0) {
            ctx->ngb_stride = SPH_NGB_STRIDE;
            ctx->csr_total = total;
            ctx->stats.ngb_total = total;
            return SPH_OK;
        }
    }
    int64_t total = 0;
    int64_t cap = ctx->idx.p ? (int64_t)(ctx->idx.bytes / 4) : 0;
    HIPCK(sph::neighbours(ctx->st, n, ctx->ngroups, I32(groups), D(x), D(y), D(z), D(h), P_<WNode>(ctx->nodes),
                          SPH_BUCKET, I32(cnt), P_<int64_t>(ctx->off), P_<int64_t>(ctx->scantmp), ovf, I32(idx), cap,
                          &total, hscale, per));
    if (total > cap) {
        CK(ensure(ctx, ctx->idx, (size_t)total * 4));
        HIPCK(sph::neighbours_fill(ctx->st, ctx->ngroups, I32(groups), D(x), D(y), D(z), D(h), P_<WNode>(ctx->nodes),
                                   SPH_BUCKET, P_<int64_t>(ctx->off), I32(idx), ovf, hscale, per));
    }
    ctx->ngb_stride = 0;
    ctx->csr_total = total;
    ctx->stats.ngb_total = total;
    return SPH_OK;
}

// The neighbour sets plus the hydro passes' packed inputs, (x, y, z, h) and (vx, vy, vz, m)
// of the sorted particles, packed from the current arrays (call after any of them changes).
// vel_only: only the velocities changed since the last call (positions, h, m and n as then).
int ngb_rows(sph_ctx *ctx, NgbRows &r, bool vel_only = false)
{
    const int64_t n = ctx->n;
    CK(ensure(ctx, ctx->p4, (size_t)n * 32));
    CK(ensure(ctx, ctx->v4, (size_t)n * 32));
    if (!vel_only) HIPCK(sph::pack4(ctx->st, n, D(x), D(y), D(z), D(h), P_<double4>(ctx->p4)));
    HIPCK(sph::pack4(ctx->st, n, D(vx), D(vy), D(vz), D(m), P_<double4>(ctx->v4)));
    r = NgbRows{P_<int64_t>(ctx->off), I32(cnt), I32(idx), ctx->ngb_stride, ctx->ngb_per,
                P_<double4>(ctx->p4), P_<double4>(ctx->v4)};
    return SPH_OK;
}

// csr: the caller needs true CSR offsets (sph_neighbours).
int ensure_csr(sph_ctx *ctx, bool csr = false)
{
    if (ctx->csr_ok && (!csr || ctx->ngb_stride == 0)) return SPH_OK;
    CK(ensure_moments(ctx));
    CK(build_ngb(ctx, csr));
    ctx->csr_ok = true;
    return SPH_OK;
}

// Common prologue of the hydro entry points.
int geom(sph_ctx *ctx, int64_t n, const double *pos, const double *h, const double *m)
{
    CK(check_particles(ctx, n, nullptr, nullptr, m, nullptr, h));
    CK(set_pos(ctx, n, pos));
    bool ch = false, cm = false;
    CK(put_scalar(ctx, h, ctx->h, &ctx->ch, &ctx->h_ok, &ch));
    if (m) CK(put_scalar(ctx, m, ctx->m, &ctx->cm, &ctx->m_ok, &cm));
    if (ch) ctx->csr_ok = false;
    if (ch || cm) ctx->moments_ok = false;
    return ensure_csr(ctx);
}

int check_common(sph_ctx *ctx, int64_t n, int precision)
{
    if (!ctx) return SPH_ERR_ARG;
    if (n < 1) return fail(ctx, SPH_ERR_ARG, "n must be >= 1");
    if (n > (int64_t)0x3fffffff) return fail(ctx, SPH_ERR_ARG, "n too large for 32-bit particle indices");
    if (precision != SPH_PREC_MIXED && precision != SPH_PREC_FP64)
        return fail(ctx, SPH_ERR_ARG, "precision must be SPH_PREC_MIXED or SPH_PREC_FP64");
    HIPCK(hipSetDevice(ctx->device));
    return SPH_OK;
}

// ---- phase timing for the resident step
void ph_begin(sph_ctx *ctx, int ph)
{
    if (ctx->timing) {
        (void)hipEventRecord(ctx->ev[2 * ph], ctx->st);
        ctx->ph_rec |= 1u << ph;
    }
}
void ph_end(sph_ctx *ctx, int ph)
{
    if (ctx->timing) (void)hipEventRecord(ctx->ev[2 * ph + 1], ctx->st);
}

}  // namespace

// =========================================================== context
extern "C" int sph_version(void) { return SPH_VERSION; }

extern "C" int sph_ctx_create(int device, sph_ctx **out)
{
    if (!out) return SPH_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SPH_ERR_HIP;
    if (device < 0 || device >= ndev) return SPH_ERR_ARG;
    sph_ctx *ctx = new sph_ctx();
    ctx->device = device;
    std::memset(&ctx->stats, 0, sizeof(ctx->stats));
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return SPH_ERR_HIP;
    }
    for (int k = 0; k < 2 * SPH_PH_COUNT; k++)
        if (hipEventCreate(&ctx->ev[k]) != hipSuccess) {
            delete ctx;
            return SPH_ERR_HIP;
        }
    ctx->scratch = resolve_scratch(device);  // the device is current here
    if (hipHostMalloc((void **)&ctx->hpin, 128 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_h, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_n, hipEventDisableTiming) != hipSuccess) {
        delete ctx;
        return SPH_ERR_HIP;
    }
    *out = ctx;
    return SPH_OK;
}

extern "C" int sph_ctx_destroy(sph_ctx *ctx)
{
    if (!ctx) return SPH_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->st);
    Buf *all[] = {&ctx->x, &ctx->y, &ctx->z, &ctx->vx, &ctx->vy, &ctx->vz, &ctx->m, &ctx->h, &ctx->eps,
                  &ctx->u, &ctx->duprev, &ctx->ax, &ctx->ay, &ctx->az, &ctx->gx, &ctx->gy, &ctx->gz,
                  &ctx->tmpd, &ctx->orig, &ctx->tmpi, &ctx->keys0, &ctx->keys1, &ctx->vals0, &ctx->vals1,
                  &ctx->hist, &ctx->bpart, &ctx->box, &ctx->nodes, &ctx->right, &ctx->parent, &ctx->ncount, &ctx->flags, &ctx->front, &ctx->groups,
                  &ctx->cnt, &ctx->rcnt0, &ctx->off, &ctx->scantmp, &ctx->idx, &ctx->rho, &ctx->P, &ctx->c, &ctx->prho2,
                  &ctx->divV, &ctx->curlV, &ctx->bal, &ctx->du, &ctx->cand, &ctx->hR, &ctx->hsteps,
                  &ctx->hstate, &ctx->hcount, &ctx->hplist, &ctx->hout, &ctx->hin, &ctx->tx, &ctx->ty, &ctx->tz,
                  &ctx->stage, &ctx->kpart, &ctx->kout, &ctx->spread, &ctx->ccidx, &ctx->cccnt,
                  &ctx->cext, &ctx->cnext, &ctx->chr, &ctx->ttab, &ctx->zmet, &ctx->j21, &ctx->ucols, &ctx->uyn,
                  &ctx->uL, &ctx->uG, &ctx->uflag, &ctx->ucnt, &ctx->hbig, &ctx->p4, &ctx->v4, &ctx->s4,
                  &ctx->colg, &ctx->coln, &ctx->tsdt, &ctx->tsdh, &ctx->phi, &ctx->tpart, &ctx->tout};
    for (Buf *b : all)
        if (b->p) (void)hipFree(b->p);
    for (Buf &b : ctx->shadow)
        if (b.p) (void)hipFree(b.p);
    Buf *cool[] = {&ctx->rf1,   &ctx->ctabL,  &ctx->ctabG,  &ctx->cin[0], &ctx->cin[1], &ctx->cin[2], &ctx->cin[3],
                   &ctx->cin[4], &ctx->cin[5], &ctx->cout[0], &ctx->cout[1], &ctx->mP,     &ctx->mG,     &ctx->mM,
                   &ctx->mV,     &ctx->mW[0],  &ctx->mW[1],  &ctx->mW[2],  &ctx->mW[3],  &ctx->mx,     &ctx->mxn,
                   &ctx->mt,     &ctx->my,     &ctx->mloss,  &ctx->memap,  &ctx->memap2, &ctx->memap3,
                   &ctx->mgate,  &ctx->mcnt,   &ctx->mfl};
    for (Buf *b : cool)
        if (b->p) (void)hipFree(b->p);
    for (int k = 0; k < 2 * SPH_PH_COUNT; k++) (void)hipEventDestroy(ctx->ev[k]);
    if (ctx->ev_h) (void)hipEventDestroy(ctx->ev_h);
    if (ctx->ev_n) (void)hipEventDestroy(ctx->ev_n);
    if (ctx->hpin) (void)hipHostFree(ctx->hpin);
    (void)hipStreamDestroy(ctx->st);
    delete ctx;
    return SPH_OK;
}

extern "C" const char *sph_last_error(const sph_ctx *ctx)
{
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

// =========================================================== h solves
static_assert(SPH_PH_COUNT <= 20, "sph_stats holds 20 phases");

namespace {

// Parameters of an h solve over m targets (scratch in the context); host counters go to
// unc (unconverged), rounds_out and work (distance tests, kernel terms).
int hsolve_params(sph_ctx *ctx, int64_t m, const double *tx, const double *ty, const double *tz, int mode,
                  double param, int n_low, int n_up, int max_steps, double tol, int64_t *unc, int64_t *rounds_out,
                  int64_t *work, bool keep_indices, sph::HSolveParams &p)
{
    CK(ensure(ctx, ctx->cand, (size_t)m * SPH_HCAP * 8));
    if (keep_indices) CK(ensure(ctx, ctx->ccidx, (size_t)m * SPH_HCAP * 4));
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
    p.cand = D(cand);
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
    p.cidx = keep_indices ? I32(ccidx) : nullptr;
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
               int64_t *work = nullptr, bool keep_indices = false)
{
    sph::HSolveParams p;
    int64_t unc = 0;
    CK(hsolve_params(ctx, m, tx, ty, tz, mode, param, n_low, n_up, max_steps, tol, &unc, rounds_out, work,
                     keep_indices, p));
    int64_t pending = 0;
    if (mode == SPH_H_REF_COUNT) HIPCK(sph::tree_escape(ctx->st, ctx->n, P_<WNode>(ctx->nodes)));  // stackless walk
    HIPCK(sph::hsolve(ctx->st, p, SPH_HROUNDS, &pending));
    if (pending) return fail(ctx, SPH_ERR_NOCONV, "smoothing-length solve: " + std::to_string(pending) +
                                                      " targets undecided after the maximum number of radius rounds");
    *n_unconv = unc;
    return SPH_OK;
}

}  // namespace

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

namespace {
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
}  // namespace

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
namespace {
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
}  // namespace

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

extern "C" int sph_kinetics(sph_ctx *ctx, int64_t n, const double *vel, const double *m, double *out5)
{
    CK(check_common(ctx, n, SPH_PREC_FP64));
    if (!vel || !m || !out5) return fail(ctx, SPH_ERR_ARG, "null argument");
    // order-independent: plain (unsorted) upload into scratch
    CK(ensure(ctx, ctx->stage, sizeof(double) * 3 * n));
    CK(ensure(ctx, ctx->tx, sizeof(double) * n));
    CK(ensure(ctx, ctx->ty, sizeof(double) * n));
    CK(ensure(ctx, ctx->tz, sizeof(double) * n));
    CK(ensure(ctx, ctx->tmpd, sizeof(double) * n));
    CK(ensure(ctx, ctx->kpart, 5 * 1024 * 8));
    CK(ensure(ctx, ctx->kout, 8 * 8));
    hipStream_t st = ctx->st;
    HIPCK(hipMemcpyAsync(ctx->stage.p, vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
    HIPCK(hipMemcpyAsync(ctx->tmpd.p, m, sizeof(double) * n, hipMemcpyHostToDevice, st));
    k_deinterleave<<<sph_blocks(n, 256), 256, 0, st>>>(n, D(stage), D(tx), D(ty), D(tz));
    HIPCK(sph::kinetics(st, n, D(tx), D(ty), D(tz), D(tmpd), D(kpart), D(kout)));
    HIPCK(hipMemcpyAsync(out5, ctx->kout.p, sizeof(double) * 5, hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    return SPH_OK;
}

// Allocation-free twin of sph_kinetics: the caller's arrays and the reduction's partials are
// staged in this device's static scratch array instead of context buffers, so the call makes
// no device allocation (for callers that cannot allocate).  n <= SPH_NOALLOC_MAX_N.  Calls
// that overlap on one GPU (two contexts, two threads) share that array and race.
extern "C" int sph_kinetics_noalloc(sph_ctx *ctx, int64_t n, const double *vel, const double *m, double *out5)
{
    CK(check_common(ctx, n, SPH_PREC_FP64));
    if (!vel || !m || !out5) return fail(ctx, SPH_ERR_ARG, "null argument");
    if (n > SPH_NOALLOC_MAX_N)
        return fail(ctx, SPH_ERR_ARG, "sph_kinetics_noalloc takes at most " + std::to_string(SPH_NOALLOC_MAX_N) +
                                          " particles (its static scratch); n = " + std::to_string(n));
    double *sp = ctx->scratch;
    if (!sp) {
        const hipError_t e = ctx->device < SPH_MAX_DEVICES ? g_scratch_err[ctx->device] : hipErrorInvalidDevice;
        return fail(ctx, SPH_ERR_STATE, std::string("no device address for the static scratch: ") +
                                            hipGetErrorString(e));
    }
    hipStream_t st = ctx->st;
    double *dv = sp, *dm = sp + 3 * n, *part = sp + SPH_NOALLOC_PART, *res = part + 5 * 1024;
    HIPCK(hipMemcpyAsync(dv, vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
    HIPCK(hipMemcpyAsync(dm, m, sizeof(double) * n, hipMemcpyHostToDevice, st));
    HIPCK(sph::kinetics(st, n, dv, dv + 1, dv + 2, dm, part, res, 3));
    HIPCK(hipMemcpyAsync(out5, res, sizeof(double) * 5, hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    return SPH_OK;
}

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

namespace {

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
    if (ctx->cool_on) {
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
                     &ctx->stats.h_rounds, ctx->hs_work, true, hp));
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
    CK(ensure(ctx, ctx->chr, (size_t)n * 16));
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
        st, n, SPH_HCAP, D(cand), I32(ccidx), I32(cccnt), D(h), D(hR), SPH_NGB_STRIDE, I32(cnt), I32(rcnt0), I32(idx),
        P_<int2>(ctx->cext), ext_cap, P_<unsigned long long>(ctx->cnext),
        P_<unsigned long long>(ctx->spread) + SPR_N * SPH_SPREAD_WORDS, P_<unsigned long long>(ctx->hcount) + 10,
        ctx->hpin + 8, ctx->ev_n, big, P_<double2>(ctx->chr)));
    ph_end(ctx, SPH_PH_NGB);
    ctx->ngb_stride = SPH_NGB_STRIDE;
    ctx->ngb_per = p->period;
    R *bal = p->balsara ? P_<R>(ctx->bal) : nullptr;
    const bool g = p->G != 0.0;
    CK(ensure(ctx, ctx->s4, (size_t)n * 4 * sizeof(R)));
    // density + EOS, fused with div/curl v and the Balsara factor when the switch is on; then
    // the acceleration (a = a_grav + a_sph)
    auto hydro_forces = [&]() -> int {
        NgbRows rows;
        CK(ngb_rows(ctx, rows));
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
    CK(ngb_rows(ctx, rows, true));  // packs the kicked velocities (positions and h as in the force passes)
    HIPCK(sph::dudt<R>(st, n, P_<R4<R>>(ctx->s4), bal != nullptr, p->alpha, rows, P_<R>(ctx->du)));
    ph_end(ctx, SPH_PH_DUDT);
    ph_begin(ctx, SPH_PH_UUPDATE);
    HIPCK(sph::u_update<R>(st, n, D(u), P_<R>(ctx->du), D(duprev), p->dt, first));
    ph_end(ctx, SPH_PH_UUPDATE);
    // cooling after the u update of each step (CR/v3t.cu:1527-1536; none after the first update)
    ctx->cool_ran = ctx->cool_on && !first;
    if (ctx->cool_ran) {
        ph_begin(ctx, SPH_PH_COOLING);
        CK(cool_run<R>(ctx, n, P_<R>(ctx->rho), D(u), D(zmet), D(j21), p->dt, ctx->cool_p, D(u),
                       P_<unsigned long long>(ctx->hcount) + 12));
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
                        I32(orig), ctx->cool_on ? D(zmet) : nullptr, ctx->cool_on ? D(j21) : nullptr};
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

}  // namespace

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

// =========================================================== cooling
namespace {

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

}  // namespace

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

namespace {
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
}  // namespace

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
namespace {

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
    return d;
}

int mlp_alloc(sph_ctx *ctx)
{
    if (!ctx) return SPH_ERR_ARG;
    HIPCK(hipSetDevice(ctx->device));
    Buf *ps[] = {&ctx->mP, &ctx->mG, &ctx->mM, &ctx->mV};
    for (Buf *b : ps) CK(ensure(ctx, *b, (size_t)MLP_NPARAM * 4));
    for (int k = 0; k < 4; k++) CK(ensure(ctx, ctx->mW[k], (size_t)MLP_HIDDEN * MLP_HIDDEN * 2));
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

}  // namespace

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

namespace {
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
}  // namespace

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
namespace {

int cool_check(sph_ctx *ctx, const sph_cool_params *cp)
{
    if (!cp) return fail(ctx, SPH_ERR_ARG, "null cooling parameters");
    if (cp->mode != SPH_COOL_EXACT && cp->mode != SPH_COOL_TABLE && cp->mode != SPH_COOL_MLP)
        return fail(ctx, SPH_ERR_ARG, "mode must be SPH_COOL_EXACT, SPH_COOL_TABLE or SPH_COOL_MLP");
    if (!(cp->unit_density_cgs > 0.0) || !(cp->unit_u_cgs > 0.0) || !(cp->unit_time_s > 0.0) ||
        !(cp->XH > 0.0 && cp->XH <= 1.0) || !(cp->u_floor_cgs >= 0.0))
        return fail(ctx, SPH_ERR_ARG, "cooling units must be > 0, 0 < XH <= 1 and u_floor_cgs >= 0");
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

}  // namespace

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

namespace {
int column_check(sph_ctx *ctx, const sph_column_params *cp)
{
    if (!(cp->unit_density_cgs > 0.0) || !(cp->unit_length_cm > 0.0) || !(cp->XH > 0.0) || !(cp->n_ngb > 0.0))
        return fail(ctx, SPH_ERR_ARG, "column density: unit_density_cgs, unit_length_cm, XH, n_ngb must be > 0");
    return SPH_OK;
}
sph::ColumnParams column_params(const sph_column_params *cp)
{
    return sph::ColumnParams{cp->unit_density_cgs, cp->unit_length_cm, cp->XH, cp->n_ngb};
}
}  // namespace

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

extern "C" int sph_state_set_cooling(sph_ctx *ctx, const double *Z, const double *J21, const sph_cool_params *cp)
{
    if (!ctx) return SPH_ERR_ARG;
    if (!ctx->resident) return fail(ctx, SPH_ERR_STATE, "no resident state (call sph_state_set first)");
    HIPCK(hipSetDevice(ctx->device));
    if (!cp) {
        ctx->cool_on = false;
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
