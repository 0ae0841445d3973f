// sph_capi.hip -- the C ABI (include/sph_amd.h): the context and the helpers its entry points
// share (device caches of the particle state in Morton order, sorting, the tree, neighbour
// sets), and the kinetic reductions including the allocation-free one.
#include "sph_ctx.h"

using namespace sphc;

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

namespace sphc {

int fail(sph_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}


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


int capacity(sph_ctx *ctx, int64_t n)
{
    n = std::max(n, ctx->n_reserve);  // room for injected particles: no state buffer ever moves
    size_t nd = (size_t)n * 8, ni = (size_t)n * 4;
    Buf *dbl[] = {&ctx->x, &ctx->y, &ctx->z, &ctx->vx, &ctx->vy, &ctx->vz, &ctx->m, &ctx->h, &ctx->eps,
                  &ctx->u, &ctx->duprev, &ctx->ax, &ctx->ay, &ctx->az, &ctx->gx, &ctx->gy, &ctx->gz,
                  &ctx->tmpd, &ctx->rho, &ctx->P, &ctx->c, &ctx->prho2, &ctx->divV, &ctx->curlV,
                  &ctx->bal, &ctx->du, &ctx->cand, &ctx->hR, &ctx->hout, &ctx->hin, &ctx->tx, &ctx->ty, &ctx->tz};
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


int put_vec3(sph_ctx *ctx, const double *src, Buf &a, Buf &b, Buf &c)
{
    int64_t n = ctx->n;
    HIPCK(hipMemcpyAsync(ctx->stage.p, src, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->st));
    Buf *t[3] = {&a, &b, &c};
    for (int d = 0; d < 3; d++) HIPCK(sph::gather<double, double>(ctx->st, n, I32(orig), D(stage), 3, d, (double *)t[d]->p));
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
int build_ngb(sph_ctx *ctx, bool csr, const double *hscale, double per)
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
// compact: also the mixed form's compact records, fixed point over the periodic box or (open
// boundaries) over the Morton bounding box of the current order (the positions must be those
// the order was built from); full = false leaves
// the FP64 records out (null in r) when no pass of the caller reads them.
int ngb_rows(sph_ctx *ctx, NgbRows &r, bool vel_only, bool compact, bool full)
{
    const int64_t n = ctx->n;
    if (full) {
        CK(ensure(ctx, ctx->p4, (size_t)n * 32));
        CK(ensure(ctx, ctx->v4, (size_t)n * 32));
        if (!vel_only) HIPCK(sph::pack4(ctx->st, n, D(x), D(y), D(z), D(h), P_<double4>(ctx->p4)));
        HIPCK(sph::pack4(ctx->st, n, D(vx), D(vy), D(vz), D(m), P_<double4>(ctx->v4)));
    }
    r = NgbRows{P_<int64_t>(ctx->off), I32(cnt), I32(idx), ctx->ngb_stride, ctx->ngb_per,
                full ? P_<double4>(ctx->p4) : nullptr, full ? P_<double4>(ctx->v4) : nullptr};
    if (compact) {
        CK(ensure(ctx, ctx->pf, (size_t)n * 16));
        CK(ensure(ctx, ctx->vf, (size_t)n * 16));
        HIPCK(sph::pack_compact(ctx->st, n, D(x), D(y), D(z), D(h), D(vx), D(vy), D(vz), D(m), ctx->ngb_per, D(box),
                                P_<uint4>(ctx->pf), P_<float4>(ctx->vf), vel_only));
        r.pf = P_<uint4>(ctx->pf);
        r.vf = P_<float4>(ctx->vf);
        r.grid = D(box);
    }
    return SPH_OK;
}

// csr: the caller needs true CSR offsets (sph_neighbours).
int ensure_csr(sph_ctx *ctx, bool csr)
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

}  // namespace sphc

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
    size_t mem_free = 0, mem_total = 0;
    if (hipMemGetInfo(&mem_free, &mem_total) == hipSuccess) ctx->r2_limit = mem_total / 3;
    if (hipHostMalloc((void **)&ctx->hpin, 128 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_h, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_n, hipEventDisableTiming) != hipSuccess) {
        delete ctx;
        return SPH_ERR_HIP;
    }
    *out = ctx;
    return SPH_OK;
}

namespace {
// Every device buffer of a context (the one list freed by sph_ctx_destroy and summed by
// sph_device_bytes).
std::vector<Buf *> all_bufs(sph_ctx *ctx)
{
    std::vector<Buf *> v = {
        &ctx->x, &ctx->y, &ctx->z, &ctx->vx, &ctx->vy, &ctx->vz, &ctx->m, &ctx->h, &ctx->eps,
        &ctx->u, &ctx->duprev, &ctx->ax, &ctx->ay, &ctx->az, &ctx->gx, &ctx->gy, &ctx->gz,
        &ctx->tmpd, &ctx->orig, &ctx->tmpi, &ctx->keys0, &ctx->keys1, &ctx->vals0, &ctx->vals1,
        &ctx->hist, &ctx->bpart, &ctx->box, &ctx->nodes, &ctx->right, &ctx->parent, &ctx->ncount, &ctx->flags,
        &ctx->front, &ctx->groups, &ctx->cnt, &ctx->rcnt0, &ctx->off, &ctx->scantmp, &ctx->idx, &ctx->rho,
        &ctx->P, &ctx->c, &ctx->prho2, &ctx->divV, &ctx->curlV, &ctx->bal, &ctx->du, &ctx->cand, &ctx->hR,
        &ctx->hsteps, &ctx->hstate, &ctx->hcount, &ctx->hplist, &ctx->hout, &ctx->hin, &ctx->tx, &ctx->ty,
        &ctx->tz, &ctx->stage, &ctx->kpart, &ctx->kout, &ctx->spread, &ctx->ccidx, &ctx->cccnt, &ctx->cext,
        &ctx->cnext, &ctx->chr, &ctx->ttab, &ctx->zmet, &ctx->j21, &ctx->ucols, &ctx->uyn, &ctx->uL, &ctx->uG,
        &ctx->uflag, &ctx->ucnt, &ctx->hbig, &ctx->p4, &ctx->v4, &ctx->s4, &ctx->pf, &ctx->vf, &ctx->colg,
        &ctx->coln, &ctx->tsdt, &ctx->tsdh, &ctx->phi, &ctx->tpart, &ctx->tout,
        // cooling and surrogate
        &ctx->rf1, &ctx->ctabL, &ctx->ctabG, &ctx->cin[0], &ctx->cin[1], &ctx->cin[2], &ctx->cin[3],
        &ctx->cin[4], &ctx->cin[5], &ctx->cout[0], &ctx->cout[1], &ctx->mP, &ctx->mG, &ctx->mM, &ctx->mV,
        &ctx->mW[0], &ctx->mW[1], &ctx->mW[2], &ctx->mW[3], &ctx->mimg, &ctx->mx, &ctx->mxn, &ctx->mt, &ctx->my,
        &ctx->mloss, &ctx->memap, &ctx->memap2, &ctx->memap3, &ctx->mgate, &ctx->mcnt, &ctx->mfl, &ctx->nhtab};
    for (Buf &b : ctx->shadow) v.push_back(&b);
    return v;
}
}  // namespace

extern "C" int sph_ctx_destroy(sph_ctx *ctx)
{
    if (!ctx) return SPH_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->st);
    for (Buf *b : all_bufs(ctx))
        if (b->p) (void)hipFree(b->p);
    for (int k = 0; k < 2 * SPH_PH_COUNT; k++) (void)hipEventDestroy(ctx->ev[k]);
    if (ctx->ev_h) (void)hipEventDestroy(ctx->ev_h);
    if (ctx->ev_n) (void)hipEventDestroy(ctx->ev_n);
    if (ctx->hpin) (void)hipHostFree(ctx->hpin);
    (void)hipStreamDestroy(ctx->st);
    delete ctx;
    return SPH_OK;
}

extern "C" int sph_set_distance_rows_limit(sph_ctx *ctx, uint64_t bytes)
{
    if (!ctx) return SPH_ERR_ARG;
    ctx->r2_limit = bytes;
    return SPH_OK;
}

extern "C" int sph_device_bytes(const sph_ctx *ctx, uint64_t *bytes)
{
    if (!ctx || !bytes) return SPH_ERR_ARG;
    uint64_t total = 0;
    for (Buf *b : all_bufs(const_cast<sph_ctx *>(ctx))) total += b->p ? (uint64_t)b->bytes : 0;
    *bytes = total;
    return SPH_OK;
}

extern "C" const char *sph_last_error(const sph_ctx *ctx)
{
    if (!ctx) return "null context";
    return ctx->err.c_str();
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

