// sph_ngb.hip -- tree neighbour search: smoothing-length solves and compact
// CSR neighbour lists.
//
// Replaces the reference's grid search with fixed 200-slot lists
// (EV/ngb.h:125-264 ngbDB_new_v3), its count-based h update
// (EV/lib.h:485-579 smoothing_h_ngb) and the CPU h searches
// (sph_code/libsx.py:752-900 do_smoothingX, h_smooth_fast).
//
// All walks are wavefront-cooperative and stackless: one lane per target,
// targets in Morton order, the wave follows one pre-order path through the
// tree (a node is opened when any active lane needs it), and each lane keeps
// a single "resume" node id at which it rejoins the walk after a subtree it
// did not need.  Node records are read with wave-uniform (scalar) loads.
//
// FP64 throughout, contraction off: membership r^2 < (h_i+h_j)^2 is decided
// exactly as the CPU oracle decides it.
#include "sph_ngb.h"

#pragma clang fp contract(off)

__device__ __forceinline__ double box_dist2(double x, double y, double z, const WNode *__restrict__ nd)
{
    double ex = 0.0, ey = 0.0, ez = 0.0;
    if (nd->lx - x > 0.0) ex = nd->lx - x;
    if (x - nd->hx > 0.0) ex = x - nd->hx;
    if (nd->ly - y > 0.0) ey = nd->ly - y;
    if (y - nd->hy > 0.0) ey = y - nd->hy;
    if (nd->lz - z > 0.0) ez = nd->lz - z;
    if (z - nd->hz > 0.0) ez = z - nd->hz;
    return ex * ex + ey * ey + ez * ez;
}

// Generic cooperative radius walk.  prune(nd) -> true skips the node for this
// lane; leaf(j) is called for every particle j of an opened bucket.
template <class Prune, class Leaf>
__device__ __forceinline__ void coop_walk(const WNode *__restrict__ nodes, bool valid, int bucket, Prune prune,
                                          Leaf leaf)
{
    int node = 0;
    int resume = valid ? 0 : -2;
    while (node >= 0) {
        const WNode *nd = nodes + node;
        // link fields as wave-uniform scalars (scalar loads, scalar select)
        const int first = wave_uniform(nd->first), cnt = wave_uniform(nd->count);
        const int left = wave_uniform(nd->left), esc = wave_uniform(nd->escape);
        bool open = false;
        if (resume == node) {
            if (prune(nd)) resume = esc;
            else open = true;
        }
        const bool anyopen = wave_ballot(open) != 0;
        int next = esc;
        if (anyopen && cnt <= bucket) {
            for (int j = first; j < first + cnt; j++)
                if (open) leaf(j);
            if (open) resume = esc;
        } else if (anyopen) {
            if (open) resume = left;
            next = left;
        }
        node = wave_uniform(next);
    }
}

// ---------------------------------------------------------------- h solves
// mode 0: Newton-Raphson root of (32/3) sum_j w(r_j/h) - n_ngb (FP64),
// mode 1: the reference's +-0.3 % count stepping (k_hstep below),
// mode 2: h = 0.5 * (k+1)-th smallest distance (libsx.py:752-774, k = 50).
// Candidates (r^2 < R_t^2) go to cand[c*M + t], c < cap.  A target whose
// candidate set cannot decide its answer grows or shrinks R_t and stays
// pending (state[t] = 0); the host relaunches until every target is done.
struct HSolveArgs {
    int64_t m;                 // number of targets
    const double *tx, *ty, *tz;  // target positions
    const double *x, *y, *z;   // sorted source positions
    const WNode *nodes;
    int bucket;
    int mode;
    double n_ngb, tol;         // mode 0
    int n_low, n_up, max_steps;  // mode 1
    int kth;                   // mode 2
    int cap;
    double *cand;              // cap * m doubles
    double *R;                 // per target search radius (in/out)
    double *h;                 // per target h (in: start, out: result)
    int32_t *steps;            // mode 1 step counter (persists over rounds)
    int32_t *state;            // 0 pending, 1 done, 2 done-unconverged
    unsigned long long *pending;  // [0] pending count, [1] unconverged count
};

__device__ __forceinline__ void nr_eval(const double *__restrict__ cand, int64_t m, int64_t t, int c, double h,
                                        double n_ngb, double &f, double &df)
{
    double s = 0.0, ds = 0.0, ih = 1.0 / h;
    for (int k = 0; k < c; k++) {
        double q = __builtin_sqrt(cand[(int64_t)k * m + t]) * ih;
        if (q < 1.0) {
            s += 1.0 - 1.5 * q * q + 0.75 * q * q * q;
            ds += (-3.0 * q + 2.25 * q * q) * (-q * ih);
        } else if (q < 2.0) {
            double u = 2.0 - q;
            s += 0.25 * u * u * u;
            ds += (-0.75 * u * u) * (-q * ih);
        }
    }
    f = (32.0 / 3.0) * s - n_ngb;
    df = (32.0 / 3.0) * ds;
}

__device__ __forceinline__ int count_below(const double *__restrict__ cand, int64_t m, int64_t t, int c, double lim)
{
    int n = 0;
    for (int k = 0; k < c; k++) n += __builtin_sqrt(cand[(int64_t)k * m + t]) < lim;
    return n;
}

__global__ void __launch_bounds__(256) k_hsolve(HSolveArgs a)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool valid = t < a.m && a.state[t] == 0;
    double xi = 0, yi = 0, zi = 0, R = 0;
    if (valid) {
        xi = a.tx[t]; yi = a.ty[t]; zi = a.tz[t];
        R = a.R[t];
    }
    const double R2 = R * R;
    int c = 0;
    const int64_t m = a.m;
    double *cand = a.cand;
    const int cap = a.cap;
    coop_walk(
        a.nodes, valid, a.bucket, [&](const WNode *nd) { return box_dist2(xi, yi, zi, nd) >= R2; },
        [&](int j) {
            double dx = a.x[j] - xi, dy = a.y[j] - yi, dz = a.z[j] - zi;
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < R2) {
                if (c < cap) cand[(int64_t)c * m + t] = r2;
                c++;
            }
        });
    if (!valid) return;
    const bool over = c > cap;
    const int cc = over ? cap : c;
    int st = 0;  // stays pending unless decided below
    double hnew = a.h[t], Rnew = R;
    if (a.mode == 0) {
        if (over) {
            Rnew = 0.75 * R;
        } else {
            double hmaxv = 0.5 * R, f, df;
            nr_eval(cand, m, t, cc, hmaxv, a.n_ngb, f, df);
            if (f < 0.0) {
                Rnew = 1.5 * R;
            } else {
                double lo = 0.0, hi = hmaxv, hh = hnew;
                if (!(hh > 0.0 && hh < hi)) hh = 0.5 * hi;
                st = 2;
                for (int it = 0; it < 100; it++) {
                    nr_eval(cand, m, t, cc, hh, a.n_ngb, f, df);
                    if (f < 0.0) lo = hh; else hi = hh;
                    double hn = (df > 0.0) ? hh - f / df : -1.0;
                    if (!(hn > lo && hn < hi)) hn = 0.5 * (lo + hi);
                    double dh = fabs(hn - hh);
                    hh = hn;
                    if (dh <= a.tol * hh) { st = 1; break; }
                }
                hnew = hh;
            }
        }
    } else {
        const int need = a.kth + 1;
        if (over) Rnew = 0.75 * R;
        else if (c < need) Rnew = 1.5 * R;
        else {
            // (k+1)-th smallest r^2 by ascending distinct values with multiplicity
            double cur = -1.0;
            int taken = 0;
            double v = 0.0;
            for (;;) {
                double nxt = INFINITY;
                int mult = 0;
                for (int k = 0; k < cc; k++) {
                    double r2 = cand[(int64_t)k * m + t];
                    if (r2 > cur) {
                        if (r2 < nxt) { nxt = r2; mult = 1; }
                        else if (r2 == nxt) mult++;
                    }
                }
                if (taken + mult >= need) { v = nxt; break; }
                taken += mult;
                cur = nxt;
            }
            hnew = __builtin_sqrt(v) * 0.5;
            st = 1;
        }
    }
    a.h[t] = hnew;
    a.R[t] = Rnew;
    a.state[t] = st;
    if (st == 0) atomicAdd(&a.pending[0], 1ull);
    if (st == 2) atomicAdd(&a.pending[1], 1ull);
}

// mode 1: the reference's count stepping, exactly (libsx.py:804-845): each
// step counts #(dist < 2h) over ALL sources with a cooperative count walk, so
// no candidate radius is involved.  dist = sqrt(r^2) correctly rounded, as in
// the reference; a node is skipped only when its box distance^2 exceeds
// (2h)^2, which cannot drop a particle with sqrt(r^2) < 2h.
__global__ void __launch_bounds__(256) k_hstep(HSolveArgs a)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool valid = t < a.m;
    double xi = 0, yi = 0, zi = 0, hh = 0;
    if (valid) { xi = a.tx[t]; yi = a.ty[t]; zi = a.tz[t]; hh = a.h[t]; }
    int steps = 0, st = valid ? 0 : 1;
    for (;;) {
        bool need = st == 0;
        if (wave_ballot(need) == 0) break;
        const double lim = 2.0 * hh, L2 = lim * lim;
        int n = 0;
        coop_walk(
            a.nodes, need, a.bucket, [&](const WNode *nd) { return box_dist2(xi, yi, zi, nd) > L2; },
            [&](int j) {
                double dx = a.x[j] - xi, dy = a.y[j] - yi, dz = a.z[j] - zi;
                n += __builtin_sqrt(dx * dx + dy * dy + dz * dz) < lim;
            });
        if (need) {
            if (n <= a.n_up && n >= a.n_low) st = 1;
            else if (steps >= a.max_steps) st = 2;
            else {
                if (n > a.n_up) hh -= 0.003 * hh;
                if (n < a.n_low) hh += 0.003 * hh;
                steps++;
            }
        }
    }
    if (!valid) return;
    a.h[t] = hh;
    a.steps[t] = steps;
    a.state[t] = st;
    if (st == 2) atomicAdd(&a.pending[1], 1ull);
}

__global__ void __launch_bounds__(256) k_hsolve_init(int64_t m, const double *__restrict__ h0, double fac,
                                                     double *__restrict__ R, double *__restrict__ h,
                                                     int32_t *__restrict__ steps, int32_t *__restrict__ state)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= m) return;
    double hv = h0[t];
    h[t] = hv;
    R[t] = 2.0 * hv * fac;
    steps[t] = 0;
    state[t] = 0;
}

// ---------------------------------------------------------------- CSR lists
// S_i = { j : r_ij^2 < (h_i + h_j)^2 }, self included, rows ascending in
// sorted index (the pre-order walk visits buckets in index order).
template <bool FILL>
__global__ void __launch_bounds__(256) k_ngb(int64_t n, const double *__restrict__ x, const double *__restrict__ y,
                                             const double *__restrict__ z, const double *__restrict__ h,
                                             const WNode *__restrict__ nodes, int bucket,
                                             int32_t *__restrict__ cnt, const int64_t *__restrict__ off,
                                             int32_t *__restrict__ idx)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool valid = i < n;
    double xi = 0, yi = 0, zi = 0, hi = 0;
    if (valid) { xi = x[i]; yi = y[i]; zi = z[i]; hi = h[i]; }
    int64_t k = (FILL && valid) ? off[i] : 0;
    int c = 0;
    coop_walk(
        nodes, valid, bucket,
        [&](const WNode *nd) {
            double t = hi + nd->hmax;
            return box_dist2(xi, yi, zi, nd) >= t * t;
        },
        [&](int j) {
            double dx = xi - x[j], dy = yi - y[j], dz = zi - z[j];
            double r2 = dx * dx + dy * dy + dz * dz, t = hi + h[j];
            if (r2 < t * t) {
                if (FILL) idx[k + c] = j;
                c++;
            }
        });
    if (valid && !FILL) cnt[i] = c;
}

namespace sph {

hipError_t hsolve(hipStream_t st, const HSolveParams &p, int max_rounds, int64_t *n_unconverged)
{
    HSolveArgs a;
    a.m = p.m; a.tx = p.tx; a.ty = p.ty; a.tz = p.tz;
    a.x = p.x; a.y = p.y; a.z = p.z; a.nodes = p.nodes; a.bucket = p.bucket;
    a.mode = p.mode; a.n_ngb = p.n_ngb; a.tol = p.tol;
    a.n_low = p.n_low; a.n_up = p.n_up; a.max_steps = p.max_steps; a.kth = p.kth;
    a.cap = p.cap; a.cand = p.cand; a.R = p.R; a.h = p.h_out; a.steps = p.steps; a.state = p.state;
    a.pending = p.counters;
    if (p.m <= 0) { *n_unconverged = 0; return hipSuccess; }
    k_hsolve_init<<<sph_blocks(p.m, 256), 256, 0, st>>>(p.m, p.h_in, p.r_fac, p.R, p.h_out, p.steps, p.state);
    unsigned long long hc[2] = {0, 0};
    if (p.mode == 1) {
        fill<unsigned long long>(st, 2, p.counters, 0ull);
        k_hstep<<<sph_blocks(p.m, 256), 256, 0, st>>>(a);
        hipError_t e = hipMemcpyAsync(hc, p.counters, sizeof(hc), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        if (p.unconverged_total) *p.unconverged_total += (int64_t)hc[1];
        *n_unconverged = 0;
        return hipGetLastError();
    }
    for (int r = 0; r < max_rounds; r++) {
        fill<unsigned long long>(st, 2, p.counters, 0ull);
        k_hsolve<<<sph_blocks(p.m, 256), 256, 0, st>>>(a);
        hipError_t e = hipMemcpyAsync(hc, p.counters, sizeof(hc), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        if (p.unconverged_total) *p.unconverged_total += (int64_t)hc[1];
        if (hc[0] == 0) { *n_unconverged = 0; return hipGetLastError(); }
    }
    *n_unconverged = (int64_t)hc[0];
    return hipGetLastError();
}

hipError_t neighbours(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                      const double *h, const WNode *nodes, int bucket, int32_t *cnt, int64_t *off,
                      int64_t *scan_tmp, int32_t *idx, int64_t idx_cap, int64_t *total)
{
    k_ngb<false><<<sph_blocks(n, 256), 256, 0, st>>>(n, x, y, z, h, nodes, bucket, cnt, nullptr, nullptr);
    hipError_t e = scan_counts(st, n, cnt, off, scan_tmp);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(total, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    if (*total > idx_cap || idx == nullptr) return hipSuccess;  // caller grows idx and calls neighbours_fill
    k_ngb<true><<<sph_blocks(n, 256), 256, 0, st>>>(n, x, y, z, h, nodes, bucket, nullptr, off, idx);
    return hipGetLastError();
}

hipError_t neighbours_fill(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                           const double *h, const WNode *nodes, int bucket, const int64_t *off, int32_t *idx)
{
    k_ngb<true><<<sph_blocks(n, 256), 256, 0, st>>>(n, x, y, z, h, nodes, bucket, nullptr, off, idx);
    return hipGetLastError();
}

}  // namespace sph
