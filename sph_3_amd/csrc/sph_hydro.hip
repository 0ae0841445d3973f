// sph_hydro.hip -- SPH sums over compact CSR neighbour lists: density with
// the fused equation of state, div/curl v with the Balsara factor, pressure
// plus artificial-viscosity acceleration, and du/dt.
//
// Reference formulas: sph_code/libsx.py:10-44 (density), :91-96 (pressure),
// :100-177 and :353-421 (acceleration, du/dt), sph_code/shear.py:6-69,
// :149-235, :332-406 (div/curl, Balsara-switched forms); GPU twins
// EV/lib.h:598-651, 713-736, 741-824, 829-925, 949-1040.
//
// Each particle is handled by a group of G consecutive lanes that stride
// through its CSR row (coalesced index reads), accumulate with Kahan
// compensation in the pass type R and combine with a TwoSum lane reduction.
// Pair arithmetic is in R; coordinate and velocity differences are taken in
// FP64 and rounded once to R.
#include "sph_hydro.h"
#include <algorithm>
#include <type_traits>

#define HG 16  // lanes per particle
// The fused density + div/curl pass and the acceleration pass run 8 lanes per particle:
// their longer pair bodies amortise the shorter lane reductions better than 16 lanes
// (measured at 4,194,304 particles: 1.25 vs 1.35 ms and 1.51 vs 1.54 ms); du/dt keeps 16.
#define HG_DENSITY 8
#define HG_ACC 8

// A particle's pair inputs as a pass reads them (its own and each neighbour's): the FP64
// records p4/v4 (CMP = false), or the mixed form's compact records pf/vf (CMP = true).
template <bool CMP> struct Nb;
template <> struct Nb<false> {
    double4 p, v;
};
template <> struct Nb<true> {
    uint4 p;  // 32-bit fixed-point x, y, z and the bits of h (FP32)
    float4 v;
};

template <bool CMP>
__device__ __forceinline__ Nb<CMP> nb_load(const NgbRows &r, int64_t j)
{
    if constexpr (CMP) return Nb<true>{r.pf[j], r.vf[j]};
    else return Nb<false>{r.p4[j], r.v4[j]};
}

// Difference of two fixed-point coordinates in steps, as FP32 (one rounding).  Periodic box
// (the period spans 2^32 steps): the wrap-around difference read as signed, i.e. the minimum
// image.  Open boundaries: the exact difference of the unsigned values whatever its size.
template <bool PER>
__device__ __forceinline__ float fx_diff(uint32_t a, uint32_t b)
{
    if constexpr (PER) return (float)(int32_t)(a - b);
    else return a >= b ? (float)(a - b) : -(float)(b - a);
}

// r_i - r_j in R and hbar_ij = (h_i + h_j) / 2.  FP64 records: the FP64 difference (minimum
// image when PER) rounded once.  Compact records: the difference of the fixed-point
// coordinates in steps (exact integer, one rounding to FP32) times the step size (a power of
// two for open boundaries: exact).  Exactly antisymmetric in i and j; the positions
// themselves carry the fixed point's rounding (half a step).
template <typename R, bool PER, bool CMP>
__device__ __forceinline__ void nb_geom(const NgbRows &r, float sc, const Nb<CMP> &a, const Nb<CMP> &b, R &dx,
                                        R &dy, R &dz, R &hij)
{
    if constexpr (CMP) {
        dx = (R)(fx_diff<PER>(a.p.x, b.p.x) * sc);
        dy = (R)(fx_diff<PER>(a.p.y, b.p.y) * sc);
        dz = (R)(fx_diff<PER>(a.p.z, b.p.z) * sc);
        hij = (R)(0.5f * (__uint_as_float(a.p.w) + __uint_as_float(b.p.w)));
    } else {
        dx = (R)wrap_d<PER>(a.p.x - b.p.x, r.L);
        dy = (R)wrap_d<PER>(a.p.y - b.p.y, r.L);
        dz = (R)wrap_d<PER>(a.p.z - b.p.z, r.L);
        hij = (R)(0.5 * (a.p.w + b.p.w));
    }
}

// v_i - v_j and m_j in R (FP64 records: the FP64 difference rounded once; compact: FP32).
template <typename R, bool CMP>
__device__ __forceinline__ void nb_vel(const Nb<CMP> &a, const Nb<CMP> &b, R &ux, R &uy, R &uz, R &mj)
{
    ux = (R)(a.v.x - b.v.x);  // FP64 or FP32 difference as the records hold them
    uy = (R)(a.v.y - b.v.y);
    uz = (R)(a.v.z - b.v.z);
    mj = (R)b.v.w;
}

// h of a particle's own record.
template <bool CMP>
__device__ __forceinline__ double nb_h(const Nb<CMP> &a)
{
    if constexpr (CMP) return (double)__uint_as_float(a.p.w);
    else return a.p.w;
}

// Step size of the compact records' fixed point (0 when they are not used).
template <bool CMP>
__device__ __forceinline__ float nb_scale(const NgbRows &r)
{
    if constexpr (CMP) return (float)r.grid[7];
    else return 0.0f;
}

// Lane l of a G-lane group visits the row entries k0 + l, k0 + l + G, ... < e in this
// order and calls body(nb_j, x[j]) for each (x may be a dummy when X_ is false).  The next
// entry's records, and the index after it, are requested before the current entry's body
// runs, so a lane keeps one neighbour gather in flight behind its arithmetic instead of
// waiting for the index and then the records on every entry.
template <int G, bool CMP, bool X_, typename X, typename Body>
__device__ __forceinline__ void pair_loop(const NgbRows &rows, int64_t k0, int64_t e, int l,
                                          const X *__restrict__ xr, Body body)
{
    int64_t k = k0 + l;
    if (k >= e) return;
    int j = rows.idx[k];
    int jn = (k + G < e) ? rows.idx[k + G] : j;
    Nb<CMP> bj = nb_load<CMP>(rows, j);
    X xj{};
    if constexpr (X_) xj = xr[j];
    for (;;) {
        const int64_t kn = k + G;
        const bool more = kn < e;
        Nb<CMP> bn = bj;
        X xn = xj;
        int jnn = jn;
        if (more) {
            bn = nb_load<CMP>(rows, jn);
            if constexpr (X_) xn = xr[jn];
            if (kn + G < e) jnn = rows.idx[kn + G];
        }
        body(bj, xj);
        if (!more) break;
        k = kn;
        jn = jnn;
        bj = bn;
        xj = xn;
    }
}

template <typename R, int G, bool PER, bool CMP>
__global__ void __launch_bounds__(256) k_density(int64_t n, const double *__restrict__ u, double gamma,
                                                 NgbRows rows,
                                                 R *__restrict__ rho, R *__restrict__ P, R *__restrict__ c,
                                                 R *__restrict__ prho2, R4<R> *__restrict__ s4)
{
    int64_t gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    int64_t i = gid / G;
    int l = (int)(gid % G);
    bool valid = i < n;
    Kahan<R> acc;
    if (valid) {
        const float cs = nb_scale<CMP>(rows);
        const Nb<CMP> bi = nb_load<CMP>(rows, i);
        int64_t k0, e;
        row_range(rows, i, k0, e);
        pair_loop<G, CMP, false, int>(rows, k0, e, l, nullptr, [&](const Nb<CMP> &bj, int) {
            R dx, dy, dz, hij, ux, uy, uz, mj;
            nb_geom<R, PER, CMP>(rows, cs, bi, bj, dx, dy, dz, hij);
            nb_vel<R, CMP>(bi, bj, ux, uy, uz, mj);
            R rr = sph_sqrt<R>(dx * dx + dy * dy + dz * dz);
            acc.add(mj * spline_W<R>(rr, hij));
        });
    }
    R s = group_sum<R, G>(acc.s, -acc.c);
    if (valid && l == 0) {
        rho[i] = s;
        if (u) {
            R ui = (R)u[i];
            R Pi = (R)(gamma - 1.0) * s * ui;
            const R ci = sph_sqrt<R>((R)(gamma * (gamma - 1.0)) * ui), pr = Pi / (s * s);
            P[i] = Pi;
            c[i] = ci;
            prho2[i] = pr;
            if (s4) s4[i] = R4<R>{s, pr, ci, (R)1};
        }
    }
}

template <typename R, int G, bool PER>
__global__ void __launch_bounds__(256) k_divcurl(int64_t n, const R *__restrict__ rho,
                                                 const R *__restrict__ c, NgbRows rows, R *__restrict__ divV,
                                                 R *__restrict__ curlV, R *__restrict__ bal)
{
    int64_t gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    int64_t i = gid / G;
    int l = (int)(gid % G);
    bool valid = i < n;
    Kahan<R> sd, scx, scy, scz;
    if (valid) {
        const double4 pi4 = rows.p4[i];
        const double xi = pi4.x, yi = pi4.y, zi = pi4.z, hi = pi4.w;
        const double4 vi4 = rows.v4[i];
        const double vxi = vi4.x, vyi = vi4.y, vzi = vi4.z;
        int64_t k0, e;
        row_range(rows, i, k0, e);
        for (int64_t k = k0 + l; k < e; k += G) {
            const int j = rows.idx[k];
            const double4 pj = rows.p4[j], vj = rows.v4[j];
            R dx = (R)wrap_d<PER>(xi - pj.x, rows.L), dy = (R)wrap_d<PER>(yi - pj.y, rows.L),
              dz = (R)wrap_d<PER>(zi - pj.z, rows.L);
            R rr = sph_sqrt<R>(dx * dx + dy * dy + dz * dz);
            R nW = spline_nW<R>(rr, (R)(0.5 * (hi + pj.w)));
            R gx = nW * dx, gy = nW * dy, gz = nW * dz;
            R ux = (R)(vxi - vj.x), uy = (R)(vyi - vj.y), uz = (R)(vzi - vj.z);
            R mj = (R)vj.w;
            sd.add(-mj * (ux * gx + uy * gy + uz * gz));
            scx.add(mj * (uy * gz - uz * gy));
            scy.add(mj * (uz * gx - ux * gz));
            scz.add(mj * (ux * gy - uy * gx));
        }
    }
    R d = group_sum<R, G>(sd.s, -sd.c);
    R cx = group_sum<R, G>(scx.s, -scx.c);
    R cy = group_sum<R, G>(scy.s, -scy.c);
    R cz = group_sum<R, G>(scz.s, -scz.c);
    if (valid && l == 0) {
        R ir = (R)1 / rho[i];
        R dv = d * ir;
        dv = dv < (R)0 ? -dv : dv;
        R cv = sph_sqrt<R>(cx * cx + cy * cy + cz * cz) * ir;
        divV[i] = dv;
        curlV[i] = cv;
        if (bal) bal[i] = dv / (dv + cv + (R)0.0001 * c[i] / (R)rows.p4[i].w);
    }
}

// Density, EOS and div/curl v in one pass over the CSR rows (the time loop's
// getDensity + getPressure + div_curlVel, sph_serial_v2.py:166-189): div and
// curl need only the particle's own density, so both sums share the
// neighbour gathers and are divided by rho_i at the end.  Per-quantity
// arithmetic and summation order are those of k_density and k_divcurl.
template <typename R, int G, bool PER, bool CMP>
__global__ void __launch_bounds__(256) k_density_divcurl(
    int64_t n, const double *__restrict__ u, double gamma, NgbRows rows, R *__restrict__ rho, R *__restrict__ P,
    R *__restrict__ c, R *__restrict__ prho2, R *__restrict__ divV, R *__restrict__ curlV, R *__restrict__ bal,
    R4<R> *__restrict__ s4)
{
    int64_t gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    int64_t i = gid / G;
    int l = (int)(gid % G);
    bool valid = i < n;
    Kahan<R> acc, sd, scx, scy, scz;
    R hown = (R)0;
    if (valid) {
        const float cs = nb_scale<CMP>(rows);
        const Nb<CMP> bi = nb_load<CMP>(rows, i);
        hown = (R)nb_h<CMP>(bi);
        int64_t k0, e;
        row_range(rows, i, k0, e);
        pair_loop<G, CMP, false, int>(rows, k0, e, l, nullptr, [&](const Nb<CMP> &bj, int) {
            R dx, dy, dz, hij, ux, uy, uz, mj;
            nb_geom<R, PER, CMP>(rows, cs, bi, bj, dx, dy, dz, hij);
            nb_vel<R, CMP>(bi, bj, ux, uy, uz, mj);
            R rr = sph_sqrt<R>(dx * dx + dy * dy + dz * dz);
            acc.add(mj * spline_W<R>(rr, hij));
            R nW = spline_nW<R>(rr, hij);
            R gx = nW * dx, gy = nW * dy, gz = nW * dz;
            sd.add(-mj * (ux * gx + uy * gy + uz * gz));
            scx.add(mj * (uy * gz - uz * gy));
            scy.add(mj * (uz * gx - ux * gz));
            scz.add(mj * (ux * gy - uy * gx));
        });
    }
    R s = group_sum<R, G>(acc.s, -acc.c);
    R d = group_sum<R, G>(sd.s, -sd.c);
    R cx = group_sum<R, G>(scx.s, -scx.c);
    R cy = group_sum<R, G>(scy.s, -scy.c);
    R cz = group_sum<R, G>(scz.s, -scz.c);
    if (valid && l == 0) {
        rho[i] = s;
        R ui = (R)u[i];
        R Pi = (R)(gamma - 1.0) * s * ui;
        R ci = sph_sqrt<R>((R)(gamma * (gamma - 1.0)) * ui);
        const R pr = Pi / (s * s);
        P[i] = Pi;
        c[i] = ci;
        prho2[i] = pr;
        R ir = (R)1 / s;
        R dv = d * ir;
        dv = dv < (R)0 ? -dv : dv;
        R cv = sph_sqrt<R>(cx * cx + cy * cy + cz * cz) * ir;
        divV[i] = dv;
        curlV[i] = cv;
        R f = (R)1;
        if (bal) {
            f = dv / (dv + cv + (R)0.0001 * ci / hown);
            bal[i] = f;
        }
        if (s4) s4[i] = R4<R>{s, pr, ci, f};
    }
}

// Artificial viscosity Pi_ij (Gadget-2 form), libsx.py:146-160.
template <typename R>
__device__ __forceinline__ R av_pi(R vr, R rr, R ci, R cj, R rhoi, R rhoj, R alpha)
{
    if (vr > (R)0) return (R)0;
    R w = vr / (rr + (R)1e-5);
    R vsig = ci + cj - (R)3 * w;
    return (R)-0.5 * alpha * vsig * w / ((R)0.5 * (rhoi + rhoj));
}

// FP32 form: hardware reciprocals (1 ulp) and a select instead of the branch.
template <>
__device__ __forceinline__ float av_pi<float>(float vr, float rr, float ci, float cj, float rhoi, float rhoj,
                                              float alpha)
{
    const float w = vr * __builtin_amdgcn_rcpf(rr + 1e-5f);
    const float vsig = ci + cj - 3.0f * w;
    const float pi = -0.5f * alpha * vsig * w * __builtin_amdgcn_rcpf(0.5f * (rhoi + rhoj));
    return vr > 0.0f ? 0.0f : pi;
}

template <typename R, int G, bool PER, bool CMP>
__global__ void __launch_bounds__(256) k_accsph(int64_t n, const R4<R> *__restrict__ s4, bool balsara,
                                                double alpha,
                                                NgbRows rows,
                                                const double *__restrict__ gx_in, const double *__restrict__ gy_in,
                                                const double *__restrict__ gz_in, double *__restrict__ ax,
                                                double *__restrict__ ay, double *__restrict__ az)
{
    int64_t gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    int64_t i = gid / G;
    int l = (int)(gid % G);
    bool valid = i < n;
    Kahan<R> sx, sy, sz;
    if (valid) {
        const float cs = nb_scale<CMP>(rows);
        const Nb<CMP> bi = nb_load<CMP>(rows, i);
        const R4<R> si = s4[i];
        const R rhoi = si.x, pi_ = si.y, ci = si.z, fi = si.w;
        const R al = (R)alpha;
        int64_t k0, e;
        row_range(rows, i, k0, e);
        pair_loop<G, CMP, true>(rows, k0, e, l, s4, [&](const Nb<CMP> &bj, const R4<R> &sj) {
            R dx, dy, dz, hij, ux, uy, uz, mj;
            nb_geom<R, PER, CMP>(rows, cs, bi, bj, dx, dy, dz, hij);
            nb_vel<R, CMP>(bi, bj, ux, uy, uz, mj);
            R rr = sph_sqrt<R>(dx * dx + dy * dy + dz * dz);
            R nW = spline_nW<R>(rr, hij);
            R vr = ux * dx + uy * dy + uz * dz;
            R PI = av_pi<R>(vr, rr, ci, sj.z, rhoi, sj.x, al);
            if (balsara) PI = (R)0.5 * (fi + sj.w) * PI;
            R t = mj * (pi_ + sj.y + PI) * nW;
            sx.add(-t * dx);
            sy.add(-t * dy);
            sz.add(-t * dz);
        });
    }
    R rx = group_sum<R, G>(sx.s, -sx.c);
    R ry = group_sum<R, G>(sy.s, -sy.c);
    R rz = group_sum<R, G>(sz.s, -sz.c);
    if (valid && l == 0) {
        double gx = gx_in ? gx_in[i] : 0.0, gy = gy_in ? gy_in[i] : 0.0, gz = gz_in ? gz_in[i] : 0.0;
        ax[i] = (double)rx + gx;
        ay[i] = (double)ry + gy;
        az[i] = (double)rz + gz;
    }
}

template <typename R, int G, bool PER, bool CMP>
__global__ void __launch_bounds__(256) k_dudt(int64_t n, const R4<R> *__restrict__ s4, bool balsara,
                                              double alpha,
                                              NgbRows rows,
                                              R *__restrict__ dudt)
{
    int64_t gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    int64_t i = gid / G;
    int l = (int)(gid % G);
    bool valid = i < n;
    Kahan<R> s;
    if (valid) {
        const float cs = nb_scale<CMP>(rows);
        const Nb<CMP> bi = nb_load<CMP>(rows, i);
        const R4<R> si = s4[i];
        const R rhoi = si.x, pi_ = si.y, ci = si.z, fi = si.w;
        const R al = (R)alpha;
        int64_t k0, e;
        row_range(rows, i, k0, e);
        pair_loop<G, CMP, true>(rows, k0, e, l, s4, [&](const Nb<CMP> &bj, const R4<R> &sj) {
            R dx, dy, dz, hij, ux, uy, uz, mj;
            nb_geom<R, PER, CMP>(rows, cs, bi, bj, dx, dy, dz, hij);
            nb_vel<R, CMP>(bi, bj, ux, uy, uz, mj);
            R rr = sph_sqrt<R>(dx * dx + dy * dy + dz * dz);
            R nW = spline_nW<R>(rr, hij);
            R vr = ux * dx + uy * dy + uz * dz;
            R PI = av_pi<R>(vr, rr, ci, sj.z, rhoi, sj.x, al);
            if (balsara) PI = (R)0.5 * (fi + sj.w) * PI;
            s.add(mj * (pi_ + (R)0.5 * PI) * nW * vr);
        });
    }
    R r = group_sum<R, G>(s.s, -s.c);
    if (valid && l == 0) dudt[i] = r;
}

// Kernel-gradient factor of the column-density pass (grad_rho_norm_ngb,
// Particles_in_BH_Tree_III/CLOUDY_reloaded_Here/run_simulation/hfvCLibs_v7_hfv_v8_gradRho_v4t.h:748-770):
// the pieces of spline_nW with q + 1e-10 in the outer piece's denominator.
template <typename R>
__device__ __forceinline__ R grad_nW(R rr, R hij)
{
    const R sig = (R)(1.0 / SPH_PI);
    R q = rr / hij, h5 = hij * hij * hij * hij * hij, nW = 0;
    if (q <= (R)1) nW = sig / h5 * ((R)-3 + (R)2.25 * q);
    else if (q <= (R)2) nW = (R)-3 * sig / ((R)4 * h5) * ((R)2 - q) * ((R)2 - q) / (q + (R)1e-10);
    return nW;
}
template <>
__device__ __forceinline__ float grad_nW<float>(float rr, float hij)
{
    const float sig = (float)(1.0 / SPH_PI);
    const float ih = __builtin_amdgcn_rcpf(hij), q = rr * ih, ih2 = ih * ih;
    const float n1 = -3.0f + 2.25f * q;
    const float t = fmaxf(2.0f - q, 0.0f);
    const float n2 = -0.75f * t * t * __builtin_amdgcn_rcpf(q + 1e-10f);
    return sig * (ih2 * ih2 * ih) * (q <= 1.0f ? n1 : n2);
}

// g_i = |sum_j m_j grad W_ij| over row i (16 lanes per particle, Kahan sums, TwoSum lane
// reduction as the other passes; the same file's :716-786), then N_H in FP64 by the
// particle's first lane (the host loop of
// Particles_in_BH_Tree_III/CLOUDY_reloaded_Here/run_simulation/hfvGpuSphBHtree_gradRho_v3t.cu:1214-1224).
template <typename R, int G, bool PER>
__global__ void __launch_bounds__(256) k_column(int64_t n, const R *__restrict__ rho, NgbRows rows,
                                                sph::ColumnParams cp, double *__restrict__ grad,
                                                double *__restrict__ log_nh)
{
    int64_t gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    int64_t i = gid / G;
    int l = (int)(gid % G);
    bool valid = i < n;
    Kahan<R> sx, sy, sz;
    double hi = 0.0;
    if (valid) {
        const double4 pi4 = rows.p4[i];
        const double xi = pi4.x, yi = pi4.y, zi = pi4.z;
        hi = pi4.w;
        int64_t k0, e;
        row_range(rows, i, k0, e);
        for (int64_t k = k0 + l; k < e; k += G) {
            const int j = rows.idx[k];
            const double4 pj = rows.p4[j];
            R dx = (R)wrap_d<PER>(pj.x - xi, rows.L), dy = (R)wrap_d<PER>(pj.y - yi, rows.L),
              dz = (R)wrap_d<PER>(pj.z - zi, rows.L);
            R rr = sph_sqrt<R>(dx * dx + dy * dy + dz * dz);
            R mnW = (R)rows.v4[j].w * grad_nW<R>(rr, (R)(0.5 * (hi + pj.w)));
            sx.add(mnW * dx);
            sy.add(mnW * dy);
            sz.add(mnW * dz);
        }
    }
    R gx = group_sum<R, G>(sx.s, -sx.c);
    R gy = group_sum<R, G>(sy.s, -sy.c);
    R gz = group_sum<R, G>(sz.s, -sz.c);
    if (valid && l == 0) {
        const double g = sqrt((double)gx * gx + (double)gy * gy + (double)gz * gz), ri = (double)rho[i];
        const double nh = ri * cp.unit_density_cgs * cp.XH / SPH_M_H;
        const double col = 0.5 * (hi / cbrt(cp.n_ngb) + ri / g) * cp.unit_length_cm * nh;
        if (grad) grad[i] = g;
        log_nh[i] = log10(col);
    }
}

// Signal-velocity maximum over row i (16 lanes, max is exact in any order), then the six
// criteria of the reference's dt_array_indiv_dt in FP64 by the particle's first lane.
template <typename R, int G, bool PER>
__global__ void __launch_bounds__(256) k_timestep(int64_t n, const R4<R> *__restrict__ s4, const R *__restrict__ divv,
                                                  NgbRows rows, const double *__restrict__ ax,
                                                  const double *__restrict__ ay, const double *__restrict__ az,
                                                  const double *__restrict__ gx, const double *__restrict__ gy,
                                                  const double *__restrict__ gz, const double *__restrict__ eps,
                                                  sph::StepCriteria sc, double *__restrict__ dt,
                                                  double *__restrict__ dhdt)
{
    int64_t gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    int64_t i = gid / G;
    int l = (int)(gid % G);
    bool valid = i < n;
    R vmax = (R)0;  // the row holds i itself (w = 0): vsig >= 2 c_i >= 0
    if (valid) {
        const double4 pi4 = rows.p4[i], vi4 = rows.v4[i];
        const R ci = s4[i].z;
        int64_t k0, e;
        row_range(rows, i, k0, e);
        for (int64_t k = k0 + l; k < e; k += G) {
            const int j = rows.idx[k];
            const double4 pj = rows.p4[j], vj = rows.v4[j];
            R dx = (R)wrap_d<PER>(pi4.x - pj.x, rows.L), dy = (R)wrap_d<PER>(pi4.y - pj.y, rows.L),
              dz = (R)wrap_d<PER>(pi4.z - pj.z, rows.L);
            R rr = sph_sqrt<R>(dx * dx + dy * dy + dz * dz);
            R vr = (R)(vi4.x - vj.x) * dx + (R)(vi4.y - vj.y) * dy + (R)(vi4.z - vj.z) * dz;
            R w;
            if constexpr (std::is_same<R, float>::value)
                w = vr * __builtin_amdgcn_rcpf(rr + 1e-5f);  // 1 ulp reciprocal
            else
                w = vr / (rr + 1e-5);
            R vsig = ci + s4[j].z - (R)3 * w;
            vmax = vsig > vmax ? vsig : vmax;
        }
    }
    for (int off = G / 2; off >= 1; off >>= 1) {
        const R o = __shfl_xor(vmax, off, G);
        vmax = o > vmax ? o : vmax;
    }
    if (valid && l == 0) {
        const double h = rows.p4[i].w, c = (double)s4[i].z, dv = (double)divv[i];
        const double4 v = rows.v4[i];
        const double av = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
        const double at = sqrt(ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i]);
        const double ag = gx ? sqrt(gx[i] * gx[i] + gy[i] * gy[i] + gz[i] * gz[i]) : at;
        const double ep = eps ? eps[i] : sc.eps_const;
        const double dt_hyd = sc.C_CFL * h / (double)vmax;
        const double dt_cour = sc.C_CFL * h / (h * fabs(dv) + fmax(c, av) * (1.0 + 0.6 * sc.alpha));
        const double dt_g1 = 0.4 * sqrt(ep / at), dt_g2 = 0.4 * ep / av;
        const double dt_f = sqrt(h / ag), dt_kin = sqrt(h / at);
        dt[i] = fmin(fmin(fmin(dt_hyd, dt_cour), fmin(dt_f, dt_kin)), fmin(dt_g1, dt_g2));
        if (dhdt) dhdt[i] = h * dv / 3.0;
    }
}

// *out = bits of min_i v[i] (v >= 0): grid-stride, wave and workgroup minima, one atomic per
// workgroup (positive doubles order as their bit patterns).
__global__ void __launch_bounds__(256) k_min_bits(int64_t n, const double *__restrict__ v,
                                                  unsigned long long *__restrict__ out)
{
    __shared__ double wmin[4];
    double m = INFINITY;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = fmin(m, v[i]);
    for (int off = 32; off >= 1; off >>= 1) m = fmin(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmin(fmin(wmin[0], wmin[1]), fmin(wmin[2], wmin[3]));
        if (m < INFINITY) atomicMin(out, (unsigned long long)__double_as_longlong(m));
    }
}

namespace sph {

__global__ void __launch_bounds__(256) k_pack4(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                                               const double *__restrict__ c, const double *__restrict__ d,
                                               double4 *__restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_double4(a[i], b[i], c[i], d[i]);
}

hipError_t pack4(hipStream_t st, int64_t n, const double *a, const double *b, const double *c, const double *d,
                 double4 *out)
{
    k_pack4<<<sph_blocks(n, 256), 256, 0, st>>>(n, a, b, c, d, out);
    return hipGetLastError();
}

// Compact records (declared in sph_hydro.h).  Open boundaries (per = 0): lo = grid[0..2], the
// largest extent E = max(grid[3+d] - grid[d]) = f 2^k with f in [0.5, 1) (frexp), so
// E < 2^k: the coordinates in the box map to [0, 2^32) in steps of sc = 2^(k-32) (the clamp
// only guards positions outside the box).  Periodic box of side per: steps of sc = per / 2^32
// from the origin, modulo 2^32 (a coordinate equal to per, which the drift's wrap can round
// to, maps to 0).  Rounding to the nearest step.  Every thread derives sc itself; thread 0
// stores it in grid[7] for the passes.
__global__ void __launch_bounds__(256) k_pack_compact(int64_t n, const double *__restrict__ x,
                                                      const double *__restrict__ y, const double *__restrict__ z,
                                                      const double *__restrict__ h, const double *__restrict__ vx,
                                                      const double *__restrict__ vy, const double *__restrict__ vz,
                                                      const double *__restrict__ m, double per, double *__restrict__ grid,
                                                      uint4 *__restrict__ pf, float4 *__restrict__ vf, bool vel_only)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    vf[i] = make_float4((float)vx[i], (float)vy[i], (float)vz[i], (float)m[i]);
    if (vel_only) return;
    double lo[3] = {0.0, 0.0, 0.0}, sc, isc;
    if (per > 0.0) {
        sc = per * 0x1p-32;
        isc = 0x1p32 / per;
    } else {
        lo[0] = grid[0];
        lo[1] = grid[1];
        lo[2] = grid[2];
        const double ext = fmax(fmax(grid[3] - grid[0], grid[4] - grid[1]), grid[5] - grid[2]);
        int k = 0;
        frexp(ext, &k);
        sc = ldexp(1.0, k - 32);
        isc = ldexp(1.0, 32 - k);
    }
    if (i == 0) grid[7] = sc;
    const double p[3] = {x[i], y[i], z[i]};
    uint32_t q[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double v = rint((p[d] - lo[d]) * isc);
        if (per > 0.0) {
            v = fmin(fmax(v, 0.0), 4294967296.0);
            q[d] = (uint32_t)(uint64_t)v;  // 2^32 wraps to 0
        } else {
            q[d] = (uint32_t)fmin(fmax(v, 0.0), 4294967295.0);
        }
    }
    pf[i] = make_uint4(q[0], q[1], q[2], __float_as_uint((float)h[i]));
}

hipError_t pack_compact(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                        const double *h, const double *vx, const double *vy, const double *vz, const double *m,
                        double per, double *grid, uint4 *pf, float4 *vf, bool vel_only)
{
    k_pack_compact<<<sph_blocks(n, 256), 256, 0, st>>>(n, x, y, z, h, vx, vy, vz, m, per, grid, pf, vf, vel_only);
    return hipGetLastError();
}

#define LAUNCH_G(K, G, ...)                                                                                        \
    do {                                                                                                           \
        if (rows.L > 0.0)                                                                                          \
            K<R, G, true><<<sph_blocks(n * G, 256), 256, 0, st>>>(__VA_ARGS__);                                     \
        else                                                                                                       \
            K<R, G, false><<<sph_blocks(n * G, 256), 256, 0, st>>>(__VA_ARGS__);                                    \
    } while (0)
#define LAUNCH(K, ...) LAUNCH_G(K, HG, __VA_ARGS__)
// Passes that read the compact records when the rows carry them (mixed form, open boundaries).
#define LAUNCH_C(K, G, ...)                                                                                        \
    do {                                                                                                           \
        constexpr bool F32 = std::is_same<R, float>::value;                                                        \
        if (rows.L > 0.0 && F32 && rows.pf)                                                                        \
            K<R, G, true, F32><<<sph_blocks(n * G, 256), 256, 0, st>>>(__VA_ARGS__);                                \
        else if (rows.L > 0.0)                                                                                     \
            K<R, G, true, false><<<sph_blocks(n * G, 256), 256, 0, st>>>(__VA_ARGS__);                              \
        else if (F32 && rows.pf)                                                                                   \
            K<R, G, false, F32><<<sph_blocks(n * G, 256), 256, 0, st>>>(__VA_ARGS__);                               \
        else                                                                                                       \
            K<R, G, false, false><<<sph_blocks(n * G, 256), 256, 0, st>>>(__VA_ARGS__);                             \
    } while (0)

template <typename R>
__global__ void __launch_bounds__(256) k_pack_s4(int64_t n, const R *__restrict__ rho, const R *__restrict__ prho2,
                                                 const R *__restrict__ c, const R *__restrict__ bal,
                                                 R4<R> *__restrict__ s4)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) s4[i] = R4<R>{rho[i], prho2[i], c[i], bal ? bal[i] : (R)1};
}

template <typename R>
hipError_t pack_s4(hipStream_t st, int64_t n, const R *rho, const R *prho2, const R *c, const R *bal, R4<R> *s4)
{
    k_pack_s4<R><<<sph_blocks(n, 256), 256, 0, st>>>(n, rho, prho2, c, bal, s4);
    return hipGetLastError();
}

template <typename R>
hipError_t density(hipStream_t st, int64_t n, const double *u, double gamma, const NgbRows &rows, R *rho, R *P,
                   R *c, R *prho2, R4<R> *s4)
{
    LAUNCH_C(k_density, HG, n, u, gamma, rows, rho, P, c, prho2, u ? s4 : nullptr);
    return hipGetLastError();
}

template <typename R>
hipError_t divcurl(hipStream_t st, int64_t n, const R *rho, const R *c, const NgbRows &rows, R *divV, R *curlV,
                   R *bal)
{
    LAUNCH(k_divcurl, n, rho, c, rows, divV, curlV, bal);
    return hipGetLastError();
}

template <typename R>
hipError_t density_divcurl(hipStream_t st, int64_t n, const double *u, double gamma, const NgbRows &rows, R *rho,
                           R *P, R *c, R *prho2, R *divV, R *curlV, R *bal, R4<R> *s4)
{
    LAUNCH_C(k_density_divcurl, HG_DENSITY, n, u, gamma, rows, rho, P, c, prho2, divV, curlV, bal, s4);
    return hipGetLastError();
}

template <typename R>
hipError_t accsph(hipStream_t st, int64_t n, const R4<R> *s4, bool balsara, double alpha, const NgbRows &rows,
                  const double *gx, const double *gy, const double *gz, double *ax, double *ay, double *az)
{
    LAUNCH_C(k_accsph, HG_ACC, n, s4, balsara, alpha, rows, gx, gy, gz, ax, ay, az);
    return hipGetLastError();
}

template <typename R>
hipError_t dudt(hipStream_t st, int64_t n, const R4<R> *s4, bool balsara, double alpha, const NgbRows &rows, R *out)
{
    LAUNCH_C(k_dudt, HG, n, s4, balsara, alpha, rows, out);
    return hipGetLastError();
}
template <typename R>
hipError_t timestep(hipStream_t st, int64_t n, const R4<R> *s4, const R *divv, const NgbRows &rows,
                    const double *ax, const double *ay, const double *az, const double *gx, const double *gy,
                    const double *gz, const double *eps, const StepCriteria &sc, double *dt, double *dhdt,
                    unsigned long long *dt_min_bits)
{
    if (n <= 0) return hipSuccess;
    if (!dt) return hipErrorInvalidValue;
    LAUNCH(k_timestep, n, s4, divv, rows, ax, ay, az, gx, gy, gz, eps, sc, dt, dhdt);
    const unsigned nb = (unsigned)std::min<int64_t>(1024, (n + 255) / 256);
    k_min_bits<<<nb, 256, 0, st>>>(n, dt, dt_min_bits);
    return hipGetLastError();
}

template <typename R>
hipError_t column_density(hipStream_t st, int64_t n, const R *rho, const NgbRows &rows, const ColumnParams &cp,
                          double *grad, double *log_nh)
{
    if (n <= 0) return hipSuccess;
    LAUNCH(k_column, n, rho, rows, cp, grad, log_nh);
    return hipGetLastError();
}
#undef LAUNCH
#undef LAUNCH_C

#define INST(R)                                                                                                    \
    template hipError_t pack_s4<R>(hipStream_t, int64_t, const R *, const R *, const R *, const R *, R4<R> *);    \
    template hipError_t density<R>(hipStream_t, int64_t, const double *, double, const NgbRows &, R *, R *, R *,  \
                                   R *, R4<R> *);                                                                  \
    template hipError_t divcurl<R>(hipStream_t, int64_t, const R *, const R *, const NgbRows &, R *, R *, R *);   \
    template hipError_t density_divcurl<R>(hipStream_t, int64_t, const double *, double, const NgbRows &, R *,    \
                                           R *, R *, R *, R *, R *, R *, R4<R> *);                                 \
    template hipError_t accsph<R>(hipStream_t, int64_t, const R4<R> *, bool, double, const NgbRows &,             \
                                  const double *, const double *, const double *, double *, double *, double *);  \
    template hipError_t dudt<R>(hipStream_t, int64_t, const R4<R> *, bool, double, const NgbRows &, R *);         \
    template hipError_t column_density<R>(hipStream_t, int64_t, const R *, const NgbRows &, const ColumnParams &,  \
                                          double *, double *);                                                     \
    template hipError_t timestep<R>(hipStream_t, int64_t, const R4<R> *, const R *, const NgbRows &,              \
                                    const double *, const double *, const double *, const double *,              \
                                    const double *, const double *, const double *, const StepCriteria &,        \
                                    double *, double *, unsigned long long *);
INST(float)
INST(double)
#undef INST

}  // namespace sph
