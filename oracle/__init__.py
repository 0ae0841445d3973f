"""FP64 CPU oracle for the SPH hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It is the checker, never the thing that
is measured or shipped: the product path (``sph_3_amd``) does not import it and
fails loudly when its HIP library is missing.

The C restatement lives in ``sph_oracle.c`` (built by ``oracle/Makefile`` into
``liboracle.so``); this module wraps it with numpy.  Functions in two groups:

* literal restatements of ``sph_code/libsx.py`` and ``sph_code/shear.py``
  (all-pairs, FP64, row range ``[nbeg, nend)`` as in the ``*_mpi`` twins);
* the build's definitions the device path must reproduce: Morton order and
  binary radix tree, symmetric CSR neighbour sets, Newton-Raphson smoothing
  length, tree gravity with a per-particle opening test, CSR-restricted sums.

Parity pins for every function live in ``tests/test_oracle*.py``; the
cooling-surrogate network's oracle is ``oracle/mlp.py`` (numpy).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64 = ctypes.c_int64


def build() -> str:
    """Compile liboracle.so in place (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.or_ke.restype = ctypes.c_double
        L.or_pe.restype = ctypes.c_double
        L.or_h_newton.restype = ctypes.c_int64
        for f in ("or_cool_rb95", "or_cool_u_to_T", "or_cool_rate_from_u", "or_do_cooling"):
            getattr(L, f).restype = ctypes.c_double
        for f in ("or_cool_rb95",):
            getattr(L, f).argtypes = [ctypes.c_double]
        L.or_cool_u_to_T.argtypes = [ctypes.c_double] * 4
        L.or_cool_rate_from_u.argtypes = [ctypes.c_double] * 5
        L.or_do_cooling.argtypes = [ctypes.c_double] * 7
        L.or_set_period.argtypes = [ctypes.c_double]
        L.or_nh_cooling.restype = ctypes.c_int64
        L.or_nh_nx0.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_float]
        _LIB = L
    return _LIB


def _p(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_d)
    if a.dtype == np.int64:
        return a.ctypes.data_as(_i64p)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_i32p)
    if a.dtype == np.uint64:
        return a.ctypes.data_as(_u64p)
    if a.dtype == np.float32:
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    raise TypeError(a.dtype)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _rows(N, nbeg, nend):
    nbeg = 0 if nbeg is None else int(nbeg)
    nend = N if nend is None else int(nend)
    return nbeg, nend


# ---------------------------------------------------------------- all-pairs
def density_allpairs(pos, m, h, nbeg=None, nend=None):
    pos = _f64(pos, (-1, 3)); N = len(pos)
    nbeg, nend = _rows(N, nbeg, nend)
    out = np.zeros(nend - nbeg)
    lib().or_density_allpairs(_i64(N), _i64(nbeg), _i64(nend), _p(pos), _p(_f64(m)), _p(_f64(h)), _p(out))
    return out


def divcurl_allpairs(pos, v, rho, m, h, nbeg=None, nend=None):
    pos = _f64(pos, (-1, 3)); N = len(pos)
    nbeg, nend = _rows(N, nbeg, nend)
    dv = np.zeros(nend - nbeg); cv = np.zeros(nend - nbeg)
    lib().or_divcurl_allpairs(_i64(N), _i64(nbeg), _i64(nend), _p(pos), _p(_f64(v, (-1, 3))),
                              _p(_f64(rho)), _p(_f64(m)), _p(_f64(h)), _p(dv), _p(cv))
    return dv, cv


def accsph_allpairs(pos, v, rho, P, c, h, m, alpha, divV=None, curlV=None, nbeg=None, nend=None):
    pos = _f64(pos, (-1, 3)); N = len(pos)
    nbeg, nend = _rows(N, nbeg, nend)
    bal = divV is not None
    dv = _f64(divV) if bal else np.zeros(N)
    cv = _f64(curlV) if bal else np.zeros(N)
    out = np.zeros((nend - nbeg, 3))
    lib().or_accsph_allpairs(_i64(N), _i64(nbeg), _i64(nend), _p(pos), _p(_f64(v, (-1, 3))), _p(_f64(rho)),
                             _p(_f64(P)), _p(_f64(c)), _p(_f64(h)), _p(_f64(m)), _p(dv), _p(cv),
                             ctypes.c_double(alpha), ctypes.c_int(int(bal)), _p(out))
    return out


def dudt_allpairs(pos, v, rho, P, c, h, m, alpha, divV=None, curlV=None, nbeg=None, nend=None):
    pos = _f64(pos, (-1, 3)); N = len(pos)
    nbeg, nend = _rows(N, nbeg, nend)
    bal = divV is not None
    dv = _f64(divV) if bal else np.zeros(N)
    cv = _f64(curlV) if bal else np.zeros(N)
    out = np.zeros(nend - nbeg)
    lib().or_dudt_allpairs(_i64(N), _i64(nbeg), _i64(nend), _p(pos), _p(_f64(v, (-1, 3))), _p(_f64(rho)),
                           _p(_f64(P)), _p(_f64(c)), _p(_f64(h)), _p(_f64(m)), _p(dv), _p(cv),
                           ctypes.c_double(alpha), ctypes.c_int(int(bal)), _p(out))
    return out


SOFT_SPLINE, SOFT_PLUMMER = 0, 1


def grav_direct(pos, m, eps, G=1.0, soft_kind=SOFT_SPLINE, eps_plummer=0.0, nbeg=None, nend=None):
    pos = _f64(pos, (-1, 3)); N = len(pos)
    nbeg, nend = _rows(N, nbeg, nend)
    eps = np.zeros(N) if eps is None else np.broadcast_to(_f64(eps), (N,)).copy()
    out = np.zeros((nend - nbeg, 3))
    lib().or_grav_direct(_i64(N), _i64(nbeg), _i64(nend), _p(pos), _p(_f64(m)), _p(eps), ctypes.c_double(G),
                         ctypes.c_int(soft_kind), ctypes.c_double(eps_plummer), _p(out))
    return out


def ke(v, m):
    v = _f64(v, (-1, 3))
    return lib().or_ke(_i64(len(v)), _p(v), _p(_f64(m)))


def pe(pos, m, G, eps, soft_kind=SOFT_SPLINE, eps_plummer=0.0):
    """Potential energy: getPE's spline form (eps_ij = (eps_i + eps_j)/2) or Plummer."""
    pos = _f64(pos, (-1, 3)); N = len(pos)
    eps = np.zeros(N) if eps is None else np.broadcast_to(_f64(eps), (N,)).copy()
    return lib().or_pe(_i64(N), _p(pos), _p(_f64(m)), ctypes.c_double(G), _p(eps), ctypes.c_int(soft_kind),
                       ctypes.c_double(eps_plummer))


def h_initial(pos, sub=None, k=50):
    pos = _f64(pos, (-1, 3))
    sub = pos if sub is None else _f64(sub, (-1, 3))
    out = np.zeros(len(sub))
    lib().or_h_initial(_i64(len(pos)), _p(pos), _i64(len(sub)), _p(sub), _i64(k), _p(out))
    return out


def h_smooth_fast(pos, h, n_low=45, n_up=55, max_iter=100000, nbeg=None, nend=None):
    pos = _f64(pos, (-1, 3)); N = len(pos)
    nbeg, nend = _rows(N, nbeg, nend)
    out = np.zeros(nend - nbeg)
    bad = np.zeros(1, np.int64)
    lib().or_h_smooth_fast(_i64(N), _i64(nbeg), _i64(nend), _p(pos), _p(_f64(h)), _i64(n_low), _i64(n_up),
                           _i64(max_iter), _p(out), _p(bad))
    return out, int(bad[0])


# ------------------------------------------------------------- build's defs
def neighbours_brute(pos, h):
    """CSR (offsets int64 [N+1], idx int32) of S_i = {j : r_ij^2 < (h_i+h_j)^2}."""
    pos = _f64(pos, (-1, 3)); N = len(pos); h = _f64(h)
    off = np.zeros(N + 1, np.int64)
    lib().or_neighbours_brute(_i64(N), _p(pos), _p(h), _p(off), None)
    idx = np.zeros(int(off[-1]), np.int32)
    lib().or_neighbours_brute(_i64(N), _p(pos), _p(h), _p(off), _p(idx))
    return off, idx


def morton_keys(pos):
    pos = _f64(pos, (-1, 3))
    keys = np.zeros(len(pos), np.uint64)
    box = np.zeros(6)
    lib().or_morton_keys(_i64(len(pos)), _p(pos), _p(keys), _p(box))
    return keys, box


def sort_keys(keys):
    keys = np.ascontiguousarray(keys, np.uint64)
    perm = np.zeros(len(keys), np.int32)
    lib().or_sort_keys(_i64(len(keys)), _p(keys), _p(perm))
    return perm


class Tree:
    """Binary radix tree over Morton-sorted particles (host FP64 twin of the
    device build).  ``perm[s]`` is the input index of sorted particle s."""

    def __init__(self, pos, m=None, h=None, eps=None, bucket=16):
        pos = _f64(pos, (-1, 3)); N = len(pos)
        self.N, self.bucket = N, int(bucket)
        keys, self.box6 = morton_keys(pos)
        self.perm = sort_keys(keys)
        self.keys = keys[self.perm]
        nn = max(2 * N - 1, 1)
        self.first = np.zeros(nn, np.int32); self.count = np.zeros(nn, np.int32)
        self.left = np.zeros(max(N - 1, 1), np.int32); self.right = np.zeros(max(N - 1, 1), np.int32)
        self.escape = np.zeros(nn, np.int32)
        lib().or_tree_topology(_i64(N), _p(self.keys), _p(self.first), _p(self.count), _p(self.left),
                               _p(self.right), _p(self.escape))
        self.spos = np.ascontiguousarray(pos[self.perm])
        self.sm = np.ascontiguousarray(_f64(m)[self.perm]) if m is not None else np.ones(N) / N
        self.sh = np.ascontiguousarray(_f64(h)[self.perm]) if h is not None else np.zeros(N)
        self.seps = (np.ascontiguousarray(np.broadcast_to(_f64(eps), (N,))[self.perm])
                     if eps is not None else np.zeros(N))
        self.nbox = np.zeros((nn, 6)); self.com = np.zeros((nn, 3)); self.mass = np.zeros(nn)
        self.hmax = np.zeros(nn); self.epsmax = np.zeros(nn); self.size = np.zeros(nn)
        self.parent = np.zeros(nn, np.int32)
        if N > 1:
            lib().or_tree_parents(_i64(N), _p(self.left), _p(self.right), _p(self.parent))
        else:
            self.parent[0] = -1
        lib().or_tree_moments(_i64(N), _p(self.spos), _p(self.sm), _p(self.sh), _p(self.seps), _p(self.first),
                              _p(self.count), _p(self.left), _p(self.right), ctypes.c_int(self.bucket),
                              _p(self.nbox), _p(self.com), _p(self.mass), _p(self.hmax), _p(self.epsmax),
                              _p(self.size))

    def gravity(self, theta=0.5, G=1.0, soft_kind=SOFT_SPLINE, eps_plummer=0.0, group=64):
        """Group-walk tree gravity (the device's interaction lists: the particles
        of each group -- a maximal tree node holding at most ``group``
        particles -- share one opening test against the group box).
        Returns (acc in INPUT order, (n_pp, n_pn))."""
        acc = np.zeros((self.N, 3)); n = np.zeros(2, np.int64)
        lib().or_grav_tree_group(_i64(self.N), _p(self.spos), _p(self.sm), _p(self.seps), _p(self.first),
                                 _p(self.count), _p(self.left), _p(self.escape), _p(self.parent), _p(self.nbox),
                                 _p(self.com),
                                 _p(self.mass), _p(self.epsmax), _p(self.size), ctypes.c_int(self.bucket),
                                 ctypes.c_int(group), ctypes.c_double(theta), ctypes.c_double(G),
                                 ctypes.c_int(soft_kind), ctypes.c_double(eps_plummer), _p(acc), _p(n))
        out = np.zeros_like(acc); out[self.perm] = acc
        return out, (int(n[0]), int(n[1]))

    def gravity_per_particle(self, theta=0.5, G=1.0, soft_kind=SOFT_SPLINE, eps_plummer=0.0):
        """Per-particle opening test (size^2 < theta^2 |com - x|^2, x outside
        the node box).  Returns (acc in INPUT order, (n_pp, n_pn))."""
        acc = np.zeros((self.N, 3)); n = np.zeros(2, np.int64)
        lib().or_grav_tree(_i64(self.N), _p(self.spos), _p(self.sm), _p(self.seps), _p(self.first),
                           _p(self.count), _p(self.left), _p(self.escape), _p(self.nbox), _p(self.com),
                           _p(self.mass), _p(self.epsmax), _p(self.size), ctypes.c_int(self.bucket),
                           ctypes.c_double(theta), ctypes.c_double(G), ctypes.c_int(soft_kind),
                           ctypes.c_double(eps_plummer), _p(acc), _p(n))
        out = np.zeros_like(acc); out[self.perm] = acc
        return out, (int(n[0]), int(n[1]))

    def neighbours(self):
        """CSR of S_i in INPUT numbering, rows ascending (uses the tree's h)."""
        N = self.N
        off = np.zeros(N + 1, np.int64)
        args = (_i64(N), _p(self.spos), _p(self.sh), _p(self.first), _p(self.count), _p(self.left),
                _p(self.escape), _p(self.nbox), _p(self.hmax), ctypes.c_int(self.bucket))
        lib().or_neighbours_tree(*args, _p(off), None)
        idx = np.zeros(int(off[-1]), np.int32)
        lib().or_neighbours_tree(*args, _p(off), _p(idx))
        return csr_permute(off, idx, self.perm)


def csr_permute(off_s, idx_s, perm):
    """Sorted-numbering CSR -> input-numbering CSR with ascending rows."""
    N = len(perm)
    cnt_s = np.diff(off_s)
    cnt = np.zeros(N, np.int64); cnt[perm] = cnt_s
    off = np.zeros(N + 1, np.int64); np.cumsum(cnt, out=off[1:])
    rows_s = np.repeat(np.arange(N), cnt_s)
    rows = perm[rows_s].astype(np.int64)
    cols = perm[idx_s].astype(np.int64)
    order = np.lexsort((cols, rows))
    return off, cols[order].astype(np.int32)


def h_newton(pos, h0, n_ngb=64.0, tol=1e-12, max_iter=200):
    pos = _f64(pos, (-1, 3))
    out = np.zeros(len(pos))
    bad = lib().or_h_newton(_i64(len(pos)), _p(pos), _p(_f64(h0)), ctypes.c_double(n_ngb), ctypes.c_double(tol),
                            _i64(max_iter), _p(out))
    return out, int(bad)


def h_newton_rows(pos, rows, h0, n_ngb=64.0, tol=1e-12, max_iter=200):
    """h_newton for the particles ``rows`` only (O(len(rows) N)); h0 per row."""
    pos = _f64(pos, (-1, 3))
    rows = np.ascontiguousarray(rows, np.int64)
    out = np.zeros(len(rows))
    lib().or_h_newton_rows.restype = ctypes.c_int64
    bad = lib().or_h_newton_rows(_i64(len(pos)), _p(pos), _i64(len(rows)), _p(rows), _p(_f64(h0)),
                                 ctypes.c_double(n_ngb), ctypes.c_double(tol), _i64(max_iter), _p(out))
    return out, int(bad)


def density_csr(pos, m, h, off, idx):
    pos = _f64(pos, (-1, 3)); out = np.zeros(len(pos))
    lib().or_density_csr(_i64(len(pos)), _p(pos), _p(_f64(m)), _p(_f64(h)), _p(off), _p(idx), _p(out))
    return out


def grad_rho_csr(pos, m, h, off, idx):
    """|grad rho| of grad_rho_norm_ngb (FP64) over CSR sets."""
    pos = _f64(pos, (-1, 3)); out = np.zeros(len(pos))
    lib().or_grad_rho_csr(_i64(len(pos)), _p(pos), _p(_f64(m)), _p(_f64(h)), _p(off), _p(idx), _p(out))
    return out


def column_density(rho, g, h, unit_density_cgs, unit_length_cm, XH=0.7, n_ngb=64.0):
    """log10 N_H [cm^-2] from rho, |grad rho| and h (the reference's host loop, FP64)."""
    rho = _f64(rho); out = np.zeros(len(rho))
    lib().or_column_density(_i64(len(rho)), _p(rho), _p(_f64(g)), _p(_f64(h)), ctypes.c_double(unit_density_cgs),
                            ctypes.c_double(unit_length_cm), ctypes.c_double(XH), ctypes.c_double(n_ngb), _p(out))
    return out


def inject_uniform(seed, step, pair, k):
    """The injection's counter-based uniform draw in [0, 1)."""
    f = lib().or_inject_uniform
    f.restype = ctypes.c_double
    return f(ctypes.c_uint64(seed), ctypes.c_uint64(step), ctypes.c_uint64(pair), ctypes.c_int(k))


def inject_pairs(pairs, seed, step, r_inject, v_in):
    """(pos, vel) of outflow_injector2's pairs with the counter-based draws."""
    pos = np.zeros((2 * pairs, 3)); vel = np.zeros((2 * pairs, 3))
    lib().or_inject_pairs(_i64(pairs), ctypes.c_uint64(seed), ctypes.c_uint64(step), ctypes.c_double(r_inject),
                          ctypes.c_double(v_in), _p(pos), _p(vel))
    return pos, vel


def halo_accel(pos, sigma, G=1.0):
    """galaxy_isothermal_potential's acceleration (FP64)."""
    pos = _f64(pos, (-1, 3)); acc = np.zeros_like(pos)
    lib().or_halo_accel(_i64(len(pos)), _p(pos), ctypes.c_double(sigma), ctypes.c_double(G), _p(acc))
    return acc


def timestep(pos, vel, h, c, divV, acc, acc_g, eps, C_CFL, alpha, off=None, idx=None):
    """(dt_i, dh/dt_i) of dt_array_indiv_dt (FP64); vsig over the CSR sets, or over all
    particles (the reference's loop) when off is None."""
    pos = _f64(pos, (-1, 3)); N = len(pos)
    dt = np.zeros(N); dh = np.zeros(N)
    eps = np.broadcast_to(_f64(eps), (N,)).copy()
    lib().or_timestep(_i64(N), _p(pos), _p(_f64(vel, (-1, 3))), _p(_f64(h)), _p(_f64(c)), _p(_f64(divV)),
                      _p(_f64(acc, (-1, 3))), _p(_f64(acc_g, (-1, 3))), _p(eps), ctypes.c_double(C_CFL),
                      ctypes.c_double(alpha), _p(off) if off is not None else None,
                      _p(idx) if idx is not None else None, _p(dt), _p(dh))
    return dt, dh


def divcurl_csr(pos, v, rho, m, h, off, idx):
    pos = _f64(pos, (-1, 3)); N = len(pos)
    dv = np.zeros(N); cv = np.zeros(N)
    lib().or_divcurl_csr(_i64(N), _p(pos), _p(_f64(v, (-1, 3))), _p(_f64(rho)), _p(_f64(m)), _p(_f64(h)),
                         _p(off), _p(idx), _p(dv), _p(cv))
    return dv, cv


def accsph_csr(pos, v, rho, P, c, h, m, alpha, off, idx, divV=None, curlV=None):
    pos = _f64(pos, (-1, 3)); N = len(pos)
    bal = divV is not None
    dv = _f64(divV) if bal else np.zeros(N)
    cv = _f64(curlV) if bal else np.zeros(N)
    out = np.zeros((N, 3))
    lib().or_accsph_csr(_i64(N), _p(pos), _p(_f64(v, (-1, 3))), _p(_f64(rho)), _p(_f64(P)), _p(_f64(c)),
                        _p(_f64(h)), _p(_f64(m)), _p(dv), _p(cv), ctypes.c_double(alpha), ctypes.c_int(int(bal)),
                        _p(off), _p(idx), _p(out))
    return out


def dudt_csr(pos, v, rho, P, c, h, m, alpha, off, idx, divV=None, curlV=None):
    pos = _f64(pos, (-1, 3)); N = len(pos)
    bal = divV is not None
    dv = _f64(divV) if bal else np.zeros(N)
    cv = _f64(curlV) if bal else np.zeros(N)
    out = np.zeros(N)
    lib().or_dudt_csr(_i64(N), _p(pos), _p(_f64(v, (-1, 3))), _p(_f64(rho)), _p(_f64(P)), _p(_f64(c)),
                      _p(_f64(h)), _p(_f64(m)), _p(dv), _p(cv), ctypes.c_double(alpha), ctypes.c_int(int(bal)),
                      _p(off), _p(idx), _p(out))
    return out


# ------------------------------------------------------------------ cooling
J21_REFERENCE = 1e-7  # the reference's fixed radiation-field amplitude (photolibs3.py:233)


def cool_rates(T):
    out = np.zeros(7)
    lib().or_cool_rates(ctypes.c_double(T), _p(out))
    return out


def cool_radfield(J21=J21_REFERENCE):
    out = np.zeros(6)
    lib().or_cool_radfield(ctypes.c_double(J21), _p(out))
    return out


def cool_abund(T, nH, gJ):
    out = np.zeros(6)
    lib().or_cool_abund(ctypes.c_double(T), ctypes.c_double(nH), _p(_f64(gJ)), _p(out))
    return out


def cool_heat_lambda(T, nH, J21=J21_REFERENCE, Z=0.0):
    out = np.zeros(2)
    lib().or_cool_heat_lambda(ctypes.c_double(T), ctypes.c_double(nH), ctypes.c_double(J21), ctypes.c_double(Z),
                              _p(out))
    return out


def cool_rb95(T):
    return lib().or_cool_rb95(float(T))


def cool_u_to_T(u, nH, XH=0.76, J21=J21_REFERENCE):
    return lib().or_cool_u_to_T(float(u), float(nH), float(XH), float(J21))


def cool_rate_from_u(u, nH, XH=0.76, J21=J21_REFERENCE, Z=0.0):
    return lib().or_cool_rate_from_u(float(u), float(nH), float(XH), float(J21), float(Z))


def do_cooling(rho, u_old, dt, XH=0.76, J21=J21_REFERENCE, Z=0.0, u_floor=1e8):
    return lib().or_do_cooling(float(rho), float(u_old), float(dt), float(XH), float(J21), float(Z), float(u_floor))


def cool_update(rho, u, Z, J21, dt, units, lo, hi):
    """Explicit per-step cooling update (or_cool_update_n): units = (rho_cgs, u_cgs,
    t_s, XH, u_floor_cgs), lo/hi = log10 box of (nH, T, Z, J21).  Returns (u', floored)."""
    rho = _f64(rho); n = len(rho)
    u = np.broadcast_to(_f64(u), (n,)).copy()
    Z = np.broadcast_to(_f64(Z), (n,)).copy()
    J21 = np.broadcast_to(_f64(J21), (n,)).copy()
    out = np.zeros(n); fl = np.zeros(n, np.int32)
    lib().or_cool_update_n(_i64(n), _p(rho), _p(u), _p(Z), _p(J21), ctypes.c_double(dt), _p(_f64(units)),
                           _p(_f64(lo)), _p(_f64(hi)), _p(out), _p(fl))
    return out, fl.astype(bool)


class periodic:
    """Context manager: minimum-image distances in a periodic box of side L for
    ``neighbours_brute``, ``h_newton`` and the ``*_csr`` functions (open
    boundaries again on exit)."""

    def __init__(self, L):
        self.L = float(L)

    def __enter__(self):
        lib().or_set_period(self.L)
        return self

    def __exit__(self, *exc):
        lib().or_set_period(0.0)
        return False


# ---------------------------------------------------------------- cooling from an N_H table
def nh_nx0(grid, x):
    """find_nx0 (CR/v4t.h:1450-1476) of float32 grid and value x."""
    g = np.ascontiguousarray(grid, np.float32)
    return int(lib().or_nh_nx0(_p(g), len(g), ctypes.c_float(float(np.float32(x)))))


def nh_cooling(table, pos, rho, u, log_nh, dt, unit_density, unit_u, unit_time, XH=0.7, u_floor=1.8e11,
               gamma=5.0 / 3.0, typ=None):
    """doCloudyHCooling (CR/v4t.h:1767-1884) over a table dict with keys dims (n_nH, n_T, n_r,
    n_NH), bounds (8), nH, T, r, NH (grids), heat, cool, mu, logu ([nH][T][r][NH]); inputs are
    rounded to float32 as the reference holds them.  Returns (u_out float32, n_floored)."""
    dims = np.asarray(table["dims"], np.int32)
    b = np.asarray(table["bounds"], np.float32)
    g = np.concatenate([np.asarray(table[k], np.float32).ravel() for k in ("nH", "T", "r", "NH")])
    tabs = [np.ascontiguousarray(np.asarray(table[k], np.float32).ravel()) for k in ("heat", "cool", "mu", "logu")]
    n = len(rho)
    pos = np.ascontiguousarray(np.asarray(pos, np.float64).reshape(n, 3).astype(np.float32))
    arr = [np.ascontiguousarray(np.asarray(v, np.float64).astype(np.float32)) for v in (rho, u, log_nh)]
    t = None if typ is None else np.ascontiguousarray(typ, np.int32)
    f = np.array([unit_density, unit_u, unit_time, XH, u_floor, gamma, dt], np.float64)
    out = np.empty(n, np.float32)
    nf = lib().or_nh_cooling(_i64(n), dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(b), _p(g),
                             *[_p(x) for x in tabs], _p(pos), *[_p(x) for x in arr], _p(t), _p(f), _p(out))
    return out, int(nf)
