"""Drop-in replacement for ``sph_code/libsx.py`` on one MI355X.

Same names and positional signatures as the reference module (numpy float64
in, fresh numpy float64 out, inputs never mutated).  Every function runs on
the GPU through libsph_amd.so; extra behaviour is reachable only through
keyword-only arguments whose defaults follow the reference:

* ``precision``: ``"fp64"`` (the default, the reference's arithmetic) or
  ``"mixed"`` (FP32 pair arithmetic with Kahan sums, FP64
  positions/velocities/centres of mass/h solve);
* ``theta`` (gravity, potential): Barnes-Hut opening angle; the default ``0`` is
  the reference's exact direct sum; ``mac="reference"`` applies the reference
  GPU walk's opening rule size^2/theta <= d^2 instead of size < theta d;
* ``device``: GPU ordinal.

Neighbour sums are taken over S_i = { j : |r_i - r_j| < h_i + h_j }, which
is exactly the support of the reference's symmetrised kernel, so the
all-pairs sums of the reference and the neighbour-list sums here add the same
non-zero terms.

The ``*_mpi(nbeg, nend, ...)`` twins return rows ``[nbeg, nend)`` like the
reference's (``getAcc_g_smth_mimj_mpi`` returns the N x 3 partial field whose
sum over a row partition is the full field, as the reference's does).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import H_KTH, H_REF_COUNT, PREC, SOFT, context, f64, ptr, theta_of

__all__ = ["validate_particles", "getDensity", "getDensity_mpi", "getPressure", "getAcc_sph", "getAcc_sph_mpi", "get_dU", "get_dU_mpi",
           "getKE", "getPE", "getAcc_g_smth", "getAcc_g_smth_mpi", "getAcc_g_smth_mimj_mpi", "do_smoothingX",
           "smoothing_length_mpi", "h_smooth_fast", "h_smooth_fast_mpi", "neighbour_lists"]


def _prec(p):
    try:
        return PREC[p]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(PREC)}") from None


# ---------------------------------------------------------------- checks
def validate_particles(pos, vel=None, m=None, u=None, h=None):
    """The library's host-side check of a particle set (runs without a GPU): raises
    ValueError with the reason when positions are not finite or all coincide, or a
    given velocity is not finite, a mass or u negative, an h not positive."""
    from ._lib import load

    pos = f64(pos, (-1, 3)); n = len(pos)
    arr = [None if a is None else (f64(a, (-1, 3)) if k == 0 else np.broadcast_to(f64(a), (n,)).copy())
           for k, a in enumerate((vel, m, u, h))]
    msg = ctypes.create_string_buffer(256)
    rc = load().sph_validate_particles(n, ptr(pos), *(ptr(a) for a in arr), msg, 256)
    if rc != 0:
        raise ValueError(msg.value.decode())


# ---------------------------------------------------------------- density
def getDensity(pos, m, h, *, precision="fp64", device=0):
    """rho_i = sum_j m_j W(r_ij, (h_i+h_j)/2) -- sph_code/libsx.py:10-44."""
    pos = f64(pos, (-1, 3)); n = len(pos)
    out = np.empty(n)
    c = context(device)
    c.check(c.lib.sph_density(c.h, n, ptr(pos), ptr(f64(m)), ptr(f64(h)), _prec(precision), ptr(out)), "getDensity")
    return out


def getDensity_mpi(nbeg, nend, pos, m, h, **kw):
    """Rows [nbeg, nend) of getDensity -- sph_code/libsx.py:50-86."""
    return getDensity(pos, m, h, **kw)[nbeg:nend].copy()


def getPressure(rho, u, gama):
    """P = (gamma - 1) rho u -- sph_code/libsx.py:91-96 (elementwise; fused
    into the density pass in the device-resident step)."""
    return (gama - 1.0) * np.asarray(rho, dtype=np.float64) * np.asarray(u, dtype=np.float64)


# ---------------------------------------------------------------- hydro
def _acc(pos, v, rho, P, c, h, m, alpha, divV, curlV, precision, device):
    pos = f64(pos, (-1, 3)); n = len(pos)
    out = np.empty((n, 3))
    cx = context(device)
    dv = None if divV is None else f64(divV)
    cv = None if curlV is None else f64(curlV)
    cx.check(cx.lib.sph_acc_hydro(cx.h, n, ptr(pos), ptr(f64(v, (-1, 3))), ptr(f64(rho)), ptr(f64(P)), ptr(f64(c)),
                                  ptr(f64(h)), ptr(f64(m)), ptr(dv), ptr(cv), float(alpha), _prec(precision),
                                  ptr(out)), "acceleration")
    return out


def _dudt(pos, v, rho, P, c, h, m, alpha, divV, curlV, precision, device):
    pos = f64(pos, (-1, 3)); n = len(pos)
    out = np.empty(n)
    cx = context(device)
    dv = None if divV is None else f64(divV)
    cv = None if curlV is None else f64(curlV)
    cx.check(cx.lib.sph_dudt(cx.h, n, ptr(pos), ptr(f64(v, (-1, 3))), ptr(f64(rho)), ptr(f64(P)), ptr(f64(c)),
                             ptr(f64(h)), ptr(f64(m)), ptr(dv), ptr(cv), float(alpha), _prec(precision), ptr(out)),
             "du/dt")
    return out


def getAcc_sph(pos, v, rho, P, c, h, m, gama, eta, alpha, beta, *, precision="fp64", device=0):
    """Pressure + Gadget-2 artificial-viscosity acceleration, no Balsara
    switch -- sph_code/libsx.py:100-177 (gama, eta, beta unused there too)."""
    return _acc(pos, v, rho, P, c, h, m, alpha, None, None, precision, device)


def getAcc_sph_mpi(nbeg, nend, pos, v, rho, P, c, h, m, gama, eta, alpha, beta, **kw):
    """Rows [nbeg, nend) -- sph_code/libsx.py:268-349."""
    return getAcc_sph(pos, v, rho, P, c, h, m, gama, eta, alpha, beta, **kw)[nbeg:nend].copy()


def get_dU(pos, v, rho, P, c, h, m, gama, eta, alpha, beta, *, precision="fp64", device=0):
    """du/dt without Balsara switch -- sph_code/libsx.py:353-421."""
    return _dudt(pos, v, rho, P, c, h, m, alpha, None, None, precision, device)


def get_dU_mpi(nbeg, nend, pos, v, rho, P, c, h, m, gama, eta, alpha, beta, **kw):
    """Rows [nbeg, nend) -- sph_code/libsx.py:496-565."""
    return get_dU(pos, v, rho, P, c, h, m, gama, eta, alpha, beta, **kw)[nbeg:nend].copy()


# ---------------------------------------------------------------- energies
def getKE(v, m, *, device=0):
    """KE = sum 1/2 m v^2 (FP64 device reduction) -- sph_code/libsx.py:569-579."""
    v = f64(v, (-1, 3)); n = len(v)
    out = np.zeros(5)
    c = context(device)
    c.check(c.lib.sph_kinetics(c.h, n, ptr(v), ptr(f64(m)), ptr(out)), "getKE")
    return float(out[0])


def getPE(pos, m, G, epsilon, *, theta=0.0, mac="geometric", softening="spline", eps_plummer=0.0,
          precision="fp64", device=0, return_phi=False):
    """PE = -G sum_{i<j} m_i m_j psi(r_ij, eps_ij), the spline potential with
    eps_ij = (eps_i + eps_j)/2 -- sph_code/libsx.py:583-620 -- computed on the device as
    -1/2 sum_i m_i phi_i over the tree walk (``theta=0``: every pair, the reference's
    double sum; FP64 total).  ``softening="plummer"`` uses the scalar ``eps_plummer``.
    ``return_phi``: also phi_i = G sum_j m_j psi(r_ij)."""
    pos = f64(pos, (-1, 3)); n = len(pos)
    eps = None if epsilon is None else np.broadcast_to(f64(epsilon), (n,)).copy()
    pe = np.zeros(1)
    phi = np.empty(n) if return_phi else None
    c = context(device)
    c.check(c.lib.sph_potential_energy(c.h, n, ptr(pos), ptr(f64(m)), ptr(eps), float(G), theta_of(theta, mac),
                                       SOFT[softening], float(eps_plummer), _prec(precision), ptr(pe), ptr(phi)),
            "getPE")
    return (float(pe[0]), phi) if return_phi else float(pe[0])


# ---------------------------------------------------------------- gravity
def getAcc_g_smth(pos, mass, G, epsilon, *, theta=0.0, mac="geometric", softening="spline", eps_plummer=0.0,
                  precision="fp64", device=0, return_counts=False):
    """Spline-softened self-gravity, eps_ij = (eps_i + eps_j)/2 --
    sph_code/libsx.py:623-662.  The default ``theta=0`` is the reference's exact
    direct sum; ``theta > 0`` is a Barnes-Hut tree walk at that opening angle
    (``mac="reference"``: the reference GPU walk's rule, EV/bh.h:505-532).
    ``softening="plummer"`` uses the scalar ``eps_plummer`` instead."""
    pos = f64(pos, (-1, 3)); n = len(pos)
    eps = None if epsilon is None else np.broadcast_to(f64(epsilon), (n,)).copy()
    out = np.empty((n, 3))
    cnt = np.zeros(2, np.int64)
    c = context(device)
    c.check(c.lib.sph_gravity(c.h, n, ptr(pos), ptr(f64(mass)), ptr(eps), float(G), theta_of(theta, mac), SOFT[softening],
                              float(eps_plummer), _prec(precision), ptr(out), ptr(cnt)), "getAcc_g_smth")
    if return_counts:
        return out, (int(cnt[0]), int(cnt[1]))
    return out


def getAcc_g_smth_mpi(nbeg, nend, pos, mass, G, epsilon, **kw):
    """Rows [nbeg, nend) -- sph_code/libsx.py:667-706."""
    return getAcc_g_smth(pos, mass, G, epsilon, **kw)[nbeg:nend].copy()


def getAcc_g_smth_mimj_mpi(nbeg, nend, pos, mass, G, epsilon, **kw):
    """N x 3 field holding rows [nbeg, nend) of the full field and zeros
    elsewhere, so a sum over a row partition gives the full field, as for
    sph_code/libsx.py:908-947."""
    full = getAcc_g_smth(pos, mass, G, epsilon, **kw)
    out = np.zeros_like(full)
    out[nbeg:nend] = full[nbeg:nend]
    return out


# ---------------------------------------------------------------- h
def _hsolve(pos, targets, h_in, mode, param, n_low=0, n_up=0, max_steps=0, tol=1e-12, device=0):
    pos = f64(pos, (-1, 3)); n = len(pos)
    tg = None if targets is None else f64(targets, (-1, 3))
    m = n if tg is None else len(tg)
    hin = None if h_in is None else f64(h_in)
    out = np.empty(m)
    unc = np.zeros(1, np.int64)
    c = context(device)
    c.check(c.lib.sph_smoothing_length(c.h, n, ptr(pos), m, ptr(tg), ptr(hin), mode, float(param), int(n_low),
                                       int(n_up), int(max_steps), float(tol), ptr(out), ptr(unc)),
            "smoothing length")
    return out, int(unc[0])


def do_smoothingX(poz, *, device=0):
    """h = 0.5 * (51st smallest distance to all particles, self included) for
    every point of ``poz[1]`` -- sph_code/libsx.py:752-774."""
    pos, subpos = poz
    pos = f64(pos, (-1, 3)); subpos = f64(subpos, (-1, 3))
    same = subpos.shape == pos.shape and np.array_equal(subpos, pos)
    h, _ = _hsolve(pos, None if same else subpos, None, H_KTH, 50, device=device)
    return h


def smoothing_length_mpi(nbeg, nend, pos, *, device=0):
    """Rows [nbeg, nend) of do_smoothingX((pos, pos)) -- sph_code/libsx.py:779-800."""
    pos = f64(pos, (-1, 3))
    h, _ = _hsolve(pos, pos[nbeg:nend], None, H_KTH, 50, device=device)
    return h


def h_smooth_fast(pos, h, *, n_low=45, n_up=55, max_steps=2000, device=0, return_unconverged=False):
    """Step h by -/+0.3 % until n_low <= #(|r_ij| < 2h) <= n_up --
    sph_code/libsx.py:804-845.  The reference loop has no iteration cap; here
    rows that reach ``max_steps`` keep their last h and are counted."""
    out, unc = _hsolve(pos, None, h, H_REF_COUNT, 0, n_low, n_up, max_steps, device=device)
    return (out, unc) if return_unconverged else out


def h_smooth_fast_mpi(nbeg, nend, pos, h, **kw):
    """Rows [nbeg, nend) -- sph_code/libsx.py:852-900."""
    pos = f64(pos, (-1, 3)); h = f64(h)
    device = kw.pop("device", 0)
    out, _ = _hsolve(pos, pos[nbeg:nend], h[nbeg:nend], H_REF_COUNT, 0, kw.get("n_low", 45), kw.get("n_up", 55),
                     kw.get("max_steps", 2000), device=device)
    return out


# ---------------------------------------------------------------- neighbours
def neighbour_lists(pos, h, *, device=0):
    """CSR (offsets int64[N+1], idx int32) of S_i = { j : |r_ij|^2 < (h_i+h_j)^2 },
    rows ascending, self included (the neighbour search of EV/ngb.h:125-264
    with the symmetric support of the kernel sums)."""
    pos = f64(pos, (-1, 3)); n = len(pos); h = f64(h)
    off = np.zeros(n + 1, np.int64)
    c = context(device)
    c.check(c.lib.sph_neighbours(c.h, n, ptr(pos), ptr(h), ptr(off), None, 0), "neighbour_lists")
    idx = np.empty(int(off[-1]), np.int32)
    c.check(c.lib.sph_neighbours(c.h, n, ptr(pos), ptr(h), ptr(off), ptr(idx), len(idx)), "neighbour_lists")
    return off, idx
