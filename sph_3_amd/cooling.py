"""Radiative cooling on the GPU: drop-in for the reference's closed-form
cooling module (Particles_in_BH_Tree_III/hcool_with_unit_tests_Here/
photolibs3.py) plus the 4D cooling table of the CR driver
(CR/v4t.h:1768-1884 doCloudyHCooling) built from the same physics.

All functions accept scalars or arrays (cgs units) and evaluate elementwise on
the device in FP64.  Keyword-only extensions: ``J21`` (UV field amplitude; the
reference's fixed value is 1e-7) and ``Z`` (metallicity in solar units, adds
the Rosen & Bregman 1995 metal-line term; 0 = the reference's primordial gas).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import context, f64, ptr

__all__ = ["coolingHeatingRates", "convert_u_to_temp_h", "coolingRateFromU_h", "DoCooling_h", "CoolingTable",
           "J21_REFERENCE", "TABLE_RANGES"]

J21_REFERENCE = 1e-7
# (lo, hi, points) per table axis: log10 nH, log10 T, log10 Z, log10 J21
TABLE_RANGES = ((-4.0, 4.0, 33), (2.0, 8.0, 61), (-3.0, 0.0, 13), (-2.0, 2.0, 17))


def _vec(*args):
    arrs = np.broadcast_arrays(*[np.asarray(a, dtype=np.float64) for a in args])
    scalar = arrs[0].ndim == 0
    return scalar, [np.ascontiguousarray(a.reshape(-1)) for a in arrs]


def _opt(v, n):
    """(array-or-None, scalar) pair for an optional per-element argument."""
    a = np.asarray(v, dtype=np.float64)
    if a.ndim == 0:
        return None, float(a)
    return np.ascontiguousarray(np.broadcast_to(a, (n,))), 0.0


def coolingHeatingRates(T, nHcgs, *, J21=J21_REFERENCE, Z=0.0, device=0):
    """(heating, cooling) per nH^2 [erg cm^3 s^-1] -- photolibs3.py:332-376."""
    scalar, (T, nH) = _vec(T, nHcgs)
    n = len(T)
    J, Jc = _opt(J21, n)
    Za, Zc = _opt(Z, n)
    h = np.empty(n); l = np.empty(n)
    c = context(device)
    c.check(c.lib.sph_cooling_rates(c.h, n, ptr(T), ptr(nH), ptr(J), Jc, ptr(Za), Zc, ptr(h), ptr(l)),
            "coolingHeatingRates")
    return (float(h[0]), float(l[0])) if scalar else (h, l)


def _from_u(what, u, nHcgs, XH, J21, Z, device):
    scalar, (u, nH) = _vec(u, nHcgs)
    n = len(u)
    J, Jc = _opt(J21, n)
    Za, Zc = _opt(Z, n)
    out = np.empty(n)
    c = context(device)
    c.check(c.lib.sph_cooling_from_u(c.h, n, what, ptr(u), ptr(nH), float(XH), ptr(J), Jc, ptr(Za), Zc, ptr(out)),
            "cooling")
    return float(out[0]) if scalar else out


def convert_u_to_temp_h(u, nHcgs, XH, *, J21=J21_REFERENCE, device=0):
    """T(u) with the ionisation state of the gas -- photolibs3.py:379-420."""
    return _from_u(0, u, nHcgs, XH, J21, 0.0, device)


def coolingRateFromU_h(u, nHcgs, XH, *, J21=J21_REFERENCE, Z=0.0, device=0):
    """Gamma - Lambda per nH^2 at T(u) clamped to [1e4, 1e8] K -- photolibs3.py:470-486."""
    return _from_u(1, u, nHcgs, XH, J21, Z, device)


def DoCooling_h(rho, u_old, dt, XH, *, J21=J21_REFERENCE, Z=0.0, u_floor=1e8, device=0):
    """Implicit u update over dt -- photolibs3.py:489-547.  The reference's
    downward bracket search has no floor (it overflows for strong cooling);
    here such elements return -1 (bracket reached ``u_floor``)."""
    scalar, (rho, u) = _vec(rho, u_old)
    n = len(rho)
    dta, dtc = _opt(dt, n)
    J, Jc = _opt(J21, n)
    Za, Zc = _opt(Z, n)
    out = np.empty(n)
    c = context(device)
    c.check(c.lib.sph_do_cooling(c.h, n, ptr(rho), ptr(u), ptr(dta), dtc, float(XH), ptr(J), Jc, ptr(Za), Zc,
                                 float(u_floor), ptr(out)), "DoCooling_h")
    return float(out[0]) if scalar else out


class CoolingTable:
    """4D table of log10 Lambda and log10 Gamma (per nH^2) over (log10 nH,
    log10 T, log10 Z, log10 J21), built on the device from the exact function;
    lookup is multilinear in log space with inputs clamped to the grid."""

    def __init__(self, ranges=TABLE_RANGES, device=0):
        self.ranges = tuple(tuple(r) for r in ranges)
        lo = np.array([r[0] for r in ranges], np.float64)
        hi = np.array([r[1] for r in ranges], np.float64)
        npts = (ctypes.c_int * 4)(*[int(r[2]) for r in ranges])
        self.ctx = context(device)
        c = self.ctx
        c.check(c.lib.sph_cooling_table_build(c.h, ptr(lo), ptr(hi), npts), "cooling table build")
        self.nbytes = 2 * 4 * int(np.prod([r[2] for r in ranges]))

    def lookup(self, lognH, logT, logZ, logJ21):
        _, (a, b, z, u) = _vec(lognH, logT, logZ, logJ21)
        n = len(a)
        L = np.empty(n); G = np.empty(n)
        c = self.ctx
        c.check(c.lib.sph_cooling_table_lookup(c.h, n, ptr(a), ptr(b), ptr(z), ptr(u), ptr(L), ptr(G)),
                "cooling table lookup")
        return L, G
