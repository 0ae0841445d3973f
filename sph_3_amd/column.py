"""Column density N_H from the density gradient, on one MI355X.

Replaces the reference's ``grad_rho_norm_ngb`` kernel
(Particles_in_BH_Tree_III/CLOUDY_reloaded_Here/run_simulation/hfvCLibs_v7_hfv_v8_gradRho_v4t.h:716-786)
and the host loop that turns it into log10 N_H
(…/run_simulation/hfvGpuSphBHtree_gradRho_v3t.cu:1214-1224)::

    g_i   = | sum_{j in S_i} m_j grad W(r_ij, (h_i + h_j) / 2) |
    n_H,i = rho_i * unit_density_cgs * XH / m_H            (m_H = 1.673534e-24 g)
    N_H,i = 0.5 * (h_i / cbrt(n_ngb) + rho_i / g_i) * unit_length_cm * n_H,i

Both run on the device (the reference copies rho, g and h to the host and
back every step).  S_i = { j : |r_i - r_j| < h_i + h_j }, the support of the
kernel the reference's loop tests (q <= 2 with hbar_ij).  Pair sums in the
precision's type (``"mixed"``: FP32 with Kahan sums; ``"fp64"``), the N_H
arithmetic in FP64.  A particle whose gradient vanishes gets N_H = inf, as in
the reference's loop.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import PREC, ColumnParams as _CP, context, f64, ptr

M_H_G = 1.673534e-24  # hydrogen atom mass [g] of the reference's formula


@dataclass
class ColumnParams:
    """Units and constants of the N_H formula (the reference's params.txt values
    UnitDensity_in_cgs and unitLength_in_cm; XH = 0.7 and Nngb_f = 64 there)."""
    unit_density_cgs: float
    unit_length_cm: float
    XH: float = 0.7
    n_ngb: float = 64.0

    @classmethod
    def from_mass_length(cls, mass_g: float, length_cm: float, **kw) -> "ColumnParams":
        """Code units of mass and length in cgs: density unit M / L^3."""
        return cls(mass_g / length_cm ** 3, length_cm, **kw)

    def c(self) -> _CP:
        return _CP(self.unit_density_cgs, self.unit_length_cm, self.XH, self.n_ngb)


def _prec(p):
    try:
        return PREC[p]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(PREC)}") from None


def column_density(pos, m, h, rho, params: ColumnParams, *, precision="fp64", device=0, with_grad=False):
    """log10 N_H [cm^-2] per particle (and |grad rho| in code units when
    ``with_grad``), caller order."""
    pos = f64(pos, (-1, 3)); n = len(pos)
    lognh = np.empty(n)
    grad = np.empty(n) if with_grad else None
    c = context(device)
    cp = params.c()
    c.check(c.lib.sph_column_density(c.h, n, ptr(pos), ptr(f64(m)), ptr(f64(h)), ptr(f64(rho)), ctypes.byref(cp),
                                     _prec(precision), ptr(grad) if with_grad else None, ptr(lognh)),
            "column_density")
    return (lognh, grad) if with_grad else lognh


def grad_rho_norm_ngb(pos, m, h, *, precision="fp64", device=0):
    """|grad rho|_i = |sum_j m_j grad W_ij| (code units), caller order -- the
    reference kernel of the same name."""
    pos = f64(pos, (-1, 3))
    rho = np.ones(len(pos))  # only enters N_H, not the gradient
    _, g = column_density(pos, m, h, rho, ColumnParams(1.0, 1.0), precision=precision, device=device,
                          with_grad=True)
    return g
