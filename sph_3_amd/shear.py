"""Drop-in replacement for ``sph_code/shear.py`` (Balsara-switched forms) on
one MI355X; conventions as in :mod:`sph_3_amd.libsx`."""
from __future__ import annotations

import numpy as np

from ._lib import context, f64, ptr
from .libsx import _acc, _dudt, _prec

__all__ = ["div_curlVel", "div_curlVel_mpi", "getAcc_sph_shear", "getAcc_sph_shear_mpi", "get_dU_shear",
           "get_dU_shear_mpi"]


def div_curlVel(pos, v, rho, m, h, *, precision="fp64", device=0):
    """(|div v|, |curl v|) -- sph_code/shear.py:6-69."""
    pos = f64(pos, (-1, 3)); n = len(pos)
    dv = np.empty(n); cv = np.empty(n)
    c = context(device)
    c.check(c.lib.sph_div_curl(c.h, n, ptr(pos), ptr(f64(v, (-1, 3))), ptr(f64(rho)), ptr(f64(m)), ptr(f64(h)),
                               _prec(precision), ptr(dv), ptr(cv)), "div_curlVel")
    return dv, cv


def div_curlVel_mpi(nbeg, nend, pos, v, rho, m, h, **kw):
    """Rows [nbeg, nend) -- sph_code/shear.py:76-141."""
    dv, cv = div_curlVel(pos, v, rho, m, h, **kw)
    return dv[nbeg:nend].copy(), cv[nbeg:nend].copy()


def getAcc_sph_shear(pos, v, rho, P, c, h, m, divV, curlV, alpha, *, precision="fp64", device=0):
    """Acceleration with the Balsara switch -- sph_code/shear.py:149-235."""
    return _acc(pos, v, rho, P, c, h, m, alpha, divV, curlV, precision, device)


def getAcc_sph_shear_mpi(nbeg, nend, pos, v, rho, P, c, h, m, divV, curlV, alpha, **kw):
    """Rows [nbeg, nend) -- sph_code/shear.py:241-328."""
    return getAcc_sph_shear(pos, v, rho, P, c, h, m, divV, curlV, alpha, **kw)[nbeg:nend].copy()


def get_dU_shear(pos, v, rho, P, c, h, m, divV, curlV, alpha, *, precision="fp64", device=0):
    """du/dt with the Balsara switch -- sph_code/shear.py:332-406."""
    return _dudt(pos, v, rho, P, c, h, m, alpha, divV, curlV, precision, device)


def get_dU_shear_mpi(nbeg, nend, pos, v, rho, P, c, h, m, divV, curlV, alpha, **kw):
    """Rows [nbeg, nend) -- sph_code/shear.py:411-489."""
    return get_dU_shear(pos, v, rho, P, c, h, m, divV, curlV, alpha, **kw)[nbeg:nend].copy()
