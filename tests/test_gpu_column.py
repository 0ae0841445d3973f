"""GPU parity of the column density (grad_rho_norm_ngb and the N_H loop of the
CLOUDY driver) against the FP64 oracle, stateless and in the step.

Tolerances: |grad rho| normalised max error (max |dg| / rms g) 1e-4 for the
mixed form, 1e-9 for FP64; log10 N_H: 99.9th percentile of |d log N_H|
1e-4 dex for the mixed form (a particle with a tiny gradient amplifies the
FP32 error of g into rho / g), 1e-9 dex for FP64.
"""
import numpy as np
import pytest

import oracle as O
from sph_3_amd import column as C
from sph_3_amd import ic
from sph_3_amd.step import DeviceSPH, StepConfig

pytestmark = pytest.mark.gpu

UNITS = C.ColumnParams.from_mass_length(1.989e43, 3.0857e22)  # 1e10 M_sun, 10 kpc


def oracle_column(pos, m, h, rho, cp):
    off, idx = O.Tree(pos, m, h=h).neighbours()
    g = O.grad_rho_csr(pos, m, h, off, idx)
    return O.column_density(rho, g, h, cp.unit_density_cgs, cp.unit_length_cm, cp.XH, cp.n_ngb), g


def check(lognh, g, lognh_ref, g_ref, prec, label):
    eg = np.max(np.abs(g - g_ref)) / np.sqrt(np.mean(g_ref ** 2))
    dl = np.abs(lognh - lognh_ref)
    print(label, prec, "grad", eg, "logNH p99.9", np.quantile(dl, 0.999), "max", dl.max())
    tol = 1e-4 if prec == "mixed" else 1e-9
    assert eg < tol
    assert np.quantile(dl, 0.999) < tol
    assert np.all(np.isfinite(lognh))


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_column_density_stateless(gpu, prec):
    n = 20000
    d = ic.evrard(n, seed=31)
    h, bad = O.h_newton(d["pos"], d["h0"], n_ngb=64.0)
    assert bad == 0
    off, idx = O.Tree(d["pos"], d["m"], h=h).neighbours()
    rho = O.density_csr(d["pos"], d["m"], h, off, idx)
    lognh, g = C.column_density(d["pos"], d["m"], h, rho, UNITS, precision=prec, with_grad=True)
    ref, g_ref = oracle_column(d["pos"], d["m"], h, rho, UNITS)
    check(lognh, g, ref, g_ref, prec, "stateless")
    assert np.allclose(C.grad_rho_norm_ngb(d["pos"], d["m"], h, precision=prec), g, rtol=0, atol=0)


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_column_density_in_step(gpu, prec):
    n = 20000
    d = ic.evrard(n, seed=32)
    sim = DeviceSPH(d["pos"], d["vel"], d["m"], d["u"], d["h0"], StepConfig(precision=prec))
    sim.set_column(UNITS)
    sim.prime()
    sim.step()
    g = sim.get()
    lognh, grad = sim.column()
    ref, g_ref = oracle_column(g["pos"], d["m"], g["h"], g["rho"], UNITS)
    check(lognh, grad, ref, g_ref, prec, "step")
    sim.reset_stats()
    sim.step(timing=True)
    assert sim.stats()["phase_calls"]["column"] == 1
    sim.set_column(None)
    sim.step(timing=True)
    assert sim.stats()["phase_calls"]["column"] == 1
