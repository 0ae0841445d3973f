"""GPU parity of the per-particle time-step criteria (dt_array_indiv_dt with the
signal velocity's maximum over S_i) against the FP64 oracle, from the state of
the priming force evaluation: c from the initial u, div v from the oracle with
the device's rho, the device's acceleration; no gravity (a_g = a_tot), Plummer
eps.  Tolerance: relative 1e-5 for the mixed form (FP32 pair terms), 1e-12 for
FP64; sph_stats.dt_min is the minimum of the per-particle values."""
import numpy as np
import pytest

import oracle as O
from sph_3_amd import ic
from sph_3_amd.step import DeviceSPH, StepConfig

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
@pytest.mark.parametrize("balsara", [True, False])
def test_timestep_criteria_match_oracle(gpu, prec, balsara):
    n = 20000
    d = ic.evrard(n, seed=61)
    rng = np.random.default_rng(6)
    vel = rng.normal(size=(n, 3)) * 0.1
    cfg = StepConfig(G=0.0, precision=prec, balsara=balsara, C_CFL=0.3, eps_plummer=1e-3)
    sim = DeviceSPH(d["pos"], vel, d["m"], d["u"], d["h0"], cfg)
    sim.prime()
    g = sim.get()
    dt, dh = sim.timestep()
    st = sim.stats()
    off, idx = O.Tree(g["pos"], d["m"], h=g["h"]).neighbours()
    c = np.sqrt(cfg.gamma * (cfg.gamma - 1) * d["u"])
    dv, _ = O.divcurl_csr(g["pos"], g["vel"], g["rho"], d["m"], g["h"], off, idx)
    ref, ref_dh = O.timestep(g["pos"], g["vel"], g["h"], c, dv, g["acc"], g["acc"], cfg.eps_plummer, cfg.C_CFL,
                             cfg.alpha, off, idx)
    rel = np.abs(dt / ref - 1)
    tol = 1e-5 if prec == "mixed" else 1e-12
    print(prec, balsara, "dt rel max", rel.max(), "dt_min", st["dt_min"], ref.min())
    assert rel.max() < tol
    assert np.max(np.abs(dh - ref_dh)) <= tol * np.max(np.abs(ref_dh))
    assert st["dt_min"] == dt.min()
    assert abs(st["dt_min"] / ref.min() - 1) < tol
