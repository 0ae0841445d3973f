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


@pytest.mark.parametrize("softening", ["plummer", "spline"])
def test_timestep_criteria_with_gravity(gpu, softening):
    """With G != 0 the dt_f criterion takes |a_grav| and the gravity criteria take the
    softening length: eps_plummer for Plummer, h_i for the spline.  a_grav comes from the
    oracle's group walk with the device's tree parameters (FP32 pair terms on the device:
    1e-5 relative).  At rest with a large Courant factor the acceleration criteria bind:
    dt_f = sqrt(h/|a_grav|) for Plummer with eps = 1 (for about
    12 % of the particles), 0.4 sqrt(h/|a|) for the spline (most of them)."""
    n = 20000
    d = ic.evrard(n, seed=62)
    vel = np.zeros((n, 3))
    cfg = StepConfig(G=1.0, softening=softening, eps_plummer=1.0, C_CFL=30.0)
    sim = DeviceSPH(d["pos"], vel, d["m"], d["u"], d["h0"], cfg)
    sim.prime()
    g = sim.get()
    dt, dh = sim.timestep()
    spline = softening == "spline"
    tree = O.Tree(g["pos"], d["m"], h=g["h"], eps=g["h"] if spline else None)
    ag, _ = tree.gravity(theta=cfg.theta, G=cfg.G, soft_kind=O.SOFT_SPLINE if spline else O.SOFT_PLUMMER,
                         eps_plummer=cfg.eps_plummer)
    off, idx = tree.neighbours()
    c = np.sqrt(cfg.gamma * (cfg.gamma - 1) * d["u"])
    dv, _ = O.divcurl_csr(g["pos"], g["vel"], g["rho"], d["m"], g["h"], off, idx)
    eps = g["h"] if spline else cfg.eps_plummer
    ref, ref_dh = O.timestep(g["pos"], g["vel"], g["h"], c, dv, g["acc"], ag, eps, cfg.C_CFL, cfg.alpha, off, idx)
    rel = np.abs(dt / ref - 1)
    print(softening, "dt rel max", rel.max())
    assert rel.max() < 1e-5
    assert np.max(np.abs(dh - ref_dh)) <= 1e-5 * np.max(np.abs(ref_dh))
    at = np.linalg.norm(g["acc"], axis=1)
    if spline:  # dt_G1 with eps = h_i binds
        assert np.mean(np.isclose(ref, 0.4 * np.sqrt(g["h"] / at), rtol=1e-12)) > 0.5
    else:  # dt_f binds, and |a_grav| differs from |a|
        assert np.mean(np.isclose(ref, np.sqrt(g["h"] / np.linalg.norm(ag, axis=1)), rtol=1e-12)) > 0.05
        ref_tot, _ = O.timestep(g["pos"], g["vel"], g["h"], c, dv, g["acc"], g["acc"], eps, cfg.C_CFL, cfg.alpha,
                                off, idx)
        assert np.max(np.abs(ref_tot / ref - 1)) > 1e-2


def test_timestep_criteria_periodic(gpu):
    """Periodic box (minimum-image pair offsets), Sedov lattice with random motion."""
    d = ic.sedov(20, seed=52)
    vel = np.random.default_rng(8).normal(size=d["pos"].shape) * 0.1
    cfg = StepConfig(dt=1e-4, G=0.0, period=1.0, C_CFL=0.3, eps_plummer=1e-3)
    sim = DeviceSPH(d["pos"], vel, d["m"], d["u"], d["h0"], cfg)
    sim.prime()
    g = sim.get()
    dt, dh = sim.timestep()
    c = np.sqrt(cfg.gamma * (cfg.gamma - 1) * d["u"])
    with O.periodic(1.0):
        off, idx = O.neighbours_brute(g["pos"], g["h"])
        dv, _ = O.divcurl_csr(g["pos"], g["vel"], g["rho"], d["m"], g["h"], off, idx)
        ref, ref_dh = O.timestep(g["pos"], g["vel"], g["h"], c, dv, g["acc"], g["acc"], cfg.eps_plummer, cfg.C_CFL,
                                 cfg.alpha, off, idx)
    rel = np.abs(dt / ref - 1)
    print("periodic dt rel max", rel.max())
    assert rel.max() < 1e-5
    assert np.max(np.abs(dh - ref_dh)) <= 1e-5 * np.max(np.abs(ref_dh))


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
@pytest.mark.parametrize("balsara", [True, False])
def test_timestep_criteria_after_a_step(gpu, prec, balsara):
    """After sph_step the criteria take v_sig from the velocities after the second kick and
    |div v| from the force evaluation's velocities (before that kick), with and without the
    Balsara switch, as the reference computes div v in the force pass (MC/drv.cu:1271) and the
    criteria after the kick (:1610).  c is the evaluation's sound speed (u before its update)."""
    n = 20000
    d = ic.evrard(n, seed=63)
    vel = np.random.default_rng(7).normal(size=(n, 3)) * 0.1
    cfg = StepConfig(G=0.0, precision=prec, balsara=balsara, C_CFL=0.3, eps_plummer=1e-3)
    sim = DeviceSPH(d["pos"], vel, d["m"], d["u"], d["h0"], cfg)
    sim.prime()
    u_eval = sim.get()["u"]  # u of the next evaluation (the step updates it after its forces)
    sim.step()
    g = sim.get()
    dt, dh = sim.timestep()
    v_half = g["vel"] - g["acc"] * (cfg.dt / 2)  # the evaluation's velocities (before the second kick)
    off, idx = O.Tree(g["pos"], d["m"], h=g["h"]).neighbours()
    c = np.sqrt(cfg.gamma * (cfg.gamma - 1) * u_eval)
    dv, _ = O.divcurl_csr(g["pos"], v_half, g["rho"], d["m"], g["h"], off, idx)
    ref, ref_dh = O.timestep(g["pos"], g["vel"], g["h"], c, dv, g["acc"], g["acc"], cfg.eps_plummer, cfg.C_CFL,
                             cfg.alpha, off, idx)
    rel = np.abs(dt / ref - 1)
    tol = 1e-5 if prec == "mixed" else 1e-10
    print(prec, balsara, "after a step: dt rel max", rel.max())
    assert rel.max() < tol
    assert np.max(np.abs(dh - ref_dh)) <= tol * np.max(np.abs(ref_dh))
