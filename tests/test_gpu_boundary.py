"""GPU checks of the drop-in boundary: reference-compatible defaults, getPE, the
reference opening rule, the step's FP64 energy and momentum totals, and input
validation before any launch.

Tolerances: FP64 form against the oracle's FP64 sums 1e-10 relative (different
summation order); the mixed form's potential 1e-5 relative (FP32 pair terms with
compensated sums, FP64 total).
"""
import numpy as np
import pytest

import oracle as O
from oracle.step import KDK
from sph_3_amd import ic
from sph_3_amd import libsx as X
from sph_3_amd.step import DeviceSPH, StepConfig

pytestmark = pytest.mark.gpu


def test_getacc_g_smth_default_is_the_reference_direct_sum(gpu):
    """getAcc_g_smth(pos, m, G, eps) with no keywords is the reference's exact symmetric
    direct sum in FP64 (sph_code/libsx.py:623-662)."""
    d = ic.evrard(6000, seed=71)
    eps = np.random.default_rng(1).uniform(0.01, 0.05, 6000)
    a = X.getAcc_g_smth(d["pos"], d["m"], 1.0, eps)
    a0 = O.grav_direct(d["pos"], d["m"], eps)
    rel = np.linalg.norm(a - a0, axis=1) / np.linalg.norm(a0, axis=1)
    print("default getAcc_g_smth vs direct sum: max rel", rel.max())
    assert rel.max() < 1e-11


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("soft", ["spline", "plummer"])
def test_getpe_matches_oracle(gpu, prec, soft):
    d = ic.evrard(6000, seed=72)
    n = len(d["m"])
    eps = np.random.default_rng(2).uniform(0.01, 0.05, n) if soft == "spline" else None
    kw = dict(soft_kind=O.SOFT_SPLINE if soft == "spline" else O.SOFT_PLUMMER, eps_plummer=0.02)
    pe0 = O.pe(d["pos"], d["m"], 0.7, eps, **kw)
    pe, phi = X.getPE(d["pos"], d["m"], 0.7, eps, softening=soft, eps_plummer=0.02, precision=prec, return_phi=True)
    tol = 1e-10 if prec == "fp64" else 1e-5
    print(prec, soft, "PE", pe, pe0, "rel", abs(pe / pe0 - 1))
    assert abs(pe / pe0 - 1) < tol
    assert abs(-0.5 * np.sum(d["m"] * phi) / pe0 - 1) < tol
    # Barnes-Hut potential at theta = 0.5: an approximation of the same sum
    pe5 = X.getPE(d["pos"], d["m"], 0.7, eps, theta=0.5, softening=soft, eps_plummer=0.02, precision=prec)
    assert abs(pe5 / pe0 - 1) < 1e-3


def test_reference_opening_rule(gpu):
    """mac="reference" is the reference walk's size^2/theta <= d^2, i.e. the geometric rule
    at sqrt(theta): same interaction counts as the oracle walk at sqrt(theta)."""
    d = ic.evrard(30000, seed=73)
    a, cnt = X.getAcc_g_smth(d["pos"], d["m"], 1.0, None, theta=0.5, mac="reference", softening="plummer",
                             eps_plummer=1e-3, return_counts=True)
    a0, cnt0 = O.Tree(d["pos"], d["m"]).gravity(theta=np.sqrt(0.5), soft_kind=O.SOFT_PLUMMER, eps_plummer=1e-3)
    _, cnt5 = O.Tree(d["pos"], d["m"]).gravity(theta=0.5, soft_kind=O.SOFT_PLUMMER, eps_plummer=1e-3)
    assert cnt == cnt0 and sum(cnt0) < sum(cnt5)
    assert np.max(np.linalg.norm(a - a0, axis=1) / np.linalg.norm(a0, axis=1)) < 1e-11


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_step_energy_and_momentum_totals(gpu, prec):
    """sph_stats energy/momentum after the priming evaluation and after steps equal the
    oracle's FP64 sums over the device's own state (theta = 0: the exact potential), and
    the oracle KDK's totals after the same steps."""
    n = 10000
    d = ic.evrard(n, seed=74)
    vel = np.random.default_rng(3).normal(size=(n, 3)) * 0.05
    cfg = StepConfig(precision=prec, theta=0.0, energy=True, h_tol=1e-12)
    sim = DeviceSPH(d["pos"], vel, d["m"], d["u"], d["h0"], cfg)
    ref = KDK(d["pos"], vel, d["m"], d["u"], d["h0"], theta=0.0, h_tol=1e-13)
    m = d["m"]
    sim.prime(); ref.prime()
    tol = 1e-10 if prec == "fp64" else 1e-5
    for k in range(4):
        if k:
            sim.step(); ref.step()
        st, g = sim.stats(), sim.get()
        e = st["energy"]
        ke0, u0 = O.ke(g["vel"], m), np.sum(m * g["u"])
        pe0 = O.pe(g["pos"], m, 1.0, None, O.SOFT_PLUMMER, cfg.eps_plummer)
        p0 = (m[:, None] * g["vel"]).sum(0)
        L0 = (m[:, None] * np.cross(g["pos"], g["vel"])).sum(0)
        print(prec, k, e, "pe rel", abs(e["potential"] / pe0 - 1))
        assert abs(e["kinetic"] / ke0 - 1) < 1e-12 and abs(e["thermal"] / u0 - 1) < 1e-12
        assert abs(e["potential"] / pe0 - 1) < tol
        pscale = np.sum(m * np.linalg.norm(g["vel"], axis=1))
        assert np.max(np.abs(st["momentum"] - p0)) < 1e-12 * pscale
        assert np.max(np.abs(st["ang_momentum"] - L0)) < 1e-12 * pscale
        # against the oracle's independent evolution
        o = ref.state()
        etot = O.ke(o["vel"], m) + np.sum(m * o["u"]) + O.pe(o["pos"], m, 1.0, None, O.SOFT_PLUMMER, cfg.eps_plummer)
        assert abs(e["total"] / etot - 1) < tol


def test_invalid_particles_are_refused_before_launch(gpu):
    d = ic.evrard(2000, seed=75)
    pos = d["pos"].copy()
    pos[17, 1] = np.nan
    with pytest.raises(RuntimeError, match="position"):
        DeviceSPH(pos, d["vel"], d["m"], d["u"], d["h0"])
    with pytest.raises(RuntimeError, match="coincide"):
        DeviceSPH(np.ones((50, 3)), np.zeros((50, 3)), np.ones(50), np.ones(50), np.ones(50))
    h = d["h0"].copy(); h[3] = 0.0
    with pytest.raises(RuntimeError, match="h"):
        X.getDensity(d["pos"], d["m"], h)
    with pytest.raises(RuntimeError, match="n_ngb"):
        DeviceSPH(d["pos"][:5], d["vel"][:5], d["m"][:5], d["u"][:5], d["h0"][:5]).prime()
    # the context stays usable after a refused call
    rho = X.getDensity(d["pos"], d["m"], d["h0"])
    assert np.all(np.isfinite(rho))


def test_device_footprint_query(gpu, monkeypatch):
    """sph_device_bytes sums every buffer a context holds: on a fresh context it grows
    in proportion to n, and the device's free memory drops by no more than that sum
    (plus 64 MiB for the runtime's own objects) while the context fills."""
    import ctypes
    import gc
    import sph_3_amd.step as S
    from sph_3_amd._lib import Context
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library is linked against

    def free_bytes():
        f = ctypes.c_size_t(0); t = ctypes.c_size_t(0)
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value

    got = {}
    for n in (50_000, 100_000):
        d = ic.evrard(n, seed=76)
        gc.collect()
        free0 = free_bytes()
        ctx = Context(0)
        monkeypatch.setattr(S, "context", lambda device=0: ctx)
        sim = S.DeviceSPH(d["pos"], d["vel"], d["m"], d["u"], d["h0"])
        sim.step()
        b = sim.device_bytes()
        drop = free0 - free_bytes()
        print(f"n={n}: {b} bytes ({b / n:.0f} per particle), free memory dropped by {drop}")
        got[n] = b
        assert 200 * n < b < 20_000 * n
        assert drop <= b + (64 << 20)
        del sim
        ctx.close()
    assert 1.6 < got[100_000] / got[50_000] < 2.4
