"""GPU parity of whole KDK steps (sph_serial_v2.py:94-203) against the oracle's own
FP64 evolution (oracle/step.py) on the 10,000-particle Evrard sphere.

Device and oracle start from the same particles and each integrates on its own:
priming evaluation, then 1 or 3 steps of kick, drift, forces (Newton h, neighbour
sets, density, div/curl, Balsara-switched acceleration, Plummer tree gravity at
theta = 0.5), kick, du/dt with the kicked velocity, trapezoidal u update.

Errors are measured against what the step changes, so that a mistake in the
integrator (a drift before the kick, a missing half kick, du/dt before the kick,
an Euler u update) fails even where it moves x or u by a tiny fraction of their
values:
  e_x = max |x_dev - x_or| / rms |x_or - x_0|      (displacement)
  e_v = max |v_dev - v_or| / rms |v_or|
  e_u = max |u_dev - u_or| / rms |u_or - u_0|      (change of u)
  e_rho = max |rho_dev / rho_or - 1|
Tolerances: FP64 form 1e-8 on these (the smoothing lengths are solved to 1e-12 on
the device and 1e-13 in the oracle), and additionally |dx|/max|x|, |dv|/max|v| and
|du|/|u| below 1e-10; mixed form 1e-4 on these, and |dx|/max|x|, |dv|/max|v| below
1e-6.
"""
import numpy as np
import pytest

from oracle.step import KDK
from sph_3_amd import ic
from sph_3_amd.step import DeviceSPH, StepConfig

pytestmark = pytest.mark.gpu


def errors(g, o, x0, u0):
    e = {}
    e["x"] = np.abs(g["pos"] - o["pos"]).max() / np.sqrt(np.mean(np.sum((o["pos"] - x0) ** 2, 1)))
    e["v"] = np.abs(g["vel"] - o["vel"]).max() / np.sqrt(np.mean(np.sum(o["vel"] ** 2, 1)))
    e["u"] = np.abs(g["u"] - o["u"]).max() / np.sqrt(np.mean((o["u"] - u0) ** 2))
    e["rho"] = np.abs(g["rho"] / o["rho"] - 1).max()
    e["x_abs"] = np.abs(g["pos"] - o["pos"]).max() / np.abs(o["pos"]).max()
    e["v_abs"] = np.abs(g["vel"] - o["vel"]).max() / np.abs(o["vel"]).max()
    e["u_abs"] = np.abs(g["u"] / o["u"] - 1).max()
    return e


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("moving", [False, True])
def test_kdk_steps_match_oracle(gpu, prec, moving):
    n = 10000
    d = ic.evrard(n, seed=31)
    vel = np.random.default_rng(7).normal(size=(n, 3)) * 0.05 if moving else d["vel"]
    cfg = StepConfig(precision=prec, h_tol=1e-12)
    sim = DeviceSPH(d["pos"], vel, d["m"], d["u"], d["h0"], cfg)
    ref = KDK(d["pos"], vel, d["m"], d["u"], d["h0"], dt=cfg.dt, gamma=cfg.gamma, alpha=cfg.alpha, G=cfg.G,
              theta=cfg.theta, softening=cfg.softening, eps_plummer=cfg.eps_plummer, n_ngb=cfg.n_ngb, h_tol=1e-13,
              balsara=cfg.balsara)
    sim.prime(); ref.prime()
    done = 0
    for steps in (1, 3):
        while done < steps:
            sim.step(); ref.step(); done += 1
        g, o = sim.get(), ref.state()
        e = errors(g, o, d["pos"], d["u"])
        print(prec, "moving" if moving else "at rest", "steps", steps, {k: f"{v:.2e}" for k, v in e.items()})
        tol = 1e-8 if prec == "fp64" else 1e-4
        assert e["x"] < tol and e["v"] < tol and e["u"] < tol and e["rho"] < tol
        if prec == "fp64":
            assert e["x_abs"] < 1e-10 and e["v_abs"] < 1e-10 and e["u_abs"] < 1e-10
        else:
            assert e["x_abs"] < 1e-6 and e["v_abs"] < 1e-6
