"""The h solve's two candidate-row layouts give the same bits.

By default the solve keeps an FP64 squared distance beside every candidate index (2 KB per
target); past the context's distance-rows limit it keeps the indices only and recomputes
each r^2 from the positions with the walk's own operations (contraction off).  These tests
set the limit to 0 on a fresh context and require bitwise equality with the default layout:
h, density, acceleration and the neighbour sets of the resident step (Evrard with gravity,
and the periodic Sedov lattice, whose minimum-image differences the recomputation must
repeat), and the stateless solves (Newton and the k-th distance rule, self and foreign
targets).  The device footprint must drop by at least the distance rows (2 KB per particle).
"""
import numpy as np
import pytest

import sph_3_amd.step as S
from sph_3_amd import ic
from sph_3_amd import libsx as X
from sph_3_amd._lib import H_NEWTON, Context

pytestmark = pytest.mark.gpu


def run_steps(monkeypatch, d, cfg, limit, steps):
    ctx = Context(0)
    monkeypatch.setattr(S, "context", lambda device=0: ctx)
    sim = S.DeviceSPH(d["pos"], d["vel"], d["m"], d["u"], d["h0"], cfg)
    if limit is not None:
        sim.set_distance_rows_limit(limit)
    for _ in range(steps):
        sim.step()
    out = sim.get()
    out["ngb"] = sim.neighbours()
    out["bytes"] = sim.device_bytes()
    del sim
    ctx.close()
    return out


@pytest.mark.parametrize("case", ["evrard", "sedov"])
def test_step_identical_without_distance_rows(gpu, monkeypatch, case):
    if case == "evrard":
        d = ic.evrard(60_000, seed=81)
        cfg = S.StepConfig(dt=1e-4)
    else:
        d = ic.sedov(24, seed=82)
        cfg = S.StepConfig(dt=1e-4, G=0.0, period=1.0)
    n = len(d["m"])
    a = run_steps(monkeypatch, d, cfg, None, 2)
    b = run_steps(monkeypatch, d, cfg, 0, 2)
    for k in ("pos", "vel", "u", "h", "rho", "acc"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["ngb"][0], b["ngb"][0]) and np.array_equal(a["ngb"][1], b["ngb"][1])
    print(f"{case}: {a['bytes'] / n:.0f} vs {b['bytes'] / n:.0f} bytes per particle")
    assert a["bytes"] - b["bytes"] >= 2048 * n


def test_stateless_solves_identical_without_distance_rows(gpu, monkeypatch):
    d = ic.evrard(20_000, seed=83)
    pos = d["pos"]
    tg = np.random.default_rng(84).uniform(-0.5, 0.5, (3000, 3))
    out = []
    for limit in (None, 0):
        ctx = Context(0)
        monkeypatch.setattr(X, "context", lambda device=0: ctx)
        if limit is not None:
            ctx.check(ctx.lib.sph_set_distance_rows_limit(ctx.h, limit), "sph_set_distance_rows_limit")
        out.append([X._hsolve(pos, None, d["h0"], H_NEWTON, 64.0)[0], X.do_smoothingX((pos, pos)),
                    X.do_smoothingX((pos, tg))])
        ctx.close()
    for r, g in zip(*out):
        assert np.array_equal(r, g)
