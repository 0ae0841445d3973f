"""Restart: a run saved after some steps and continued from the file matches the
uninterrupted run (same positions, velocities, u and h to rounding of the sort's
key ties, which this workload does not have: bit-identical), also with outflow
injection (same pairs, same bookkeeping); and the reference-layout snapshot of
the resident state reads back."""
import numpy as np
import pytest

from sph_3_amd import ic
from sph_3_amd import snapshot as S
from sph_3_amd.step import AgnParams, DeviceSPH, StepConfig

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("with_agn", [False, True])
def test_restart_continues_the_run(gpu, tmp_path, with_agn):
    n0 = 20000
    d = ic.evrard(n0, seed=51)
    cfg = StepConfig(dt=1e-3)
    m = d["m"][0]
    agn = AgnParams(sigma=0.3, M_dot=3.5 * 2 * m / cfg.dt, v_in=0.2, m_inject=m, u_inject=0.05, r_inject=1e-3,
                    seed=9, h_ref_index=3) if with_agn else None
    n_max = n0 + 200
    sim = DeviceSPH(d["pos"], d["vel"], d["m"], d["u"], d["h0"], cfg, n_max=n_max)
    if agn:
        sim.set_agn(agn)
    sim.prime()
    for _ in range(2):
        sim.step()
    fn = tmp_path / "run.npz"
    S.save_restart(fn, sim, t=2 * cfg.dt)
    for _ in range(2):
        sim.step()
    ga = sim.get()
    ca = sim.counts()
    del sim
    sim2, t = S.load_restart(fn, cfg, n_max=n_max, agn=agn)
    assert t == 2 * cfg.dt
    for _ in range(2):
        sim2.step()
    gb = sim2.get()
    cb = sim2.counts()
    assert ca[0] == cb[0] and ca[1] == cb[1] and ca[2:] == cb[2:]
    for k in ["pos", "vel", "u", "h", "acc"]:
        diff = np.max(np.abs(ga[k] - gb[k]))
        print(k, diff)
        assert np.array_equal(ga[k], gb[k]), k


def test_snapshot_of_resident_state(gpu, tmp_path):
    n = 5000
    d = ic.evrard(n, seed=52)
    sim = DeviceSPH(d["pos"], d["vel"], d["m"], d["u"], d["h0"], StepConfig())
    sim.prime()
    sim.step()
    g = sim.get()
    S.write_snapshot(tmp_path / "snap.bin", S.state_arrays(g["pos"], g["vel"], sim.masses(), g["u"], g["h"], g["rho"]))
    back = S.read_snapshot(tmp_path / "snap.bin")
    pos, vel = S.positions(back)
    assert np.array_equal(pos, g["pos"].astype(np.float32).astype(np.float64))
    assert np.array_equal(back["rho"], g["rho"].astype(np.float32))
    assert np.all(back["Typ"] == 0)
