"""GPU parity of the N_H-table cooling update (doCloudyHCooling, CR/v4t.h:1767-1884) against
the oracle's FP32 restatement (oracle/sph_nhcool.c, pinned in test_oracle_nhcool.py), and of
the resident step's SPH_COOL_NH mode against the stateless update applied to the step's own
positions, density, u and log N_H.

Both sides run the same FP32 operations with correctly rounded log10, pow and sqrt, so the
results are expected to be bitwise equal; a double-precision library call that rounds the
other way at a float boundary may move a particle by one T or u row, so at most 1 in 1,000 may
differ (none did when this was written), and all within 1e-2 relative."""
import numpy as np
import pytest

import oracle as O
from _nhtab import UNITS, particles, table
from sph_3_amd import ic
from sph_3_amd.column import ColumnParams
from sph_3_amd.cooling import CoolingParams, NHCoolingTable
from sph_3_amd.step import DeviceSPH, StepConfig

pytestmark = pytest.mark.gpu

DT = 1e-5


def params():
    return CoolingParams(UNITS["unit_density"], UNITS["unit_u"], UNITS["unit_time"], mode="nh")


def nh_table(t):
    return NHCoolingTable([t["nH"], t["T"], t["r"], t["NH"]], t["bounds"], t["heat"], t["cool"], t["mu"], t["logu"])


@pytest.mark.parametrize("n", [1, 1000, 300_000])
def test_stateless_update_matches_oracle(gpu, n):
    t = table()
    pos, rho, u, lnh = particles(n, 11 + n)
    typ = (np.arange(n) % 5 == 3).astype(np.int32)
    got, nf = nh_table(t).cooling_update(pos, rho, u, lnh, DT, params(), typ=typ)
    want, nf0 = O.nh_cooling(t, pos, rho, u, lnh, DT, UNITS["unit_density"], UNITS["unit_u"], UNITS["unit_time"],
                             typ=typ)
    # gas (typ 0): the FP32 result; the others keep their FP64 u
    gas = typ == 0
    want = want.astype(np.float64)
    same = got[gas] == want[gas]
    rel = np.abs(got[gas] - want[gas]) / np.abs(want[gas])
    print(f"n={n}: {np.count_nonzero(~same)} differ, max rel {rel.max():.3e}, floored {nf} / {nf0}")
    assert np.count_nonzero(~same) <= n // 1000
    assert rel.max() < 1e-2
    assert abs(nf - nf0) <= n // 1000
    assert np.all(got[~gas] == u[~gas])


def test_step_nh_cooling_is_the_update_of_the_step_state(gpu):
    """One step with SPH_COOL_NH against the same step without cooling followed by the
    stateless update on its positions, density, u and column density."""
    d = ic.evrard(20_000, seed=12, u0=1.0)
    t = nh_table(table())
    mass_g, length_cm = 1.0e9 * 1.989e33, 3.086e21
    col = ColumnParams(UNITS["unit_density"], length_cm)
    out = {}
    for mode in ("nh", None):
        sim = DeviceSPH(d["pos"], d["vel"], d["m"], d["u"], d["h0"], StepConfig(dt=DT))
        sim.set_column(col)
        if mode:
            sim.set_cooling(None, None, params(), surrogate=t)
        sim.step()
        g = sim.get()
        g["lognh"] = sim.column()[0]
        out[mode] = g
    a, b = out["nh"], out[None]
    for k in ("pos", "rho", "lognh"):
        assert np.array_equal(a[k], b[k]), k
    want, nf = t.cooling_update(b["pos"], b["rho"], b["u"], b["lognh"], DT, params())
    print("floored", nf, "of", len(want), "u range", want.min(), want.max())
    assert np.array_equal(a["u"], want)
    assert not np.array_equal(a["u"], b["u"])


def test_nh_cooling_needs_table_and_column(gpu):
    d = ic.evrard(2000, seed=13, u0=1.0)
    sim = DeviceSPH(d["pos"], d["vel"], d["m"], d["u"], d["h0"], StepConfig(dt=DT))
    sim.set_cooling(None, None, params(), surrogate=nh_table(table()))
    with pytest.raises(RuntimeError, match="column density"):
        sim.step()
