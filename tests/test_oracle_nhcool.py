"""Pins of the oracle's doCloudyHCooling restatement (oracle/sph_nhcool.c) by what the
mathematics fixes: find_nx0 is the floor index clamped to [0, N-2]; the hypercube interpolation
reproduces tables that are linear in n_H and N_H; a table that varies along T only is
interpolated with the r fraction and one that varies along r only with the T fraction (the
reference's own weights, CR/v4t.h:1716-1727), which a textbook multilinear interpolation would
not do; equal heating and cooling leave u unchanged; strong cooling hits the floor; typ != 0 is
left alone; mu comes from the first T row at the upper (n_H, r, N_H) corner whose log u
reaches u.  Tolerances: 1e-5 relative (FP32 arithmetic) unless stated."""
import numpy as np
import pytest

import oracle as O
from _nhtab import GRIDS, KB, MH, UNITS, particles, table

DT = 1e-5


def run(tab, pos, rho, u, lnh, dt=DT, typ=None):
    return O.nh_cooling(tab, pos, rho, u, lnh, dt, UNITS["unit_density"], UNITS["unit_u"], UNITS["unit_time"],
                        typ=typ)


def clampv(v, lo, hi):
    return np.where(v < lo, lo + 1e-4, np.where(v > hi, hi - 1e-4, v))


def cell(grid, v):
    k = np.clip(np.searchsorted(grid, v, side="right") - 1, 0, len(grid) - 2)
    return k, (v - grid[k]) / (grid[k + 1] - grid[k])


def point(pos, rho, u, lnh, mu=0.6, gamma=5.0 / 3.0):
    """(log n_H, log T, r, log N_H) of each particle after the clamps, in FP64, for a constant mu."""
    nH = clampv(np.log10(rho * UNITS["unit_density"] * 0.7 / MH), -4.0, 4.0)
    r = clampv(np.linalg.norm(pos, axis=1), 0.0, 2.0)
    NH = clampv(lnh, 18.0, 24.0)
    T = clampv(np.log10((gamma - 1.0) * mu * MH / KB * u * UNITS["unit_u"]), 2.0, 8.0)
    return nH, T, r, NH


def test_nx0_is_the_clamped_floor_index():
    rng = np.random.default_rng(1)
    for _ in range(20):
        g = np.sort(rng.uniform(-5, 5, rng.integers(2, 40))).astype(np.float32)
        g = np.unique(g)
        if len(g) < 2:
            continue
        xs = np.concatenate([g, rng.uniform(-7, 7, 200).astype(np.float32)])
        for x in xs:
            want = min(max(int(np.searchsorted(g, x, side="right")) - 1, 0), len(g) - 2)
            assert O.nh_nx0(g, x) == want


def test_tables_linear_in_nh_and_column_are_reproduced():
    a, b, c = -25.0, 0.4, 0.15
    nH, _, _, NH = np.meshgrid(*[np.asarray(x, np.float64) for x in GRIDS], indexing="ij")
    tab = table(heat=a + b * nH + c * (NH - 21.0), cool=-100.0, mu=0.6)
    pos, rho, u, lnh = particles(4000, 2)
    out, nf = run(tab, pos, rho, u, lnh)
    p = point(pos, rho, u, lnh)
    rhocgs = rho * UNITS["unit_density"]
    want = u + 10.0 ** (a + b * p[0] + c * (p[3] - 21.0)) / rhocgs * DT * UNITS["unit_time"] / UNITS["unit_u"]
    want = np.maximum(want, 1.8e11 / UNITS["unit_u"])
    np.testing.assert_allclose(out, want, rtol=2e-5)


@pytest.mark.parametrize("axis", ["T", "r"])
def test_interpolation_keeps_the_reference_axis_weights(axis):
    """heat varying only along T is interpolated with the r fraction, and vice versa."""
    rng = np.random.default_rng(3)
    vals = rng.uniform(-26.0, -22.0, len(GRIDS[1] if axis == "T" else GRIDS[2]))
    shape = (1, -1, 1, 1) if axis == "T" else (1, 1, -1, 1)
    tab = table(heat=vals.reshape(shape), cool=-100.0, mu=0.6)
    pos, rho, u, lnh = particles(4000, 4, spill=0.0)
    out, _ = run(tab, pos, rho, u, lnh)
    nH, T, r, NH = point(pos, rho, u, lnh)
    g32 = [np.asarray(x, np.float32).astype(np.float64) for x in GRIDS]
    jT, fT = cell(g32[1], T)
    kr, fr = cell(g32[2], r)
    if axis == "T":
        gam = vals[jT] + fr * (vals[jT + 1] - vals[jT])
        textbook = vals[jT] + fT * (vals[jT + 1] - vals[jT])
    else:
        gam = vals[kr] + fT * (vals[kr + 1] - vals[kr])
        textbook = vals[kr] + fr * (vals[kr + 1] - vals[kr])
    rhocgs = rho * UNITS["unit_density"]
    inc = lambda g: 10.0 ** g / rhocgs * DT * UNITS["unit_time"] / UNITS["unit_u"]
    got = out.astype(np.float64) - u.astype(np.float32).astype(np.float64)
    # increments the FP32 sum resolves, above the u floor
    floor = 1.8e11 / UNITS["unit_u"]
    sig = (inc(gam) > 1e-3 * u) & (inc(gam) < 1e3 * u) & (u + inc(gam) > 1.01 * floor) & (u + inc(textbook) > 1.01 * floor)
    assert sig.sum() > 1000
    ok = np.abs(got / inc(gam) - 1.0) < 1e-3
    assert ok[sig].mean() > 0.999
    # the textbook weights give a different answer for most particles
    assert (np.abs(got / inc(textbook) - 1.0) > 1e-2)[sig].mean() > 0.5


def test_equal_rates_leave_u_and_strong_cooling_floors():
    tab = table()
    pos, rho, u, lnh = particles(3000, 5)
    same = dict(tab, cool=tab["heat"].copy())
    out, nf = run(same, pos, rho, u, lnh)
    floor = np.float32(np.float32(1.8e11) / np.float32(UNITS["unit_u"]))
    low = u * UNITS["unit_u"] < 1.8e11
    assert nf == low.sum() and np.all(out[low] == floor)
    np.testing.assert_allclose(out[~low], u[~low].astype(np.float32), rtol=3e-7)
    floored = dict(tab, heat=np.full_like(tab["heat"], -100.0), cool=np.full_like(tab["cool"], 10.0))
    out, nf = run(floored, pos, rho, u, lnh)
    assert nf == len(u)
    assert np.all(out == floor)


def test_other_types_are_left_alone():
    tab = table()
    pos, rho, u, lnh = particles(1000, 6)
    typ = (np.arange(1000) % 3 == 1).astype(np.int32)
    out, _ = run(tab, pos, rho, u, lnh, typ=typ)
    ref, _ = run(tab, pos, rho, u, lnh)
    assert np.array_equal(out[typ == 1], u[typ == 1].astype(np.float32))
    assert np.array_equal(out[typ == 0], ref[typ == 0])


def test_mu_from_the_first_row_at_the_upper_corner():
    """mu differs per T row and log u also depends on the n_H index: the chosen row (and so T,
    and the T cell of the heat that varies along T only) follows the upper (n_H, r, N_H)
    corner of the particle's cell."""
    g = [np.asarray(x, np.float64) for x in GRIDS]
    nHg, Tg, rg, NHg = np.meshgrid(*g, indexing="ij")
    mu = 0.55 + 0.05 * (np.arange(len(g[1])) % 7)[None, :, None, None] + 0 * nHg
    logu = np.log10(KB * 10.0 ** Tg / ((2.0 / 3.0) * mu * MH)) + 0.02 * nHg
    hT = np.linspace(-26.0, -22.0, len(g[1]))
    hv = hT[None, :, None, None] + 0 * nHg
    tab = table(heat=hv, cool=-100.0, mu=mu, logu=logu)
    pos, rho, u, lnh = particles(12000, 7, spill=0.0)
    # r a quarter into its cell (well away from the nodes, where the r fraction decides which T
    # corner dominates)
    rr = np.linalg.norm(pos, axis=1)
    k = np.clip(np.floor(rr / 0.25), 0, 7)
    pos = pos / rr[:, None] * ((k + 0.25) * 0.25)[:, None]
    out, _ = run(tab, pos, rho, u, lnh)
    nH, _, r, NH = point(pos, rho, u, lnh)
    g32 = [np.asarray(x, np.float32).astype(np.float64) for x in GRIDS]
    i0, _ = cell(g32[0], nH)
    k0, fr = cell(g32[2], r)
    l0, _ = cell(g32[3], NH)
    up = np.log10(u * UNITS["unit_u"])
    lu = tab["logu"].astype(np.float64)
    rows = lu[i0 + 1, :, k0 + 1, l0 + 1]
    j = np.where((up[:, None] <= rows).any(1), np.argmax(up[:, None] <= rows, axis=1), -1)
    mu_p = np.where(j >= 0, tab["mu"].astype(np.float64)[i0 + 1, np.maximum(j, 0), k0 + 1, l0 + 1],
                    tab["mu"].astype(np.float64).ravel()[-1])
    T = clampv(np.log10((2.0 / 3.0) * mu_p * MH / KB * u * UNITS["unit_u"]), 2.0, 8.0)
    jT, fT = cell(g32[1], T)
    # heat along T weighted by the r fraction
    want = hT[jT] + fr * (hT[jT + 1] - hT[jT])
    rhocgs = rho * UNITS["unit_density"]
    got = (out.astype(np.float64) - u.astype(np.float32)) / (DT * UNITS["unit_time"] / UNITS["unit_u"]) * rhocgs
    # away from T cell edges (where a 1-ulp T decides the cell), increments the FP32 sum resolves
    inc = 10.0 ** want / rhocgs * DT * UNITS["unit_time"] / UNITS["unit_u"]
    inner = (np.abs(fT - 0.5) < 0.45) & (inc > 1e-3 * u) & (inc < 1e3 * u) & (u + inc > 1.01 * 1.8e11 / UNITS["unit_u"])
    assert inner.sum() > 1000
    assert np.mean(np.abs(np.log10(got[inner]) - want[inner]) < 1e-3) > 0.999


def test_table_file_layout(tmp_path):
    """NHCoolingTable writes and reads the layout HCoolMu.bin has for the reference's reader
    (CR/v3t.cu:57-85): four int32 sizes, eight float32 bounds, the four grids, then heat, cool,
    mu and u, float32 in [nH][T][r][NH] order."""
    from sph_3_amd.cooling import NHCoolingTable

    t = NHCoolingTable.synthetic()
    p = tmp_path / "HCoolMu.bin"
    t.write(p)
    raw = p.read_bytes()
    dims = np.frombuffer(raw[:16], np.int32)
    assert tuple(dims) == t.dims == (33, 61, 21, 25)
    f = np.frombuffer(raw[16:], np.float32)
    assert np.array_equal(f[:8], t.bounds)
    ng = int(dims.sum())
    cells = int(np.prod(dims))
    assert len(f) == 8 + ng + 4 * cells
    assert np.array_equal(f[8:8 + 33], t.grids[0]) and np.array_equal(f[8 + 33 + 61 + 21:8 + ng], t.grids[3])
    heat = f[8 + ng:8 + ng + cells].reshape(dims)
    assert np.array_equal(heat, t.tables[0]) and heat[1, 2, 3, 4] == t.tables[0][1, 2, 3, 4]
    u = NHCoolingTable.read(p)
    assert u.dims == t.dims and all(np.array_equal(a, b) for a, b in zip(u.tables, t.tables))
