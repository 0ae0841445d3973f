"""Synthetic tables in the layout of the reference's N_H cooling table (HCoolMu.bin) for the
oracle pins and the GPU parity tests of doCloudyHCooling."""
import numpy as np

KB = 1.3807e-16
MH = 1.673534e-24
GRIDS = (np.linspace(-4.0, 4.0, 17), np.linspace(2.0, 8.0, 25), np.linspace(0.0, 2.0, 9), np.linspace(18.0, 24.0, 13))
# code units of the CLOUDY run (CR/v3t.cu output: UnitDensity, Unit_u, unitTime)
UNITS = dict(unit_density=2.83261e-24, unit_u=1.80046e12, unit_time=2.29987e15)


def table(heat=None, cool=None, mu=None, logu=None, gamma=5.0 / 3.0):
    """A table dict on GRIDS; defaults: mu falling from 1.2 to 0.6 across log T = 4 and the
    log u of each T row for that mu (increasing along T), heat and cool smooth in all axes."""
    g = [np.asarray(a, np.float32) for a in GRIDS]
    shape = tuple(len(a) for a in g)
    nH, T, r, NH = np.meshgrid(*[a.astype(np.float64) for a in g], indexing="ij")
    if mu is None:
        mu = 0.6 + 0.6 / (1.0 + np.exp((T - 4.0) / 0.3))
    mu = np.broadcast_to(mu, shape)
    if logu is None:
        logu = np.log10(KB * 10.0 ** T / ((gamma - 1.0) * mu * MH))
    if heat is None:
        heat = -24.0 + 0.3 * np.sin(nH) + 0.2 * np.cos(1.3 * T) - 0.1 * r + 0.05 * (NH - 21.0)
    if cool is None:
        cool = -23.0 + 0.5 * nH + 0.4 * np.sin(T) + 0.15 * r - 0.1 * (NH - 21.0)
    bounds = [g[0][0], g[0][-1], g[1][0], g[1][-1], g[2][0], g[2][-1], g[3][0], g[3][-1]]
    f32 = lambda a: np.ascontiguousarray(np.broadcast_to(a, shape), np.float32)
    return dict(dims=shape, bounds=np.asarray(bounds, np.float32), nH=g[0], T=g[1], r=g[2], NH=g[3],
                heat=f32(heat), cool=f32(cool), mu=f32(mu), logu=f32(logu))


def particles(n, seed, spill=0.1):
    """Particles whose (n_H, r, N_H, T) cover the table, a fraction ``spill`` outside it."""
    rng = np.random.default_rng(seed)
    lo, hi = -4.0 - 8 * spill, 4.0 + 8 * spill
    nH = rng.uniform(lo, hi, n)
    rho = 10.0 ** nH * MH / (0.7 * UNITS["unit_density"])
    r = rng.uniform(0.0, 2.0 + 2 * spill, n)
    d = rng.normal(size=(n, 3))
    pos = d / np.linalg.norm(d, axis=1)[:, None] * r[:, None]
    log_nh = rng.uniform(18.0 - 6 * spill, 24.0 + 6 * spill, n)
    logT = rng.uniform(2.0 - 6 * spill, 8.0 + 6 * spill, n)
    u = KB * 10.0 ** logT / ((2.0 / 3.0) * 0.9 * MH) / UNITS["unit_u"]
    return pos, rho, u, log_nh
