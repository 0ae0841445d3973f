"""Seeded synthetic initial conditions.

Evrard collapse (the reference's test problem, Particles_in_BH_Tree_III/
Evrard_Collapse_2024/IC_evrard.py:6-155): total mass M, radius R, rho ~ 1/r,
cold (u = 0.05), at rest.  The reference stretches a 161^3 lattice with an
unseeded jitter; here radii are drawn as r = R sqrt(U) (so M(<r) ~ r^2, i.e.
rho ~ 1/r) with isotropic directions from a seeded PCG64 stream, which gives
any N exactly.
"""
from __future__ import annotations

import numpy as np


def evrard(n: int, seed: int = 42, R: float = 1.0, M: float = 1.0, u0: float = 0.05, n_ngb: float = 64.0):
    """Returns dict(pos, vel, m, u, h0) with h0 the smoothing length of the
    analytic density rho(r) = M / (2 pi R^2 r) at the target neighbour number
    (a starting guess for the h solve)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = R * np.sqrt(rng.random(n))
    ct = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    st = np.sqrt(1.0 - ct * ct)
    pos = np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * ct], axis=1)
    m = np.full(n, M / n)
    rho = M / (2.0 * np.pi * R * R * np.maximum(r, 1e-3 * R))
    h0 = 0.5 * np.cbrt(3.0 * n_ngb * m / (4.0 * np.pi * rho))
    return dict(pos=pos, vel=np.zeros((n, 3)), m=m, u=np.full(n, u0), h0=h0)


def uniform_cube(n_side: int, seed: int = 0, jitter: float = 0.1, L: float = 1.0):
    """Jittered cubic lattice in [0, L)^3 (jitter in lattice spacings)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a = L / n_side
    g = (np.arange(n_side) + 0.5) * a
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    pos += rng.uniform(-jitter, jitter, pos.shape) * a
    return pos


def cooling_fields(n: int, seed: int = 43):
    """Per-particle metallicity Z [Z_sun], log-uniform in [1e-3, 1], and UV field
    amplitude J21, log-uniform in [1e-2, 1e2], from a seeded PCG64 stream (the
    Evrard-plus-cooling workload)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Z = 10.0 ** rng.uniform(-3.0, 0.0, n)
    J21 = 10.0 ** rng.uniform(-2.0, 2.0, n)
    return Z, J21
