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


def gaussian_field(n_grid: int, slope: float = -11.0 / 3.0, seed: int = 0):
    """Seeded Gaussian random field on an n_grid^3 periodic grid with an isotropic
    power-law spectrum P(k) ~ k^slope (the k = 0 mode removed), normalised to zero
    mean and unit variance; float32."""
    rng = np.random.Generator(np.random.PCG64(seed))
    white = rng.standard_normal((n_grid,) * 3).astype(np.float32)
    f = np.fft.rfftn(white)
    k1 = np.fft.fftfreq(n_grid) * n_grid
    kz = np.fft.rfftfreq(n_grid) * n_grid
    k2 = (k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2).astype(np.float32)
    k2[0, 0, 0] = 1.0
    f *= k2 ** np.float32(slope / 4.0)  # amplitude ~ sqrt(P) = k^(slope/2)
    f[0, 0, 0] = 0.0
    g = np.fft.irfftn(f, s=(n_grid,) * 3, axes=(0, 1, 2)).astype(np.float32)
    g -= g.mean()
    g /= g.std()
    return g


def agn_cloud(n_cloud: int, n_wind: int, n_grid: int = 256, sigma_lnrho: float = 1.5, R: float = 1.0,
              M: float = 1.0, r_wind: float = 0.05, v_wind: float = 0.1, u_cloud: float = 0.0,
              u_wind: float = 0.0, slope: float = -11.0 / 3.0, seed: int = 46, n_ngb: float = 64.0):
    """AGN-outflow-like initial conditions: a spherical cloud of radius R and mass M
    whose density is lognormal, rho ~ exp(sigma g - sigma^2/2) with g the Gaussian
    field of ``gaussian_field`` on n_grid^3 cells over [-R, R]^3 (particles are
    drawn cell by cell with probability ~ rho and placed uniformly in the cell), at
    rest with specific energy u_cloud; plus n_wind equal-mass particles in the
    shell 0.9 r_wind < r < r_wind (uniform in volume, isotropic) moving radially
    outward at v_wind with u_wind.  Returns dict(pos, vel, m, u, h0, wind) with h0
    from the cell density at the target neighbour number (a starting guess)."""
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    g = gaussian_field(n_grid, slope, seed)
    a = 2.0 * R / n_grid
    c = (np.arange(n_grid, dtype=np.float32) + 0.5) * a - R
    r2 = c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2
    w = np.exp(np.float32(sigma_lnrho) * g - np.float32(0.5 * sigma_lnrho ** 2))
    w[r2 >= R * R] = 0.0
    w = w.ravel().astype(np.float64)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    cell = np.searchsorted(cdf, rng.random(n_cloud), side="right")
    cell = np.minimum(cell, len(w) - 1)
    ix, iy, iz = np.unravel_index(cell, (n_grid,) * 3)
    pos_c = np.stack([c[ix], c[iy], c[iz]], axis=1).astype(np.float64) + rng.uniform(-0.5, 0.5, (n_cloud, 3)) * a
    m_p = M / n_cloud
    rho_cell = M * w[cell] / (w.sum() * a ** 3)  # mass density of the particle's cell
    # wind shell
    rr = r_wind * np.cbrt(rng.uniform(0.9 ** 3, 1.0, n_wind))
    ct = rng.uniform(-1.0, 1.0, n_wind)
    ph = rng.uniform(0.0, 2.0 * np.pi, n_wind)
    st = np.sqrt(1.0 - ct * ct)
    nhat = np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)
    pos_w = rr[:, None] * nhat
    rho_w = n_wind * m_p / (4.0 / 3.0 * np.pi * r_wind ** 3 * (1.0 - 0.9 ** 3))
    pos = np.concatenate([pos_c, pos_w])
    vel = np.concatenate([np.zeros((n_cloud, 3)), v_wind * nhat])
    u = np.concatenate([np.full(n_cloud, u_cloud), np.full(n_wind, u_wind)])
    rho0 = np.concatenate([rho_cell, np.full(n_wind, rho_w)])
    n = n_cloud + n_wind
    m = np.full(n, m_p)
    h0 = 0.5 * np.cbrt(3.0 * n_ngb * m_p / (4.0 * np.pi * rho0))
    wind = np.zeros(n, bool)
    wind[n_cloud:] = True
    return dict(pos=pos, vel=vel, m=m, u=u, h0=h0, wind=wind)


def u_of_T(T: float, unit_u_cgs: float, mu: float = 0.6, gamma: float = 5.0 / 3.0) -> float:
    """Specific energy in code units of gas at temperature T [K] with mean molecular weight mu."""
    return 1.3807e-16 * T / ((gamma - 1.0) * mu * 1.6726e-24) / unit_u_cgs


def sedov(n_side: int, jitter: float = 0.1, E: float = 1.0, n_deposit: int = 64, u0: float = 1e-5,
          rho0: float = 1.0, seed: int = 50, n_ngb: float = 64.0):
    """3D Sedov blast in the unit periodic box [0, 1)^3: a cubic lattice of n_side^3
    particles with a seeded uniform jitter of +-jitter lattice spacings per axis
    (wrapped into the box), density rho0 (m = rho0 / N), specific energy u0, and
    the energy E deposited kernel-weighted on the n_deposit particles nearest the
    centre: u_i += E w_i / (m sum_j w_j) with w = W(r, h_d) of the cubic spline and
    2 h_d the distance of the (n_deposit+1)-th nearest particle.  Returns dict(pos,
    vel, m, u, h0) with h0 the lattice value at the target neighbour number."""
    pos = uniform_cube(n_side, seed=seed, jitter=jitter) % 1.0
    n = len(pos)
    m = np.full(n, rho0 / n)
    u = np.full(n, u0)
    d = np.linalg.norm(pos - 0.5, axis=1)
    order = np.argsort(d, kind="stable")
    near = order[:n_deposit]
    hd = 0.5 * d[order[n_deposit]]
    q = d[near] / hd
    w = np.where(q < 1.0, 1.0 - 1.5 * q * q + 0.75 * q ** 3, np.where(q < 2.0, 0.25 * (2.0 - q) ** 3, 0.0))
    u[near] += E * w / (m[near] * w.sum())
    h0 = np.full(n, 0.5 * np.cbrt(3.0 * n_ngb * (rho0 / n) / (4.0 * np.pi * rho0)))
    return dict(pos=pos, vel=np.zeros((n, 3)), m=m, u=u, h0=h0)
