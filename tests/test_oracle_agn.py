"""The oracle's AGN source terms pinned on the CPU: the injection's counter-based
draws are uniform and independent, the pairs have the geometry of
outflow_injector2 (mirror pairs, radii within r_inject, radial speed v_in,
directions uniform on the sphere), and the isothermal halo's pull has the
closed form of M(r) = 2 sigma^2 r / G (v_circ^2 = 2 sigma^2 at every radius,
independent of G)."""
import numpy as np

import oracle as O


def test_inject_uniform_is_uniform_and_independent():
    u = np.array([[O.inject_uniform(3, s, j, k) for k in range(3)] for s in range(4) for j in range(2500)])
    assert np.all((u >= 0) & (u < 1))
    assert np.allclose(u.mean(0), 0.5, atol=0.01)
    assert np.allclose(u.var(0), 1 / 12, atol=0.003)
    c = np.corrcoef(u.T)
    assert np.all(np.abs(c[np.triu_indices(3, 1)]) < 0.03)
    assert O.inject_uniform(3, 1, 0, 0) != O.inject_uniform(4, 1, 0, 0) != O.inject_uniform(3, 2, 0, 0)


def test_inject_pairs_geometry():
    pairs, r_in, v_in = 4000, 1e-3, 0.1
    pos, vel = O.inject_pairs(pairs, 11, 5, r_in, v_in)
    assert np.array_equal(pos[0::2], -pos[1::2]) and np.array_equal(vel[0::2], -vel[1::2])
    r = np.linalg.norm(pos, axis=1)
    assert np.all(r <= r_in)
    assert np.allclose(np.linalg.norm(vel, axis=1), v_in, rtol=1e-14)
    e = vel / v_in
    assert np.allclose(np.cross(pos, e), 0.0, atol=1e-18)  # radial
    assert np.all(np.einsum("ij,ij->i", pos, e) >= 0)      # outward
    # uniform on the sphere: mean direction ~ 0, <e_z^2> = 1/3; radii uniform in [0, r_in)
    e1 = e[0::2]
    assert np.all(np.abs(e1.mean(0)) < 0.04)
    assert np.allclose((e1 ** 2).mean(0), 1 / 3, atol=0.02)
    assert abs(r[0::2].mean() / r_in - 0.5) < 0.02


def test_halo_closed_form():
    rng = np.random.default_rng(1)
    pos = rng.normal(size=(500, 3))
    sigma = 0.7
    a = O.halo_accel(pos, sigma, G=1.0)
    r = np.linalg.norm(pos, axis=1)
    assert np.allclose(np.linalg.norm(a, axis=1) * r, 2 * sigma ** 2, rtol=1e-13)  # v_circ^2 = 2 sigma^2
    assert np.allclose(a / np.linalg.norm(a, axis=1)[:, None], -pos / r[:, None], atol=1e-14)
    assert np.allclose(O.halo_accel(pos, sigma, G=3.0), a, rtol=1e-13)
    assert np.array_equal(O.halo_accel(np.zeros((1, 3)), sigma), np.zeros((1, 3)))
