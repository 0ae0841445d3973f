"""The oracle's time-step criteria (dt_array_indiv_dt) pinned by closed forms on
the CPU: particles at rest (vsig = 2c, no divergence) give each criterion in
closed form and dt is their minimum; an approaching pair raises the signal
velocity by -3 w; and when every particle is a neighbour of every other, the
CSR maximum equals the reference's all-particle maximum."""
import numpy as np

import oracle as O


def test_at_rest_closed_forms():
    rng = np.random.default_rng(3)
    n = 200
    pos = rng.random((n, 3))
    vel = np.zeros((n, 3))
    h = rng.uniform(0.05, 0.1, n); c = np.full(n, 0.7); divV = np.zeros(n)
    acc = rng.normal(size=(n, 3)); acc_g = rng.normal(size=(n, 3)); eps = 0.01
    C, alpha = 0.3, 1.0
    dt, dh = O.timestep(pos, vel, h, c, divV, acc, acc_g, eps, C, alpha)
    at = np.linalg.norm(acc, axis=1); ag = np.linalg.norm(acc_g, axis=1)
    crit = np.stack([C * h / (2 * 0.7), C * h / (0.7 * (1 + 0.6 * alpha)), np.sqrt(h / ag), np.sqrt(h / at),
                     0.4 * np.sqrt(eps / at)])  # dt_G2 = 0.4 eps / |v| is infinite at rest
    assert np.allclose(dt, crit.min(0), rtol=1e-14)
    assert np.array_equal(dh, np.zeros(n))


def test_approaching_pair_signal_velocity():
    pos = np.array([[0.0, 0, 0], [0.1, 0, 0]])
    vel = np.array([[0.5, 0, 0], [-0.5, 0, 0]])  # closing at 1.0
    h = np.full(2, 1.0); c = np.full(2, 0.2); divV = np.full(2, -2.0)
    acc = np.full((2, 3), 1e-30); eps = 1e30     # switch off the force criteria
    dt, dh = O.timestep(pos, vel, h, c, divV, acc, acc, eps, 1.0, 0.0)
    w = (1.0 * -0.1) / (0.1 + 1e-5)            # v_ij . r_ij / (r + 1e-5), both orders
    vsig = 0.4 - 3 * w
    cour = 1.0 / (2.0 + 0.5)                   # h / (h |divV| + max(c, |v|))
    assert np.allclose(dt, min(1.0 / vsig, cour), rtol=1e-14)
    assert np.allclose(dh, -2.0 / 3.0, rtol=1e-15)


def test_csr_maximum_equals_all_particles_when_all_are_neighbours():
    rng = np.random.default_rng(4)
    n = 40
    pos = rng.random((n, 3)) * 0.1
    vel = rng.normal(size=(n, 3)); h = np.full(n, 0.2); c = rng.uniform(0.1, 1, n)
    divV = rng.normal(size=n); acc = rng.normal(size=(n, 3))
    off, idx = O.Tree(pos, np.ones(n), h=h).neighbours()
    assert np.all(np.diff(off) == n)
    a = O.timestep(pos, vel, h, c, divV, acc, acc, 0.01, 0.3, 1.0, off, idx)
    b = O.timestep(pos, vel, h, c, divV, acc, acc, 0.01, 0.3, 1.0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_periodic_pair_through_the_boundary_equals_the_unwrapped_pair():
    """A pair approaching through x = 0 of the unit box has the signal velocity of the
    same pair placed side by side without the box (minimum image)."""
    h = np.array([0.05, 0.05]); c = np.array([1.0, 1.0]); dv = np.zeros(2)
    vel = np.array([[-0.5, 0, 0], [0.5, 0, 0]])
    a = np.ones((2, 3))
    wrapped = np.array([[0.01, 0.5, 0.5], [0.99, 0.5, 0.5]])
    plain = np.array([[0.01, 0.5, 0.5], [-0.01, 0.5, 0.5]])
    ref, _ = O.timestep(plain, vel, h, c, dv, a, a, 1.0, 0.3, 1.0)  # eps = 1: dt_hyd is the minimum
    open_box, _ = O.timestep(wrapped, vel, h, c, dv, a, a, 1.0, 0.3, 1.0)
    with O.periodic(1.0):
        per, _ = O.timestep(wrapped, vel, h, c, dv, a, a, 1.0, 0.3, 1.0)
    np.testing.assert_allclose(per, ref, rtol=1e-13)
    assert np.all(open_box > per)  # without the wrap the pair looks 0.98 apart and receding
