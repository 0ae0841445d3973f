"""The oracle's column density (grad_rho_norm_ngb and the N_H loop of the
CLOUDY driver) pinned by what the mathematics fixes, on the CPU.

* With one h for every particle the pair sum sum_j m_j nW_ij (x_j - x_i) is
  minus the gradient of the density field rho(x) = sum_j m_j W(|x - x_j|, h)
  at x_i, so its norm must match central finite differences of that field
  (evaluated here with the cubic spline W itself, not with nW): a dropped
  piece, a wrong power of h or a wrong sign in nW fails.
* N_H = 0.5 (h / cbrt(n_ngb) + rho / g) L n_H: exact scalings in L and XH,
  and the two limits g -> infinity (only the h term) and h -> 0 (only the
  rho / g term), each against the other leaving one term.
"""
import numpy as np

import oracle as O


def spline_W(r, h):
    q = r / h
    w = np.where(q <= 1, 1 - 1.5 * q ** 2 + 0.75 * q ** 3, np.where(q <= 2, 0.25 * (2 - q) ** 3, 0.0))
    return w / (np.pi * h ** 3)


def test_grad_rho_matches_finite_differences_of_the_density_field():
    rng = np.random.default_rng(5)
    n = 3000
    pos = rng.random((n, 3))
    m = rng.uniform(0.5, 1.5, n) / n
    h0 = 0.06
    h = np.full(n, h0)
    off, idx = O.Tree(pos, m, h=h).neighbours()
    g = O.grad_rho_csr(pos, m, h, off, idx)
    tgt = rng.choice(np.where(np.all((pos > 0.2) & (pos < 0.8), axis=1))[0], 40, replace=False)
    d = 1e-5 * h0
    for i in tgt:
        grad = np.zeros(3)
        for k in range(3):
            e = np.zeros(3); e[k] = d
            rp = spline_W(np.linalg.norm(pos - (pos[i] + e), axis=1), h0) @ m
            rm = spline_W(np.linalg.norm(pos - (pos[i] - e), axis=1), h0) @ m
            grad[k] = (rp - rm) / (2 * d)
        assert abs(g[i] - np.linalg.norm(grad)) <= 1e-6 * np.linalg.norm(grad) + 1e-9


def test_column_density_scalings_and_limits():
    rng = np.random.default_rng(6)
    n = 200
    rho = rng.uniform(0.1, 10, n); g = rng.uniform(0.1, 10, n); h = rng.uniform(0.01, 0.1, n)
    ud, L = 6.77e-25, 3.0857e22
    base = O.column_density(rho, g, h, ud, L)
    assert np.allclose(O.column_density(rho, g, h, ud, 10 * L) - base, 1.0, atol=1e-12)
    assert np.allclose(O.column_density(rho, g, h, ud, L, XH=1.4) - base, np.log10(2.0), atol=1e-12)
    # g -> infinity keeps the h term, h -> 0 keeps the rho / g term; their sum is N_H
    nh_h = 10 ** O.column_density(rho, np.full(n, 1e300), h, ud, L)
    nh_g = 10 ** O.column_density(rho, g, np.zeros(n), ud, L)
    assert np.allclose(10 ** base, nh_h + nh_g, rtol=1e-12)
    # the h term is linear in h / cbrt(n_ngb), the gradient term in rho^2 / g
    assert np.allclose(10 ** O.column_density(rho, np.full(n, 1e300), 2 * h, ud, L, n_ngb=512.0), nh_h, rtol=1e-12)
    assert np.allclose(10 ** O.column_density(2 * rho, 4 * g, np.zeros(n), ud, L), nh_g, rtol=1e-12)
