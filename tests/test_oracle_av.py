"""Closed-form pins of the oracle's artificial viscosity and Balsara switch (CPU only).

The oracle's acceleration and du/dt (``or_accsph_*``, ``or_dudt_*``) restate
sph_code/libsx.py:146-160 (Gadget-2 viscosity)

    w_ij = v_ij . r_ij / (|r_ij| + 1e-5),  v_sig = c_i + c_j - 3 w_ij,
    Pi_ij = -alpha/2 v_sig w_ij / rho_bar  if v_ij . r_ij <= 0, else 0,

and sph_code/shear.py:205-209 (Balsara factor)

    f_i = |div v|_i / (|div v|_i + |curl v|_i + 1e-4 c_i / h_i),  Pi_ij <- (f_i + f_j)/2 Pi_ij.

Conservation identities hold for any symmetric Pi_ij and f_ij, so they cannot
catch a wrong coefficient.  These tests evaluate two-particle configurations
whose results follow by hand from the formulas above and the textbook cubic
spline (Monaghan & Lattanzio 1985; its radial derivative is written out
below, independently of the oracle's gradient code), and lattice limits that
fix the switch: a shear flow has no divergence (switch -> 0), a uniform
compression has no curl (switch -> 1 when c = 0).
"""
import numpy as np
import pytest

import oracle as O


def dWdr(r, h):
    """Radial derivative of the 3D cubic spline W(r, h) = (1/(pi h^3)) w(r/h)."""
    q = r / h
    if q <= 1.0:
        dw = -3.0 * q + 2.25 * q * q
    elif q <= 2.0:
        dw = -0.75 * (2.0 - q) ** 2
    else:
        dw = 0.0
    return dw / (np.pi * h ** 4)


def pair(r, v0, v1, h=0.5):
    """Particle 0 at the origin, particle 1 at (r, 0, 0)."""
    pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    vel = np.array([v0, v1], dtype=float)
    return pos, vel, np.full(2, h)


@pytest.mark.parametrize("r", [1e-3, 0.3, 0.7])
def test_av_approaching_pair_closed_form(r):
    """Pressure-free approaching pair: a_0 and du_0/dt are pure viscosity.
    alpha = 2, c = 0, rho_bar = 1 gives Pi = -(1/2)(2)(-3w)w = 3 w^2."""
    s = 0.8  # particle 0 moves toward particle 1 at speed s
    pos, vel, h = pair(r, [s, 0, 0], [0, 0, 0])
    m = np.array([1.0, 1.0])
    rho = np.ones(2)
    zero = np.zeros(2)
    vr = (s - 0.0) * (0.0 - r)  # v_01 . (x_0 - x_1) < 0: approaching
    w = vr / (r + 1e-5)
    Pi = 3.0 * w * w
    gx = dWdr(r, h[0]) * (0.0 - r) / r  # d/dx_0 of W(|x_0 - x_1|)
    a = O.accsph_allpairs(pos, vel, rho, zero, zero, h, m, 2.0)
    assert abs(a[0, 0] - (-m[1] * Pi * gx)) <= 1e-12 * abs(m[1] * Pi * gx)
    assert a[0, 0] < 0.0  # the approaching particle is decelerated
    assert abs(a[1, 0] + a[0, 0]) <= 1e-14 * abs(a[0, 0]) and np.all(a[:, 1:] == 0.0)
    du = O.dudt_allpairs(pos, vel, rho, zero, zero, h, m, 2.0)
    du_exp = m[1] * 0.5 * Pi * (s * gx)  # (P/rho^2 = 0 + Pi/2) v_01 . grad W
    assert abs(du[0] - du_exp) <= 1e-12 * abs(du_exp) and du[0] > 0.0
    assert abs(du[1] - du_exp) <= 1e-12 * abs(du_exp)  # both particles heat equally


def test_av_sound_speeds_alpha_and_mean_density():
    """c_i + c_j enters v_sig, alpha scales Pi linearly, rho_bar is the arithmetic mean;
    the pressure term P_i/rho_i^2 + P_j/rho_j^2 adds to Pi."""
    r, s, h = 0.4, 0.5, 0.45
    pos, vel, hh = pair(r, [0, 0, 0], [-s, 0, 0], h)  # particle 1 moves toward particle 0
    m = np.array([0.7, 1.3])
    rho = np.array([2.0, 6.0])
    c = np.array([0.3, 0.9])
    P = np.array([0.5, 4.0])
    alpha = 1.5
    vr = (0.0 - (-s)) * (0.0 - r)
    w = vr / (r + 1e-5)
    vsig = 0.3 + 0.9 - 3.0 * w
    Pi = -0.5 * alpha * vsig * w / 4.0
    gx = dWdr(r, h) * (0.0 - r) / r
    t = P[0] / 4.0 + P[1] / 36.0 + Pi
    a = O.accsph_allpairs(pos, vel, rho, P, c, hh, m, alpha)
    assert abs(a[0, 0] - (-m[1] * t * gx)) <= 1e-12 * abs(m[1] * t * gx)
    du = O.dudt_allpairs(pos, vel, rho, P, c, hh, m, alpha)
    vg = s * gx  # v_01 = (s, 0, 0)
    assert abs(du[0] - m[1] * (P[0] / 4.0 + 0.5 * Pi) * vg) <= 1e-12 * abs(m[1] * (P[0] / 4.0 + 0.5 * Pi) * vg)
    assert abs(du[1] - m[0] * (P[1] / 36.0 + 0.5 * Pi) * vg) <= 1e-12 * abs(m[0] * (P[1] / 36.0 + 0.5 * Pi) * vg)


@pytest.mark.parametrize("balsara", [False, True])
def test_av_receding_pair_has_no_viscosity(balsara):
    """v_ij . r_ij > 0: Pi = 0, so a and du/dt are the pressure terms alone."""
    r, h = 0.3, 0.5
    pos, vel, hh = pair(r, [-0.6, 0.1, 0], [0.2, 0.0, 0.3])  # separating along x
    m = np.array([1.0, 2.0]); rho = np.array([1.5, 2.5]); c = np.array([1.0, 1.0]); P = np.array([0.3, 0.2])
    kw = dict(divV=np.ones(2), curlV=np.zeros(2)) if balsara else {}
    gx = dWdr(r, h) * (0.0 - r) / r
    t = P[0] / 1.5 ** 2 + P[1] / 2.5 ** 2
    a = O.accsph_allpairs(pos, vel, rho, P, c, hh, m, 1.0, **kw)
    assert abs(a[0, 0] - (-m[1] * t * gx)) <= 1e-13 * abs(m[1] * t * gx)
    du = O.dudt_allpairs(pos, vel, rho, P, c, hh, m, 1.0, **kw)
    vg = (-0.6 - 0.2) * gx
    assert abs(du[0] - m[1] * P[0] / 1.5 ** 2 * vg) <= 1e-13 * abs(m[1] * P[0] / 1.5 ** 2 * vg)


def test_balsara_factor_two_particles():
    """f_i = |div|/(|div| + |curl| + 1e-4 c/h) chosen per particle through the inputs:
    (div, curl, c, h) = (1, 0, 0, .) -> 1; (1, 3, 0, .) -> 1/4; (1, 0, 1, 1e-4) -> 1/2;
    (0, 1, 0, .) -> 0.  The pair's viscosity is scaled by the mean of the two factors
    (compared with the unswitched sums at the same c and h, which enter v_sig and W)."""
    r, s = 0.35, 0.7
    pos, vel, hh = pair(r, [s, 0, 0], [0, 0, 0], 0.5)
    m = np.ones(2); rho = np.ones(2); zero = np.zeros(2)
    cases = [  # divV, curlV, c, h, expected mean factor
        ([1.0, 1.0], [0.0, 3.0], [0.0, 0.0], hh, 0.5 * (1.0 + 0.25)),
        ([2.0, 1.0], [6.0, 0.0], [0.0, 0.0], hh, 0.5 * (0.25 + 1.0)),
        ([1.0, 1.0], [0.0, 0.0], [1.0, 0.0], [1e-4, 0.5], 0.5 * (0.5 + 1.0)),
        ([0.0, 1.0], [1.0, 1.0], [0.0, 0.0], hh, 0.5 * (0.0 + 0.5)),
    ]
    for dv, cv, c, hx, f in cases:
        dv, cv, c, hx = (np.asarray(a, float) for a in (dv, cv, c, hx))
        a1 = O.accsph_allpairs(pos, vel, rho, zero, c, hx, m, 2.0)
        d1 = O.dudt_allpairs(pos, vel, rho, zero, c, hx, m, 2.0)
        a = O.accsph_allpairs(pos, vel, rho, zero, c, hx, m, 2.0, divV=dv, curlV=cv)
        d = O.dudt_allpairs(pos, vel, rho, zero, c, hx, m, 2.0, divV=dv, curlV=cv)
        assert abs(a1[0, 0]) > 0 and abs(d1[0]) > 0
        assert abs(a[0, 0] - f * a1[0, 0]) <= 1e-13 * abs(a1[0, 0])
        assert abs(d[0] - f * d1[0]) <= 1e-13 * abs(d1[0])


def _lattice(k):
    g = np.arange(k, dtype=float)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


def test_balsara_limits_on_lattice():
    """Shear flow v = (0.5 y, 0, 0): div v = 0 in the interior, so the switch removes the
    viscosity there; uniform compression v = -0.2 x with c = 0: curl v = 0, so the switch is
    exactly 1 and the switched forms equal the unswitched ones."""
    p = _lattice(13)
    n = len(p)
    h = np.full(n, 1.3)
    m = np.ones(n)
    rho = O.density_allpairs(p, m, h)
    zero = np.zeros(n)
    c = np.full(n, 0.1)
    centre = int(np.argmin(np.sum((p - 6.0) ** 2, 1)))
    # --- shear: the unswitched viscosity is non-zero, the switched one vanishes
    v = np.zeros_like(p); v[:, 0] = 0.5 * p[:, 1]
    dv, cv = O.divcurl_allpairs(p, v, rho, m, h)
    assert dv[centre] < 1e-12 and abs(cv[centre] - 0.5) < 5e-3
    d_plain = O.dudt_allpairs(p, v, rho, zero, c, h, m, 1.0, nbeg=centre, nend=centre + 1)[0]
    d_sw = O.dudt_allpairs(p, v, rho, zero, c, h, m, 1.0, dv, cv, nbeg=centre, nend=centre + 1)[0]
    assert d_plain > 1e-3 and abs(d_sw) < 1e-9 * d_plain
    # --- compression: no curl, c = 0 -> f = 1 for every particle with div != 0
    v = -0.2 * (p - 6.0)
    dv, cv = O.divcurl_allpairs(p, v, rho, m, h)
    assert abs(dv[centre] - 0.6) < 6e-3 and cv[centre] < 1e-12
    a_plain = O.accsph_allpairs(p, v, rho, zero, zero, h, m, 1.0, nbeg=centre, nend=centre + 1)
    a_sw = O.accsph_allpairs(p, v, rho, zero, zero, h, m, 1.0, dv, cv, nbeg=centre, nend=centre + 1)
    d_plain = O.dudt_allpairs(p, v, rho, zero, zero, h, m, 1.0, nbeg=centre, nend=centre + 1)[0]
    d_sw = O.dudt_allpairs(p, v, rho, zero, zero, h, m, 1.0, dv, cv, nbeg=centre, nend=centre + 1)[0]
    assert d_plain > 1e-3 and abs(d_sw - d_plain) < 1e-12 * d_plain
    assert np.all(np.abs(a_sw - a_plain) <= 1e-12 * max(np.abs(a_plain).max(), 1e-300))
