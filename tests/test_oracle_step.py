"""Pins of the oracle's KDK step (oracle/step.py) -- CPU only.

The integrator is checked with closed-form fields: for a harmonic force
a = -w^2 x one kick-drift-kick step is the linear map

    [x']   [ 1 - w^2 dt^2/2          dt            ] [x]
    [v'] = [ -w^2 dt (1 - w^2 dt^2/4)  1 - w^2 dt^2/2 ] [v],

which conserves v^2 + w^2 (1 - w^2 dt^2/4) x^2 exactly (textbook leapfrog
results, written independently of the step's sequential updates), and u
follows the trapezoid of sph_serial_v2.py:201 with du/dt taken at the velocity
after the second kick.  A drift before the first kick, a missing half kick, du/dt
taken before the second kick or an Euler u update each fail one of these.
The SPH force composition is checked against the all-pairs restatements
(libsx.py / shear.py) and the direct-sum gravity at theta = 0.
"""
import numpy as np
import pytest

import oracle as O
from oracle.step import KDK


class Harmonic(KDK):
    w = 3.0

    def forces(self):
        return -self.w ** 2 * self.pos

    def dudt(self):
        return np.sum(self.vel ** 2, axis=1)  # any function of the kicked velocity


def test_kdk_harmonic_closed_form():
    rng = np.random.default_rng(0)
    n, dt, w, steps = 5, 0.01, 3.0, 40
    x0 = rng.normal(size=(n, 3)); v0 = rng.normal(size=(n, 3)); u0 = rng.uniform(1, 2, n)
    sim = Harmonic(x0, v0, np.ones(n), u0, np.ones(n), dt=dt)
    a2 = (w * dt) ** 2
    M = np.array([[1 - a2 / 2, dt], [-w * w * dt * (1 - a2 / 4), 1 - a2 / 2]])
    xs, vs = [x0], [v0]
    for _ in range(steps):
        x, v = xs[-1], vs[-1]
        xs.append(M[0, 0] * x + M[0, 1] * v)
        vs.append(M[1, 0] * x + M[1, 1] * v)
    g = [np.sum(v * v, 1) for v in vs]
    u_exp = u0 + dt * g[0]  # priming update: u += du dt with du at rest positions/velocities
    sim.prime()
    assert np.allclose(sim.u, u_exp, rtol=0, atol=1e-14)
    inv0 = np.sum(v0 ** 2, 1) + w * w * (1 - a2 / 4) * np.sum(x0 ** 2, 1)
    for k in range(1, steps + 1):
        sim.step()
        u_exp = u_exp + 0.5 * dt * (g[k] + g[k - 1])
        assert np.abs(sim.pos - xs[k]).max() < 1e-12
        assert np.abs(sim.vel - vs[k]).max() < 1e-12
        assert np.abs(sim.u - u_exp).max() < 1e-12
        inv = np.sum(sim.vel ** 2, 1) + w * w * (1 - a2 / 4) * np.sum(sim.pos ** 2, 1)
        assert np.abs(inv - inv0).max() < 1e-12 * inv0.max()


def _setup(n=600, seed=3):
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.random(n)); ct = rng.uniform(-1, 1, n); ph = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - ct * ct)
    pos = np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * ct], 1)
    vel = rng.normal(size=(n, 3)) * 0.2
    m = np.full(n, 1.0 / n)
    u = rng.uniform(0.03, 0.07, n)
    h0 = O.h_initial(pos)
    return pos, vel, m, u, h0


@pytest.mark.parametrize("balsara", [True, False])
def test_kdk_prime_is_the_reference_composition(balsara):
    """Priming evaluation = the libsx/shear all-pairs functions at the solved h, with
    P and c from u before its update, plus the direct-sum gravity (theta = 0)."""
    pos, vel, m, u, h0 = _setup()
    sim = KDK(pos, vel, m, u, h0, theta=0.0, softening="plummer", eps_plummer=0.05, alpha=1.0, balsara=balsara,
              n_ngb=40.0)
    sim.prime()
    h = sim.h
    d = np.sqrt(((pos[:, None] - pos[None]) ** 2).sum(-1))
    q = d / h[:, None]
    wq = np.where(q < 1, 1 - 1.5 * q ** 2 + 0.75 * q ** 3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0))
    assert np.abs((32.0 / 3.0) * wq.sum(1) - 40.0).max() < 1e-8  # h solves its equation
    rho = O.density_allpairs(pos, m, h)
    P = (2.0 / 3.0) * rho * u
    c = np.sqrt(10.0 / 9.0 * u)
    kw = {}
    if balsara:
        dv, cv = O.divcurl_allpairs(pos, vel, rho, m, h)
        kw = dict(divV=dv, curlV=cv)
    a = O.accsph_allpairs(pos, vel, rho, P, c, h, m, 1.0, **kw)
    a = a + O.grav_direct(pos, m, None, soft_kind=O.SOFT_PLUMMER, eps_plummer=0.05)
    du = O.dudt_allpairs(pos, vel, rho, P, c, h, m, 1.0, **kw)
    assert np.abs(sim.ev["rho"] - rho).max() < 1e-13 * rho.max()
    assert np.abs(sim.acc - a).max() < 1e-12 * np.abs(a).max()
    assert np.abs(sim.u - (u + du * sim.dt)).max() < 1e-15


def test_kdk_sph_step_conserves_momentum_and_energy():
    """Exact direct-sum gravity (theta = 0) and pairwise hydro forces: total momentum is
    conserved to round-off; total energy KE + U + PE (Plummer) drifts only at O(dt^2)."""
    pos, vel, m, u, h0 = _setup(500, 5)
    vel -= (m[:, None] * vel).sum(0) / m.sum()
    eps = 0.05
    sim = KDK(pos, vel, m, u, h0, theta=0.0, softening="plummer", eps_plummer=eps, dt=2e-4, n_ngb=40.0)

    def energy(s):
        return O.ke(s.vel, m) + np.sum(m * s.u) + O.pe(s.pos, m, 1.0, np.zeros(len(m)), soft_kind=O.SOFT_PLUMMER,
                                                        eps_plummer=eps)

    sim.prime()
    e0 = energy(sim)
    for _ in range(5):
        sim.step()
    p = (m[:, None] * sim.vel).sum(0)
    assert np.abs(p).max() < 1e-14 * (m[:, None] * np.abs(sim.vel)).sum()
    assert abs(energy(sim) - e0) < 1e-5 * abs(e0)
