"""FP64 restatement of the reference time loop -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` (and ``bench.py``'s cpu_baseline leg) may use this module; the
product path never imports the oracle.

``KDK`` restates ``sph_code/sph_serial_v2.py``: the priming evaluation before
the loop (lines 94-143: forces at t = 0, du/dt with v at rest, then
u += du dt, u_prev = u, du_prev = du) and one pass of the loop per ``step()``
(lines 148-203):

    v += a dt/2;  x += v dt;                      # kick, drift
    h, rho, P, c, div/curl v, a = a_g + a_sph;     # forces at the new positions
    v += a dt/2;                                   # second kick
    du = du/dt(v after the kick, the forces' rho, P, c, div/curl);
    u = u_prev + dt/2 (du + du_prev);  u_prev = u;  du_prev = du.

The per-step force evaluation follows the device step's definitions, each an
oracle function pinned elsewhere: Newton-Raphson h from the previous h
(``or_h_newton``), the symmetric neighbour sets S_i (``Tree.neighbours``), the
CSR density, div/curl, acceleration and du/dt sums (restatements of
libsx.py / shear.py), and self-gravity by the group tree walk
(``Tree.gravity``; ``theta = 0`` is the exact direct sum) with Plummer or spline
softening.  ``forces`` and ``dudt`` are the two hooks the integrator calls, so
a test can replace them with closed-form fields and pin the integrator alone.
"""
from __future__ import annotations

import numpy as np

from . import (SOFT_PLUMMER, SOFT_SPLINE, Tree, accsph_csr, density_csr, divcurl_csr, dudt_csr, grav_direct,
               h_newton)


class KDK:
    def __init__(self, pos, vel, m, u, h, *, dt=1e-3, gamma=5.0 / 3.0, alpha=1.0, G=1.0, theta=0.5,
                 softening="plummer", eps_plummer=1e-3, n_ngb=64.0, h_tol=1e-13, balsara=True):
        self.pos = np.array(pos, dtype=np.float64).reshape(-1, 3)
        n = len(self.pos)
        self.vel = np.array(vel, dtype=np.float64).reshape(n, 3)
        self.m = np.broadcast_to(np.asarray(m, np.float64), (n,)).copy()
        self.u = np.broadcast_to(np.asarray(u, np.float64), (n,)).copy()
        self.h = np.broadcast_to(np.asarray(h, np.float64), (n,)).copy()
        self.dt, self.gamma, self.alpha, self.G, self.theta = dt, gamma, alpha, G, theta
        self.soft = {"plummer": SOFT_PLUMMER, "spline": SOFT_SPLINE}[softening]
        self.eps_plummer, self.n_ngb, self.h_tol, self.balsara = eps_plummer, n_ngb, h_tol, balsara
        self.acc = np.zeros((n, 3))
        self.du_prev = np.zeros(n)
        self.primed = False
        self.ev = {}

    # ------------------------------------------------------------ hooks
    def forces(self):
        """Force evaluation at the current positions and velocities; returns the total
        acceleration and keeps what du/dt needs in self.ev (sph_serial_v2.py:160-189)."""
        pos, vel, m = self.pos, self.vel, self.m
        h, bad = h_newton(pos, self.h, n_ngb=self.n_ngb, tol=self.h_tol)
        if bad:
            raise RuntimeError(f"oracle h solve: {bad} rows did not converge")
        self.h = h
        off, idx = Tree(pos, m, h=h).neighbours()
        rho = density_csr(pos, m, h, off, idx)
        P = (self.gamma - 1.0) * rho * self.u
        c = np.sqrt(self.gamma * (self.gamma - 1.0) * self.u)
        dv = cv = None
        if self.balsara:
            dv, cv = divcurl_csr(pos, vel, rho, m, h, off, idx)
        a = accsph_csr(pos, vel, rho, P, c, h, m, self.alpha, off, idx, dv, cv)
        if self.G != 0.0:
            eps = h if self.soft == SOFT_SPLINE else None
            if self.theta == 0.0:
                a = a + grav_direct(pos, m, eps, G=self.G, soft_kind=self.soft, eps_plummer=self.eps_plummer)
            else:
                g, _ = Tree(pos, m, eps=eps).gravity(theta=self.theta, G=self.G, soft_kind=self.soft,
                                                     eps_plummer=self.eps_plummer)
                a = a + g
        self.ev = dict(off=off, idx=idx, rho=rho, P=P, c=c, divV=dv, curlV=cv)
        return a

    def dudt(self):
        """du/dt with the current (kicked) velocities and the last evaluation's rho, P, c,
        div/curl v (sph_serial_v2.py:197, shear.py:332-406)."""
        e = self.ev
        return dudt_csr(self.pos, self.vel, e["rho"], e["P"], e["c"], self.h, self.m, self.alpha, e["off"], e["idx"],
                        e["divV"], e["curlV"])

    # ------------------------------------------------------------ integrator
    def prime(self):
        """sph_serial_v2.py:94-143: forces at t = 0, then u += du dt (first-order)."""
        self.acc = self.forces()
        du = self.dudt()
        self.u = self.u + du * self.dt
        self.du_prev = du
        self.primed = True

    def step(self):
        """One pass of sph_serial_v2.py:148-203."""
        if not self.primed:
            self.prime()
        dt = self.dt
        self.vel = self.vel + self.acc * (dt / 2.0)
        self.pos = self.pos + self.vel * dt
        self.acc = self.forces()
        self.vel = self.vel + self.acc * (dt / 2.0)
        du = self.dudt()
        self.u = self.u + 0.5 * dt * (du + self.du_prev)
        self.du_prev = du

    def state(self):
        out = dict(pos=self.pos.copy(), vel=self.vel.copy(), u=self.u.copy(), h=self.h.copy(), acc=self.acc.copy())
        if "rho" in self.ev:
            out["rho"] = self.ev["rho"].copy()
        return out
