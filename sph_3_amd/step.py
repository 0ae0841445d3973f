"""Device-resident time loop: the step of ``sph_code/sph_serial_v2.py``
(lines 94-143 before the loop, 148-203 per step) with every array kept on the
GPU between steps.

Per step (KDK leapfrog):  v += a dt/2; r += v dt; Morton re-sort; tree;
FP64 Newton-Raphson h solve; CSR neighbour lists; density + EOS; tree gravity;
div/curl v (Balsara); pressure + AV acceleration (a = a_grav + a_sph);
v += a dt/2; du/dt with the kicked v; u = u_prev + dt/2 (du + du_prev);
optionally the explicit cooling update of CR/v3t.cu:1527-1536 (``set_cooling``).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import PHASES, PREC, SOFT, Stats, StepParams, context, f64, ptr, theta_of


@dataclass
class AgnParams:
    """AGN source terms of the outflow driver (code units): isothermal halo of velocity
    dispersion ``sigma`` and outflow pairs bought with ``M_dot`` per unit time, of mass
    ``m_inject``, speed ``v_in``, internal energy ``u_inject``, within ``r_inject`` of the
    origin; ``h_ref_index`` (caller index) lends its h to the first pairs."""
    sigma: float = 0.0
    M_dot: float = 0.0
    v_in: float = 0.0
    m_inject: float = 0.0
    u_inject: float = 0.0
    r_inject: float = 1e-3
    Z_inject: float = 0.0
    J21_inject: float = 1e-7  # photolibs3's fixed UV amplitude
    seed: int = 0
    h_ref_index: int = 0

    def c(self):
        from ._lib import AgnParams as _AP

        return _AP(self.sigma, self.M_dot, self.v_in, self.m_inject, self.u_inject, self.r_inject, self.Z_inject,
                   self.J21_inject, int(self.seed) & 0xFFFFFFFFFFFFFFFF, int(self.h_ref_index))


@dataclass
class StepConfig:
    dt: float = 1e-3
    gamma: float = 5.0 / 3.0
    alpha: float = 1.0
    G: float = 1.0
    theta: float = 0.5
    softening: str = "plummer"
    eps_plummer: float = 1e-3
    n_ngb: float = 64.0
    h_tol: float = 1e-10
    precision: str = "mixed"
    balsara: bool = True
    period: float = 0.0  # > 0: periodic box [0, period)^3 (minimum image; needs G = 0)
    C_CFL: float = 0.0   # > 0: per-particle time-step criteria after every force evaluation
    energy: bool = False  # FP64 totals (KE, U, PE, momentum, angular momentum) after every evaluation
    mac: str = "geometric"  # "reference": the reference walk's opening rule size^2/theta <= d^2

    def params(self, timing=False) -> StepParams:
        return StepParams(self.dt, self.gamma, self.alpha, self.G, theta_of(self.theta, self.mac), self.eps_plummer,
                          self.n_ngb, self.h_tol, self.period, self.C_CFL, SOFT[self.softening], PREC[self.precision],
                          int(self.balsara), int(timing), int(self.energy))


class DeviceSPH:
    """Resident particle state on one GPU plus the step driver."""

    def __init__(self, pos, vel, m, u, h, config: StepConfig | None = None, device: int = 0,
                 n_max: int | None = None):
        """``n_max``: room for particles injected later (``set_agn``)."""
        self.cfg = config or StepConfig()
        self.ctx = context(device)
        pos = f64(pos, (-1, 3))
        self.n = len(pos)
        n = self.n
        u = np.broadcast_to(f64(u), (n,)).copy()
        m = np.broadcast_to(f64(m), (n,)).copy()
        h = np.broadcast_to(f64(h), (n,)).copy()
        c = self.ctx
        if n_max is not None:
            c.check(c.lib.sph_state_reserve(c.h, int(n_max)), "sph_state_reserve")
        c.check(c.lib.sph_state_set(c.h, n, ptr(pos), ptr(f64(vel, (-1, 3))), ptr(m), ptr(u), ptr(h)),
                "sph_state_set")
        self.primed = False
        self._m0 = m
        self._agn = None

    def set_agn(self, params):
        """AGN source terms (``AgnParams``): the isothermal halo in every force evaluation
        and the outflow pairs injected after every step; ``None`` switches them off."""
        c = self.ctx
        if params is None:
            c.check(c.lib.sph_state_set_agn(c.h, None), "sph_state_set_agn")
            return
        ap = params.c()
        c.check(c.lib.sph_state_set_agn(c.h, ctypes.byref(ap)), "sph_state_set_agn")
        self._agn = params

    def timestep(self):
        """(dt_i, dh/dt_i) of the last force evaluation (StepConfig.C_CFL > 0), caller order."""
        dt = np.empty(self.n); dh = np.empty(self.n)
        c = self.ctx
        c.check(c.lib.sph_state_get_timestep(c.h, ptr(dt), ptr(dh)), "sph_state_get_timestep")
        return dt, dh

    def device_bytes(self) -> int:
        """Device memory the state's context holds (every buffer it allocated), in bytes."""
        b = ctypes.c_uint64(0)
        c = self.ctx
        c.check(c.lib.sph_device_bytes(c.h, ctypes.byref(b)), "sph_device_bytes")
        return int(b.value)

    def set_distance_rows_limit(self, nbytes: int):
        """Largest size of the h solve's FP64 candidate distance rows (2 KB per particle);
        beyond it the solve keeps candidate indices only and recomputes the distances (the
        same results, less device memory, a slower solve).  Default: a third of the device's
        memory."""
        c = self.ctx
        c.check(c.lib.sph_set_distance_rows_limit(c.h, int(nbytes)), "sph_set_distance_rows_limit")

    def leftover_mass(self) -> float:
        lo = np.zeros(1)
        c = self.ctx
        c.check(c.lib.sph_state_count(c.h, None, ptr(lo), None, None), "sph_state_count")
        return float(lo[0])

    def counts(self):
        """(n, leftover mass, steps since set_agn, particles injected since the state was set)."""
        n = np.zeros(1, np.int64); lo = np.zeros(1); st = np.zeros(1, np.uint64); inj = np.zeros(1, np.int64)
        c = self.ctx
        c.check(c.lib.sph_state_count(c.h, ptr(n), ptr(lo), st.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                      ptr(inj)), "sph_state_count")
        return int(n[0]), float(lo[0]), int(st[0]), int(inj[0])

    def masses(self):
        """Particle masses in caller order (the state's plus the injected particles')."""
        extra = self.n - len(self._m0)
        if extra <= 0:
            return self._m0.copy()
        return np.concatenate([self._m0, np.full(extra, self._agn.m_inject if self._agn else 0.0)])

    def set_cooling(self, Z, J21, params, surrogate=None, table_ranges=None):
        """Per-particle metallicity Z [Z_sun] and UV amplitude J21 (caller order) and the
        cooling update (``cooling.CoolingParams``) of every later step.  The cooling
        table is built in this state's context; mode "mlp" also loads ``surrogate``
        (a calibrated ``cooling.CoolingSurrogate``), mode "nh" the reference's
        ``cooling.NHCoolingTable`` passed in its place (Z and J21 unused; the column
        density must be on).  ``params=None`` switches it off."""
        from . import cooling as C

        c = self.ctx
        if params is None:
            c.check(c.lib.sph_state_set_cooling(c.h, None, None, None), "sph_state_set_cooling")
            return
        if params.mode == "nh":
            if not isinstance(surrogate, C.NHCoolingTable):
                raise ValueError('mode "nh" needs its NHCoolingTable (passed as surrogate)')
            surrogate.install(c)
            cp = params.c()
            c.check(c.lib.sph_state_set_cooling(c.h, None, None, ctypes.byref(cp)), "sph_state_set_cooling")
            return
        if params.mode == "mlp":
            if surrogate is None:
                raise ValueError('mode "mlp" needs a calibrated CoolingSurrogate')
            surrogate.install(c)
        else:
            C.install_table(c, table_ranges or (surrogate.table.ranges if surrogate else C.TABLE_RANGES))
        n = self.n
        Z = np.broadcast_to(f64(Z), (n,)).copy()
        J = np.broadcast_to(f64(J21), (n,)).copy()
        cp = params.c()
        c.check(c.lib.sph_state_set_cooling(c.h, ptr(Z), ptr(J), ctypes.byref(cp)), "sph_state_set_cooling")

    def set_column(self, params):
        """Column density (``column.ColumnParams``) computed by every later force
        evaluation after its density pass; ``params=None`` switches it off."""
        c = self.ctx
        if params is None:
            c.check(c.lib.sph_state_set_column(c.h, None), "sph_state_set_column")
            return
        cp = params.c()
        c.check(c.lib.sph_state_set_column(c.h, ctypes.byref(cp)), "sph_state_set_column")

    def column(self):
        """(log10 N_H [cm^-2], |grad rho|) of the last force evaluation, caller order."""
        n = self.n
        lognh = np.empty(n); g = np.empty(n)
        c = self.ctx
        c.check(c.lib.sph_state_get_column(c.h, ptr(g), ptr(lognh)), "sph_state_get_column")
        return lognh, g

    def prime(self):
        """Forces at the initial positions and the first u update."""
        p = self.cfg.params()
        c = self.ctx
        c.check(c.lib.sph_prime(c.h, ctypes.byref(p)), "sph_prime")
        self.primed = True

    def step(self, timing=False):
        if not self.primed:
            self.prime()
        p = self.cfg.params(timing)
        c = self.ctx
        c.check(c.lib.sph_step(c.h, ctypes.byref(p)), "sph_step")
        cnt = np.zeros(1, np.int64)
        c.check(c.lib.sph_state_count(c.h, ptr(cnt), None, None, None), "sph_state_count")
        self.n = int(cnt[0])  # grows when outflow pairs are injected

    def get(self):
        n = self.n
        pos = np.empty((n, 3)); vel = np.empty((n, 3)); acc = np.empty((n, 3))
        u = np.empty(n); h = np.empty(n); rho = np.empty(n)
        c = self.ctx
        c.check(c.lib.sph_state_get(c.h, ptr(pos), ptr(vel), ptr(u), ptr(h), ptr(rho) if self.primed else None,
                                    ptr(acc)), "sph_state_get")
        out = dict(pos=pos, vel=vel, u=u, h=h, acc=acc)
        if self.primed:
            out["rho"] = rho
        return out

    def neighbours(self):
        """(offsets[N+1], idx) of the neighbour sets the last force evaluation used (caller order)."""
        n = self.n
        off = np.empty(n + 1, np.int64)
        c = self.ctx
        c.check(c.lib.sph_state_neighbours(c.h, ptr(off), None, 0), "sph_state_neighbours")
        idx = np.empty(int(off[-1]), np.int32)
        c.check(c.lib.sph_state_neighbours(c.h, ptr(off), ptr(idx), len(idx)), "sph_state_neighbours")
        return off, idx

    def stats(self) -> dict:
        s = Stats()
        c = self.ctx
        c.check(c.lib.sph_get_stats(c.h, ctypes.byref(s)), "sph_get_stats")
        return dict(n=s.n, steps=s.steps, ngb_total=s.ngb_total, grav_pp=s.grav_pp, grav_pn=s.grav_pn,
                    h_unconverged=s.h_unconverged, h_rounds=s.h_rounds, h_tests=s.h_tests, h_terms=s.h_terms, ngb_extra=s.ngb_extra,
                    cool_floor=s.cool_floor, cool_fallback=s.cool_fallback, h_big=s.h_big, dt_min=s.dt_min,
                    energy=dict(kinetic=s.energy[0], thermal=s.energy[1], potential=s.energy[2],
                                total=s.energy[0] + s.energy[1] + s.energy[2]),
                    momentum=np.array(s.momentum[:]), ang_momentum=np.array(s.ang_momentum[:]),
                    phase_ms={k: s.phase_ms[i] for i, k in enumerate(PHASES)},
                    phase_calls={k: s.phase_calls[i] for i, k in enumerate(PHASES)})

    def reset_stats(self):
        c = self.ctx
        c.check(c.lib.sph_reset_stats(c.h), "sph_reset_stats")
