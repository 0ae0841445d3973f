"""Radiative cooling on the GPU: drop-in for the reference's closed-form
cooling module (Particles_in_BH_Tree_III/hcool_with_unit_tests_Here/
photolibs3.py) plus the 4D cooling table of the CR driver
(CR/v4t.h:1768-1884 doCloudyHCooling) built from the same physics.

All functions accept scalars or arrays (cgs units) and evaluate elementwise on
the device in FP64.  Keyword-only extensions: ``J21`` (UV field amplitude; the
reference's fixed value is 1e-7) and ``Z`` (metallicity in solar units, adds
the Rosen & Bregman 1995 metal-line term; 0 = the reference's primordial gas).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import COOL_MODE, MLP_NPARAM, MLP_REF, Context, CoolParams, context, f64, ptr

__all__ = ["coolingHeatingRates", "convert_u_to_temp_h", "coolingRateFromU_h", "DoCooling_h", "CoolingTable",
           "CoolingSurrogate", "CoolingParams", "J21_REFERENCE", "TABLE_RANGES"]

J21_REFERENCE = 1e-7
# (lo, hi, points) per table axis: log10 nH, log10 T, log10 Z, log10 J21
TABLE_RANGES = ((-4.0, 4.0, 33), (2.0, 8.0, 61), (-3.0, 0.0, 13), (-2.0, 2.0, 17))


def _vec(*args):
    arrs = np.broadcast_arrays(*[np.asarray(a, dtype=np.float64) for a in args])
    scalar = arrs[0].ndim == 0
    return scalar, [np.ascontiguousarray(a.reshape(-1)) for a in arrs]


def _opt(v, n):
    """(array-or-None, scalar) pair for an optional per-element argument."""
    a = np.asarray(v, dtype=np.float64)
    if a.ndim == 0:
        return None, float(a)
    return np.ascontiguousarray(np.broadcast_to(a, (n,))), 0.0


def coolingHeatingRates(T, nHcgs, *, J21=J21_REFERENCE, Z=0.0, device=0):
    """(heating, cooling) per nH^2 [erg cm^3 s^-1] -- photolibs3.py:332-376."""
    scalar, (T, nH) = _vec(T, nHcgs)
    n = len(T)
    J, Jc = _opt(J21, n)
    Za, Zc = _opt(Z, n)
    h = np.empty(n); l = np.empty(n)
    c = context(device)
    c.check(c.lib.sph_cooling_rates(c.h, n, ptr(T), ptr(nH), ptr(J), Jc, ptr(Za), Zc, ptr(h), ptr(l)),
            "coolingHeatingRates")
    return (float(h[0]), float(l[0])) if scalar else (h, l)


def _from_u(what, u, nHcgs, XH, J21, Z, device):
    scalar, (u, nH) = _vec(u, nHcgs)
    n = len(u)
    J, Jc = _opt(J21, n)
    Za, Zc = _opt(Z, n)
    out = np.empty(n)
    c = context(device)
    c.check(c.lib.sph_cooling_from_u(c.h, n, what, ptr(u), ptr(nH), float(XH), ptr(J), Jc, ptr(Za), Zc, ptr(out)),
            "cooling")
    return float(out[0]) if scalar else out


def convert_u_to_temp_h(u, nHcgs, XH, *, J21=J21_REFERENCE, device=0):
    """T(u) with the ionisation state of the gas -- photolibs3.py:379-420."""
    return _from_u(0, u, nHcgs, XH, J21, 0.0, device)


def coolingRateFromU_h(u, nHcgs, XH, *, J21=J21_REFERENCE, Z=0.0, device=0):
    """Gamma - Lambda per nH^2 at T(u) clamped to [1e4, 1e8] K -- photolibs3.py:470-486."""
    return _from_u(1, u, nHcgs, XH, J21, Z, device)


def DoCooling_h(rho, u_old, dt, XH, *, J21=J21_REFERENCE, Z=0.0, u_floor=1e8, device=0):
    """Implicit u update over dt -- photolibs3.py:489-547.  The reference's
    downward bracket search has no floor (it overflows for strong cooling);
    here such elements return -1 (bracket reached ``u_floor``)."""
    scalar, (rho, u) = _vec(rho, u_old)
    n = len(rho)
    dta, dtc = _opt(dt, n)
    J, Jc = _opt(J21, n)
    Za, Zc = _opt(Z, n)
    out = np.empty(n)
    c = context(device)
    c.check(c.lib.sph_do_cooling(c.h, n, ptr(rho), ptr(u), ptr(dta), dtc, float(XH), ptr(J), Jc, ptr(Za), Zc,
                                 float(u_floor), ptr(out)), "DoCooling_h")
    return float(out[0]) if scalar else out


def install_table(ctx: Context, ranges=TABLE_RANGES):
    """Build the cooling table in a given context (the one of a resident state)."""
    lo = np.array([r[0] for r in ranges], np.float64)
    hi = np.array([r[1] for r in ranges], np.float64)
    npts = (ctypes.c_int * 4)(*[int(r[2]) for r in ranges])
    ctx.check(ctx.lib.sph_cooling_table_build(ctx.h, ptr(lo), ptr(hi), npts), "cooling table build")


class CoolingTable:
    """4D table of log10 Lambda and log10 Gamma (per nH^2) over (log10 nH,
    log10 T, log10 Z, log10 J21), built on the device from the exact function;
    lookup is multilinear in log space with inputs clamped to the grid."""

    def __init__(self, ranges=TABLE_RANGES, device=0):
        self.ranges = tuple(tuple(r) for r in ranges)
        self.device = device
        self.ctx = Context(device)  # one context (stream + buffers) per table
        install_table(self.ctx, self.ranges)
        self.nbytes = 2 * 4 * int(np.prod([r[2] for r in ranges]))

    def lookup(self, lognH, logT, logZ, logJ21):
        _, (a, b, z, u) = _vec(lognH, logT, logZ, logJ21)
        n = len(a)
        L = np.empty(n); G = np.empty(n)
        c = self.ctx
        c.check(c.lib.sph_cooling_table_lookup(c.h, n, ptr(a), ptr(b), ptr(z), ptr(u), ptr(L), ptr(G)),
                "cooling table lookup")
        return L, G

    def temperature(self, u, nHcgs, J21, XH=0.7):
        """T(u) [K] by the temperature table of the "table" and "mlp" update modes
        (log10 T over log10 nH, log10 u, log10 J21; the table's counterpart of
        convert_u_to_temp_h)."""
        _, (u, nH, J) = _vec(u, nHcgs, J21)
        T = np.empty(len(u))
        c = self.ctx
        c.check(c.lib.sph_cooling_temperature(c.h, len(u), ptr(u), ptr(nH), ptr(J), float(XH), ptr(T)),
                "cooling temperature")
        return T

    def cooling_update(self, rho, u, Z, J21, dt, params: CoolingParams):
        """One step of the explicit cooling update (doCloudyHCooling, CR/v4t.h:1768-1884)
        in code units; mode "mlp" uses the surrogate trained and calibrated on this
        table.  Returns (u_new, info) with info = {"floored", "fallback" (counts),
        "fallback_mask"}."""
        _, (rho, u, Z, J) = _vec(rho, u, Z, J21)
        n = len(rho)
        out = np.empty(n)
        fb = np.zeros(n, np.int32)
        nfl = ctypes.c_int64(); nfb = ctypes.c_int64()
        cp = params.c()
        c = self.ctx
        c.check(c.lib.sph_cooling_update(c.h, n, ptr(rho), ptr(u), ptr(Z), ptr(J), float(dt), ctypes.byref(cp),
                                         ptr(out), ptr(fb), ctypes.byref(nfl), ctypes.byref(nfb)), "cooling update")
        return out, dict(floored=int(nfl.value), fallback=int(nfb.value), fallback_mask=fb.astype(bool))


G_CGS = 6.674e-8
M_SUN_G = 1.989e33
KPC_CM = 3.0857e21


@dataclass
class CoolingParams:
    """Units and switches of the per-step cooling update (the explicit update of
    CR/v4t.h:1768-1884 doCloudyHCooling; defaults XH = 0.7 and u floor 1.8e11 erg/g
    are that kernel's).  mode: "exact" (FP64 T(u) and rates), "table" (temperature
    table + 4D rate table), "mlp" (as "table" with log Lambda from the gated
    surrogate) or "nh" (the reference's own table over (log n_H, log T, r, log N_H),
    ``NHCoolingTable``; resident step only, with the column density on); tol / ref: the
    surrogate's gate (CoolingSurrogate.rates)."""
    unit_density_cgs: float
    unit_u_cgs: float
    unit_time_s: float
    XH: float = 0.7
    u_floor_cgs: float = 1.8e11
    mode: str = "mlp"
    tol: float = 0.05
    ref: str = "exact"

    @classmethod
    def from_mass_length(cls, mass_g: float, length_cm: float, **kw) -> "CoolingParams":
        """Code units with G = 1 for a unit mass and length in cgs: density M/L^3,
        specific energy G M / L, time sqrt(L^3 / (G M))."""
        v2 = G_CGS * mass_g / length_cm
        return cls(mass_g / length_cm ** 3, v2, length_cm / np.sqrt(v2), **kw)

    def c(self) -> CoolParams:
        if self.mode not in COOL_MODE:
            raise ValueError(f"mode must be one of {sorted(COOL_MODE)}")
        if self.ref not in MLP_REF:
            raise ValueError(f"ref must be one of {sorted(MLP_REF)}")
        return CoolParams(self.unit_density_cgs, self.unit_u_cgs, self.unit_time_s, self.XH, self.u_floor_cgs,
                          self.tol, COOL_MODE[self.mode], MLP_REF[self.ref])


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(shape) if shape is not None else a


class CoolingSurrogate:
    """Learned log10 Lambda: an MLP 4 -> 128 -> 128 -> 128 -> 1 (ReLU) over
    (log10 nH, log10 T, log10 Z, log10 J21), BF16 weights/activations with FP32
    accumulation on the matrix cores, trained on the device by its own forward
    and backward kernels on samples of the exact cooling function.

    Design point: the reference's offline Keras surrogate
    (Particles_in_BH_Tree_III/CLOUDY_replace_CHIMES/HCooling_Here/ML/prog2.py:24-52).
    ``rates`` gates every input: outside the table's box, or in a cell of the
    error map where the network's measured error exceeds ``tol`` [dex] and the
    table measured better, the result is the table's interpolation.
    ``calibrate`` measures both errors per cell against the exact FP64 function
    the table is built from, and also the network's deviation from the table
    (``ref="table"`` gates on that one alone)."""

    NPARAM = MLP_NPARAM

    def __init__(self, table: CoolingTable, seed: int = 0):
        self.table = table
        self.ctx = table.ctx
        self.cells = None
        c = self.ctx
        c.check(c.lib.sph_mlp_init(c.h, int(seed)), "surrogate init")

    # ---- parameters
    def get_params(self):
        p = np.empty(MLP_NPARAM, np.float32)
        nm = np.empty(10)
        c = self.ctx
        c.check(c.lib.sph_mlp_get_params(c.h, ptr(p), ptr(nm)), "surrogate get_params")
        return p, nm

    def set_params(self, params, norm):
        p = _f32(params)
        nm = f64(norm)
        if p.size != MLP_NPARAM or nm.size != 10:
            raise ValueError(f"params must have {MLP_NPARAM} entries and norm 10")
        c = self.ctx
        c.check(c.lib.sph_mlp_set_params(c.h, ptr(p), ptr(nm)), "surrogate set_params")
        self.cells = None

    @property
    def nbytes(self):
        """Device bytes the inference reads: FP32 small layers + BF16 hidden weights."""
        return 4 * (MLP_NPARAM - 2 * 128 * 128) + 2 * 2 * 128 * 128

    # ---- raw network (normalised inputs / outputs)
    def forward(self, x_norm):
        x = _f32(x_norm).reshape(-1, 4)
        y = np.empty(len(x), np.float32)
        c = self.ctx
        c.check(c.lib.sph_mlp_forward(c.h, len(x), ptr(x), ptr(y)), "surrogate forward")
        return y

    def gradient(self, x_norm, t_norm):
        x = _f32(x_norm).reshape(-1, 4)
        t = _f32(t_norm)
        g = np.empty(MLP_NPARAM, np.float32)
        loss = ctypes.c_double()
        c = self.ctx
        c.check(c.lib.sph_mlp_gradient(c.h, len(x), ptr(x), ptr(t), ptr(g), ctypes.byref(loss)), "surrogate gradient")
        return g, loss.value

    def train_batch(self, x_norm, t_norm, lr):
        x = _f32(x_norm).reshape(-1, 4)
        t = _f32(t_norm)
        loss = ctypes.c_double()
        c = self.ctx
        c.check(c.lib.sph_mlp_train_batch(c.h, len(x), ptr(x), ptr(t), float(lr), ctypes.byref(loss)),
                "surrogate train_batch")
        self.cells = None
        return loss.value

    # ---- training and calibration
    def train(self, n_samples=2_000_000, epochs=200, batch=4096, lr=3e-3, seed=1):
        """Train on n_samples uniform in the box (exact FP64 targets); returns the per-epoch losses."""
        losses = np.empty(int(epochs))
        c = self.ctx
        c.check(c.lib.sph_mlp_train(c.h, int(n_samples), int(epochs), int(batch), float(lr), int(seed), ptr(losses)),
                "surrogate train")
        self.cells = None
        return losses

    def calibrate(self, cells=(16, 24, 6, 8), per=16, seed=2):
        """Build the error map; returns per-cell maxima [dex] of |MLP - table|,
        |MLP - exact| and |table - exact|, each shaped ``cells``."""
        cl = (ctypes.c_int * 4)(*[int(v) for v in cells])
        ncell = int(np.prod(cells))
        et, ex, te = (np.empty(ncell, np.float32) for _ in range(3))
        c = self.ctx
        c.check(c.lib.sph_mlp_calibrate(c.h, cl, int(per), int(seed), ptr(et), ptr(ex), ptr(te)),
                "surrogate calibrate")
        self.cells = tuple(int(v) for v in cells)
        self.calibration = (self.cells, int(per), int(seed))
        return et.reshape(cells), ex.reshape(cells), te.reshape(cells)

    def install(self, ctx: Context):
        """Load this table, network and error map into another context (the one of a
        resident state, for its per-step cooling): the table is rebuilt and the map
        re-measured with the same cells, samples and seed, so the gate is identical."""
        if self.cells is None:
            raise RuntimeError("calibrate the surrogate first")
        params, norm = self.get_params()
        install_table(ctx, self.table.ranges)
        ctx.check(ctx.lib.sph_mlp_set_params(ctx.h, ptr(params), ptr(norm)), "surrogate set_params")
        cells, per, seed = self.calibration
        cl = (ctypes.c_int * 4)(*cells)
        ctx.check(ctx.lib.sph_mlp_calibrate(ctx.h, cl, per, seed, None, None, None), "surrogate calibrate")

    # ---- gated inference
    def rates(self, lognH, logT, logZ, logJ21, tol=0.02, ref="exact"):
        """(log10 Lambda, fallback mask, fallback count): the table's value wherever the
        input leaves the box or its cell fails the gate (ref="exact": network error
        > tol and > the table's error; ref="table": |network - table| > tol)."""
        if ref not in ("exact", "table"):
            raise ValueError("ref must be 'exact' or 'table'")
        _, (a, b, z, u) = _vec(lognH, logT, logZ, logJ21)
        n = len(a)
        L = np.empty(n)
        fb = np.empty(n, np.int32)
        nf = ctypes.c_int64()
        c = self.ctx
        c.check(c.lib.sph_mlp_cooling(c.h, n, ptr(a), ptr(b), ptr(z), ptr(u), float(tol), 0 if ref == "exact" else 1,
                                      ptr(L), ptr(fb), ctypes.byref(nf)), "surrogate rates")
        return L, fb.astype(bool), int(nf.value)

    def timing(self, n=1 << 22, reps=20, seed=3, tol=0.02):
        """Device ms per launch: gated network, table lookup, bare network; and the fallback fraction."""
        out = np.empty(4)
        c = self.ctx
        c.check(c.lib.sph_mlp_timing(c.h, int(n), int(reps), int(seed), float(tol), ptr(out)), "surrogate timing")
        return dict(ms_gated=out[0], ms_table=out[1], ms_network=out[2], fallback_fraction=out[3])


NH_TABLE_KEYS = ("nH", "T", "r", "NH")


class NHCoolingTable:
    """The cooling table of doCloudyHCooling (CR/v4t.h:1767-1884) over (log n_H, log T, r,
    log N_H): the four ascending grids, their clamp bounds, and log10 heating, log10
    cooling, mu and log10 u on [nH][T][r][NH], all float32 as HCoolMu.bin holds them
    (layout as CR/v3t.cu:45-122 reads it: four int32 sizes, eight float32 bounds, the
    grids, then heat, cool, mu, u)."""

    def __init__(self, grids, bounds, heat, cool, mu, logu):
        self.grids = [np.ascontiguousarray(g, np.float32).ravel() for g in grids]
        self.dims = tuple(len(g) for g in self.grids)
        self.bounds = np.ascontiguousarray(bounds, np.float32).ravel()
        if len(self.grids) != 4 or self.bounds.shape != (8,):
            raise ValueError("four grids and eight bounds")
        self.tables = [np.ascontiguousarray(t, np.float32).reshape(self.dims) for t in (heat, cool, mu, logu)]

    @classmethod
    def synthetic(cls, gamma=5.0 / 3.0):
        """A smooth stand-in on axes like the reference run's (log n_H -4..4, log T 2..8, r 0..2,
        log N_H 18..24), for timing and tests where HCoolMu.bin is absent: mu falls from 1.2 to
        0.6 across log T = 4, log u is that mu's at each T row, the rates scale as n_H^2 with
        cooling above heating beyond log T = 4 and below it under, smooth on every axis.  Not a
        physical table."""
        kb, mh = 1.3807e-16, 1.673534e-24
        grids = [np.linspace(-4.0, 4.0, 33), np.linspace(2.0, 8.0, 61), np.linspace(0.0, 2.0, 21),
                 np.linspace(18.0, 24.0, 25)]
        nH, T, r, NH = np.meshgrid(*grids, indexing="ij")
        mu = 0.6 + 0.6 / (1.0 + np.exp((T - 4.0) / 0.3))
        logu = np.log10(kb * 10.0 ** T / ((gamma - 1.0) * mu * mh))
        heat = -27.0 + 2.0 * nH + 0.1 * np.sin(1.3 * T) - 0.05 * r + 0.05 * (NH - 21.0)
        cool = heat + 0.3 * np.tanh((T - 4.0) / 0.5) + 0.05 * np.cos(nH)
        bounds = [g[i] for g in grids for i in (0, -1)]
        return cls(grids, bounds, heat, cool, mu, logu)

    @classmethod
    def read(cls, path):
        """A table file in the HCoolMu.bin layout."""
        with open(path, "rb") as f:
            dims = np.fromfile(f, np.int32, 4)
            bounds = np.fromfile(f, np.float32, 8)
            grids = [np.fromfile(f, np.float32, int(d)) for d in dims]
            cells = int(np.prod(dims.astype(np.int64)))
            tabs = [np.fromfile(f, np.float32, cells) for _ in range(4)]
        if any(len(t) != cells for t in tabs):
            raise ValueError(f"{path}: truncated table")
        return cls(grids, bounds, *tabs)

    def write(self, path):
        with open(path, "wb") as f:
            np.asarray(self.dims, np.int32).tofile(f)
            self.bounds.tofile(f)
            for g in self.grids:
                g.tofile(f)
            for t in self.tables:
                t.tofile(f)

    def as_dict(self):
        d = dict(zip(NH_TABLE_KEYS, self.grids))
        d.update(dims=self.dims, bounds=self.bounds, heat=self.tables[0], cool=self.tables[1], mu=self.tables[2],
                 logu=self.tables[3])
        return d

    def install(self, ctx: Context):
        f = ctypes.POINTER(ctypes.c_float)
        dims = np.asarray(self.dims, np.int32)
        args = [a.ctypes.data_as(f) for a in (self.bounds, *self.grids, *self.tables)]
        ctx.check(ctx.lib.sph_nh_table_set(ctx.h, ptr(dims), *args), "sph_nh_table_set")

    def cooling_update(self, pos, rho, u, log_nh, dt, params: CoolingParams, gamma=5.0 / 3.0, typ=None,
                       device=0):
        """u after one step dt (code units) of doCloudyHCooling for particles (pos, rho, u,
        log10 N_H); ``typ`` (optional) leaves particles with typ != 0 unchanged.  Returns
        (u_out, number floored)."""
        c = context(device)
        self.install(c)
        pos = f64(pos, (-1, 3))
        n = len(pos)
        rho, u, lnh = (np.broadcast_to(f64(v), (n,)).copy() for v in (rho, u, log_nh))
        t = None if typ is None else np.ascontiguousarray(typ, np.int32)
        out = np.empty(n)
        nf = np.zeros(1, np.int64)
        cp = params.c()
        c.check(c.lib.sph_nh_cooling(c.h, n, ptr(pos), ptr(rho), ptr(u), ptr(lnh), ptr(t), float(gamma), float(dt),
                                     ctypes.byref(cp), ptr(out), ptr(nf)), "sph_nh_cooling")
        return out, int(nf[0])
