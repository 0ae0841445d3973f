"""ctypes binding of libsph_amd.so (declared in include/sph_amd.h).

The library is the only compute path: if it is missing or cannot be loaded
this module raises, it never substitutes a host implementation.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsph_amd.so")

SPH_OK, SPH_ERR_ARG, SPH_ERR_HIP, SPH_ERR_NOCONV, SPH_ERR_CAPACITY, SPH_ERR_STATE = range(6)
MLP_NPARAM = 33793
PREC = {"mixed": 0, "fp64": 1}
SOFT = {"spline": 0, "plummer": 1}
# opening rule: "geometric" accepts a node when size < theta d; "reference" is the
# reference walk's rule size^2 / theta <= d^2 (EV/bh.h:505-532), i.e. size < sqrt(theta) d
MAC = ("geometric", "reference")


def theta_of(theta, mac):
    """Opening angle the device's geometric test gets for the rule ``mac``."""
    if mac not in MAC:
        raise ValueError(f"mac must be one of {MAC}")
    theta = float(theta)
    return theta ** 0.5 if mac == "reference" else theta
H_NEWTON, H_REF_COUNT, H_KTH = 0, 1, 2

PHASES = ["kick_drift", "order", "topology", "hsolve", "moments", "ngb", "density", "gravity", "divcurl",
          "accsph", "kick", "dudt", "u_update", "moments_h", "cooling", "column", "timestep"]

_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


class StepParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("gamma", ctypes.c_double), ("alpha", ctypes.c_double),
                ("G", ctypes.c_double), ("theta", ctypes.c_double), ("eps_plummer", ctypes.c_double),
                ("n_ngb", ctypes.c_double), ("h_tol", ctypes.c_double), ("period", ctypes.c_double),
                ("C_CFL", ctypes.c_double), ("soft_kind", ctypes.c_int),
                ("precision", ctypes.c_int), ("balsara", ctypes.c_int), ("timing", ctypes.c_int),
                ("energy", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("steps", ctypes.c_int64), ("ngb_total", ctypes.c_int64),
                ("grav_pp", ctypes.c_int64), ("grav_pn", ctypes.c_int64), ("h_unconverged", ctypes.c_int64),
                ("h_rounds", ctypes.c_int64), ("phase_ms", ctypes.c_double * 20),
                ("phase_calls", ctypes.c_int64 * 20), ("h_tests", ctypes.c_int64), ("h_terms", ctypes.c_int64),
                ("ngb_extra", ctypes.c_int64), ("cool_floor", ctypes.c_int64), ("cool_fallback", ctypes.c_int64),
                ("h_big", ctypes.c_int64), ("dt_min", ctypes.c_double), ("energy", ctypes.c_double * 3),
                ("momentum", ctypes.c_double * 3), ("ang_momentum", ctypes.c_double * 3)]


class CoolParams(ctypes.Structure):
    _fields_ = [("unit_density_cgs", ctypes.c_double), ("unit_u_cgs", ctypes.c_double),
                ("unit_time_s", ctypes.c_double), ("XH", ctypes.c_double), ("u_floor_cgs", ctypes.c_double),
                ("tol", ctypes.c_double), ("mode", ctypes.c_int), ("ref", ctypes.c_int)]


class ColumnParams(ctypes.Structure):
    _fields_ = [("unit_density_cgs", ctypes.c_double), ("unit_length_cm", ctypes.c_double),
                ("XH", ctypes.c_double), ("n_ngb", ctypes.c_double)]


class AgnParams(ctypes.Structure):
    _fields_ = [("sigma", ctypes.c_double), ("M_dot", ctypes.c_double), ("v_in", ctypes.c_double),
                ("m_inject", ctypes.c_double), ("u_inject", ctypes.c_double), ("r_inject", ctypes.c_double),
                ("Z_inject", ctypes.c_double), ("J21_inject", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("h_ref_index", ctypes.c_int64)]


COOL_MODE = {"exact": 0, "table": 1, "mlp": 2, "nh": 3}
MLP_REF = {"exact": 0, "table": 1}


_SIGS = {
    "sph_version": ([], ctypes.c_int),
    "sph_ctx_create": ([ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "sph_ctx_destroy": ([_vp], ctypes.c_int),
    "sph_last_error": ([_vp], ctypes.c_char_p),
    "sph_device_bytes": ([_vp, ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
    "sph_set_distance_rows_limit": ([_vp, ctypes.c_uint64], ctypes.c_int),
    "sph_smoothing_length": ([_vp, _i64, _d, _i64, _d, _d, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, ctypes.c_double, _d, _i64p], ctypes.c_int),
    "sph_neighbours": ([_vp, _i64, _d, _d, _i64p, _i32p, _i64], ctypes.c_int),
    "sph_density": ([_vp, _i64, _d, _d, _d, ctypes.c_int, _d], ctypes.c_int),
    "sph_div_curl": ([_vp, _i64, _d, _d, _d, _d, _d, ctypes.c_int, _d, _d], ctypes.c_int),
    "sph_acc_hydro": ([_vp, _i64, _d, _d, _d, _d, _d, _d, _d, _d, _d, ctypes.c_double, ctypes.c_int, _d],
                      ctypes.c_int),
    "sph_dudt": ([_vp, _i64, _d, _d, _d, _d, _d, _d, _d, _d, _d, ctypes.c_double, ctypes.c_int, _d], ctypes.c_int),
    "sph_gravity": ([_vp, _i64, _d, _d, _d, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                     ctypes.c_int, _d, _i64p], ctypes.c_int),
    "sph_kinetics": ([_vp, _i64, _d, _d, _d], ctypes.c_int),
    "sph_kinetics_noalloc": ([_vp, _i64, _d, _d, _d], ctypes.c_int),
    "sph_potential_energy": ([_vp, _i64, _d, _d, _d, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                              ctypes.c_int, _d, _d], ctypes.c_int),
    "sph_validate_particles": ([_i64, _d, _d, _d, _d, _d, ctypes.c_char_p, _i64], ctypes.c_int),
    "sph_cooling_rates": ([_vp, _i64, _d, _d, _d, ctypes.c_double, _d, ctypes.c_double, _d, _d], ctypes.c_int),
    "sph_cooling_from_u": ([_vp, _i64, ctypes.c_int, _d, _d, ctypes.c_double, _d, ctypes.c_double, _d,
                            ctypes.c_double, _d], ctypes.c_int),
    "sph_do_cooling": ([_vp, _i64, _d, _d, _d, ctypes.c_double, ctypes.c_double, _d, ctypes.c_double, _d,
                        ctypes.c_double, ctypes.c_double, _d], ctypes.c_int),
    "sph_cooling_table_build": ([_vp, _d, _d, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "sph_cooling_table_lookup": ([_vp, _i64, _d, _d, _d, _d, _d, _d], ctypes.c_int),
    "sph_mlp_init": ([_vp, ctypes.c_uint64], ctypes.c_int),
    "sph_mlp_set_params": ([_vp, _f, _d], ctypes.c_int),
    "sph_mlp_get_params": ([_vp, _f, _d], ctypes.c_int),
    "sph_mlp_forward": ([_vp, _i64, _f, _f], ctypes.c_int),
    "sph_mlp_gradient": ([_vp, _i64, _f, _f, _f, _d], ctypes.c_int),
    "sph_mlp_train_batch": ([_vp, _i64, _f, _f, ctypes.c_double, _d], ctypes.c_int),
    "sph_mlp_train": ([_vp, _i64, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _d], ctypes.c_int),
    "sph_mlp_calibrate": ([_vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_uint64, _f, _f, _f],
                          ctypes.c_int),
    "sph_mlp_cooling": ([_vp, _i64, _d, _d, _d, _d, ctypes.c_double, ctypes.c_int, _d, _i32p, _i64p], ctypes.c_int),
    "sph_mlp_timing": ([_vp, _i64, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, _d], ctypes.c_int),
    "sph_state_set": ([_vp, _i64, _d, _d, _d, _d, _d], ctypes.c_int),
    "sph_state_get": ([_vp, _d, _d, _d, _d, _d, _d], ctypes.c_int),
    "sph_prime": ([_vp, ctypes.POINTER(StepParams)], ctypes.c_int),
    "sph_step": ([_vp, ctypes.POINTER(StepParams)], ctypes.c_int),
    "sph_state_neighbours": ([_vp, _i64p, _i32p, _i64], ctypes.c_int),
    "sph_cooling_update": ([_vp, _i64, _d, _d, _d, _d, ctypes.c_double, ctypes.POINTER(CoolParams), _d, _i32p,
                            _i64p, _i64p], ctypes.c_int),
    "sph_cooling_temperature": ([_vp, _i64, _d, _d, _d, ctypes.c_double, _d], ctypes.c_int),
    "sph_state_set_cooling": ([_vp, _d, _d, ctypes.POINTER(CoolParams)], ctypes.c_int),
    "sph_nh_table_set": ([_vp, _i32p, _f, _f, _f, _f, _f, _f, _f, _f, _f], ctypes.c_int),
    "sph_nh_cooling": ([_vp, _i64, _d, _d, _d, _d, _i32p, ctypes.c_double, ctypes.c_double,
                        ctypes.POINTER(CoolParams), _d, _i64p], ctypes.c_int),
    "sph_column_density": ([_vp, _i64, _d, _d, _d, _d, ctypes.POINTER(ColumnParams), ctypes.c_int, _d, _d],
                           ctypes.c_int),
    "sph_state_set_column": ([_vp, ctypes.POINTER(ColumnParams)], ctypes.c_int),
    "sph_state_reserve": ([_vp, _i64], ctypes.c_int),
    "sph_state_set_agn": ([_vp, ctypes.POINTER(AgnParams)], ctypes.c_int),
    "sph_state_count": ([_vp, _i64p, _d, ctypes.POINTER(ctypes.c_uint64), _i64p], ctypes.c_int),
    "sph_state_get_duprev": ([_vp, _d], ctypes.c_int),
    "sph_state_get_timestep": ([_vp, _d, _d], ctypes.c_int),
    "sph_state_set_restart": ([_vp, _d, _d, ctypes.c_double, ctypes.c_uint64, _i64], ctypes.c_int),
    "sph_state_get_column": ([_vp, _d, _d], ctypes.c_int),
    "sph_get_stats": ([_vp, ctypes.POINTER(Stats)], ctypes.c_int),
    "sph_reset_stats": ([_vp], ctypes.c_int),
}

EXPORTS = tuple(_SIGS)

_LIB = None
_LOCK = threading.Lock()


def load():
    """Load libsph_amd.so and declare its signatures (raises if absent)."""
    global _LIB
    if _LIB is None:
        with _LOCK:
            if _LIB is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(f"{LIB_PATH} is not built: run `python -m sph_3_amd.build` "
                                      "(there is no host fallback for the SPH kernels)")
                L = ctypes.CDLL(LIB_PATH)
                for name, (args, res) in _SIGS.items():
                    f = getattr(L, name)
                    f.argtypes = args
                    f.restype = res
                _LIB = L
    return _LIB


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def ptr(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_d)
    if a.dtype == np.float32:
        return a.ctypes.data_as(_f)
    if a.dtype == np.int64:
        return a.ctypes.data_as(_i64p)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_i32p)
    raise TypeError(a.dtype)


class Context:
    """One device context (one HIP stream, cached device state)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = _vp()
        rc = self.lib.sph_ctx_create(int(device), ctypes.byref(h))
        if rc != SPH_OK:
            raise RuntimeError(f"sph_ctx_create(device={device}) failed with status {rc} (no usable GPU?)")
        self.h = h
        self.device = device

    def check(self, rc: int, what: str):
        if rc != SPH_OK:
            msg = self.lib.sph_last_error(self.h)
            msg = msg.decode() if msg else ""
            raise RuntimeError(f"{what} failed (status {rc}): {msg}")

    def close(self):
        if self.h:
            self.lib.sph_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_CTX: dict = {}


def context(device: int = 0) -> Context:
    """Per-thread default context on `device`."""
    key = (threading.get_ident(), device)
    c = _CTX.get(key)
    if c is None:
        c = Context(device)
        _CTX[key] = c
    return c
