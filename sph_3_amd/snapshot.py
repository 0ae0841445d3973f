"""Initial-condition and snapshot files of the reference, and restart files.

Binary layouts (little-endian int32 / float32, no padding), as the reference
writes and reads them:

* ``"evrard"`` initial conditions -- written by
  ``Particles_in_BH_Tree_III/Evrard_Collapse_2024/IC_evrard.py:137-155`` and read by
  ``…/Evrard_Collapse_2024/hfvGpuSphBHtree_newCool_v8_KH.cu:70-91``: no header (N from the
  file size, 44 bytes per particle), then Typ, x, y, z, vx, vy, vz, mass, h, eps, u.
* ``"header"`` initial conditions -- read by ``readVectorsFromFile``
  (``…/Evrard_Collapse_2024/hfvCLibs_v7_hfv_v8.1.h:58-103``, used by the AGN and CLOUDY
  drivers): int32 N, then Typ, x, y, z, vx, vy, vz, u, h, eps, mass.
* snapshots -- ``saveArraysToBinary`` (same file :110-141; reader
  ``…/Evrard_Collapse_2024/pplot3.py:52-70``): int32 N, then Typ, x, y, z, vx, vy, vz,
  rho, h, u, mass; the CLOUDY driver's variant
  (``…/CLOUDY_reloaded_Here/run_simulation/hfvCLibs_v7_hfv_v8_gradRho_v4t.h:111-143``) appends
  N_H (log10 column density).

The reference cannot feed a snapshot back as initial conditions (different
order, no softening); ``ic_from_snapshot`` converts (eps = h, as the IC
generator sets it).  Arrays are streamed with ``numpy.fromfile`` / ``tofile``
(no whole-file buffers beyond the arrays themselves).

Restart files are this package's own (the reference has no resume): a numpy
``.npz`` (no pickled objects) with the resident state and what the next step
needs besides it (acceleration, the previous du/dt, the injection's
bookkeeping); ``load_restart`` continues the run where ``save_restart`` left it.
"""
from __future__ import annotations

import os

import numpy as np

IC_FIELDS = {
    "evrard": ["Typ", "x", "y", "z", "vx", "vy", "vz", "mass", "h", "eps", "u"],
    "header": ["Typ", "x", "y", "z", "vx", "vy", "vz", "u", "h", "eps", "mass"],
}
SNAP_FIELDS = ["Typ", "x", "y", "z", "vx", "vy", "vz", "rho", "h", "u", "mass"]


def _dtype(name):
    return np.int32 if name == "Typ" else np.float32


def _write(f, fields, data, n):
    for k in fields:
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(data[k]), (n,)), dtype=_dtype(k))
        a.astype(a.dtype.newbyteorder("<"), copy=False).tofile(f)


def _read(f, fields, n):
    out = {}
    for k in fields:
        a = np.fromfile(f, dtype=np.dtype(_dtype(k)).newbyteorder("<"), count=n)
        if len(a) != n:
            raise ValueError(f"file ends inside field {k} ({len(a)} of {n} values)")
        out[k] = a.astype(_dtype(k))
    return out


def write_ic(filename, data: dict, layout="header"):
    """Write initial conditions (dict of per-particle arrays named as ``IC_FIELDS``; Typ
    defaults to 0 and eps to h) in one of the reference's two layouts."""
    if layout not in IC_FIELDS:
        raise ValueError(f"layout must be one of {sorted(IC_FIELDS)}")
    n = len(np.asarray(data["x"]))
    d = {"Typ": 0, "eps": data.get("eps", data["h"]), **data}
    with open(filename, "wb") as f:
        if layout == "header":
            np.array([n], dtype="<i4").tofile(f)
        _write(f, IC_FIELDS[layout], d, n)


def read_ic(filename, layout="header") -> dict:
    """Read initial conditions; ``"evrard"`` files carry no header (N from the size)."""
    if layout not in IC_FIELDS:
        raise ValueError(f"layout must be one of {sorted(IC_FIELDS)}")
    size = os.path.getsize(filename)
    with open(filename, "rb") as f:
        if layout == "header":
            n = int(np.fromfile(f, dtype="<i4", count=1)[0])
            if size != 4 + 44 * n:
                raise ValueError(f"size {size} does not match N = {n} ({4 + 44 * n} bytes expected)")
        else:
            if size % 44:
                raise ValueError(f"size {size} is not a whole number of 44-byte particles")
            n = size // 44
        return _read(f, IC_FIELDS[layout], n)


def write_snapshot(filename, data: dict, log_nh=None):
    """saveArraysToBinary: N, Typ, x, y, z, vx, vy, vz, rho, h, u, mass (and N_H when given)."""
    n = len(np.asarray(data["x"]))
    d = {"Typ": 0, **data}
    with open(filename, "wb") as f:
        np.array([n], dtype="<i4").tofile(f)
        _write(f, SNAP_FIELDS, d, n)
        if log_nh is not None:
            _write(f, ["NH"], {"NH": log_nh}, n)


def read_snapshot(filename) -> dict:
    """Read a snapshot of either variant (N_H present when the size says so)."""
    size = os.path.getsize(filename)
    with open(filename, "rb") as f:
        n = int(np.fromfile(f, dtype="<i4", count=1)[0])
        if size not in (4 + 44 * n, 4 + 48 * n):
            raise ValueError(f"size {size} matches neither snapshot variant for N = {n}")
        out = _read(f, SNAP_FIELDS, n)
        if size == 4 + 48 * n:
            out["NH"] = _read(f, ["NH"], n)["NH"]
    return out


def state_arrays(pos, vel, m, u, h, rho=None, Typ=0) -> dict:
    """Per-field dict of a particle state (positions and velocities as N x 3)."""
    pos = np.asarray(pos); vel = np.asarray(vel)
    d = {"Typ": Typ, "x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2], "vx": vel[:, 0], "vy": vel[:, 1],
         "vz": vel[:, 2], "mass": m, "u": u, "h": h}
    if rho is not None:
        d["rho"] = rho
    return d


def ic_from_snapshot(snap: dict) -> dict:
    """Initial conditions from a snapshot (eps = h, as IC_evrard.py sets the softening)."""
    return {k: snap[k] for k in ["Typ", "x", "y", "z", "vx", "vy", "vz", "mass", "h", "u"]} | {"eps": snap["h"]}


def positions(d: dict):
    """(pos, vel) as float64 N x 3 from a field dict."""
    return (np.stack([d["x"], d["y"], d["z"]], 1).astype(np.float64),
            np.stack([d["vx"], d["vy"], d["vz"]], 1).astype(np.float64))


# ---------------------------------------------------------------- restart
def save_restart(filename, sim, t=0.0):
    """Everything ``load_restart`` needs to continue ``sim`` (a primed ``DeviceSPH``)."""
    from ._lib import ptr

    g = sim.get()
    duprev = np.empty(sim.n)
    c = sim.ctx
    c.check(c.lib.sph_state_get_duprev(c.h, ptr(duprev)), "sph_state_get_duprev")
    _, lo, step, injected = sim.counts()
    np.savez(filename, pos=g["pos"], vel=g["vel"], m=sim.masses(), u=g["u"], h=g["h"], acc=g["acc"],
             duprev=duprev, t=np.float64(t), agn_leftover=np.float64(lo), agn_step=np.uint64(step),
             agn_injected=np.int64(injected))


def load_restart(filename, config=None, device=0, n_max=None, agn=None, cooling=None):
    """A ``DeviceSPH`` holding the saved state, primed, so ``step()`` continues the run;
    ``agn`` (``AgnParams``) and ``cooling`` (callable(sim)) re-apply the run's options."""
    from ._lib import ptr
    from .step import DeviceSPH

    with np.load(filename, allow_pickle=False) as z:
        r = {k: z[k] for k in z.files}
    sim = DeviceSPH(r["pos"], r["vel"], r["m"], r["u"], r["h"], config, device=device, n_max=n_max)
    if cooling is not None:
        cooling(sim)
    if agn is not None:
        sim.set_agn(agn)
    c = sim.ctx
    acc = np.ascontiguousarray(r["acc"], dtype=np.float64)
    dup = np.ascontiguousarray(r["duprev"], dtype=np.float64)
    c.check(c.lib.sph_state_set_restart(c.h, ptr(acc), ptr(dup), float(r["agn_leftover"]), int(r["agn_step"]),
                                        int(r["agn_injected"])), "sph_state_set_restart")
    sim.primed = True
    return sim, float(r["t"])
