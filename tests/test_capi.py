"""The C-ABI library loads and exports every entry point include/sph_amd.h
declares (CPU only: nothing here launches a kernel)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import sph_3_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "sph_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(sph_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_expected_surface():
    names = declared()
    for must in ["sph_ctx_create", "sph_density", "sph_div_curl", "sph_acc_hydro", "sph_dudt", "sph_gravity",
                 "sph_smoothing_length", "sph_neighbours", "sph_step", "sph_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_python_binding_covers_header():
    assert sorted(L.EXPORTS) == declared()


def test_version_without_gpu():
    assert L.load().sph_version() >= 1


def test_ctx_create_rejects_bad_args_without_gpu():
    lib = L.load()
    # a null out-pointer is rejected before any device call
    assert lib.sph_ctx_create(0, None) == L.SPH_ERR_ARG
    assert lib.sph_get_stats(None, None) == L.SPH_ERR_ARG
    assert lib.sph_last_error(None) == b"null context"


def test_struct_layouts_match_header(tmp_path):
    """ctypes structures vs the header as gcc lays it out: size and every field offset."""
    structs = {"sph_step_params": L.StepParams, "sph_stats": L.Stats, "sph_cool_params": L.CoolParams,
               "sph_column_params": L.ColumnParams, "sph_agn_params": L.AgnParams}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT}/include/sph_amd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        c, f, v = line.split()
        got[(c, f)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def _validate(pos, vel=None, m=None, u=None, h=None):
    lib = L.load()
    n = len(pos)
    msg = ctypes.create_string_buffer(256)
    arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (vel, m, u, h)]
    rc = lib.sph_validate_particles(n, L.ptr(np.ascontiguousarray(pos, np.float64)), *(L.ptr(a) for a in arrs),
                                    msg, 256)
    return rc, msg.value.decode()


def test_particle_validation_without_gpu():
    """Degenerate particle sets are refused by the host check every entry point runs before
    uploading: NaN or infinite coordinates, all particles coincident, negative masses or u,
    non-positive h, non-finite velocities.  Valid sets pass."""
    rng = np.random.default_rng(0)
    pos = rng.random((100, 3)); vel = rng.normal(size=(100, 3)); m = np.full(100, 0.01)
    u = np.full(100, 0.05); h = np.full(100, 0.1)
    assert _validate(pos, vel, m, u, h) == (L.SPH_OK, "")
    bad = pos.copy(); bad[7, 2] = np.nan
    rc, msg = _validate(bad, vel, m, u, h)
    assert rc == L.SPH_ERR_ARG and "position" in msg and "particle 7" in msg
    bad = pos.copy(); bad[3, 0] = np.inf
    assert _validate(bad)[0] == L.SPH_ERR_ARG
    rc, msg = _validate(np.tile([0.3, 0.2, 0.1], (100, 1)), vel, m, u, h)
    assert rc == L.SPH_ERR_ARG and "coincide" in msg
    assert _validate(np.array([[0.3, 0.2, 0.1]]))[0] == L.SPH_OK  # a single particle has a box of its own
    for k, arr in enumerate((m, u, h)):
        b = arr.copy(); b[5] = -1.0 if k < 2 else 0.0
        args = [m, u, h]; args[k] = b
        rc, msg = _validate(pos, vel, *args)
        assert rc == L.SPH_ERR_ARG and "particle 5" in msg, msg
    b = vel.copy(); b[9, 1] = np.nan
    rc, msg = _validate(pos, b, m, u, h)
    assert rc == L.SPH_ERR_ARG and "velocity" in msg and "particle 9" in msg
    assert _validate(pos[:0])[0] == L.SPH_ERR_ARG  # n = 0
    # the Python face of the same check
    from sph_3_amd import libsx as X
    X.validate_particles(pos, vel, m, u, h)
    with pytest.raises(ValueError, match="coincide"):
        X.validate_particles(np.zeros((5, 3)))
