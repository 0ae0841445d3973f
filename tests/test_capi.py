"""The C-ABI library loads and exports every entry point include/sph_amd.h
declares (CPU only: nothing here launches a kernel)."""
import ctypes
import os
import re

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


def test_step_params_layout_matches_header():
    # 9 doubles + 4 ints
    assert ctypes.sizeof(L.StepParams) == 10 * 8 + 4 * 4
    assert ctypes.sizeof(L.Stats) == 7 * 8 + 20 * 8 + 20 * 8 + 6 * 8 + 8
    # 6 doubles + 2 ints
    assert ctypes.sizeof(L.CoolParams) == 6 * 8 + 2 * 4
    assert ctypes.sizeof(L.ColumnParams) == 4 * 8
    assert ctypes.sizeof(L.AgnParams) == 10 * 8
