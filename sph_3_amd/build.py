"""Build libsph_amd.so (gfx950) in place with hipcc.

Each .hip source is compiled to an object for gfx950 only, then linked into
``sph_3_amd/libsph_amd.so``.  Packed-FP32 instruction selection is disabled
(``-packed-fp32-ops``); the ordering, tree and neighbour sources also turn FP
contraction off with a file-level pragma so their FP64 decisions match the CPU
oracle bit for bit.

Usage: ``python -m sph_3_amd.build`` (or ``build()`` from Python).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsph_amd.so")
SOURCES = ["sph_order.hip", "sph_tree.hip", "sph_ngb.hip", "sph_hydro.hip", "sph_grav.hip",
           "sph_integrate.hip", "sph_cool.hip", "sph_nhcool.hip", "sph_mlp.hip", "sph_agn.hip", "sph_capi.hip",
           "sph_capi_api.hip", "sph_capi_step.hip", "sph_capi_cool.hip"]
ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-parameter",
          "-Wno-unused-command-line-argument"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
