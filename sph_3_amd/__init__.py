"""sph_3_amd -- MI355X-native SPH per-step hot path (gfx950 HIP kernels).

Python call surface of the reference (``sph_code/libsx.py``,
``sph_code/shear.py``) backed by libsph_amd.so through ctypes:

    from sph_3_amd.libsx import *     # getDensity, getAcc_g_smth, h_smooth_fast, ...
    from sph_3_amd.shear import *     # div_curlVel, getAcc_sph_shear, get_dU_shear

plus a device-resident time loop (:class:`sph_3_amd.step.DeviceSPH`).
"""
from ._lib import EXPORTS, LIB_PATH, load  # noqa: F401

__version__ = "0.1.0"
