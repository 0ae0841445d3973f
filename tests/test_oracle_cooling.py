"""Pins of the oracle's cooling restatement against golden vectors written by
the reference module itself (tests/golden/make_cooling_vectors.py imports
Particles_in_BH_Tree_III/hcool_with_unit_tests_Here/photolibs3.py), plus the
closed-form checks of the build's metal term (Rosen & Bregman 1995)."""
import json
import os

import numpy as np
import pytest

import oracle as O

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cooling_photolibs3.json")))


def rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def test_rates():
    for e in G["RandCIRates"]:
        assert rel(O.cool_rates(e["T"]), e["out"]) < 1e-13, e["T"]


def test_radiation_field():
    assert rel(O.cool_radfield(O.J21_REFERENCE), G["RadiationField"]) < 1e-12
    # the field amplitude enters linearly
    assert rel(O.cool_radfield(3.0), 3.0 / O.J21_REFERENCE * np.array(G["RadiationField"])) < 1e-12


def test_abundances():
    gJ = G["RadiationField"][:3]
    for e in G["Abundance_hX"]:
        assert rel(O.cool_abund(e["T"], e["nH"], gJ), e["out"]) < 1e-12, (e["T"], e["nH"])


def test_heating_and_cooling_rates():
    for e in G["coolingHeatingRates"]:
        h, l = O.cool_heat_lambda(e["T"], e["nH"], O.J21_REFERENCE, 0.0)
        assert rel(h, e["heat"]) < 1e-11 and rel(l, e["lambda"]) < 1e-11, (e["T"], e["nH"])


def test_u_to_T_and_net_rate():
    for e in G["u_to_T_and_rate"]:
        assert rel(O.cool_u_to_T(e["u"], e["nH"]), e["T"]) < 1e-11
        assert rel(O.cool_rate_from_u(e["u"], e["nH"]), e["rate"]) < 1e-10


def test_do_cooling():
    n = 0
    for e in G["DoCooling_h"]:
        if "error" in e:
            continue
        assert rel(O.do_cooling(e["rho"], e["u"], e["dt"]), e["u_new"]) < 1e-10, e
        n += 1
    assert n >= 10


def test_metal_term_rb95():
    # continuity of the published piecewise fit at its break points
    for Tb in [2000.0, 8000.0, 1e5, 4e7]:
        assert abs(O.cool_rb95(Tb * (1 - 1e-9)) / O.cool_rb95(Tb * (1 + 1e-9)) - 1) < 2e-3
    assert O.cool_rb95(200.0) == 0.0
    # Z enters additively and linearly
    h0, l0 = O.cool_heat_lambda(2e5, 1.0, 1.0, 0.0)
    h1, l1 = O.cool_heat_lambda(2e5, 1.0, 1.0, 0.3)
    assert h1 == h0 and abs((l1 - l0) - 0.3 * O.cool_rb95(2e5)) < 1e-12 * l1


UNITS = (6.77e-25, 4.3e13, 4.7e15, 0.7, 1.8e11)  # rho_cgs, u_cgs, t_s, XH, u_floor
BOX_LO, BOX_HI = (-4.0, 2.0, -3.0, -2.0), (4.0, 8.0, 0.0, 2.0)


def test_cool_update_rate_is_photolibs3_rate():
    """Inside the box with T(u) in [1e4, 1e8] K, (u' - u) rho / (nH^2 dt) is
    coolingRateFromU_h (photolibs3.py:470-486, pinned by the reference's vectors)."""
    XH = UNITS[3]
    for rho_code, u_code, J, Z in [(1.0, 0.5, 1.0, 0.1), (30.0, 2.0, 0.05, 1.0), (0.2, 8.0, 50.0, 0.01)]:
        dt = 1e-9
        un, fl = O.cool_update([rho_code], [u_code], [Z], [J], dt, UNITS, BOX_LO, BOX_HI)
        rc = rho_code * UNITS[0]; nH = XH * rc / 1.6726e-24; uc = u_code * UNITS[1]
        T = O.cool_u_to_T(uc, nH, XH, J)
        assert 1e4 < T < 1e8 and not fl[0]
        rate = (un[0] - u_code) * UNITS[1] * rc / (nH * nH * dt * UNITS[2])
        ref = O.cool_rate_from_u(uc, nH, XH, J, Z)
        assert abs(rate / ref - 1) < 1e-6, (rate, ref)


def test_cool_update_clamps_floor_and_dt():
    XH = UNITS[3]
    # dt = 0: u unchanged above the floor, the floor below it
    u = np.array([1.0, 1e-9])
    un, fl = O.cool_update([1.0, 1.0], u, [0.1, 0.1], [1.0, 1.0], 0.0, UNITS, BOX_LO, BOX_HI)
    assert un[0] == 1.0 and not fl[0]
    assert un[1] == UNITS[4] / UNITS[1] and fl[1]
    # n_H above the box: the rate per nH^2 is the box edge's for any density there
    nH_top = 10 ** BOX_HI[0]
    rho = np.array([2.0, 50.0]) * nH_top * 1.6726e-24 / (XH * UNITS[0])
    dt = 1e-14
    un, fl = O.cool_update(rho, [3.0, 3.0], [0.1, 0.1], [1.0, 1.0], dt, UNITS, BOX_LO, BOX_HI)
    rc = rho * UNITS[0]; nH = XH * rc / 1.6726e-24
    per = (un - 3.0) * rc / (nH * nH)
    assert not fl.any() and abs(per[0] / per[1] - 1) < 1e-9
    # linear in dt while no particle floors
    d1, _ = O.cool_update([1.0], [0.5], [0.1], [1.0], 1e-9, UNITS, BOX_LO, BOX_HI)
    d2, _ = O.cool_update([1.0], [0.5], [0.1], [1.0], 2e-9, UNITS, BOX_LO, BOX_HI)
    assert abs((d2[0] - 0.5) / (d1[0] - 0.5) - 2.0) < 1e-6
