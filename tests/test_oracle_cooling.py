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
