"""GPU cooling against the oracle and the reference's golden vectors.

Tolerances: exact FP64 device function vs oracle and vs photolibs3 values:
relative 1e-9 (device libm vs host libm, iterative solves stop at the same
1e-4 criteria); table lookup vs the exact function at grid nodes: 1e-4 dex (FP32 storage of
log10 values and FP32 cell coordinates: ~5e-6 of a cell times slopes of up to
~5 dex per cell), off-grid: rms interpolation error < 0.05 dex.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from sph_3_amd import cooling as C

pytestmark = pytest.mark.gpu
G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cooling_photolibs3.json")))


def rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def test_rates_vs_reference_vectors(gpu):
    T = np.array([e["T"] for e in G["coolingHeatingRates"]])
    nH = np.array([e["nH"] for e in G["coolingHeatingRates"]])
    h, l = C.coolingHeatingRates(T, nH)
    assert rel(h, [e["heat"] for e in G["coolingHeatingRates"]]) < 1e-9
    assert rel(l, [e["lambda"] for e in G["coolingHeatingRates"]]) < 1e-9


def test_u_functions_vs_reference_vectors(gpu):
    e = G["u_to_T_and_rate"]
    u = np.array([x["u"] for x in e]); nH = np.array([x["nH"] for x in e])
    assert rel(C.convert_u_to_temp_h(u, nH, 0.76), [x["T"] for x in e]) < 1e-9
    assert rel(C.coolingRateFromU_h(u, nH, 0.76), [x["rate"] for x in e]) < 1e-9
    ok = [x for x in G["DoCooling_h"] if "error" not in x]
    un = C.DoCooling_h(np.array([x["rho"] for x in ok]), np.array([x["u"] for x in ok]),
                       np.array([x["dt"] for x in ok]), 0.76)
    assert rel(un, [x["u_new"] for x in ok]) < 1e-8


def test_rates_vs_oracle_with_field_and_metals(gpu):
    rng = np.random.default_rng(0)
    n = 400
    T = 10 ** rng.uniform(2, 8, n); nH = 10 ** rng.uniform(-4, 4, n)
    J = 10 ** rng.uniform(-2, 2, n); Z = 10 ** rng.uniform(-3, 0, n)
    h, l = C.coolingHeatingRates(T, nH, J21=J, Z=Z)
    ref = np.array([O.cool_heat_lambda(T[i], nH[i], J[i], Z[i]) for i in range(n)])
    assert rel(h, ref[:, 0]) < 1e-9 and rel(l, ref[:, 1]) < 1e-9


def test_table_nodes_and_interpolation(gpu):
    tab = C.CoolingTable()
    # nodes reproduce the exact function (FP32 storage of log10 values)
    (a0, a1, an), (b0, b1, bn), (z0, z1, zn), (u0, u1, un) = C.TABLE_RANGES
    rng = np.random.default_rng(1)
    ia, ib, iz, iu = (rng.integers(0, k, 300) for k in (an, bn, zn, un))
    la = a0 + ia * (a1 - a0) / (an - 1); lb = b0 + ib * (b1 - b0) / (bn - 1)
    lz = z0 + iz * (z1 - z0) / (zn - 1); lu = u0 + iu * (u1 - u0) / (un - 1)
    L, Gm = tab.lookup(la, lb, lz, lu)
    h, l = C.coolingHeatingRates(10 ** lb, 10 ** la, J21=10 ** lu, Z=10 ** lz)
    assert np.max(np.abs(L - np.log10(np.maximum(l, 1e-40)))) < 1e-4
    assert np.max(np.abs(Gm - np.log10(np.maximum(h, 1e-40)))) < 1e-4
    # multilinear interpolation is exact on a function multilinear in the
    # table coordinates: check it on the table's own nodes' neighbours mixing
    # (midpoint of two adjacent nodes along T equals the mean of their values)
    la2 = la.copy(); lb2 = np.minimum(lb + (b1 - b0) / (bn - 1), b1)
    Lm, _ = tab.lookup(la, 0.5 * (lb + lb2), lz, lu)
    L2, _ = tab.lookup(la2, lb2, lz, lu)
    assert np.max(np.abs(Lm - 0.5 * (L + L2))) < 1e-4
    # off-grid accuracy vs the exact function
    n = 2000
    la = rng.uniform(a0, a1, n); lb = rng.uniform(b0, b1, n); lz = rng.uniform(z0, z1, n); lu = rng.uniform(u0, u1, n)
    L, _ = tab.lookup(la, lb, lz, lu)
    _, l = C.coolingHeatingRates(10 ** lb, 10 ** la, J21=10 ** lu, Z=10 ** lz)
    err = np.abs(L - np.log10(l))
    print("table interpolation error [dex]: rms", np.sqrt(np.mean(err ** 2)), "max", err.max())
    assert np.sqrt(np.mean(err ** 2)) < 0.05
