"""GPU parity of the per-function entry points against the FP64 oracle.

Tolerances (stated per result; mixed = FP32 pair arithmetic + Kahan FP32 sums):
  * neighbour sets, h_smooth_fast, do_smoothingX: bit-exact;
  * Newton h: relative 1e-9 (FP64, different summation order);
  * density: max relative 2e-6 (mixed), 1e-12 (fp64);
  * div/curl, acceleration, du/dt: max |err| / rms(|ref|) <= 1e-4 (mixed),
    1e-10 (fp64) -- normalised by the rms because individual values pass
    through zero;
  * tree gravity vs the FP64 oracle walk over the same tree: identical
    interaction counts, rel. acc. error max 1e-4 / rms 1e-5 (mixed), 1e-10
    (fp64); theta = 0 vs the direct sum: 1e-4 (mixed).
"""
import numpy as np
import pytest

import oracle as O
from sph_3_amd import ic
from sph_3_amd import libsx as X
from sph_3_amd import shear as S
from sph_3_amd._lib import H_NEWTON

pytestmark = pytest.mark.gpu

TOL = {"mixed": 1e-4, "fp64": 1e-10}


def nerr(a, b):
    a = np.asarray(a).reshape(len(b), -1); b = np.asarray(b).reshape(len(b), -1)
    ref = np.sqrt(np.mean(np.sum(b * b, 1)))
    return np.max(np.linalg.norm(a - b, axis=1)) / ref


def case(n, seed=42):
    d = ic.evrard(n, seed=seed)
    pos = d["pos"]
    h, bad = O.h_newton(pos, d["h0"], n_ngb=64.0)
    assert bad == 0
    rng = np.random.default_rng(seed + 1)
    v = rng.normal(size=(n, 3)) * 0.2
    u = rng.uniform(0.03, 0.07, n)
    m = d["m"] * rng.uniform(0.8, 1.2, n)
    return pos, v, m, h, u


@pytest.fixture(scope="module")
def small():
    pos, v, m, h, u = case(4000)
    rho = O.density_allpairs(pos, m, h)
    P = (2.0 / 3.0) * rho * u
    c = np.sqrt(10.0 / 9.0 * u)
    dv, cv = O.divcurl_allpairs(pos, v, rho, m, h)
    return dict(pos=pos, v=v, m=m, h=h, u=u, rho=rho, P=P, c=c, dv=dv, cv=cv)


def test_neighbours_bit_exact(gpu, small):
    off0, idx0 = O.neighbours_brute(small["pos"], small["h"])
    off, idx = X.neighbour_lists(small["pos"], small["h"])
    assert np.array_equal(off, off0) and np.array_equal(idx, idx0)


def test_neighbours_bit_exact_large(gpu):
    d = ic.evrard(120000, seed=7)
    h = d["h0"] * np.random.default_rng(3).uniform(0.8, 1.25, len(d["h0"]))
    T = O.Tree(d["pos"], d["m"], h=h)
    off0, idx0 = T.neighbours()
    off, idx = X.neighbour_lists(d["pos"], h)
    assert np.array_equal(off, off0) and np.array_equal(idx, idx0)
    print("mean neighbours", np.diff(off).mean())


def test_h_smooth_fast_exact(gpu):
    pos, v, m, h, u = case(3000, seed=5)
    h0 = O.h_initial(pos) * np.random.default_rng(9).uniform(0.85, 1.2, len(pos))
    ho, bad = O.h_smooth_fast(pos, h0)
    hg, badg = X.h_smooth_fast(pos, h0, return_unconverged=True)
    assert bad == badg == 0
    assert np.array_equal(hg, ho)
    # row slices
    assert np.array_equal(X.h_smooth_fast_mpi(100, 900, pos, h0), ho[100:900])


def test_do_smoothing_exact(gpu):
    pos, v, m, h, u = case(3000, seed=6)
    assert np.array_equal(X.do_smoothingX((pos, pos)), O.h_initial(pos))
    sub = pos[::7] + 1e-3
    assert np.array_equal(X.do_smoothingX((pos, sub)), O.h_initial(pos, sub))
    assert np.array_equal(X.smoothing_length_mpi(10, 500, pos), O.h_initial(pos, pos[10:500]))


def test_h_newton(gpu):
    pos, v, m, h, u = case(4000, seed=8)
    hg, unc = X._hsolve(pos, None, h * 1.1, H_NEWTON, 64.0, tol=1e-12)
    assert unc == 0
    assert np.max(np.abs(hg / h - 1)) < 1e-9


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_density(gpu, small, prec):
    r = X.getDensity(small["pos"], small["m"], small["h"], precision=prec)
    err = np.max(np.abs(r / small["rho"] - 1))
    print(prec, "density max rel", err)
    assert err < (2e-6 if prec == "mixed" else 1e-12)
    assert np.array_equal(X.getDensity_mpi(5, 50, small["pos"], small["m"], small["h"], precision=prec), r[5:50])


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_divcurl(gpu, small, prec):
    dv, cv = S.div_curlVel(small["pos"], small["v"], small["rho"], small["m"], small["h"], precision=prec)
    e1, e2 = nerr(dv, small["dv"]), nerr(cv, small["cv"])
    print(prec, "divV", e1, "curlV", e2)
    assert e1 < TOL[prec] and e2 < TOL[prec]


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
@pytest.mark.parametrize("balsara", [False, True])
def test_acc_and_dudt(gpu, small, prec, balsara):
    s = small
    args = (s["pos"], s["v"], s["rho"], s["P"], s["c"], s["h"], s["m"])
    if balsara:
        a = S.getAcc_sph_shear(*args, s["dv"], s["cv"], 1.0, precision=prec)
        du = S.get_dU_shear(*args, s["dv"], s["cv"], 1.0, precision=prec)
        a0 = O.accsph_allpairs(*args, 1.0, s["dv"], s["cv"])
        du0 = O.dudt_allpairs(*args, 1.0, s["dv"], s["cv"])
    else:
        a = X.getAcc_sph(*args, 5 / 3, 0.1, 1.0, 2.0, precision=prec)
        du = X.get_dU(*args, 5 / 3, 0.1, 1.0, 2.0, precision=prec)
        a0 = O.accsph_allpairs(*args, 1.0)
        du0 = O.dudt_allpairs(*args, 1.0)
    e1, e2 = nerr(a, a0), nerr(du, du0)
    print(prec, balsara, "acc", e1, "dudt", e2)
    assert e1 < TOL[prec] and e2 < TOL[prec]


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
@pytest.mark.parametrize("soft", ["spline", "plummer"])
def test_gravity_vs_oracle_tree(gpu, prec, soft):
    d = ic.evrard(60000, seed=11)
    n = len(d["m"])
    eps = np.full(n, 0.01) if soft == "spline" else None
    T = O.Tree(d["pos"], d["m"], eps=eps)
    kw = dict(soft_kind=O.SOFT_SPLINE if soft == "spline" else O.SOFT_PLUMMER, eps_plummer=1e-3)
    a0, n0 = T.gravity(theta=0.5, **kw)
    a, ng = X.getAcc_g_smth(d["pos"], d["m"], 1.0, eps, theta=0.5, softening=soft, eps_plummer=1e-3,
                            precision=prec, return_counts=True)
    assert ng == n0
    rel = np.linalg.norm(a - a0, axis=1) / np.linalg.norm(a0, axis=1)
    print(prec, soft, "interactions", n0, "max rel", rel.max(), "rms", np.sqrt(np.mean(rel ** 2)))
    if prec == "mixed":
        assert rel.max() < 1e-4 and np.sqrt(np.mean(rel ** 2)) < 1e-5
    else:
        assert rel.max() < 1e-10


def test_gravity_theta0_direct_sum(gpu):
    d = ic.evrard(5000, seed=12)
    eps = np.full(5000, 0.02)
    a0 = O.grav_direct(d["pos"], d["m"], eps)
    a = X.getAcc_g_smth(d["pos"], d["m"], 1.0, eps, theta=0.0)
    rel = np.linalg.norm(a - a0, axis=1) / np.linalg.norm(a0, axis=1)
    assert rel.max() < 1e-4
    a64 = X.getAcc_g_smth(d["pos"], d["m"], 1.0, eps, theta=0.0, precision="fp64")
    assert np.max(np.linalg.norm(a64 - a0, axis=1) / np.linalg.norm(a0, axis=1)) < 1e-11
    part = sum(X.getAcc_g_smth_mimj_mpi(b, e, d["pos"], d["m"], 1.0, eps, theta=0.0, precision="fp64")
               for b, e in [(0, 1700), (1700, 3300), (3300, 5000)])
    assert np.max(np.abs(part - a64)) == 0.0


def test_ke(gpu):
    rng = np.random.default_rng(0)
    v = rng.normal(size=(10000, 3)); m = rng.random(10000)
    assert abs(X.getKE(v, m) - O.ke(v, m)) < 1e-12 * O.ke(v, m)


def test_errors_are_reported(gpu):
    pos = np.random.default_rng(1).random((64, 3))
    with pytest.raises(ValueError):
        X.getDensity(pos, np.ones(64), np.ones(64), precision="bf16")
    with pytest.raises(RuntimeError, match="divV and curlV"):
        S.getAcc_sph_shear(pos, pos, np.ones(64), np.ones(64), np.ones(64), np.ones(64), np.ones(64),
                           np.ones(64), None, 1.0)
    with pytest.raises(RuntimeError, match="kth"):
        X._hsolve(pos, None, None, 2, 64)
