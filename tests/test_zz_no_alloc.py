"""GPU check of the allocation-free entry point sph_kinetics_noalloc against
sph_kinetics (same kernels, same summation order: bit-identical results), called
twice in a row with the second input smaller, so the second call runs over the
static scratch array still holding the first call's data; plus the size limit."""
import numpy as np
import pytest

import oracle as O
import sph_3_amd._lib as L

pytestmark = pytest.mark.gpu

MAX_N = 1 << 20


def kin(c, name, v, m):
    out = np.full(5, np.nan)
    c.check(getattr(c.lib, name)(c.h, len(m), L.ptr(v), L.ptr(m), L.ptr(out)), name)
    return out


def test_noalloc_kinetics_matches_kinetics(gpu):
    rng = np.random.default_rng(11)
    for n in (300000, 1000):  # the second call is smaller than the first
        v = rng.normal(size=(n, 3)); m = rng.random(n)
        ref = kin(gpu, "sph_kinetics", v, m)
        got = kin(gpu, "sph_kinetics_noalloc", v, m)
        print(n, got)
        assert np.array_equal(got, ref)
        assert abs(got[0] / O.ke(v, m) - 1) < 1e-12 and abs(got[4] / m.sum() - 1) < 1e-12
        p = (m[:, None] * v).sum(0)
        assert np.max(np.abs(got[1:4] - p)) < 1e-12 * np.sum(m * np.linalg.norm(v, axis=1))


def test_noalloc_kinetics_size_limit(gpu):
    rng = np.random.default_rng(12)
    v = rng.normal(size=(MAX_N + 1, 3)); m = rng.random(MAX_N + 1)
    got = kin(gpu, "sph_kinetics_noalloc", v[:MAX_N], m[:MAX_N])
    assert np.array_equal(got, kin(gpu, "sph_kinetics", v[:MAX_N], m[:MAX_N]))
    out = np.zeros(5)
    rc = gpu.lib.sph_kinetics_noalloc(gpu.h, MAX_N + 1, L.ptr(v), L.ptr(m), L.ptr(out))
    assert rc == L.SPH_ERR_ARG
    assert b"at most 1048576" in gpu.lib.sph_last_error(gpu.h)
