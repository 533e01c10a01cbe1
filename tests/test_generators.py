"""Seeded host generators (set_uniform / set_normal_seed fill order, darray.h:513-544).
CPU only: the generators are host code of libdarray_hip.so."""
import math

import numpy as np

import darray_optimizer_amd as dh


def test_mt19937_standard_known_answer():
    # The C++ standard ([rand.predef]) requires the 10000th output of a
    # default-constructed std::mt19937 (seed 5489) to be 4123659995.
    assert int(dh.mt19937_raw(5489, 10000)[-1]) == 4123659995


def test_mt19937_matches_numpy_legacy_seeding():
    # numpy's RandomState(seed) uses the same init_genrand seeding.
    for seed in (0, 1, 12345):
        raw = np.random.MT19937()
        raw._legacy_seeding(seed)
        ref = raw.random_raw(64).astype(np.uint32)
        np.testing.assert_array_equal(dh.mt19937_raw(seed, 64), ref)


def _canonical(raw):
    # generate_canonical<double, 53> over a 32-bit engine consumes two draws:
    # (lo + hi * 2^32) / 2^64 (libstdc++ <bits/random.tcc>).
    lo = raw[0::2].astype(np.float64)
    hi = raw[1::2].astype(np.float64)
    return (lo + hi * 2.0**32) / 2.0**64


def test_fill_uniform_order_and_formula():
    m, n = 5, 7
    A = dh.fill_uniform(0, m, n, 0.0, 1.0)
    u = _canonical(dh.mt19937_raw(0, 2 * m * n))
    # row-outer / column-inner fill (darray.h:516-521)
    np.testing.assert_array_equal(A, u.reshape(m, n))
    B = dh.fill_uniform(3, m, n, -1.0, 1.0)
    u3 = _canonical(dh.mt19937_raw(3, 2 * m * n))
    np.testing.assert_allclose(B, (u3 * 2.0 - 1.0).reshape(m, n), rtol=0, atol=1e-15)


def test_fill_normal_polar_method():
    # libstdc++ normal_distribution: Marsaglia polar on 2U-1 pairs, returns
    # y*mult first and caches x*mult.
    m, n = 4, 9
    A = dh.fill_normal(1, m, n)
    u = _canonical(dh.mt19937_raw(1, 4000))
    out, k = [], 0
    while len(out) < m * n:
        x, y = 2.0 * u[k] - 1.0, 2.0 * u[k + 1] - 1.0
        k += 2
        r2 = x * x + y * y
        if r2 > 1.0 or r2 == 0.0:
            continue
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        out += [y * mult, x * mult]
    np.testing.assert_allclose(A, np.array(out[: m * n]).reshape(m, n), rtol=1e-15, atol=1e-15)


def test_fill_normal_moments():
    A = dh.fill_normal(7, 400, 500)
    assert abs(A.mean()) < 0.01 and abs(A.std() - 1) < 0.01


def test_fill_mixture_shape_and_labels():
    X, y = dh.fill_mixture(0, 8, 2000, 4, 0.3, labels=True)
    assert X.shape == (2000, 8) and y.shape == (2000,)
    assert set(np.unique(y)) <= {-1.0, 1.0}
    X2, y2 = dh.fill_mixture(0, 8, 2000, 4, 0.3, labels=True)
    np.testing.assert_array_equal(X, X2)
    np.testing.assert_array_equal(y, y2)
    assert np.all(np.abs(X) < 1 + 6 * 0.3)
