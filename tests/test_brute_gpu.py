"""GPU parity of the brute-force entry points (bruteK, bruteSearch, bruteRangeCount,
bruteCPU) against the CPU oracle: indices and FP32 distances bit-exact."""
import numpy as np
import pytest

import oracle
from conftest import datagen, rbc

pytestmark = pytest.mark.gpu


def data(n, m, d, kind="uniform", seed=0):
    if kind == "uniform":
        return datagen.uniform(n, d, seed), datagen.uniform(m, d, seed + 1)
    if kind == "gmm":
        return datagen.gmm(n, d, seed), datagen.gmm(m, d, seed + 1)
    return datagen.lowrank(n, d, seed), datagen.lowrank(m, d, seed + 1)


@pytest.mark.parametrize("n,m,d,kind,metric", [
    (4096, 300, 16, "uniform", 0),
    (4096, 300, 16, "uniform", 1),
    (5000, 129, 17, "gmm", 0),     # ragged n, m and d
    (3000, 64, 64, "gmm", 1),
    (20000, 70, 128, "lowrank", 0),
])
def test_bruteK_matches_oracle(n, m, d, kind, metric):
    x, q = data(n, m, d, kind)
    ids, ds = rbc.bruteK(x, q, 32, metric)
    oi, od = oracle.brute_k_f32(x, q, 32, metric)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(ds, od)


def test_bruteK_fewer_points_than_k():
    x, q = data(20, 9, 16)
    ids, ds = rbc.bruteK(x, q, 32)
    oi, od = oracle.brute_k_f32(x, q, 32)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(ds, od)
    assert (ids[:, 20:] == rbc.DUMMY_IDX).all()


def test_bruteK_duplicates_break_ties_by_lowest_index():
    x, q = data(512, 40, 16)
    x = np.concatenate([x, x[::-1], x])  # every point three times
    ids, ds = rbc.bruteK(x, q, 32)
    oi, od = oracle.brute_k_f32(x, q, 32)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(ds, od)


@pytest.mark.parametrize("metric", [0, 1])
def test_bruteSearch_matches_oracle(metric):
    x, q = data(30000, 500, 32, "gmm")
    ids, ds = rbc.bruteSearch(x, q, metric)
    oi, od = oracle.nn_f32(x, q, metric)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(ds, od)


def test_bruteRangeCount_matches_oracle():
    x, q = data(8000, 200, 24)
    _, od = oracle.brute_k_f32(x, q, 32)
    ranges = od[:, 31]
    cnt = rbc.bruteRangeCount(x, q, ranges)
    np.testing.assert_array_equal(cnt, oracle.range_count(x, q, ranges))
    assert (cnt <= 31).all()


def test_bruteCPU_matches_oracle():
    x, q = data(2000, 50, 16)
    np.testing.assert_array_equal(rbc.bruteCPU(x, q), oracle.brute_cpu(x, q))
