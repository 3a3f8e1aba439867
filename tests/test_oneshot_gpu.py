"""GPU parity of the one-shot RBC (buildRBC / queryRBC / kqueryRBC semantics) against
the oracle's restatement of rbc.cu: ranges, lists, counts and results bit-exact."""
import numpy as np
import pytest

import oracle
from conftest import datagen, rbc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("gen,n,d,R,s,metric", [
    ("uniform", 16384, 16, 128, 128, 0),   # the driver's s = numReps (driver.cu:76)
    ("gmm", 20000, 24, 100, 300, 0),       # R not a multiple of 4, ragged d
    ("gmm", 12000, 32, 64, 500, 1),        # L1
])
def test_oneshot_build_and_kquery_match_oracle(gen, n, d, R, s, metric):
    g = getattr(datagen, gen)
    x, q = g(n, d, 0), g(700, d, 1)
    reps = np.random.default_rng(11).choice(n, R, replace=False).astype(np.uint32)
    rbc.set_metric(metric)
    try:
        idx = rbc.OneShotRBC(x, R, s=s, rep_ids=reps)
        omap, ocnt, orng, bad = oracle.oneshot_build(x, reps, s, metric)
        assert bad == 0
        np.testing.assert_array_equal(idx.group_counts(), ocnt)
        ids, ds = idx.kquery(q)
        oi, od = oracle.oneshot_kquery(x, reps, omap, ocnt, q, 32, metric)
        np.testing.assert_array_equal(ids, oi)
        np.testing.assert_array_equal(ds, od)
        nid, nd = idx.query(q)
        np.testing.assert_array_equal(nid, oi[:, 0])
        np.testing.assert_array_equal(nd, od[:, 0])
        idx.close()
    finally:
        rbc.set_metric(0)


def test_oneshot_sample_shaped_matches_oracle():
    # The reference readme's example shape: 1024 x 16 database, 128 queries, -r 128.
    x = datagen.gmm(1024, 16, 0, components=8)
    q = datagen.gmm(128, 16, 1, components=8)
    reps = np.random.default_rng(0).choice(1024, 128, replace=False).astype(np.uint32)
    idx = rbc.OneShotRBC(x, 128, rep_ids=reps)
    omap, ocnt, _, _ = oracle.oneshot_build(x, reps, 128)
    np.testing.assert_array_equal(idx.group_counts(), ocnt)
    ids, ds = idx.kquery(q)
    oi, od = oracle.oneshot_kquery(x, reps, omap, ocnt, q, 32)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(ds, od)
