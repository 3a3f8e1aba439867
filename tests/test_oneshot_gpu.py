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


@pytest.mark.parametrize("kind,metric", [("random", 0), ("two_nearest", 0), ("full", 1), ("empty_rows", 0)])
def test_oneshot_plan_queries_match_oracle(kind, metric):
    # queryRBC / kqueryRBC under a general computation matrix (initCompPlan, rbc.cu:375-431):
    # ragged plans with more planned lists per group than one merge pass holds, the
    # readme's "two closest representatives" plan, fullIntersection, and empty rows.
    n, d, R, s = 12000, 20, 72, 400
    x, q = datagen.gmm(n, d, 2), datagen.gmm(900, d, 3)
    reps = np.random.default_rng(12).choice(n, R, replace=False).astype(np.uint32)
    rng = np.random.default_rng(13)
    if kind == "random":
        plan = (rng.random((R, R)) < 0.15).astype(np.uint8)
        np.fill_diagonal(plan, 1)
    elif kind == "two_nearest":
        dr = ((x[reps][:, None, :].astype(np.float64) - x[reps][None, :, :]) ** 2).sum(-1)
        plan = np.zeros((R, R), np.uint8)
        for i in range(R):
            plan[i, np.argsort(dr[i], kind="stable")[:2]] = 1
    elif kind == "full":
        plan = np.ones((R, R), np.uint8)
    else:
        plan = np.eye(R, dtype=np.uint8)
        plan[rng.choice(R, 10, replace=False)] = 0
    rbc.set_metric(metric)
    try:
        idx = rbc.OneShotRBC(x, R, s=s, rep_ids=reps)
        omap, ocnt, _, bad = oracle.oneshot_build(x, reps, s, metric)
        assert bad == 0
        oi, od = oracle.oneshot_kquery_plan(x, reps, omap, ocnt, plan, q, 32, metric)
        ids, ds = idx.kquery_plan(q, plan)
        np.testing.assert_array_equal(ids, oi)
        np.testing.assert_array_equal(ds, od)
        nid, nd = idx.query_plan(q, plan)
        np.testing.assert_array_equal(nid, oi[:, 0])
        np.testing.assert_array_equal(nd, od[:, 0])
        if kind == "empty_rows":
            assert (oi[:, 0] == 0xFFFFFFFF).any()
        idx.close()
    finally:
        rbc.set_metric(0)


def test_oneshot_identity_plan_equals_kquery():
    n, d, R = 8000, 12, 50
    x, q = datagen.uniform(n, d, 4), datagen.uniform(500, d, 5)
    idx = rbc.OneShotRBC(x, R, s=R * 3)
    ids, ds = idx.kquery(q)
    pid, pd = idx.kquery_plan(q, np.eye(R, dtype=np.uint8))
    np.testing.assert_array_equal(pid, ids)
    np.testing.assert_array_equal(pd, ds)
    idx.close()
