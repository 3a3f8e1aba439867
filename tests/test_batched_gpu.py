"""Batched exact search (many independent databases in one persistent, graph-captured
launch): every instance's indices bit-exact against the FP64 oracle, distances within
1e-6 relative; graph replay gives the same answers."""
import numpy as np
import pytest

import oracle
from conftest import datagen, rbc

pytestmark = pytest.mark.gpu


def make_batch(count, nmax_log2, mmax, seed=4242):
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** nmax_log2), size=count)).astype(int)
    d = rng.choice([32, 96], size=count)
    m = np.maximum(1, np.exp(rng.uniform(0, np.log(mmax), size=count)).astype(int))
    xs = [datagen.uniform(int(n[i]), int(d[i]), 100 + i) for i in range(count)]
    qs = [datagen.uniform(int(m[i]), int(d[i]), 500 + i) for i in range(count)]
    reps = [int(np.ceil(np.sqrt(n[i]))) for i in range(count)]
    return xs, qs, reps


def test_batched_matches_fp64_oracle_per_instance():
    k = 16
    xs, qs, reps = make_batch(24, 14, 200)
    b = rbc.BatchedRBC(xs, reps)
    ids, ds, st = b.kquery(qs, k)
    assert st["tile_queries"] in (16, 32, 64)
    for i, (x, q) in enumerate(zip(xs, qs)):
        oi, od = oracle.brute_k_f64(x, q, k)
        np.testing.assert_array_equal(ids[i], oi, err_msg=f"instance {i}")
        rel = np.abs(ds[i].astype(np.float64) - od) / np.maximum(od, 1e-300)
        assert rel.max() <= 1e-6, f"instance {i}: {rel.max()}"


def test_batched_graph_replay_is_stable():
    import torch
    k = 16
    xs, qs, reps = make_batch(12, 13, 300, seed=7)
    b = rbc.BatchedRBC(xs, reps)
    dq = [torch.from_numpy(q).cuda() for q in qs]
    ids = [torch.empty((q.shape[0], k), dtype=torch.int32, device="cuda") for q in qs]
    ds = [torch.empty((q.shape[0], k), dtype=torch.float32, device="cuda") for q in qs]
    s1 = b.kquery_device(dq, k, ids, ds)
    first = [(i.cpu().numpy().copy(), d.cpu().numpy().copy()) for i, d in zip(ids, ds)]
    for rep in range(3):
        s = b.kquery_device(dq, k, ids, ds)
        assert s["graph_replayed"] == 1
        for (fi, fd), i, d in zip(first, ids, ds):
            np.testing.assert_array_equal(fi, i.cpu().numpy())
            np.testing.assert_array_equal(fd, d.cpu().numpy())
    for i in range(0, 12, 3):
        oi, _ = oracle.brute_k_f64(xs[i], qs[i], k)
        np.testing.assert_array_equal(first[i][0].astype(np.uint32), oi)
