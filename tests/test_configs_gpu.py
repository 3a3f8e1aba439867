"""Parity at the full sizes of the five measured configurations.

Every query of a configuration is answered; the FP64 oracle checks a sample of them
directly, and size-independent properties check all of them: distinct ids in rows
ordered by distance, the FP32 and BF16 filters returning identical results, and the
learned proposal leaving results unchanged.

* C1: n 65,536, d 16, U[0,1) (seed 0), m 8,192 (seed 1), 256 reps, k 1 -- one-shot
  query against the oracle's restatement with the same reps, exact query against FP64.
* C2: n 2^20, d 64, 64-component mixture, m 65,536, 1,024 reps, k 32, L2 and L1.
* C3: n 2^22, d 128, low-rank model, m 131,072, 2,048 reps, k 10; every query is also
  checked against the FP32 brute force (another kernel path) through FP64 distances.
* C4: 256 databases (n log-uniform in [2^10, 2^18], d in {32, 96}, U[0,1)), query
  batches log-uniform in [1, 4096], ceil(sqrt(n)) reps, k 16, one batched launch.
* C5: n 2^22, d 96, 1,024-component Dirichlet mixture, m 262,144 held-out queries,
  2,048 reps, k 10, proposal MLP 96-256-256-2048 trained on 2^20 pairs.
"""
import numpy as np
import pytest

import oracle
from conftest import datagen, rbc

pytestmark = pytest.mark.gpu

REL_TOL = 1e-6


def check_sample(x, q, ids, ds, k, metric, rows):
    oi, od = oracle.brute_k_f64(x, q[rows], k, metric)
    np.testing.assert_array_equal(ids[rows], oi)
    rel = np.abs(ds[rows].astype(np.float64) - od) / np.maximum(od, 1e-300)
    rel[od == 0] = np.abs(ds[rows][od == 0])
    assert rel.max() <= REL_TOL, rel.max()


def check_rows(ids, ds, n):
    """Rows are ordered by the exact distance (FP64 settles near ties) and report the
    FP32 values, so adjacent values may step back only within the FP32 error."""
    assert (ids < n).all()
    d64 = ds.astype(np.float64)
    assert (np.diff(d64, axis=1) >= -REL_TOL * d64[:, 1:]).all()
    s = np.sort(ids, axis=1)
    assert (np.diff(s.astype(np.int64), axis=1) > 0).all()           # distinct


def sample_rows(m, count, seed=0):
    return np.sort(np.random.default_rng(seed).choice(m, count, replace=False))


def test_c1_oneshot_and_exact_k1():
    x, q = datagen.uniform(65536, 16, 0), datagen.uniform(8192, 16, 1)
    reps = np.sort(np.random.default_rng(1).choice(65536, 256, replace=False)).astype(np.uint32)
    one = rbc.OneShotRBC(x, 256, rep_ids=reps)
    omap, ocnt, _, bad = oracle.oneshot_build(x, reps, 256)
    assert bad == 0
    np.testing.assert_array_equal(one.group_counts(), ocnt)
    nid, nd = one.query(q)
    oi, od = oracle.oneshot_kquery(x, reps, omap, ocnt, q, 32)
    np.testing.assert_array_equal(nid, oi[:, 0])
    np.testing.assert_array_equal(nd, od[:, 0])
    one.close()
    idx = rbc.ExactRBC(x, 256, rep_ids=reps)
    ids, ds = idx.kquery(q, 1, precision=rbc.PREC_FP32)
    bi, bd = oracle.brute_k_f64(x, q, 1)
    np.testing.assert_array_equal(ids, bi)
    ib, db = idx.kquery(q, 1, precision=rbc.PREC_BF16)
    np.testing.assert_array_equal(ib, ids)
    np.testing.assert_array_equal(db, ds)


@pytest.mark.parametrize("metric", [0, 1])
def test_c2_exact_k32(metric):
    n, m = 1 << 20, 65536
    x, q = datagen.gmm(n, 64, 0), datagen.gmm(m, 64, 1)
    idx = rbc.ExactRBC(x, 1024, metric=metric, seed=0)
    ids, ds = idx.kquery(q, 32, precision=rbc.PREC_FP32)
    check_rows(ids, ds, n)
    check_sample(x, q, ids, ds, 32, metric, sample_rows(m, 256))
    if metric == 0:
        ib, db = idx.kquery(q, 32, precision=rbc.PREC_BF16)
        np.testing.assert_array_equal(ib, ids)
        np.testing.assert_array_equal(db, ds)


def test_c3_exact_k10_bf16_and_fp32():
    n, m = 1 << 22, 131072
    x, q, _, _ = datagen.workload("lowrank16_n4M_d128")
    idx = rbc.ExactRBC(x, 2048, seed=0)
    ids, ds = idx.kquery(q, 10, precision=rbc.PREC_BF16)
    check_rows(ids, ds, n)
    i32, d32 = idx.kquery(q, 10, precision=rbc.PREC_FP32)
    np.testing.assert_array_equal(i32, ids)
    np.testing.assert_array_equal(d32, ds)
    check_sample(x, q, ids, ds, 10, 0, sample_rows(m, 96))
    # Every query against a path that shares nothing with the exact search (no lists, no
    # filters, no re-rank kernel): the FP32 brute force's 10 nearest.  Their FP64 distances
    # must equal the exact result's, query by query; ids may differ only among points
    # whose distances tie to within FP32 rounding.
    bi, _ = rbc.bruteK(x, q, 10)
    mism = 0
    for c0 in range(0, m, 8192):
        qc = q[c0:c0 + 8192].astype(np.float64)[:, None, :]
        e64 = np.sqrt(((x[ids[c0:c0 + 8192]].astype(np.float64) - qc) ** 2).sum(-1))
        b64 = np.sort(np.sqrt(((x[bi[c0:c0 + 8192]].astype(np.float64) - qc) ** 2).sum(-1)), axis=1)
        assert (np.diff(e64, axis=1) >= -1e-12 * e64[:, 1:]).all()
        assert (e64 <= b64 * (1 + REL_TOL)).all(), "the brute force found a closer point"
        assert (np.abs(e64 - b64) <= REL_TOL * b64).all()
        mism += int((np.sort(ids[c0:c0 + 8192], 1) != np.sort(bi[c0:c0 + 8192], 1)).any(1).sum())
    assert mism <= m // 1000, f"{mism} queries differ in ids from the brute force"


def test_c4_batched_256_databases():
    ns, dims, ms = datagen.batched_shapes()
    xs = [datagen.uniform(int(ns[i]), int(dims[i]), 100 + i) for i in range(len(ns))]
    qs = [datagen.uniform(int(ms[i]), int(dims[i]), 10000 + i) for i in range(len(ns))]
    reps = [int(np.ceil(np.sqrt(v))) for v in ns]
    b = rbc.BatchedRBC(xs, reps)
    ids, ds, st = b.kquery(qs, 16)
    assert st["tile_queries"] in (16, 32, 64)
    for i in range(len(ns)):
        check_rows(ids[i], ds[i], int(ns[i]))
    # FP64 oracle on every instance whose brute force is small, plus the largest database.
    cost = ns * ms * dims
    picks = list(np.argsort(cost)[:40]) + [int(np.argmax(ns))]
    for i in picks:
        rows = np.arange(ms[i]) if ms[i] <= 64 else sample_rows(int(ms[i]), 64, i)
        check_sample(xs[i], qs[i], ids[i], ds[i], 16, 0, rows)


def test_c5_learned_proposal_full_size():
    """The learned-proposal form at full size: the MLP 96 -> 256 -> 256 -> 2048 trained on
    2^20 training queries' exact 10-NN lists (from exact RBC itself) picks each query's
    probe lists; the probe work falls by more than a third, the total of probe and scan
    work falls, and the results are identical to the search without it."""
    n, m, d, R, k = 1 << 22, 262144, 96, 2048, 10
    x, q, _, _ = datagen.workload("gmm1024_n4M_d96")
    idx = rbc.ExactRBC(x, R, seed=0)
    i0, d0, s0 = idx.kquery(q, k, precision=rbc.PREC_BF16, stats=True)
    check_rows(i0, d0, n)
    qt = datagen.training_queries("gmm1024_n4M_d96", 1 << 20)
    nn, _ = idx.kquery(qt, k, precision=rbc.PREC_BF16)
    labels = idx.lists()["owner"][nn]
    net = rbc.ProposalMLP(d, R, seed=0)
    assert net.num_params == 96 * 256 + 256 + 256 * 256 + 256 + 256 * 2048 + 2048
    losses = net.train(qt, labels, epochs=2, batch=1024, lr=1e-3)
    assert losses[-1] < losses[0]
    idx.set_proposal(net)
    i1, d1, s1 = idx.kquery(q, k, precision=rbc.PREC_BF16, stats=True)
    np.testing.assert_array_equal(i1, i0)
    np.testing.assert_array_equal(d1, d0)
    assert s1["pairs_probe"] < 0.65 * s0["pairs_probe"]
    assert s1["pairs_probe"] + s1["pairs_scan"] < s0["pairs_probe"] + s0["pairs_scan"]
    check_sample(x, q, i0, d0, k, 0, sample_rows(m, 96))
    print(f"C5 with the proposal: probe pairs {s0['pairs_probe']:.4e} -> {s1['pairs_probe']:.4e}, "
          f"scan pairs {s0['pairs_scan']:.4e} -> {s1['pairs_scan']:.4e}")


def _device_rows(ptr, rows, ld, count):
    """Copies rows x ld uint32 from a device pointer (the rbcStruct's dxMap)."""
    import ctypes
    import torch  # noqa: F401  (its HIP runtime serves this process)
    hip = ctypes.CDLL("libamdhip64.so.7")
    out = np.empty((rows, ld), np.uint32)
    assert hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr),
                         ctypes.c_size_t(out.nbytes), 2) == 0
    return out[:, :count]


def test_oneshot_build_at_c3_scale():
    """buildRBC at n = 2^22, d = 128, 2,048 reps, s = 2,048 (sizes at which the reference's
    32-bit arithmetic overflows): counts within the bisection's target band, and three
    4-rep blocks (the bisection couples the reps of a block) equal to the oracle's
    restatement run on just those reps."""
    x, _, _, _ = datagen.workload("lowrank16_n4M_d128", m=1)
    n, R, s = x.shape[0], 2048, 2048
    reps = np.sort(np.random.default_rng(7).choice(n, R, replace=False)).astype(np.uint32)
    one = rbc.OneShotRBC(x, R, s=s, rep_ids=reps)
    cnt = one.group_counts()
    assert (cnt <= s).all() and (cnt >= int(0.9 * s)).all()
    mp = _device_rows(one.rbcs.dxMap.mat, R, one.rbcs.dxMap.ld, s)
    for b0 in (0, 1000, 2044):
        omap, ocnt, _, bad = oracle.oneshot_build(x, reps[b0:b0 + 4], s)
        assert bad == 0
        np.testing.assert_array_equal(cnt[b0:b0 + 4], ocnt)
        for r in range(4):
            np.testing.assert_array_equal(mp[b0 + r, :ocnt[r]], omap[r, :ocnt[r]])
    one.close()
