"""GPU parity of the exact RBC (build and k-query) against the CPU oracle.

Indices must equal the FP64 brute-force oracle's (ties to the lowest index) and the
returned distances must be within 1e-6 relative of FP64."""
import numpy as np
import pytest

import oracle
from conftest import datagen, rbc

pytestmark = pytest.mark.gpu

REL_TOL = 1e-6  # distances vs FP64, relative


def check_against_fp64(x, q, ids, ds, k, metric=0):
    oi, od = oracle.brute_k_f64(x, q, k, metric)
    bad = np.nonzero((ids != oi).any(1))[0]
    assert bad.size == 0, f"{bad.size} queries differ, first {bad[:5]}: {ids[bad[0]]} vs {oi[bad[0]]}"
    rel = np.abs(ds.astype(np.float64) - od) / np.maximum(od, 1e-300)
    rel[od == 0] = np.abs(ds[od == 0])
    assert rel.max() <= REL_TOL, f"max relative error {rel.max()}"


@pytest.mark.parametrize("gen,n,m,d,R,k,metric", [
    ("uniform", 20000, 512, 16, 141, 1, 0),
    ("gmm", 30000, 400, 64, 173, 32, 0),
    ("gmm", 30000, 400, 64, 173, 32, 1),
    ("lowrank", 40000, 300, 128, 200, 10, 0),
    ("uniform", 5000, 257, 33, 70, 7, 1),
])
def test_exact_fp32_matches_fp64_oracle(gen, n, m, d, R, k, metric):
    g = getattr(datagen, gen)
    x, q = g(n, d, 0), g(m, d, 1)
    idx = rbc.ExactRBC(x, R, metric=metric, seed=3)
    ids, ds, st = idx.kquery(q, k, precision=rbc.PREC_FP32, stats=True)
    check_against_fp64(x, q, ids, ds, k, metric)
    assert st["pairs_scan"] + st["pairs_probe"] > 0


def test_exact_build_lists_match_oracle():
    x = datagen.gmm(20000, 32, 0)
    reps = np.sort(np.random.default_rng(5).choice(20000, 100, replace=False)).astype(np.uint32)
    idx = rbc.ExactRBC(x, 100, rep_ids=reps)
    L = idx.lists()
    np.testing.assert_array_equal(L["rep_ids"], reps)
    own = oracle.exact_owners(x, reps)
    np.testing.assert_array_equal(L["owner"], own)
    np.testing.assert_array_equal(np.diff(L["offsets"]), np.bincount(own, minlength=100))
    # Lists are the points of each owner in ascending original index.
    np.testing.assert_array_equal(L["perm"], np.argsort(own, kind="stable"))
    np.testing.assert_array_equal(L["radii"], oracle.exact_radii(x, reps, own))


def test_exact_duplicates_and_ties():
    base = datagen.uniform(3000, 16, 0)
    x = np.concatenate([base, base, base[:1000]])
    q = np.concatenate([base[:50], datagen.uniform(50, 16, 1)])  # queries equal to points
    idx = rbc.ExactRBC(x, 60, seed=1)
    ids, ds = idx.kquery(q, 8)
    check_against_fp64(x, q, ids, ds, 8)


@pytest.mark.parametrize("n,R,m,k", [(40, 1, 5, 3), (40, 40, 5, 40), (100, 7, 1, 64), (65, 9, 130, 1)])
def test_exact_edge_shapes(n, R, m, k):
    x = datagen.gmm(n, 24, 0)
    q = datagen.gmm(m, 24, 1)
    idx = rbc.ExactRBC(x, R, seed=2)
    ids, ds = idx.kquery(q, k)
    check_against_fp64(x, q, ids, ds, k)


def test_exact_rejects_bad_arguments():
    x = datagen.uniform(100, 8, 0)
    with pytest.raises(rbc.RBCError):
        rbc.ExactRBC(x, 101)
    with pytest.raises(rbc.RBCError):
        rbc.ExactRBC(x, 10, rep_ids=np.array([1, 1, 2, 3, 4, 5, 6, 7, 8, 9]))
    idx = rbc.ExactRBC(x, 10)
    with pytest.raises(rbc.RBCError):
        idx.kquery(datagen.uniform(3, 8, 1), 65)


@pytest.mark.parametrize("gen,n,m,d,R,k", [
    ("lowrank", 40000, 300, 128, 200, 10),
    ("gmm", 30000, 300, 64, 173, 32),
    ("uniform", 20000, 300, 16, 141, 1),
    ("gmm_dirichlet", 30000, 300, 96, 180, 10),
])
def test_exact_bf16_matches_fp64_oracle(gen, n, m, d, R, k):
    g = getattr(datagen, gen)
    x, q = g(n, d, 0), g(m, d, 1)
    idx = rbc.ExactRBC(x, R, seed=4)
    ids, ds, st = idx.kquery(q, k, precision=rbc.PREC_BF16, stats=True)
    check_against_fp64(x, q, ids, ds, k)
    fi, fd = idx.kquery(q, k, precision=rbc.PREC_FP32)
    np.testing.assert_array_equal(ids, fi)
    np.testing.assert_array_equal(ds, fd)


@pytest.mark.parametrize("n,R,m,k", [(40, 1, 5, 3), (300, 5, 40, 10), (100, 7, 1, 64), (65, 11, 130, 1),
                                     (2000, 13, 257, 20)])
def test_exact_bf16_edge_shapes(n, R, m, k):
    # Fewer reps than the probe's lists and the plan's plane reps, k above the probe's K',
    # item counts that are not multiples of the tiles.
    x = datagen.lowrank(n, 128, 0)
    q = datagen.lowrank(m, 128, 1)
    idx = rbc.ExactRBC(x, R, seed=2)
    ids, ds = idx.kquery(q, k, precision=rbc.PREC_BF16)
    check_against_fp64(x, q, ids, ds, k)


def test_exact_bf16_very_long_lists():
    # Two reps over 200,000 points: a list of more than 65,536 points, so the BF16 probe's
    # keys spend 15 or more bits on the point slot and keep at most 8 of the value's
    # mantissa bits; the probe then orders its candidates coarsely, and the results must
    # still be exact.
    x = datagen.lowrank(200000, 32, 7)
    q = datagen.lowrank(300, 32, 8)
    idx = rbc.ExactRBC(x, 2, seed=1)
    assert idx.info()["max_list"] > 65536
    ids, ds = idx.kquery(q, 10, precision=rbc.PREC_BF16)
    check_against_fp64(x, q, ids, ds, 10)


def test_exact_bf16_coincident_reps():
    # Rows repeated many times: some reps coincide (zero rep-rep distance, no plane).
    base = datagen.lowrank(300, 128, 3)
    x = np.concatenate([base] * 8)
    q = np.concatenate([base[:30], datagen.lowrank(30, 128, 4)])
    idx = rbc.ExactRBC(x, 200, seed=5)
    ids, ds = idx.kquery(q, 12, precision=rbc.PREC_BF16)
    check_against_fp64(x, q, ids, ds, 12)


def test_exact_bf16_duplicates():
    base = datagen.lowrank(4000, 128, 0)
    x = np.concatenate([base, base[::-1]])
    q = np.concatenate([base[:40], datagen.lowrank(40, 128, 1)])
    idx = rbc.ExactRBC(x, 64, seed=1)
    ids, ds = idx.kquery(q, 10, precision=rbc.PREC_BF16)
    check_against_fp64(x, q, ids, ds, 10)


@pytest.mark.parametrize("precision,metric", [(rbc.PREC_FP32, 0), (rbc.PREC_BF16, 0), (rbc.PREC_FP32, 1)])
def test_exact_pitched_device_queries_ignore_padding(precision, metric):
    """Device queries with ldq > d, d not a multiple of 4 and ldq a multiple of 4 (the q->reps
    distances then stage queries with 16-byte loads that reach past d), and 1e6 in the
    padding columns: only the d coordinates enter any distance."""
    import torch
    n, d, m, k, ldq = 20000, 33, 3000, 7, 36
    x = datagen.uniform(n, d, 0)
    q = datagen.uniform(m, d, 1)
    qp = np.full((m, ldq), 1e6, np.float32)
    qp[:, :d] = q
    idx = rbc.ExactRBC(x, 141, metric=metric, seed=3)
    qd = torch.from_numpy(qp).cuda()[:, :d]
    assert qd.stride(0) == ldq
    gi = torch.empty((m, k), dtype=torch.int32, device="cuda")
    gd = torch.empty((m, k), dtype=torch.float32, device="cuda")
    idx.kquery_device(qd, k, gi, gd, precision=precision)
    check_against_fp64(x, q, gi.cpu().numpy().view(np.uint32), gd.cpu().numpy(), k, metric)


def test_exact_strided_host_buffers_match_device_path():
    """Host inputs with ldq > d and outputs with ldo > k (re-pitched copies) give the same
    answers as the device path, and the output padding columns are left untouched."""
    import ctypes
    import torch
    n, d, m, k, ldq, ldo = 20000, 24, 70000, 5, 32, 8
    x = datagen.uniform(n, d, 0)
    q = datagen.uniform(m, d, 1)
    qp = np.zeros((m, ldq), np.float32)
    qp[:, :d] = q
    idx = rbc.ExactRBC(x, 141, seed=3)
    ids = np.full((m, ldo), 0xFFFFFFFE, np.uint32)
    ds = np.full((m, ldo), -7.0, np.float32)
    st = rbc.QueryStats()
    rc = rbc.lib().rbc_exact_kquery(idx.h, qp.ctypes.data, m, ldq, k, rbc.PREC_BF16, ids.ctypes.data,
                                    ds.ctypes.data, ldo, rbc.MEM_HOST, None, ctypes.byref(st))
    assert rc == 0, rbc.lib().rbc_last_error().decode()
    assert (ids[:, k:] == 0xFFFFFFFE).all() and (ds[:, k:] == -7.0).all()
    qd = torch.from_numpy(q).cuda()
    gi = torch.empty((m, k), dtype=torch.int32, device="cuda")
    gd = torch.empty((m, k), dtype=torch.float32, device="cuda")
    idx.kquery_device(qd, k, gi, gd, precision=rbc.PREC_BF16)
    np.testing.assert_array_equal(ids[:, :k], gi.cpu().numpy().view(np.uint32))
    np.testing.assert_array_equal(ds[:, :k], gd.cpu().numpy())
    check_against_fp64(x, q[:500], ids[:500, :k], ds[:500, :k], k)


def _rep_d64(x, rep_ids, q, metric):
    r = x[rep_ids].astype(np.float64)
    q = q.astype(np.float64)
    if metric == 1:
        return np.abs(q[:, None, :] - r[None, :, :]).sum(2)
    return ((q[:, None, :] - r[None, :, :]) ** 2).sum(2)


@pytest.mark.parametrize("gen,n,d,R,metric,prec", [
    ("lowrank", 20000, 128, 300, 0, "bf16"),
    ("gmm", 20000, 96, 257, 0, "bf16"),
    ("uniform", 8000, 33, 100, 0, "bf16"),
    ("gmm", 20000, 64, 173, 0, "fp32"),
    ("uniform", 8000, 33, 100, 1, "fp32"),
])
def test_rep_distances_within_their_stated_bound(gen, n, d, R, metric, prec):
    # The query-rep distances the exact query prunes with (split-BF16 MFMA for the BF16
    # filter, FP32 VALU otherwise) against FP64: within rel_err * D + abs_err[q] for every
    # pair, and the bound not loose by more than the operands' scale warrants.
    x = getattr(datagen, gen)(n, d, 0)
    q = getattr(datagen, gen)(500, d, 1)
    q[:3] = x[:3]  # queries on top of points (and reps): D near 0, where cancellation bites
    idx = rbc.ExactRBC(x, R, metric=metric, seed=5)
    p = rbc.PREC_BF16 if prec == "bf16" else rbc.PREC_FP32
    D, ae, rel = idx.rep_dists(q, p)
    D64 = _rep_d64(x, idx.lists()["rep_ids"], q, metric)
    err = np.abs(D.astype(np.float64) - D64)
    bound = rel * D64 + ae.astype(np.float64)[:, None]
    assert (err <= bound).all(), f"max excess {(err - bound).max()}"
    if prec == "bf16":
        assert rel == 0.0 and (ae > 0).all()
        scale = (q.astype(np.float64) ** 2).sum(1) + (x[idx.lists()["rep_ids"]].astype(np.float64) ** 2).sum(1).max()
        assert (ae <= 1e-3 * scale).all()
        assert err.max() > 0  # a real rounding error exists for the bound to cover
    else:
        assert (ae == 0).all() and 0 < rel < 1e-5
