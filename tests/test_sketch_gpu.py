"""The projected lower-bound filter (sketch) of the BF16 exact query.

The bound the filter rests on is checked directly against FP64: every stored sketch
coordinate is within its stated error of P (x - mu) evaluated exactly, P has orthonormal
rows to within the stated norm bound, and for every (query, point) pair the sketch
distance is at most s_P |q - x| + e_q + e_x.  The k-query results with the sketch equal
the FP64 oracle's (ids bit-exact, distances within 1e-6 relative) and are identical to
those of the full-dimension BF16 filter and of the FP32 path."""
import numpy as np
import pytest

from conftest import datagen, rbc
from test_exact_gpu import check_against_fp64

pytestmark = pytest.mark.gpu


def bf16_to_f64(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def test_sketch_rows_within_their_stated_error():
    x = datagen.lowrank(30000, 128, 0)
    q = datagen.lowrank(200, 128, 1)
    idx = rbc.ExactRBC(x, 150, seed=4)
    info = idx.sketch_info()
    assert info["dims"] == 28 and info["active"] == 1
    P, mu = idx.sketch_basis()
    P64, mu64 = P.astype(np.float64), mu.astype(np.float64)
    # |P|_2 <= s_P, and P is orthonormal to FP32 rounding.
    sv = np.linalg.svd(P64, compute_uv=False)
    assert sv.max() <= info["norm_bound"] and info["norm_bound"] < 1 + 1e-5
    assert np.abs(P64 @ P64.T - np.eye(28)).max() < 1e-6
    # Leading principal directions: the variance kept is nearly all of it.
    assert info["resid_var"] < 1e-3 * info["total_var"]
    for rows in (x[:5000], q):
        bits, err, n2 = idx.sketch_rows(rows)
        y = bf16_to_f64(bits[:, :28])
        exact = (rows.astype(np.float64) - mu64) @ P64.T
        dev = np.linalg.norm(y - exact, axis=1)
        assert (dev <= err.astype(np.float64)).all(), f"worst {np.max(dev / err)}"
        # The stated error is dominated by the BF16 rounding, not a loose bound.
        assert np.median(err) < 4 * np.median(dev) + 1e-3
        np.testing.assert_allclose(n2, (y * y).sum(1), rtol=1e-6)
        # Query form: coordinates 28..31 are 1, 1, 0, 0.
        assert (bits[:, 28:] == np.array([0x3F80, 0x3F80, 0, 0], np.uint16)[None, :]).all()


def test_sketch_distance_bounds_true_distance_for_every_pair():
    x = datagen.lowrank(8000, 96, 0)
    q = np.concatenate([datagen.lowrank(100, 96, 1), x[:20]])  # queries on top of points too
    idx = rbc.ExactRBC(x, 80, seed=2)
    idx.set_sketch(rbc.SKETCH_ON)
    sP = idx.sketch_info()["norm_bound"]
    bx, ex, _ = idx.sketch_rows(x)
    bq, eq, _ = idx.sketch_rows(q)
    yx, yq = bf16_to_f64(bx[:, :28]), bf16_to_f64(bq[:, :28])
    x64, q64 = x.astype(np.float64), q.astype(np.float64)
    for i in range(len(q)):
        true = np.sqrt(((x64 - q64[i]) ** 2).sum(1))
        sk = np.sqrt(((yx - yq[i]) ** 2).sum(1))
        slack = sP * true + eq[i] + ex - sk
        assert slack.min() >= 0, f"query {i}: bound violated by {-slack.min()}"


@pytest.mark.parametrize("gen,n,m,d,R,k,mode", [
    ("lowrank", 60000, 700, 128, 240, 10, rbc.SKETCH_AUTO),
    ("lowrank", 30000, 300, 96, 170, 32, rbc.SKETCH_AUTO),
    ("lowrank", 20000, 1, 128, 100, 1, rbc.SKETCH_AUTO),
    ("uniform", 20000, 300, 64, 141, 16, rbc.SKETCH_ON),   # loose bound, still exact
    ("gmm", 30000, 400, 64, 173, 32, rbc.SKETCH_ON),
    ("gmm_dirichlet", 30000, 300, 96, 173, 10, rbc.SKETCH_ON),
])
def test_sketch_query_matches_fp64_and_the_full_filter(gen, n, m, d, R, k, mode):
    g = getattr(datagen, gen)
    x, q = g(n, d, 0), g(m, d, 1)
    idx = rbc.ExactRBC(x, R, seed=3)
    idx.set_sketch(mode)
    ids, ds, st = idx.kquery(q, k, precision=rbc.PREC_BF16, stats=True)
    assert st["sketch"] == 1.0 and st["flops_scan"] == 64.0
    check_against_fp64(x, q, ids, ds, k)
    idx.set_sketch(rbc.SKETCH_OFF)
    ids2, ds2, st2 = idx.kquery(q, k, precision=rbc.PREC_BF16, stats=True)
    assert st2["sketch"] == 0.0
    np.testing.assert_array_equal(ids, ids2)
    np.testing.assert_array_equal(ds, ds2)
    # The probe runs on the sketches too, so the bound (and the plan) may differ; where the
    # sketch is chosen automatically it is nearly exact and the scan stays the same size.
    if mode == rbc.SKETCH_AUTO:
        assert st["pairs_scan"] <= 1.02 * st2["pairs_scan"]


def test_sketch_auto_mode_follows_the_data():
    # Low intrinsic dimension: on; a full-rank mixture: off; d <= 32: no sketch at all.
    lo = rbc.ExactRBC(datagen.lowrank(20000, 128, 0), 100, seed=1)
    assert lo.sketch_info()["active"] == 1
    hi = rbc.ExactRBC(datagen.gmm_dirichlet(20000, 96, 0), 100, seed=1)
    inf = hi.sketch_info()
    assert inf["dims"] == 28 and inf["active"] == 0
    sm = rbc.ExactRBC(datagen.uniform(5000, 32, 0), 50, seed=1)
    assert sm.sketch_info()["dims"] == 0
    with pytest.raises(rbc.RBCError):
        sm.set_sketch(rbc.SKETCH_ON)
    sm.set_sketch(rbc.SKETCH_OFF)
    l1 = rbc.ExactRBC(datagen.lowrank(5000, 64, 0), 50, metric=1, seed=1)
    assert l1.sketch_info()["dims"] == 0


def test_sketch_duplicates_ragged_lists_and_many_candidates():
    base = datagen.lowrank(4000, 128, 0)
    # Exact duplicates and a cluster of near-identical points: many candidates per query.
    x = np.concatenate([base, base, base[:1500], base[:1] + 1e-4 * datagen.lowrank(3000, 128, 5)])
    q = np.concatenate([base[:40], x[-30:], datagen.lowrank(60, 128, 1)])
    idx = rbc.ExactRBC(x, 97, seed=9)
    assert idx.sketch_info()["active"] == 1
    ids, ds, st = idx.kquery(q, 32, precision=rbc.PREC_BF16, stats=True)
    assert st["sketch"] == 1.0
    check_against_fp64(x, q, ids, ds, 32)


def test_sketch_survives_save_and_load(tmp_path):
    x = datagen.lowrank(20000, 128, 0)
    q = datagen.lowrank(300, 128, 1)
    idx = rbc.ExactRBC(x, 120, seed=6)
    ids, ds = idx.kquery(q, 10, precision=rbc.PREC_BF16)
    path = tmp_path / "idx.rbc"
    idx.save(str(path))
    ld = rbc.ExactRBC.load(str(path))
    assert ld.sketch_info() == idx.sketch_info()
    np.testing.assert_array_equal(ld.sketch_rows(x[:500])[0], idx.sketch_rows(x[:500])[0])
    ids2, ds2 = ld.kquery(q, 10, precision=rbc.PREC_BF16)
    np.testing.assert_array_equal(ids, ids2)
    np.testing.assert_array_equal(ds, ds2)


def test_sketch_probe_under_a_confident_proposal():
    # A proposal that often gives a query a single list (low mass threshold): that list's
    # probe entry alone must still hold k candidates, or the query falls back to the loose
    # representative bound and floods the re-rank.  Results stay identical either way.
    n, d, R, k = 1 << 17, 128, 512, 10
    x = datagen.lowrank(n, d, 0)
    idx = rbc.ExactRBC(x, R, seed=0)
    assert idx.sketch_info()["active"] == 1
    owner = idx.lists()["owner"]
    qt = datagen.lowrank(1 << 15, d, 2)
    nn, _ = idx.kquery(qt, k, precision=rbc.PREC_BF16)
    net = rbc.ProposalMLP(d, R, seed=0)
    net.train(qt, owner[nn], epochs=4, batch=1024, lr=1e-3)
    q = datagen.lowrank(4096, d, 1)
    i0, d0, s0 = idx.kquery(q, k, precision=rbc.PREC_BF16, stats=True)
    idx.set_proposal(net, mass=0.3)
    i1, d1, s1 = idx.kquery(q, k, precision=rbc.PREC_BF16, stats=True)
    idx.set_proposal(None)
    _, counts = net.propose_lists(q, 8, 0.3)
    assert (counts == 1).mean() > 0.1, "the proposal should give many queries one list"
    np.testing.assert_array_equal(i0, i1)
    np.testing.assert_array_equal(d0, d1)
    # A single list's bound is looser than eight lists' (about twice the candidates here);
    # the representative bound would give orders of magnitude more.
    assert s1["pairs_rerank"] < 4.0 * s0["pairs_rerank"]
    assert s1["pairs_scan"] < 2.0 * s0["pairs_scan"]
