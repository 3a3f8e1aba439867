"""Pins the CPU oracle (oracle/rbc_oracle.c) before it is trusted as the checker.

Each oracle function is checked against something it does not share code with:
closed forms on exactly representable inputs, an independent numpy FP64 k-NN, the
error bounds the engine relies on, invariants of the reference's algorithms, and
hand-traced known answers for the range bisection (kernels.cu:322-414)."""
import os

import numpy as np
import pytest

import oracle
from conftest import datagen

U = 2.0 ** -24


def test_fma_distance_closed_form_on_integers():
    # Small integers: every difference, square and partial sum is exact in FP32, so the
    # FMA-order sum must equal the exact integer distance (a dropped or repeated term,
    # a wrong index or a sign slip would show).
    rng = np.random.default_rng(0)
    for d in (1, 3, 16, 17, 64, 128, 256):
        a = rng.integers(-20, 20, size=d).astype(np.float32)
        b = rng.integers(-20, 20, size=d).astype(np.float32)
        exact_l2 = float(((a.astype(np.int64) - b.astype(np.int64)) ** 2).sum())
        exact_l1 = float(np.abs(a.astype(np.int64) - b.astype(np.int64)).sum())
        assert oracle.dist_f32_fma(a, b, 0) == exact_l2
        assert oracle.dist_f32_fma(a, b, 1) == exact_l1
        assert oracle.dist_f32_host(a, b, 0) == exact_l2
        assert oracle.dist_f32_tree(a, b, 0) == exact_l2
        assert oracle.dist_f32_tree(a, b, 1) == exact_l1
        assert oracle.dist_f64(a, b, 0) == exact_l2


def test_fp32_orders_stay_inside_the_engine_error_bounds():
    # The engine's pruning and candidate cut-offs assume |fma - exact| <= (d+4)u exact
    # and |tree - exact| <= (log2 P + 4)u exact; the reported distances need <= 1e-6.
    rng = np.random.default_rng(1)
    for d in (16, 64, 96, 128):
        P = 64
        while P < d:
            P *= 2
        e_fma = (d + 4) * U
        e_tree = (np.log2(P) + 4) * U
        for _ in range(300):
            a = rng.standard_normal(d).astype(np.float32) * 3
            scale = np.float32(rng.choice([1e-3, 0.1, 1.0]))
            b = (a + rng.standard_normal(d).astype(np.float32) * scale).astype(np.float32)
            ex = sum((float(x) - float(y)) ** 2 for x, y in zip(a, b))
            f = oracle.dist_f32_fma(a, b)
            t = oracle.dist_f32_tree(a, b)
            assert abs(f - ex) <= e_fma * ex
            assert abs(t - ex) <= e_tree * ex
            assert abs(t - ex) <= 1e-6 * ex


def test_tree_order_matches_independent_pairwise_sum():
    rng = np.random.default_rng(2)
    for d in (16, 50, 64, 100, 128, 200, 256):
        a = rng.standard_normal(d).astype(np.float32)
        b = rng.standard_normal(d).astype(np.float32)
        P = 64
        while P < d:
            P *= 2
        t = np.zeros(P, np.float32)
        diff = (a - b).astype(np.float32)
        t[:d] = diff * diff
        s = P // 2
        while s >= 1:
            t[:s] = t[:s] + t[s:2 * s]
            s //= 2
        assert oracle.dist_f32_tree(a, b) == t[0]


def test_fp64_knn_matches_independent_numpy():
    x = datagen.gmm(3000, 24, 0)
    q = datagen.gmm(40, 24, 1)
    x = np.concatenate([x, x[:100]])  # exact duplicates -> lowest index first
    for metric in (0, 1):
        oi, od = oracle.brute_k_f64(x, q, 20, metric)
        ni, nd = oracle.knn_f64_numpy(x, q, 20, metric)
        np.testing.assert_array_equal(oi, ni)
        np.testing.assert_allclose(od, nd, rtol=1e-13)


def test_fp32_bruteK_agrees_with_fp64_on_tie_free_data():
    x = datagen.uniform(4000, 16, 0)
    q = datagen.uniform(60, 16, 1)
    fi, fd = oracle.brute_k_f32(x, q, 32)
    di, dd = oracle.brute_k_f64(x, q, 32)
    np.testing.assert_array_equal(fi, di)
    for i in range(0, 60, 7):
        for j in range(32):
            assert fd[i, j] == oracle.dist_f32_fma(x[fi[i, j]], q[i])
    # Unfilled slots when n < k.
    fi, fd = oracle.brute_k_f32(x[:10], q, 32)
    assert (fi[:, 10:] == 0xFFFFFFFF).all() and (fd[:, 10:] == np.finfo(np.float32).max).all()


def test_nn_and_bruteCPU_lowest_index_on_ties():
    x = datagen.uniform(500, 8, 0)
    x = np.concatenate([x, x])
    q = x[[3, 17, 499]] + 0.0
    ids, ds = oracle.nn_f32(x, q)
    np.testing.assert_array_equal(ids, [3, 17, 499])
    np.testing.assert_array_equal(oracle.brute_cpu(x, q), [3, 17, 499])
    assert (ds == 0).all()


def test_range_count_closed_form():
    x = datagen.uniform(3000, 12, 0)
    q = datagen.uniform(50, 12, 1)
    d64 = ((x[None, :, :].astype(np.float64) - q[:, None, :]) ** 2).sum(-1)
    # Ranges halfway between consecutive sorted distances: no rounding can flip a count.
    srt = np.sort(d64, 1)
    ranges = ((srt[:, 99] + srt[:, 100]) / 2).astype(np.float32)
    np.testing.assert_array_equal(oracle.range_count(x, q, ranges), 100)


def simulate_find_range(D, cnt_want, max_iters=1000):
    """Independent restatement of findRangeKernel's loop (kernels.cu:329-413) over
    blocks of four rows, in float32 numpy arithmetic."""
    R, n = D.shape
    LB, UB = (90 * cnt_want) // 100, cnt_want
    out = np.zeros(R, np.float32)
    for r0 in range(0, R, 4):
        rows = [r0 + i for i in range(4)]
        valid = [r < R for r in rows]
        smin = [np.float32(D[r].min()) if v else np.float32(np.finfo(np.float32).max) for r, v in zip(rows, valid)]
        smax = [np.float32(max(0.0, D[r].max())) if v else np.float32(0) for r, v in zip(rows, valid)]
        rg = [np.float32(0)] * 4
        cont = [1, 1, 1, 1]
        it = 0
        while True:
            it += 1
            for i in range(4):
                if cont[i]:
                    rg[i] = np.float32(np.float32(smax[i] + smin[i]) / np.float32(2))
                c = int((D[rows[i]] <= rg[i]).sum()) if valid[i] else 0
                if c < cnt_want:
                    smin[i] = rg[i]
                else:
                    smax[i] = rg[i]
                cont[i] = int(valid[i] and (c < LB or c > UB))
            cont[0] |= cont[2]
            cont[1] |= cont[3]
            cont[0] |= cont[1]
            if not cont[0] or it >= max_iters:
                break
        for i in range(4):
            if valid[i]:
                out[rows[i]] = rg[i]
    return out


def test_find_range_hand_traced_coupling():
    # Three reps over four points, s = 2 (LB = 1, UB = 2).  Rows 0 and 1 settle at the
    # first midpoint (2.5 and 25); row 2 needs nine halvings of [1, 100] to reach
    # 1.193359375.  The reference's OR-reduction (kernels.cu:403-407) keeps cont[0] set
    # while row 2 runs, so row 0 keeps re-bisecting (2.5, 1.75, 2.125, 1.9375, 2.03125,
    # 1.984375, 2.0078125, 1.99609375, 2.001953125); row 1's flag is only OR-ed with
    # row 3 (padding), so it stays at 25.
    D = np.array([[1, 2, 3, 4], [10, 20, 30, 40], [1, 1.1, 1.2, 100]], np.float32)
    expected = np.array([2.001953125, 25.0, 1.193359375], np.float32)
    np.testing.assert_array_equal(simulate_find_range(D, 2), expected)


def one_d_points(sq_dists):
    """Points on a line at the origin's squared distances (exact for perfect squares)."""
    return np.sqrt(np.asarray(sq_dists, np.float64)).astype(np.float32)[:, None]


def test_oracle_build_reproduces_hand_traced_ranges():
    # Rep at 0 with points at squared distances 1, 4, 9, 16, 25 (exact in FP32), s = 2:
    # first midpoint (25 + 0)/2 = 12.5 counts 4 points (0, 1, 4, 9) -> too many;
    # 6.25 -> 3 (0, 1, 4) too many; 3.125 -> 2 (0, 1) -> settles at 3.125.
    x = one_d_points([0, 1, 4, 9, 16, 25])
    mp, counts, ranges, bad = oracle.oneshot_build(x, np.array([0], np.uint32), 2)
    assert bad == 0 and ranges[0] == np.float32(3.125)
    assert counts[0] == 2 and list(mp[0, :2]) == [0, 1]


def test_oneshot_build_matches_restatement_and_invariants():
    x = datagen.gmm(6000, 16, 0)
    reps = np.random.default_rng(3).choice(6000, 70, replace=False).astype(np.uint32)
    s = 300
    mp, counts, ranges, bad = oracle.oneshot_build(x, reps, s)
    assert bad == 0
    D = np.empty((70, 6000), np.float32)
    for i, r in enumerate(reps):
        D[i] = [oracle.dist_f32_fma(x[r], x[j]) for j in range(6000)]
    np.testing.assert_array_equal(simulate_find_range(D, s), ranges)
    lo = (90 * s) // 100
    for i in range(70):
        lst = np.nonzero(D[i] < ranges[i])[0]
        np.testing.assert_array_equal(mp[i, :counts[i]], lst)
        assert counts[i] == lst.size <= s
        assert lo <= int((D[i] <= ranges[i]).sum()) <= s


def test_oneshot_kquery_is_brute_force_over_the_nearest_list():
    x = datagen.uniform(3000, 8, 0)
    q = datagen.uniform(40, 8, 1)
    reps = np.random.default_rng(4).choice(3000, 40, replace=False).astype(np.uint32)
    mp, counts, ranges, bad = oracle.oneshot_build(x, reps, 200)
    ids, ds = oracle.oneshot_kquery(x, reps, mp, counts, q, 32)
    near, _ = oracle.nn_f32(x[reps], q)
    for i in range(40):
        lst = mp[near[i], :counts[near[i]]]
        si, sd = oracle.brute_k_f32(x[lst], q[i:i + 1], 32)
        valid = si[0] != 0xFFFFFFFF
        exp_ids = np.where(valid, lst[np.minimum(si[0], len(lst) - 1)], 0xFFFFFFFF)
        np.testing.assert_array_equal(ids[i], exp_ids)
        np.testing.assert_array_equal(ds[i], sd[0])


def test_oneshot_plan_identity_is_the_one_shot_query():
    x = datagen.uniform(3000, 8, 0)
    q = datagen.uniform(40, 8, 1)
    reps = np.random.default_rng(4).choice(3000, 40, replace=False).astype(np.uint32)
    mp, counts, _, _ = oracle.oneshot_build(x, reps, 200)
    ids, ds = oracle.oneshot_kquery(x, reps, mp, counts, q, 32)
    pids, pds = oracle.oneshot_kquery_plan(x, reps, mp, counts, np.eye(40, dtype=np.uint8), q, 32)
    np.testing.assert_array_equal(pids, ids)
    np.testing.assert_array_equal(pds, ds)


def test_oneshot_plan_is_the_k_best_of_the_planned_lists_as_a_multiset():
    # Sort-based restatement: every point of every planned list (repeats kept), ordered by
    # (distance, id); the one-shot lists overlap, so repeats do occur.
    x = datagen.uniform(2000, 6, 3)
    q = datagen.uniform(30, 6, 4)
    R = 24
    reps = np.random.default_rng(5).choice(2000, R, replace=False).astype(np.uint32)
    mp, counts, _, _ = oracle.oneshot_build(x, reps, 150)
    plan = (np.random.default_rng(6).random((R, R)) < 0.2).astype(np.uint8)
    plan[3] = 0  # a group with an empty plan row
    ids, ds = oracle.oneshot_kquery_plan(x, reps, mp, counts, plan, q, 32)
    near, _ = oracle.nn_f32(x[reps], q)
    repeats = 0
    for i in range(30):
        pts = np.concatenate([mp[j, :counts[j]] for j in np.flatnonzero(plan[near[i]])] +
                             [np.zeros(0, np.uint32)]).astype(np.uint32)
        repeats += len(pts) - len(np.unique(pts))
        dv = np.array([oracle.dist_f32_fma(x[j], q[i]) for j in pts], np.float32)
        order = np.lexsort((pts, dv))[:32]
        exp_i = np.full(32, 0xFFFFFFFF, np.uint32)
        exp_d = np.full(32, np.finfo(np.float32).max, np.float32)
        exp_i[:len(order)] = pts[order]
        exp_d[:len(order)] = dv[order]
        np.testing.assert_array_equal(ids[i], exp_i)
        np.testing.assert_array_equal(ds[i], exp_d)
    assert repeats > 0


def test_exact_owners_and_radii():
    x = datagen.gmm(5000, 20, 0)
    reps = np.random.default_rng(6).choice(5000, 50, replace=False).astype(np.uint32)
    own = oracle.exact_owners(x, reps)
    ids, _ = oracle.nn_f32(x[reps], x)
    np.testing.assert_array_equal(own, ids)
    rad = oracle.exact_radii(x, reps, own)
    d64 = np.sqrt(((x.astype(np.float64) - x[reps[own]].astype(np.float64)) ** 2).sum(1))
    exact = np.zeros(50)
    np.maximum.at(exact, own, d64)
    assert (rad.astype(np.float64) >= exact * (1 - 1e-15)).all()
    below = np.nextafter(rad, np.float32(0)).astype(np.float64)
    assert (below < exact * (1 + 1e-15)).all()


def test_exact_owners_against_fp64_argmin():
    # Independent of the oracle's FP32 arithmetic: NumPy's FP64 distances.  Where the two
    # nearest reps are farther apart than the FP32 error bound, the owner is the FP64
    # argmin; elsewhere its FP64 distance is within that bound of the minimum.
    x = datagen.gmm(4000, 20, 0)
    reps = np.random.default_rng(6).choice(4000, 50, replace=False).astype(np.uint32)
    own = oracle.exact_owners(x, reps)
    X, Rr = x.astype(np.float64), x[reps].astype(np.float64)
    D = ((X[:, None, :] - Rr[None, :, :]) ** 2).sum(-1)
    best, arg = D.min(1), D.argmin(1)
    second = np.partition(D, 1, axis=1)[:, 1]
    e = 4 * (20 + 4) * 2.0 ** -24
    clear = second > best * (1 + e)
    assert clear.mean() > 0.99
    np.testing.assert_array_equal(own[clear], arg[clear])
    assert (D[np.arange(len(x)), own] <= best * (1 + e)).all()
    # A small-integer lattice: FP32 distances are exact, ties are common (repeated rows),
    # and the owner is the FP64 argmin with the lowest rep index among equals.
    xi = np.random.default_rng(3).integers(-3, 4, (3000, 4)).astype(np.float32)
    ri = np.random.default_rng(4).choice(3000, 40, replace=False).astype(np.uint32)
    oi = oracle.exact_owners(xi, ri)
    Di = ((xi[:, None, :].astype(np.float64) - xi[ri][None, :, :].astype(np.float64)) ** 2).sum(-1)
    ties = (Di == Di.min(1)[:, None]).sum(1) > 1
    assert ties.mean() > 0.1
    np.testing.assert_array_equal(oi, Di.argmin(1))


def test_golden_fixture_reproduces():
    path = os.path.join(os.path.dirname(__file__), "golden", "sample_shaped_fp64_knn.npz")
    g = np.load(path)
    ids, ds = oracle.brute_k_f64(g["x"], g["q"], 32)
    np.testing.assert_array_equal(ids, g["ids"])
    np.testing.assert_array_equal(ds, g["dists"])


@pytest.mark.parametrize("data", ["gmm", "lattice"])
def test_cpu_exact_rbc_equals_brute_force(data):
    # The CPU exact RBC (the CPU baseline's search) returns orc_brute_k_f32's results exactly,
    # ties included (the lattice has many equal distances), while skipping lists; lists
    # have ragged lengths (partial 8-point blocks).
    rng = np.random.default_rng(11)
    if data == "gmm":
        x = datagen.gmm(4000, 12, 3)
        q = datagen.gmm(300, 12, 4)
    else:
        x = rng.integers(0, 4, size=(4000, 6)).astype(np.float32)
        q = rng.integers(0, 4, size=(300, 6)).astype(np.float32)
    R, k = 50, 7
    reps = np.sort(rng.choice(x.shape[0], R, replace=False)).astype(np.uint32)
    owner = oracle.exact_owners(x, reps)
    radii = oracle.exact_radii(x, reps, owner)
    perm = np.argsort(owner, kind="stable").astype(np.uint32)
    off = np.concatenate([[0], np.cumsum(np.bincount(owner, minlength=R))]).astype(np.uint32)
    bi, bd = oracle.brute_k_f32(x, q, k)
    seen = []
    for planes in (False, True):
        ids, ds, pairs = oracle.rbc_exact_knn(x, q, reps, perm, off, radii, k, planes=planes)
        np.testing.assert_array_equal(ids, bi)
        np.testing.assert_array_equal(ds, bd)
        seen.append(pairs)
    # Both bounds skip lists, and the plane bound skips more.
    assert seen[1] <= seen[0] < 0.9 * x.shape[0] * q.shape[0]
