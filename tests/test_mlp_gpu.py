"""The proposal network on the GPU: gradients of the library's FP16 forward/backward
kernels against the FP64 oracle, training, and what the trained network gains the exact
query (with identical results)."""
import numpy as np
import pytest

import mlp_oracle
import oracle
from conftest import datagen, rbc

pytestmark = pytest.mark.gpu


def test_gradients_match_fp64_oracle():
    dims = (96, 256, 256, 512)
    net = rbc.ProposalMLP(96, 512, seed=3)
    x = datagen.gmm_dirichlet(300, 96, 5)
    labels = np.random.default_rng(0).integers(0, 512, size=300).astype(np.uint32)
    loss, g = net.loss_grad(x, labels)
    oloss, og = mlp_oracle.loss_grad(net.get_params().astype(np.float64), x, labels, dims)
    assert abs(loss - oloss) <= 2e-3 * abs(oloss)
    sizes = [256 * 96, 256, 256 * 256, 256, 512 * 256, 512]
    o = 0
    for name, sz in zip(["W1", "b1", "W2", "b2", "W3", "b3"], sizes):
        a, b = g[o:o + sz].astype(np.float64), og[o:o + sz]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel < 2e-2, f"{name}: relative gradient error {rel}"
        o += sz


def test_logits_match_fp64_oracle():
    net = rbc.ProposalMLP(128, 300, seed=1)
    x = datagen.lowrank(100, 128, 2)
    z = net.logits(x)
    oz = mlp_oracle.forward(net.get_params().astype(np.float64), x, (128, 256, 256, 300))["z3"]
    assert np.abs(z - oz).max() <= 1e-2 * np.abs(oz).max()


@pytest.mark.parametrize("rows,n_out", [(70000, 2048), (300, 300), (1, 65)])
def test_proposals_are_the_logit_argmax(rows, n_out):
    # The fused GEMM + argmax of the query path against the argmax of the logits (lowest
    # index among equal logits), over more rows than one inference chunk holds.
    net = rbc.ProposalMLP(128, n_out, seed=4)
    x = datagen.lowrank(rows, 128, 6)
    p = net.propose(x)
    z = net.logits(x)
    np.testing.assert_array_equal(p, np.argmax(z, axis=1).astype(np.uint32))
    # Equal logits: a network with all-zero weights proposes list 0 for every row.
    net.set_params(np.zeros_like(net.get_params()))
    assert (net.propose(x[:100]) == 0).all()


def test_soft_target_gradients_match_fp64_oracle():
    # Ten labels per row (the lists of a query's 10 NNs), with repeats, labels past the
    # output width (no label) and one row without any.
    dims = (96, 256, 256, 512)
    net = rbc.ProposalMLP(96, 512, seed=5)
    x = datagen.gmm_dirichlet(300, 96, 7)
    rng = np.random.default_rng(1)
    labels = rng.integers(0, 512, size=(300, 10)).astype(np.uint32)
    labels[:, 5] = labels[:, 4]
    labels[3, 2] = 5000
    labels[7, :] = 0xFFFFFFFF
    loss, g = net.loss_grad(x, labels)
    oloss, og = mlp_oracle.loss_grad(net.get_params().astype(np.float64), x, labels, dims)
    assert abs(loss - oloss) <= 2e-3 * abs(oloss)
    sizes = [256 * 96, 256, 256 * 256, 256, 512 * 256, 512]
    o = 0
    for name, sz in zip(["W1", "b1", "W2", "b2", "W3", "b3"], sizes):
        a, b = g[o:o + sz].astype(np.float64), og[o:o + sz]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel < 2e-2, f"{name}: relative gradient error {rel}"
        o += sz


@pytest.mark.parametrize("rows,n_out,P,mass", [(70000, 2048, 8, 0.9), (300, 300, 5, 0.5), (33, 65, 16, 0.99)])
def test_proposed_lists_are_the_top_logits(rows, n_out, P, mass):
    # The GPU's top-P lists and mass counts against the oracle on the GPU's own logits
    # (more rows than one inference chunk holds).  A count may differ only where the
    # cumulative mass is within FP32 rounding of the threshold.
    net = rbc.ProposalMLP(128, n_out, seed=4)
    # Sharpen the untrained network's softmax (last layer x 40) so that the counts vary.
    w = net.get_params()
    w[-(n_out * 256 + n_out):] *= 40.0
    net.set_params(w)
    x = datagen.lowrank(rows, 128, 6)
    lists, counts = net.propose_lists(x, P, mass)
    z = net.logits(x)
    ol, oc, cum = mlp_oracle.propose_lists(z, P, mass)
    # The fused kernel accumulates the last layer with its MFMA operands swapped, so its
    # logits may differ from rbc_mlp_logits' in the last bits: the chosen lists must carry
    # the oracle's top-P logit values (within that rounding), in order, and be the oracle's
    # lists wherever no near-tie makes the choice ambiguous.
    tol = 1e-5 * np.abs(z).max()
    assert (lists < n_out).all()
    got = np.take_along_axis(z, lists.astype(np.int64), 1)
    want = np.take_along_axis(z, ol.astype(np.int64), 1)
    np.testing.assert_allclose(got, want, rtol=0, atol=tol)
    srt = -np.sort(-z, 1)
    gap = np.abs(np.diff(srt[:, :P + 1], axis=1)).min(1) > 2 * tol
    np.testing.assert_array_equal(lists[gap], ol[gap])
    assert gap.mean() > 0.5
    close = (np.abs(cum - mass) < 1e-4).any(1)
    assert (counts[~close & gap] == oc[~close & gap]).all()
    assert (counts <= P).all() and (counts > 0).any() and (counts == 0).any() or P == 16


def test_training_on_knn_lists_and_proposal_gain():
    # Learned proposal on a 1,024-component mixture: trained on the lists holding the exact
    # 10 NNs of training queries (from exact RBC itself), it lets most held-out queries
    # probe fewer lists than the geometric set of 8 (the probe work falls by well over a
    # third) while the scan grows by far less, and the results are identical.
    n, d, R, k = 1 << 18, 96, 512, 10
    x = datagen.gmm_dirichlet(n, d, 0)
    idx = rbc.ExactRBC(x, R, seed=0)
    owner = idx.lists()["owner"]
    qt = datagen.gmm_dirichlet(1 << 16, d, 2)
    nn, _ = idx.kquery(qt, k, precision=rbc.PREC_BF16)
    labels = owner[nn]
    net = rbc.ProposalMLP(d, R, seed=0)
    losses = net.train(qt, labels, epochs=6, batch=1024, lr=1e-3)
    assert losses[-1] < losses[0]
    q = datagen.gmm_dirichlet(8192, d, 1)
    i0, d0, s0 = idx.kquery(q, k, precision=rbc.PREC_BF16, stats=True)
    idx.set_proposal(net)
    i1, d1, s1 = idx.kquery(q, k, precision=rbc.PREC_BF16, stats=True)
    idx.set_proposal(None)
    np.testing.assert_array_equal(i0, i1)
    np.testing.assert_array_equal(d0, d1)
    _, counts = net.propose_lists(q, 8, 0.9)
    print(f"mean lists probed with the proposal {np.where(counts > 0, counts, 8).mean():.2f} of 8; "
          f"probe pairs {s0['pairs_probe']:.3e} -> {s1['pairs_probe']:.3e}, scan pairs "
          f"{s0['pairs_scan']:.3e} -> {s1['pairs_scan']:.3e}")
    assert s1["pairs_probe"] < 0.65 * s0["pairs_probe"]
    assert s1["pairs_scan"] - s0["pairs_scan"] < 0.25 * (s0["pairs_probe"] - s1["pairs_probe"])
    check = oracle.brute_k_f64(x, q[:64], k)[0]
    np.testing.assert_array_equal(i1[:64], check)
