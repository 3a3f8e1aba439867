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


def test_training_reduces_loss_and_proposal_gain():
    n, d, R, k = 1 << 17, 96, 256, 10
    x = datagen.gmm_dirichlet(n, d, 0)
    idx = rbc.ExactRBC(x, R, seed=0)
    owner = idx.lists()["owner"]
    # Training pairs from exact RBC itself: query -> owner list of its exact 1-NN.
    qt = datagen.gmm_dirichlet(1 << 15, d, 2)
    nn, _ = idx.kquery(qt, 1, precision=rbc.PREC_BF16)
    labels = owner[nn[:, 0]]
    net = rbc.ProposalMLP(d, R, seed=0)
    losses = net.train(qt, labels, epochs=8, batch=1024, lr=2e-3)
    assert losses[-1] < losses[0]
    # Held-out queries: results identical with and without the proposal; pairs scanned
    # can only go down (the extra probe only tightens the bound).
    q = datagen.gmm_dirichlet(4096, d, 1)
    i0, d0, s0 = idx.kquery(q, k, precision=rbc.PREC_BF16, stats=True)
    idx.set_proposal(net)
    i1, d1, s1 = idx.kquery(q, k, precision=rbc.PREC_BF16, stats=True)
    idx.set_proposal(None)
    np.testing.assert_array_equal(i0, i1)
    np.testing.assert_array_equal(d0, d1)
    assert s1["pairs_scan"] <= s0["pairs_scan"]
    nnq, _ = oracle.nn_f32(x, q[:512])
    acc_mlp = float(np.mean(np.argmax(net.logits(q[:512]), 1) == owner[nnq]))
    print(f"proposal top-1 accuracy {acc_mlp:.3f}; pairs scanned {s0['pairs_scan']:.3e} -> "
          f"{s1['pairs_scan']:.3e}, probe {s0['pairs_probe']:.3e} -> {s1['pairs_probe']:.3e}, "
          f"re-ranked {s0['pairs_rerank']:.3e} -> {s1['pairs_rerank']:.3e}")
