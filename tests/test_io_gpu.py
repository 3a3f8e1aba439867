"""Saving and loading: a loaded exact index and a loaded proposal network answer exactly
like the ones that were saved; damaged files are rejected with an error."""
import numpy as np
import pytest

from conftest import datagen, rbc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("metric,d", [(0, 128), (1, 33)])
def test_exact_index_roundtrip(tmp_path, metric, d):
    x, q = datagen.gmm(30000, d, 0), datagen.gmm(500, d, 1)
    idx = rbc.ExactRBC(x, 173, metric=metric, seed=5)
    path = tmp_path / "index.rbc"
    idx.save(path)
    back = rbc.ExactRBC.load(path)
    a_inf, b_inf = idx.info(), back.info()
    a_inf.pop("device_bytes"), b_inf.pop("device_bytes")   # includes the build's workspace
    assert a_inf == b_inf
    a = idx.lists()
    b = back.lists()
    for key in a:
        np.testing.assert_array_equal(a[key], b[key])
    precs = [rbc.PREC_FP32] + ([rbc.PREC_BF16] if metric == 0 else [])
    for prec in precs:
        i0, d0 = idx.kquery(q, 16, precision=prec)
        i1, d1 = back.kquery(q, 16, precision=prec)
        np.testing.assert_array_equal(i0, i1)
        np.testing.assert_array_equal(d0, d1)


def test_damaged_index_files_are_rejected(tmp_path):
    x = datagen.uniform(5000, 16, 0)
    idx = rbc.ExactRBC(x, 70)
    path = tmp_path / "index.rbc"
    idx.save(path)
    raw = path.read_bytes()
    (tmp_path / "short.rbc").write_bytes(raw[: len(raw) // 2])
    (tmp_path / "magic.rbc").write_bytes(b"XXXXXXXX" + raw[8:])
    for bad in ("short.rbc", "magic.rbc", "missing.rbc"):
        with pytest.raises(rbc.RBCError):
            rbc.ExactRBC.load(tmp_path / bad)


def test_proposal_network_roundtrip(tmp_path):
    d, R = 32, 64
    x = datagen.gmm(20000, d, 0)
    idx = rbc.ExactRBC(x, R, seed=1)
    qt = datagen.gmm(4096, d, 2)
    nn, _ = idx.kquery(qt, 1)
    labels = idx.lists()["owner"][nn[:, 0]]
    net = rbc.ProposalMLP(d, R, h1=64, h2=64, seed=3)
    net.train(qt, labels, epochs=2, batch=256, lr=2e-3)
    path = tmp_path / "net.mlp"
    net.save(path)
    back = rbc.ProposalMLP.load(path)
    assert back.dims == (d, 64, 64, R)
    np.testing.assert_array_equal(back.get_params(), net.get_params())
    np.testing.assert_array_equal(back.logits(qt[:100]), net.logits(qt[:100]))
    # Training continues from the saved Adam state identically (the reported loss is a
    # float atomic sum over rows, so it may differ in the last bits).
    l0 = net.train(qt, labels, epochs=1, batch=256, lr=2e-3, seed=9)
    l1 = back.train(qt, labels, epochs=1, batch=256, lr=2e-3, seed=9)
    np.testing.assert_allclose(l0, l1, rtol=1e-5)
    np.testing.assert_array_equal(back.get_params(), net.get_params())
