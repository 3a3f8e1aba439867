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


_ARRAYS = ("xp", "xb", "xb_norm", "list_qerr", "list_nmax", "perm", "inv", "owner", "list_off",
           "list_len", "reps", "reps_b", "reps_bnorm", "reps_qerr", "rep_ids", "radius", "rr")
_ELEM = {"xb": 2, "reps_b": 2}


def _checksum(pieces):
    """The index file's checksum as its format defines it: FNV-1a over 64-bit words, piece by
    piece (a last partial word zero-padded), each piece followed by its length in bytes."""
    h, P, M = 0xCBF29CE484222325, 0x100000001B3, (1 << 64) - 1
    for b in pieces:
        n = len(b)
        for w in np.frombuffer(b + b"\0" * (-n % 8), "<u8").tolist():
            h = ((h ^ w) * P) & M
        h = ((h ^ n) * P) & M
    return h


def _parse(raw):
    """Splits an index file into header and arrays (small files: one chunk per array)."""
    assert raw[:8] == b"RBCXIDX2"
    hdr = raw[8:40]
    pos, arrays = 40, {}
    for name in _ARRAYS:
        count = int(np.frombuffer(raw[pos:pos + 8], "<u8")[0])
        nbytes = count * _ELEM.get(name, 4)
        arrays[name] = bytearray(raw[pos + 8:pos + 8 + nbytes])
        pos += 8 + nbytes
    return hdr, arrays, raw[pos:]


def _assemble(hdr, arrays):
    pieces, out = [hdr], [b"RBCXIDX2", hdr]
    for name in _ARRAYS:
        data = bytes(arrays[name])
        cnt = np.array([len(data) // _ELEM.get(name, 4)], "<u8").tobytes()
        pieces += [cnt, data]
        out += [cnt, data]
    return b"".join(out) + np.array([_checksum(pieces)], "<u8").tobytes()


def test_damaged_index_files_are_rejected(tmp_path):
    x = datagen.uniform(5000, 16, 0)
    idx = rbc.ExactRBC(x, 70)
    path = tmp_path / "index.rbc"
    idx.save(path)
    raw = path.read_bytes()
    hdr, arrays, tail = _parse(raw)
    assert len(tail) == 8
    # The test's restatement of the checksum reproduces the file exactly.
    assert _assemble(hdr, arrays) == raw
    bad = {"short.rbc": raw[: len(raw) // 2], "magic.rbc": b"XXXXXXXX" + raw[8:],
           "trailing.rbc": raw + b"\0"}
    # One flipped byte in each structural array: caught by the checksum.
    for name in ("perm", "inv", "list_off", "list_len", "owner"):
        a = {k: bytearray(v) for k, v in arrays.items()}
        a[name][5] ^= 0x40
        data = bytearray(_assemble(hdr, a))
        data[-8:] = raw[-8:]  # keep the original checksum
        bad[f"flip_{name}.rbc"] = bytes(data)

    # Structurally invalid files with a correct checksum: caught by the host checks.
    def mutate(name, fn):
        a = {k: bytearray(v) for k, v in arrays.items()}
        v = np.frombuffer(bytes(a[name]), "<u4").copy()
        fn(v)
        a[name] = bytearray(v.tobytes())
        bad[f"struct_{name}.rbc"] = _assemble(hdr, a)

    mutate("perm", lambda v: v.__setitem__(slice(0, 2), v[[1, 0]]))   # a permutation, inv stale
    mutate("inv", lambda v: v.__setitem__(3, 5000))                     # out of range
    mutate("list_off", lambda v: v.__setitem__(7, v[9] + 1))            # decreasing offsets
    mutate("list_len", lambda v: v.__setitem__(4, v[4] + 1))            # length != offsets
    mutate("owner", lambda v: v.__setitem__(11, 70))                    # owner out of range
    mutate("rep_ids", lambda v: v.__setitem__(2, 5000))                 # rep id out of range
    for fname, data in bad.items():
        (tmp_path / fname).write_bytes(data)
    for fname in list(bad) + ["missing.rbc"]:
        with pytest.raises(rbc.RBCError):
            rbc.ExactRBC.load(tmp_path / fname)
    # The untouched file still loads.
    rbc.ExactRBC.load(path).close()


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
