"""testRBC end to end on the GPU: the reference's file formats (driver.cu:201-245,
352-393), its evaluation output (driver.cu:283-349), one-shot results equal to the
library's buildRBC/kqueryRBC, and exact results equal to the FP64 oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from conftest import PKG, datagen, rbc

pytestmark = pytest.mark.gpu

EXE = os.path.join(PKG, "testRBC")
N, M, D, R, KMAX = 1024, 128, 16, 128, 32   # the shape of the reference's sample_input


@pytest.fixture(scope="module")
def data(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    x = datagen.uniform(N, D, 0)
    q = datagen.uniform(M, D, 1)
    x.tofile(d / "db.bin")
    q.tofile(d / "q.bin")
    np.savetxt(d / "db.txt", x, fmt="%.9g")
    np.savetxt(d / "q.txt", q, fmt="%.9g")
    return d, x, q


def run(*args):
    r = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def read_bin(path):
    raw = np.fromfile(path, dtype=np.uint8)
    assert raw.size == M * KMAX * 8
    ids = raw[: M * KMAX * 4].view(np.uint32).reshape(M, KMAX)
    ds = raw[M * KMAX * 4:].view(np.float32).reshape(M, KMAX)
    return ids, ds


def read_txt(path):
    lines = open(path).read().strip().split("\n")
    assert len(lines) == 2 * M
    ids = np.array([[int(v) for v in l.split()] for l in lines[:M]], dtype=np.uint32)
    ds = np.array([[float(v) for v in l.split()] for l in lines[M:]], dtype=np.float64)
    return ids, ds


def test_oneshot_matches_library_and_formats(data):
    d, x, q = data
    out = run("-x", d / "db.bin", "-q", d / "q.bin", "-n", N, "-m", M, "-d", D, "-r", R,
              "-o", d / "o.bin", "-O", d / "o.txt", "-b", "-e")
    assert "build time" in out and "query time for krbc" in out and "running k-brute force" in out
    assert re.search(r"avg overlap = [0-9.]+/32", out)
    assert re.search(r"avg actual rank of 32nd NN returned by the RBC = [0-9.]+", out)
    ids, ds = read_bin(d / "o.bin")
    tids, tds = read_txt(d / "o.txt")
    np.testing.assert_array_equal(ids, tids)
    np.testing.assert_allclose(tds, ds, rtol=1e-6, atol=1e-6)   # "%f" text
    lib_ids, lib_ds = rbc.OneShotRBC(x, R).kquery(q)
    np.testing.assert_array_equal(ids, lib_ids[:, :KMAX])
    np.testing.assert_array_equal(ds, lib_ds[:, :KMAX])


def test_text_input_equals_binary_input(data):
    d, x, q = data
    run("-X", d / "db.txt", "-Q", d / "q.txt", "-n", N, "-m", M, "-d", D, "-r", R, "-o", d / "t.bin")
    run("-x", d / "db.bin", "-q", d / "q.bin", "-n", N, "-m", M, "-d", D, "-r", R, "-o", d / "b.bin")
    a, b = read_bin(d / "t.bin"), read_bin(d / "b.bin")
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("metric,flags", [(0, ["-P", "bf16"]), (0, ["-P", "fp32"]), (1, ["-L"])])
def test_exact_mode_matches_fp64_oracle(data, metric, flags):
    d, x, q = data
    out = run("-x", d / "db.bin", "-q", d / "q.bin", "-n", N, "-m", M, "-d", D, "-r", R, "-E", *flags,
              "-o", d / "e.bin", "-e")
    ids, ds = read_bin(d / "e.bin")
    oi, od = oracle.brute_k_f64(x, q, KMAX, metric)
    np.testing.assert_array_equal(ids, oi)
    rel = np.abs(ds.astype(np.float64) - od) / np.maximum(od, 1e-300)
    assert rel.max() <= 1e-6
    assert re.search(r"avg overlap = 32\.0000/32", out)
