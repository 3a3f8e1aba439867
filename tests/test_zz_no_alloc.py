"""The allocation-free brute force (rbc_brute_k_noalloc, through the C ABI) returns what the
allocating rbc_brute_k returns: ids and FP32 distances bit-exact, on two calls in a row
whose second is smaller (the shared scratch array keeps the first call's contents), and
sizes beyond its scratch array are rejected before any work is queued."""
import numpy as np
import pytest

from conftest import datagen, rbc

pytestmark = pytest.mark.gpu


def _device(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.mark.parametrize("metric", [rbc.L2, rbc.L1])
def test_noalloc_matches_bruteK_on_two_calls(metric):
    import torch
    cases = [(60000, 3000, 40, 32, 0), (7000, 300, 16, 5, 1)]  # the second call is smaller
    for n, m, d, k, seed in cases:
        print(f"noalloc case n={n} m={m} d={d} k={k} metric={metric}", flush=True)
        x = datagen.gmm(n, d, 10 + seed)
        q = datagen.gmm(m, d, 20 + seed)
        ref_i, ref_d = rbc.bruteK(x, q, k, metric)
        xd, qd = _device(x), _device(q)
        ids = torch.full((m, k), -7, dtype=torch.int32, device="cuda:0")
        ds = torch.full((m, k), -7.0, dtype=torch.float32, device="cuda:0")
        rbc.bruteK_noalloc(xd, qd, k, ids, ds, metric)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), ref_i)
        np.testing.assert_array_equal(ds.cpu().numpy(), ref_d)


def test_noalloc_rejects_what_its_scratch_cannot_hold():
    import torch
    d, k = 4, 32
    x = _device(datagen.uniform(1000, d, 0))
    q = _device(datagen.uniform(300000, d, 1))  # 300000 x 32 partial entries > 64 MiB
    ids = torch.empty((300000, k), dtype=torch.int32, device="cuda:0")
    ds = torch.empty((300000, k), dtype=torch.float32, device="cuda:0")
    with pytest.raises(rbc.RBCError, match="scratch"):
        rbc.bruteK_noalloc(x, q, k, ids, ds)
    xs = torch.zeros((1000, 6), dtype=torch.float32, device="cuda:0")[:, :5]  # row stride 6
    with pytest.raises(rbc.RBCError, match="multiples of 4"):
        rbc.bruteK_noalloc(xs, _device(datagen.uniform(10, 5, 2)), 3, ids[:10], ds[:10])
