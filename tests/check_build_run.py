"""Runs the main entry points through the device bounds-checking build of the library
(librbc_hip_check.so, selected with RBC_LIB) on small inputs: each must return normally,
with results equal to the oracle's, and the checker itself must catch a planted
out-of-range index.  Started by tests/test_zy_check_build_gpu.py in a child process (the
library is chosen when it is first loaded)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from conftest import datagen, rbc  # noqa: E402
import oracle  # noqa: E402  (the checker; conftest put oracle/ on the path)


def step(msg):
    print(msg, flush=True)


def main():
    L = rbc.lib()
    assert L.rbc_check_enabled() == 1, "RBC_LIB does not name the check build"
    v = rbc._u32()
    step("self-test: index inside and outside its buffer")
    assert L.rbc_check_selftest(3, 8, rbc.ctypes.byref(v)) == 0 and v.value == 1003
    st = L.rbc_check_selftest(8, 8, rbc.ctypes.byref(v))
    assert st == 6 and v.value == 1000, (st, v.value)
    assert b"bounds check failed" in L.rbc_last_error()
    assert L.rbc_check_selftest(2, 8, rbc.ctypes.byref(v)) == 0  # the word was cleared
    x = datagen.lowrank(30000, 128, 0)
    q = datagen.lowrank(300, 128, 1)
    k = 10
    oi, _ = oracle.brute_k_f64(x, q, k)
    idx = rbc.ExactRBC(x, 160, seed=3)
    for prec in (rbc.PREC_FP32, rbc.PREC_BF16):
        step(f"exact kquery, precision {prec}")
        ids, _ = idx.kquery(q, k, precision=prec)
        assert np.array_equal(ids, oi)
    idx.close()
    step("one-shot build and kquery")
    xs = datagen.uniform(8000, 16, 0)
    qs = datagen.uniform(200, 16, 1)
    os_idx = rbc.OneShotRBC(xs, 64, s=400, seed=1)
    ids, _ = os_idx.kquery(qs)
    oi1, _ = oracle.brute_k_f32(xs, qs, 1)
    assert ids.shape[0] == 200 and (ids[:, 0] < 8000).all()
    os_idx.close()
    step("batched kquery")
    xb = [datagen.uniform(n, 32, 50 + n) for n in (1500, 3000)]
    qb = [datagen.uniform(m, 32, 90 + m) for m in (40, 7)]
    b = rbc.BatchedRBC(xb, [39, 55])
    bi, _, _ = b.kquery(qb, 16)
    for x_, q_, i_ in zip(xb, qb, bi):
        assert np.array_equal(i_, oracle.brute_k_f64(x_, q_, 16)[0])
    b.close()
    step("check build ok")


if __name__ == "__main__":
    main()
