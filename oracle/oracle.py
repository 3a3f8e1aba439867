"""Python access to the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only to check results or to time the CPU baseline.  It loads
oracle/_build/librbc_oracle.so (built by oracle/Makefile) and wraps each C function
with numpy arrays.  See rbc_oracle.h for what each function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "librbc_oracle.so")
_lib = None

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32 = ctypes.c_uint32
_int = ctypes.c_int


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_dist_f32_fma.restype = ctypes.c_float
        L.orc_dist_f32_fma.argtypes = [_f32p, _f32p, _u32, _int]
        L.orc_dist_f32_host.restype = ctypes.c_float
        L.orc_dist_f32_host.argtypes = [_f32p, _f32p, _u32, _int]
        L.orc_dist_f32_tree.restype = ctypes.c_float
        L.orc_dist_f32_tree.argtypes = [_f32p, _f32p, _u32, _int]
        L.orc_dist_f64.restype = ctypes.c_double
        L.orc_dist_f64.argtypes = [_f32p, _f32p, _u32, _int]
        L.orc_brute_k_f32.argtypes = [_f32p, _u32, _u32, _f32p, _u32, _u32, _u32, _int, _u32,
                                      _u32p, _f32p, _u32, _int]
        L.orc_brute_k_f64.argtypes = [_f32p, _u32, _u32, _f32p, _u32, _u32, _u32, _int, _u32,
                                      _u32p, _f64p, _u32, _int]
        L.orc_brute_cpu.argtypes = [_f32p, _u32, _u32, _f32p, _u32, _u32, _u32, _int, _u32p, _int]
        L.orc_nn_f32.argtypes = [_f32p, _u32, _u32, _f32p, _u32, _u32, _u32, _int, _u32p, _f32p,
                                 _int]
        L.orc_range_count.argtypes = [_f32p, _u32, _u32, _f32p, _u32, _u32, _u32, _int, _f32p,
                                      _u32p, _int]
        L.orc_oneshot_build.restype = _int
        L.orc_oneshot_build.argtypes = [_f32p, _u32, _u32, _u32, _int, _u32p, _u32, _u32, _u32p,
                                        _u32, _u32p, _f32p, _u32, _int]
        L.orc_oneshot_kquery.argtypes = [_f32p, _u32, _u32, _int, _u32p, _u32, _u32p, _u32, _u32p,
                                         _f32p, _u32, _u32, _u32, _u32p, _f32p, _u32, _int]
        L.orc_oneshot_kquery_plan.argtypes = [_f32p, _u32, _u32, _int, _u32p, _u32, _u32p, _u32, _u32p,
                                              _u8p, _u32, _f32p, _u32, _u32, _u32, _u32p, _f32p, _u32,
                                              _int]
        L.orc_exact_owners.argtypes = [_f32p, _u32, _u32, _u32, _int, _u32p, _u32, _u32p, _int]
        L.orc_exact_radii.argtypes = [_f32p, _u32, _u32, _u32, _int, _u32p, _u32, _u32p, _f32p]
        L.orc_rbc_soa_size.argtypes = [_u32p, _u32, _u32]
        L.orc_rbc_soa_size.restype = ctypes.c_size_t
        L.orc_rbc_soa.argtypes = [_f32p, _u32, _u32, _u32p, _u32p, _u32, _f32p]
        L.orc_rep_dist2.argtypes = [_f32p, _u32, _u32, _u32, _f64p, _int]
        L.orc_rbc_exact_knn.argtypes = [_f32p, _f32p, _u32, _u32, _u32, _f32p, _u32, _u32, _u32p,
                                        _u32p, _f32p, ctypes.c_void_p, _u32, _u32p, _f32p, _u32, _int,
                                        ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def dist_f32_fma(a, b, metric=0):
    a, b = _c(a, np.float32), _c(b, np.float32)
    return lib().orc_dist_f32_fma(a, b, a.shape[-1], metric)


def dist_f32_host(a, b, metric=0):
    a, b = _c(a, np.float32), _c(b, np.float32)
    return lib().orc_dist_f32_host(a, b, a.shape[-1], metric)


def dist_f32_tree(a, b, metric=0):
    a, b = _c(a, np.float32), _c(b, np.float32)
    return lib().orc_dist_f32_tree(a, b, a.shape[-1], metric)


def dist_f64(a, b, metric=0):
    a, b = _c(a, np.float32), _c(b, np.float32)
    return lib().orc_dist_f64(a, b, a.shape[-1], metric)


def brute_k_f32(x, q, k, metric=0, threads=0):
    x, q = _c(x, np.float32), _c(q, np.float32)
    m, d = q.shape
    ids = np.empty((m, k), np.uint32)
    dists = np.empty((m, k), np.float32)
    lib().orc_brute_k_f32(x, x.shape[0], d, q, m, d, d, metric, k, ids, dists, k, threads)
    return ids, dists


def brute_k_f64(x, q, k, metric=0, threads=0):
    x, q = _c(x, np.float32), _c(q, np.float32)
    m, d = q.shape
    ids = np.empty((m, k), np.uint32)
    dists = np.empty((m, k), np.float64)
    lib().orc_brute_k_f64(x, x.shape[0], d, q, m, d, d, metric, k, ids, dists, k, threads)
    return ids, dists


def brute_cpu(x, q, metric=0, threads=0):
    x, q = _c(x, np.float32), _c(q, np.float32)
    ids = np.empty(q.shape[0], np.uint32)
    lib().orc_brute_cpu(x, x.shape[0], x.shape[1], q, q.shape[0], q.shape[1], q.shape[1], metric,
                        ids, threads)
    return ids


def nn_f32(x, q, metric=0, threads=0):
    x, q = _c(x, np.float32), _c(q, np.float32)
    ids = np.empty(q.shape[0], np.uint32)
    dists = np.empty(q.shape[0], np.float32)
    lib().orc_nn_f32(x, x.shape[0], x.shape[1], q, q.shape[0], q.shape[1], q.shape[1], metric,
                     ids, dists, threads)
    return ids, dists


def range_count(x, q, ranges, metric=0, threads=0):
    x, q = _c(x, np.float32), _c(q, np.float32)
    counts = np.empty(q.shape[0], np.uint32)
    lib().orc_range_count(x, x.shape[0], x.shape[1], q, q.shape[0], q.shape[1], q.shape[1],
                          metric, _c(ranges, np.float32), counts, threads)
    return counts


def pad16(v: int) -> int:
    return v if v % 16 == 0 else (v // 16 + 1) * 16


def oneshot_build(x, rep_ids, s, metric=0, max_iters=1000, threads=0):
    x = _c(x, np.float32)
    rep_ids = _c(rep_ids, np.uint32)
    R = rep_ids.shape[0]
    ldmap = pad16(s)
    mp = np.zeros((R, ldmap), np.uint32)
    counts = np.zeros(R, np.uint32)
    ranges = np.zeros(R, np.float32)
    bad = lib().orc_oneshot_build(x, x.shape[0], x.shape[1], x.shape[1], metric, rep_ids, R, s,
                                  mp, ldmap, counts, ranges, max_iters, threads)
    return mp, counts, ranges, bad


def oneshot_kquery(x, rep_ids, mp, counts, q, k=32, metric=0, threads=0):
    x, q = _c(x, np.float32), _c(q, np.float32)
    rep_ids = _c(rep_ids, np.uint32)
    mp = _c(mp, np.uint32)
    m = q.shape[0]
    ids = np.empty((m, k), np.uint32)
    dists = np.empty((m, k), np.float32)
    lib().orc_oneshot_kquery(x, x.shape[1], x.shape[1], metric, rep_ids, rep_ids.shape[0], mp,
                             mp.shape[1], _c(counts, np.uint32), q, m, q.shape[1], k, ids, dists,
                             k, threads)
    return ids, dists


def oneshot_kquery_plan(x, rep_ids, mp, counts, plan, q, k=32, metric=0, threads=0):
    x, q = _c(x, np.float32), _c(q, np.float32)
    rep_ids = _c(rep_ids, np.uint32)
    mp = _c(mp, np.uint32)
    plan = _c(plan, np.uint8)
    m = q.shape[0]
    ids = np.empty((m, k), np.uint32)
    dists = np.empty((m, k), np.float32)
    lib().orc_oneshot_kquery_plan(x, x.shape[1], x.shape[1], metric, rep_ids, rep_ids.shape[0], mp,
                                  mp.shape[1], _c(counts, np.uint32), plan, plan.shape[1], q, m,
                                  q.shape[1], k, ids, dists, k, threads)
    return ids, dists


def exact_owners(x, rep_ids, metric=0, threads=0):
    x = _c(x, np.float32)
    rep_ids = _c(rep_ids, np.uint32)
    owner = np.empty(x.shape[0], np.uint32)
    lib().orc_exact_owners(x, x.shape[0], x.shape[1], x.shape[1], metric, rep_ids,
                           rep_ids.shape[0], owner, threads)
    return owner


def exact_radii(x, rep_ids, owner, metric=0):
    x = _c(x, np.float32)
    rep_ids = _c(rep_ids, np.uint32)
    rad = np.empty(rep_ids.shape[0], np.float32)
    lib().orc_exact_radii(x, x.shape[0], x.shape[1], x.shape[1], metric, rep_ids,
                          rep_ids.shape[0], _c(owner, np.uint32), rad)
    return rad


class CpuRBC:
    """CPU exact RBC search over the lists of an exact index (rep_ids, perm, offsets, radii):
    the points are laid out list-major in 8-point blocks once, then ``knn`` answers queries
    with triangle-inequality pruning, AVX2 + OpenMP (the CPU baseline)."""

    def __init__(self, x, rep_ids, perm, offsets, radii, planes=True, threads=0):
        x = _c(x, np.float32)
        self.d = x.shape[1]
        self.reps = _c(x[_c(rep_ids, np.int64)], np.float32)
        self.perm = _c(perm, np.uint32)
        self.off = _c(offsets, np.uint32)
        self.radii = _c(radii, np.float32)
        R = self.reps.shape[0]
        self.soa = np.empty(lib().orc_rbc_soa_size(self.off, R, self.d), np.float32)
        lib().orc_rbc_soa(x, x.shape[1], self.d, self.perm, self.off, R, self.soa)
        self.rr = None
        if planes:
            self.rr = np.empty((R, R), np.float64)
            lib().orc_rep_dist2(self.reps, self.d, R, self.d, self.rr, threads)

    def knn(self, q, k, threads=0):
        """(ids, dists, pairs evaluated) of the k nearest points of each query."""
        q = _c(q, np.float32)
        m = q.shape[0]
        ids = np.empty((m, k), np.uint32)
        dists = np.empty((m, k), np.float32)
        pairs = ctypes.c_double()
        lib().orc_rbc_exact_knn(self.soa, q, m, q.shape[1], self.d, self.reps, self.d,
                                self.reps.shape[0], self.perm, self.off, self.radii,
                                None if self.rr is None else self.rr.ctypes.data, k, ids, dists,
                                k, threads, ctypes.byref(pairs))
        return ids, dists, pairs.value


def rbc_exact_knn(x, q, rep_ids, perm, offsets, radii, k, threads=0, planes=True):
    """CPU exact RBC k-NN over the lists of an exact index: (ids, dists, pairs evaluated)."""
    return CpuRBC(x, rep_ids, perm, offsets, radii, planes, threads).knn(q, k, threads)


def knn_f64_numpy(x, q, k, metric=0):
    """Independent FP64 k-NN (numpy, not the C oracle): exact distances by direct
    differences, order by (distance, index).  Used to pin orc_brute_k_f64."""
    x64 = np.asarray(x, np.float64)
    q64 = np.asarray(q, np.float64)
    ids = np.empty((q64.shape[0], k), np.int64)
    ds = np.empty((q64.shape[0], k), np.float64)
    for i in range(q64.shape[0]):
        diff = x64 - q64[i]
        dd = (diff * diff).sum(1) if metric == 0 else np.abs(diff).sum(1)
        order = np.lexsort((np.arange(dd.shape[0]), dd))[:k]
        ids[i] = order
        ds[i] = dd[order]
    return ids, ds
