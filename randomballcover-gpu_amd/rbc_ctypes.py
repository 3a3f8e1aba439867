"""ctypes binding of librbc_hip.so (the engine's C ABI, include/rbc_hip.h).

The Python names mirror the reference's C entry points (bruteK, bruteSearch,
bruteRangeCount, bruteCPU, buildRBC, queryRBC, kqueryRBC, destroyRBC) plus the exact
search handle.  Host arrays are numpy; device arrays are anything with a
``data_ptr()`` (torch tensors on the GPU).  Every call raises RBCError on a non-zero
status.  Loading fails loudly when the library has not been built: there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# RBC_LIB selects another build of the library (librbc_hip_check.so: device bounds checks).
LIB_PATH = os.environ.get("RBC_LIB") or os.path.join(HERE, "librbc_hip.so")

KMAX = 32
DUMMY_IDX = 0xFFFFFFFF
L2, L1 = 0, 1
MEM_HOST, MEM_DEVICE = 0, 1
PREC_FP32, PREC_BF16 = 0, 1
SKETCH_AUTO, SKETCH_OFF, SKETCH_ON = 0, 1, 2

_u32 = ctypes.c_uint32
_int = ctypes.c_int
_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64


class RBCError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"rbc status {status}: {msg}")
        self.status = status


class Matrix(ctypes.Structure):  # defs.h:56-63
    _fields_ = [("mat", _vp), ("r", _u32), ("c", _u32), ("pr", _u32), ("pc", _u32), ("ld", _u32)]


class IntMatrix(ctypes.Structure):  # defs.h:76-83
    _fields_ = [("mat", _vp), ("r", _u32), ("c", _u32), ("pr", _u32), ("pc", _u32), ("ld", _u32)]


class RBCStruct(ctypes.Structure):  # defs.h:95-100
    _fields_ = [("dx", Matrix), ("dxMap", IntMatrix), ("dr", Matrix), ("groupCount", _vp)]


class QueryStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "pairs_probe", "pairs_scan", "pairs_rerank", "items_scan", "lists_scanned", "bytes_scan",
        "guard_fp64", "fallbacks", "ms_qreps", "ms_probe", "ms_plan", "ms_scan", "ms_select",
        "ms_total", "flops_scan", "sketch")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class SketchInfo(ctypes.Structure):
    _fields_ = [("dims", _u32), ("active", _u32), ("norm_bound", ctypes.c_double),
                ("resid_var", ctypes.c_double), ("total_var", ctypes.c_double),
                ("mean_owner_d2", ctypes.c_double)]


class ExactInfo(ctypes.Structure):
    _fields_ = [("n", _u32), ("d", _u32), ("ld", _u32), ("num_reps", _u32), ("metric", _u32),
                ("max_list", _u32), ("min_list", _u32), ("device_bytes", ctypes.c_size_t)]


_lib = None

_SIGS = {
    "rbc_last_error": (ctypes.c_char_p, []),
    "rbc_check_enabled": (_int, []),
    "rbc_check_selftest": (_int, [_u32, _u32, ctypes.POINTER(_u32)]),
    "rbc_version": (_int, []),
    "rbc_set_device": (_int, [_int]),
    "rbc_set_metric": (_int, [_int]),
    "rbc_get_metric": (_int, []),
    "rbc_measure_peaks": (_int, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), _vp]),
    "rbc_brute_k": (_int, [_vp, _u32, _u32, _vp, _u32, _u32, _u32, _int, _u32, _vp, _vp, _u32, _int,
                           _vp]),
    "rbc_brute_k_noalloc": (_int, [_vp, _u32, _u32, _vp, _u32, _u32, _u32, _int, _u32, _vp, _vp, _u32,
                                    _vp]),
    "rbc_brute_search": (_int, [_vp, _u32, _u32, _vp, _u32, _u32, _u32, _int, _vp, _vp, _int, _vp]),
    "rbc_brute_range_count": (_int, [_vp, _u32, _u32, _vp, _u32, _u32, _u32, _int, _vp, _vp, _int,
                                     _vp]),
    "rbc_brute_cpu": (_int, [_vp, _u32, _u32, _vp, _u32, _u32, _u32, _int, _vp]),
    "rbc_build": (_int, [_vp, _u32, _u32, _u32, _u32, _u32, _vp, _u64, _int,
                         ctypes.POINTER(RBCStruct)]),
    "rbc_query": (_int, [_vp, _u32, _u32, ctypes.POINTER(RBCStruct), _vp, _vp, _int, _vp]),
    "rbc_kquery": (_int, [_vp, _u32, _u32, ctypes.POINTER(RBCStruct), _vp, _vp, _u32, _int, _vp]),
    "rbc_query_plan": (_int, [_vp, _u32, _u32, ctypes.POINTER(RBCStruct), _vp, _u32, _vp, _vp, _int, _vp]),
    "rbc_kquery_plan": (_int, [_vp, _u32, _u32, ctypes.POINTER(RBCStruct), _vp, _u32, _vp, _vp, _u32, _int,
                               _vp]),
    "rbc_destroy": (_int, [ctypes.POINTER(RBCStruct)]),
    "rbc_exact_build": (_int, [_vp, _u32, _u32, _u32, _u32, _int, _vp, _u64, _int,
                               ctypes.POINTER(_vp)]),
    "rbc_exact_kquery": (_int, [_vp, _vp, _u32, _u32, _u32, _int, _vp, _vp, _u32, _int, _vp,
                                ctypes.POINTER(QueryStats)]),
    "rbc_exact_get_info": (_int, [_vp, ctypes.POINTER(ExactInfo)]),
    "rbc_exact_rep_dists": (_int, [_vp, _vp, _u32, _u32, _int, _vp, _u32, _vp, ctypes.POINTER(ctypes.c_double),
                                   _int, _vp]),
    "rbc_exact_get_lists": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "rbc_exact_destroy": (_int, [_vp]),
    "rbc_exact_set_sketch": (_int, [_vp, _int]),
    "rbc_exact_set_search_params": (_int, [_vp, _u32, _u32]),
    "rbc_exact_sketch_info": (_int, [_vp, ctypes.POINTER(SketchInfo)]),
    "rbc_exact_sketch_rows": (_int, [_vp, _vp, _u32, _u32, _int, _vp, _vp, _vp]),
    "rbc_exact_sketch_basis": (_int, [_vp, _vp, _vp]),
    "rbc_exact_save": (_int, [_vp, ctypes.c_char_p]),
    "rbc_exact_load": (_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "rbc_exact_set_proposal": (_int, [_vp, _vp]),
    "rbc_exact_set_proposal_mass": (_int, [_vp, ctypes.c_float]),
    "rbc_mlp_create": (_int, [_u32, _u32, _u32, _u32, _u64, ctypes.POINTER(_vp)]),
    "rbc_mlp_destroy": (_int, [_vp]),
    "rbc_mlp_num_params": (_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "rbc_mlp_get_params": (_int, [_vp, _vp]),
    "rbc_mlp_set_params": (_int, [_vp, _vp]),
    "rbc_mlp_save": (_int, [_vp, ctypes.c_char_p]),
    "rbc_mlp_load": (_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "rbc_mlp_loss_grad": (_int, [_vp, _vp, _u32, _u32, _vp, ctypes.POINTER(ctypes.c_float), _vp]),
    "rbc_mlp_loss_grad_multi": (_int, [_vp, _vp, _u32, _u32, _vp, _u32, ctypes.POINTER(ctypes.c_float), _vp]),
    "rbc_mlp_propose_lists": (_int, [_vp, _vp, _u32, _u32, _u32, ctypes.c_float, _vp, _vp]),
    "rbc_mlp_train_multi": (_int, [_vp, _vp, _u32, _u32, _vp, _u32, _int, _u32, _u32, ctypes.c_float, _u64,
                                   _vp]),
    "rbc_mlp_logits": (_int, [_vp, _vp, _u32, _u32, _vp]),
    "rbc_mlp_propose": (_int, [_vp, _vp, _u32, _u32, _vp]),
    "rbc_mlp_train": (_int, [_vp, _vp, _u32, _u32, _vp, _int, _u32, _u32, ctypes.c_float, _u64, _vp]),
    "rbc_batched_build": (_int, [_u32, _vp, _vp, _vp, _vp, _vp, _int, _u64, _int, ctypes.POINTER(_vp)]),
    "rbc_batched_kquery": (_int, [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _int, _vp, _vp]),
    "rbc_batched_get_index": (_int, [_vp, _u32, ctypes.POINTER(_vp)]),
    "rbc_batched_destroy": (_int, [_vp]),
}


def lib():
    """Loads librbc_hip.so; raises if it is missing (build it with __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RBCError(-1, f"{LIB_PATH} is not built")
        # torch bundles its own libamdhip64.so.7.  Both it and this library resolve the
        # HIP runtime by that soname, so whichever loads first serves the whole process.
        # Loading torch first keeps one runtime (torch's) for torch tensors and our kernels;
        # the other order leaves torch unable to initialise its devices.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        raise RBCError(status, lib().rbc_last_error().decode())


def measure_peaks():
    """Measured dense BF16 MFMA and FP32 FMA throughput of the current device (TFLOP/s)."""
    a, b = ctypes.c_double(), ctypes.c_double()
    _check(lib().rbc_measure_peaks(ctypes.byref(a), ctypes.byref(b), None))
    return {"bf16_mfma_tflops": a.value, "fp32_fma_tflops": b.value}


def _host(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data


def _ptr(a):
    return a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data


def set_metric(metric):
    _check(lib().rbc_set_metric(metric))


# ------------------------------------------------------------------ brute force --

def bruteK(x, q, k=KMAX, metric=L2):
    """bruteK (brute.cu:65-88) on host arrays: (ids [m,k] uint32, dists [m,k] float32)."""
    x, xp = _host(x, np.float32)
    q, qp = _host(q, np.float32)
    m, d = q.shape
    ids = np.empty((m, k), np.uint32)
    ds = np.empty((m, k), np.float32)
    _check(lib().rbc_brute_k(xp, x.shape[0], x.shape[1], qp, m, q.shape[1], d, metric, k,
                             ids.ctypes.data, ds.ctypes.data, k, MEM_HOST, None))
    return ids, ds


def bruteK_noalloc(x, q, k, ids, ds, metric=L2, stream=None):
    """rbc_brute_k_noalloc on device tensors (x [n,d], q [m,d] float32, row strides multiples of
    4; ids [m,>=k] int32/uint32 and ds [m,>=k] float32 outputs): the allocation-free bruteK.
    Asynchronous on `stream` (a raw hipStream_t or None for the default stream)."""
    m, d = q.shape
    _check(lib().rbc_brute_k_noalloc(_ptr(x), x.shape[0], x.stride(0), _ptr(q), m, q.stride(0), d, metric,
                                     k, _ptr(ids), _ptr(ds), ids.stride(0), stream))


def bruteSearch(x, q, metric=L2):
    x, xp = _host(x, np.float32)
    q, qp = _host(q, np.float32)
    m, d = q.shape
    ids = np.empty(m, np.uint32)
    ds = np.empty(m, np.float32)
    _check(lib().rbc_brute_search(xp, x.shape[0], x.shape[1], qp, m, q.shape[1], d, metric,
                                  ids.ctypes.data, ds.ctypes.data, MEM_HOST, None))
    return ids, ds


def bruteRangeCount(x, q, ranges, metric=L2):
    x, xp = _host(x, np.float32)
    q, qp = _host(q, np.float32)
    r, rp = _host(ranges, np.float32)
    m, d = q.shape
    cnt = np.empty(m, np.uint32)
    _check(lib().rbc_brute_range_count(xp, x.shape[0], x.shape[1], qp, m, q.shape[1], d, metric,
                                       rp, cnt.ctypes.data, MEM_HOST, None))
    return cnt


def bruteCPU(x, q, metric=L2):
    x, xp = _host(x, np.float32)
    q, qp = _host(q, np.float32)
    ids = np.empty(q.shape[0], np.uint32)
    _check(lib().rbc_brute_cpu(xp, x.shape[0], x.shape[1], qp, q.shape[0], q.shape[1], q.shape[1],
                               metric, ids.ctypes.data))
    return ids


# ------------------------------------------------------------- one-shot RBC ----

class OneShotRBC:
    """buildRBC / queryRBC / kqueryRBC / destroyRBC (rbc.cu) over an rbcStruct."""

    def __init__(self, x, num_reps, s=None, rep_ids=None, seed=0):
        x, xp = _host(x, np.float32)
        self.n, self.d = x.shape
        self.s = num_reps if s is None else s
        self.rbcs = RBCStruct()
        rp = None
        if rep_ids is not None:
            self._reps, rp = _host(rep_ids, np.uint32)
        _check(lib().rbc_build(xp, self.n, self.d, self.d, num_reps, self.s, rp, seed, MEM_HOST,
                               ctypes.byref(self.rbcs)))
        self.num_reps = num_reps

    def group_counts(self):
        n = self.num_reps
        return np.ctypeslib.as_array(ctypes.cast(self.rbcs.groupCount,
                                                 ctypes.POINTER(ctypes.c_uint32)), (n,)).copy()

    def query(self, q):
        q, qp = _host(q, np.float32)
        m = q.shape[0]
        ids = np.empty(m, np.uint32)
        ds = np.empty(m, np.float32)
        _check(lib().rbc_query(qp, m, q.shape[1], ctypes.byref(self.rbcs), ids.ctypes.data,
                               ds.ctypes.data, MEM_HOST, None))
        return ids, ds

    def kquery(self, q):
        q, qp = _host(q, np.float32)
        m = q.shape[0]
        ids = np.empty((m, KMAX), np.uint32)
        ds = np.empty((m, KMAX), np.float32)
        _check(lib().rbc_kquery(qp, m, q.shape[1], ctypes.byref(self.rbcs), ids.ctypes.data,
                                ds.ctypes.data, KMAX, MEM_HOST, None))
        return ids, ds

    def query_plan(self, q, plan):
        """queryRBC under a computation matrix (num_reps x num_reps, nonzero = scan)."""
        q, qp = _host(q, np.float32)
        plan = np.ascontiguousarray(plan, dtype=np.uint8)
        m = q.shape[0]
        ids = np.empty(m, np.uint32)
        ds = np.empty(m, np.float32)
        _check(lib().rbc_query_plan(qp, m, q.shape[1], ctypes.byref(self.rbcs), plan.ctypes.data, plan.shape[1],
                                    ids.ctypes.data, ds.ctypes.data, MEM_HOST, None))
        return ids, ds

    def kquery_plan(self, q, plan):
        """kqueryRBC under a computation matrix (num_reps x num_reps, nonzero = scan)."""
        q, qp = _host(q, np.float32)
        plan = np.ascontiguousarray(plan, dtype=np.uint8)
        m = q.shape[0]
        ids = np.empty((m, KMAX), np.uint32)
        ds = np.empty((m, KMAX), np.float32)
        _check(lib().rbc_kquery_plan(qp, m, q.shape[1], ctypes.byref(self.rbcs), plan.ctypes.data, plan.shape[1],
                                     ids.ctypes.data, ds.ctypes.data, KMAX, MEM_HOST, None))
        return ids, ds

    def close(self):
        if self.rbcs.dx.mat:
            _check(lib().rbc_destroy(ctypes.byref(self.rbcs)))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- exact RBC ----

class ExactRBC:
    """Exact k-NN index (rbc_exact_*).  x may be a host numpy array or a device tensor."""

    def __init__(self, x, num_reps, metric=L2, rep_ids=None, seed=0, device=False):
        self.h = _vp()
        if device:
            n, d = x.shape
            xp, ld, mem = _ptr(x), x.stride(0), MEM_DEVICE
        else:
            x, xp = _host(x, np.float32)
            n, d = x.shape
            ld, mem = d, MEM_HOST
        rp = None
        if rep_ids is not None:
            self._reps, rp = _host(rep_ids, np.uint32)
        _check(lib().rbc_exact_build(xp, n, d, ld, num_reps, metric, rp, seed, mem,
                                     ctypes.byref(self.h)))
        self.n, self.d, self.num_reps, self.metric = n, d, num_reps, metric

    def save(self, path):
        """rbc_exact_save: the whole index to a file."""
        _check(lib().rbc_exact_save(self.h, os.fsencode(path)))

    @classmethod
    def load(cls, path):
        """rbc_exact_load: an index saved by save(), ready to query."""
        self = cls.__new__(cls)
        self.h = _vp()
        _check(lib().rbc_exact_load(os.fsencode(path), ctypes.byref(self.h)))
        inf = self.info()
        self.n, self.d, self.num_reps, self.metric = inf["n"], inf["d"], inf["num_reps"], inf["metric"]
        return self

    def kquery(self, q, k, precision=PREC_FP32, stats=False):
        q, qp = _host(q, np.float32)
        m = q.shape[0]
        ids = np.empty((m, k), np.uint32)
        ds = np.empty((m, k), np.float32)
        st = QueryStats()
        _check(lib().rbc_exact_kquery(self.h, qp, m, q.shape[1], k, precision, ids.ctypes.data,
                                      ds.ctypes.data, k, MEM_HOST, None, ctypes.byref(st)))
        return (ids, ds, st.as_dict()) if stats else (ids, ds)

    def kquery_device(self, q, k, ids, ds, precision=PREC_FP32, stream=None, stats=True):
        """q, ids, ds are device tensors ([m,d] float32, [m,k] int32/uint32, [m,k] float32).
        Returns the query's statistics (per-stage hipEvent times, pair counts), or None with
        stats=False (no events recorded, no statistics copied back)."""
        st = QueryStats() if stats else None
        _check(lib().rbc_exact_kquery(self.h, _ptr(q), q.shape[0], q.stride(0), k, precision,
                                      _ptr(ids), _ptr(ds), ids.stride(0), MEM_DEVICE, stream,
                                      ctypes.byref(st) if stats else None))
        return st.as_dict() if stats else None

    def set_proposal(self, mlp, mass=None):
        """rbc_exact_set_proposal (and, when given, rbc_exact_set_proposal_mass)."""
        _check(lib().rbc_exact_set_proposal(self.h, mlp.h if mlp is not None else None))
        if mass is not None:
            _check(lib().rbc_exact_set_proposal_mass(self.h, mass))
        self._proposal = mlp

    def rep_dists(self, q, precision=PREC_FP32):
        """rbc_exact_rep_dists on host arrays: (D [m, num_reps], abs_err [m], rel_err) with
        |D - exact| <= rel_err * exact + abs_err[i]."""
        q, qp = _host(q, np.float32)
        m = q.shape[0]
        D = np.empty((m, self.num_reps), np.float32)
        ae = np.empty(m, np.float32)
        rel = ctypes.c_double()
        _check(lib().rbc_exact_rep_dists(self.h, qp, m, q.shape[1], precision, D.ctypes.data, self.num_reps,
                                         ae.ctypes.data, ctypes.byref(rel), MEM_HOST, None))
        return D, ae, rel.value

    def set_sketch(self, mode):
        """rbc_exact_set_sketch: SKETCH_AUTO, SKETCH_OFF or SKETCH_ON."""
        _check(lib().rbc_exact_set_sketch(self.h, mode))

    def set_search_params(self, probe_lists=8, plane_reps=12):
        """rbc_exact_set_search_params (BF16 query; results do not depend on them)."""
        _check(lib().rbc_exact_set_search_params(self.h, probe_lists, plane_reps))

    def sketch_info(self):
        inf = SketchInfo()
        _check(lib().rbc_exact_sketch_info(self.h, ctypes.byref(inf)))
        return {n: getattr(inf, n) for n, _ in inf._fields_}

    def sketch_rows(self, x):
        """rbc_exact_sketch_rows on a host array: (BF16 bits [m, 32] as uint16, err [m], n2 [m])."""
        x, xp = _host(x, np.float32)
        m = x.shape[0]
        out = np.empty((m, 32), np.uint16)
        err = np.empty(m, np.float32)
        n2 = np.empty(m, np.float32)
        _check(lib().rbc_exact_sketch_rows(self.h, xp, m, x.shape[1], MEM_HOST, out.ctypes.data,
                                           err.ctypes.data, n2.ctypes.data))
        return out, err, n2

    def sketch_basis(self):
        """rbc_exact_sketch_basis: (P [dims, d], mu [d]) as stored (FP32)."""
        dims = self.sketch_info()["dims"]
        P = np.empty((dims, self.d), np.float32)
        mu = np.empty(self.d, np.float32)
        _check(lib().rbc_exact_sketch_basis(self.h, P.ctypes.data, mu.ctypes.data))
        return P, mu

    def info(self):
        inf = ExactInfo()
        _check(lib().rbc_exact_get_info(self.h, ctypes.byref(inf)))
        return {n: getattr(inf, n) for n, _ in inf._fields_}

    def lists(self):
        R, n = self.num_reps, self.n
        reps = np.empty(R, np.uint32)
        off = np.empty(R + 1, np.uint32)
        perm = np.empty(n, np.uint32)
        owner = np.empty(n, np.uint32)
        radii = np.empty(R, np.float32)
        _check(lib().rbc_exact_get_lists(self.h, reps.ctypes.data, off.ctypes.data,
                                         perm.ctypes.data, owner.ctypes.data, radii.ctypes.data))
        return {"rep_ids": reps, "offsets": off, "perm": perm, "owner": owner, "radii": radii}

    def close(self):
        if self.h:
            _check(lib().rbc_exact_destroy(self.h))
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ proposal network ----

class ProposalMLP:
    """rbc_mlp_*: d_in -> h1 -> h2 -> n_out, GELU, FP16 weights, trained on the GPU."""

    def __init__(self, d_in, n_out, h1=256, h2=256, seed=0):
        self.h = _vp()
        _check(lib().rbc_mlp_create(d_in, h1, h2, n_out, seed, ctypes.byref(self.h)))
        self.dims = (d_in, h1, h2, n_out)
        n = ctypes.c_size_t()
        _check(lib().rbc_mlp_num_params(self.h, ctypes.byref(n)))
        self.num_params = n.value

    def save(self, path):
        """rbc_mlp_save: layer sizes, parameters and Adam state to a file."""
        _check(lib().rbc_mlp_save(self.h, os.fsencode(path)))

    @classmethod
    def load(cls, path):
        self = cls.__new__(cls)
        self.h = _vp()
        _check(lib().rbc_mlp_load(os.fsencode(path), ctypes.byref(self.h)))
        with open(path, "rb") as f:
            self.dims = tuple(int(v) for v in np.frombuffer(f.read(24)[8:], dtype=np.uint32))
        n = ctypes.c_size_t()
        _check(lib().rbc_mlp_num_params(self.h, ctypes.byref(n)))
        self.num_params = n.value
        return self

    def get_params(self):
        out = np.empty(self.num_params, np.float32)
        _check(lib().rbc_mlp_get_params(self.h, out.ctypes.data))
        return out

    def set_params(self, p):
        p, pp = _host(p, np.float32)
        _check(lib().rbc_mlp_set_params(self.h, pp))

    def loss_grad(self, x, labels):
        """Loss and gradient; labels [rows] (one label) or [rows, L] (soft targets)."""
        x, xp = _host(x, np.float32)
        lab, lp = _host(labels, np.uint32)
        L = 1 if lab.ndim == 1 else lab.shape[1]
        loss = ctypes.c_float()
        g = np.empty(self.num_params, np.float32)
        _check(lib().rbc_mlp_loss_grad_multi(self.h, xp, x.shape[0], x.shape[1], lp, L, ctypes.byref(loss),
                                             g.ctypes.data))
        return loss.value, g

    def logits(self, x):
        x, xp = _host(x, np.float32)
        out = np.empty((x.shape[0], self.dims[3]), np.float32)
        _check(lib().rbc_mlp_logits(self.h, xp, x.shape[0], x.shape[1], out.ctypes.data))
        return out

    def propose(self, x):
        """Index of each row's largest logit (lowest index among equals)."""
        x, xp = _host(x, np.float32)
        out = np.empty(x.shape[0], np.uint32)
        _check(lib().rbc_mlp_propose(self.h, xp, x.shape[0], x.shape[1], out.ctypes.data))
        return out

    def propose_lists(self, x, P, mass):
        """rbc_mlp_propose_lists: (lists [rows, P], counts [rows])."""
        x, xp = _host(x, np.float32)
        lists = np.empty((x.shape[0], P), np.uint32)
        counts = np.empty(x.shape[0], np.uint32)
        _check(lib().rbc_mlp_propose_lists(self.h, xp, x.shape[0], x.shape[1], P, mass, lists.ctypes.data,
                                           counts.ctypes.data))
        return lists, counts

    def train(self, x, labels, epochs=5, batch=1024, lr=1e-3, seed=0):
        """Adam training; labels [rows] (one label) or [rows, L] (soft targets over L)."""
        x, xp = _host(x, np.float32)
        lab, lp = _host(labels, np.uint32)
        L = 1 if lab.ndim == 1 else lab.shape[1]
        losses = np.zeros(epochs, np.float32)
        _check(lib().rbc_mlp_train_multi(self.h, xp, x.shape[0], x.shape[1], lp, L, MEM_HOST, epochs, batch,
                                         lr, seed, losses.ctypes.data))
        return losses

    def close(self):
        if self.h:
            _check(lib().rbc_mlp_destroy(self.h))
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- batched exact ----

class BatchedStats(ctypes.Structure):
    _fields_ = [("ms_graph", ctypes.c_double), ("tile_queries", _u32), ("graph_replayed", _u32),
                ("uncertified", _u32), ("blocks_per_cu", _u32), ("tune_ms", ctypes.c_double * 3),
                ("guard_fp64", ctypes.c_double), ("pairs_scan", ctypes.c_double),
                ("flops", ctypes.c_double), ("steps", ctypes.c_double), ("longest_tile_ms", ctypes.c_double)]


def _ptr_array(ptrs, ctype=_vp):
    arr = (ctype * len(ptrs))(*ptrs)
    return arr


class BatchedRBC:
    """rbc_batched_*: many independent exact indexes queried in one persistent launch."""

    def __init__(self, xs, num_reps, metric=L2, seed=0):
        self._xs = [np.ascontiguousarray(x, np.float32) for x in xs]
        c = len(self._xs)
        n = np.array([x.shape[0] for x in self._xs], np.uint32)
        d = np.array([x.shape[1] for x in self._xs], np.uint32)
        r = np.ascontiguousarray(num_reps, np.uint32)
        self.h = _vp()
        _check(lib().rbc_batched_build(c, _ptr_array([x.ctypes.data for x in self._xs]), n.ctypes.data,
                                       d.ctypes.data, d.ctypes.data, r.ctypes.data, metric, seed,
                                       MEM_HOST, ctypes.byref(self.h)))
        self.count, self.dims = c, d

    def kquery(self, qs, k):
        qs = [np.ascontiguousarray(q, np.float32) for q in qs]
        m = np.array([q.shape[0] for q in qs], np.uint32)
        ldq = np.array([q.shape[1] for q in qs], np.uint32)
        ids = [np.empty((q.shape[0], k), np.uint32) for q in qs]
        ds = [np.empty((q.shape[0], k), np.float32) for q in qs]
        ldo = np.full(len(qs), k, np.uint32)
        st = BatchedStats()
        _check(lib().rbc_batched_kquery(self.h, _ptr_array([q.ctypes.data for q in qs]), m.ctypes.data,
                                        ldq.ctypes.data, k, _ptr_array([i.ctypes.data for i in ids]),
                                        _ptr_array([x.ctypes.data for x in ds]), ldo.ctypes.data,
                                        MEM_HOST, None, ctypes.byref(st)))
        stats = {n: getattr(st, n) for n, _ in st._fields_ if n != "tune_ms"}
        stats["tune_ms"] = list(st.tune_ms)
        return ids, ds, stats

    def lists(self, i):
        """Instance i's exact index lists (as ExactRBC.lists)."""
        h = _vp()
        _check(lib().rbc_batched_get_index(self.h, i, ctypes.byref(h)))
        inf = ExactInfo()
        _check(lib().rbc_exact_get_info(h, ctypes.byref(inf)))
        R, n = inf.num_reps, inf.n
        reps = np.empty(R, np.uint32)
        off = np.empty(R + 1, np.uint32)
        perm = np.empty(n, np.uint32)
        owner = np.empty(n, np.uint32)
        radii = np.empty(R, np.float32)
        _check(lib().rbc_exact_get_lists(h, reps.ctypes.data, off.ctypes.data, perm.ctypes.data,
                                         owner.ctypes.data, radii.ctypes.data))
        return {"rep_ids": reps, "offsets": off, "perm": perm, "owner": owner, "radii": radii}

    def kquery_device(self, qs, k, ids, ds):
        """qs, ids, ds: lists of device tensors; returns the stats dict."""
        m = np.array([q.shape[0] for q in qs], np.uint32)
        ldq = np.array([q.stride(0) for q in qs], np.uint32)
        ldo = np.array([i.stride(0) for i in ids], np.uint32)
        st = BatchedStats()
        _check(lib().rbc_batched_kquery(self.h, _ptr_array([q.data_ptr() for q in qs]), m.ctypes.data,
                                        ldq.ctypes.data, k, _ptr_array([i.data_ptr() for i in ids]),
                                        _ptr_array([x.data_ptr() for x in ds]), ldo.ctypes.data,
                                        MEM_DEVICE, None, ctypes.byref(st)))
        stats = {n: getattr(st, n) for n, _ in st._fields_ if n != "tune_ms"}
        stats["tune_ms"] = list(st.tune_ms)
        return stats

    def close(self):
        if self.h:
            _check(lib().rbc_batched_destroy(self.h))
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
