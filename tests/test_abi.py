"""The C-ABI library loads and exports every symbol include/*.h declares (CPU only:
no compute calls)."""
import ctypes
import os
import re

from conftest import ROOT, rbc

HEADERS = ["rbc_hip.h", "rbc.h", "brute.h"]


def declared_c_symbols():
    names = set()
    src = open(os.path.join(ROOT, "include", "rbc_hip.h")).read()
    for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(rbc_\w+)\s*\(", src, re.M):
        names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_c_symbols()
    for must in ["rbc_brute_k", "rbc_brute_search", "rbc_brute_range_count", "rbc_brute_cpu",
                 "rbc_exact_build", "rbc_exact_kquery", "rbc_exact_destroy", "rbc_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    L = rbc.lib()
    missing = [n for n in declared_c_symbols() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_version_and_metric_setting():
    L = rbc.lib()
    assert L.rbc_version() >= 100
    assert L.rbc_get_metric() == rbc.L2
    assert L.rbc_set_metric(7) == 1  # RBC_ERR_ARG
    assert b"metric" in L.rbc_last_error()
    assert L.rbc_set_metric(rbc.L1) == 0 and L.rbc_get_metric() == rbc.L1
    assert L.rbc_set_metric(rbc.L2) == 0


def test_struct_layouts_match_reference():
    # defs.h:56-100 on LP64
    assert ctypes.sizeof(rbc.Matrix) == 32
    assert ctypes.sizeof(rbc.IntMatrix) == 32
    assert ctypes.sizeof(rbc.RBCStruct) == 104


def test_argument_errors_do_not_touch_the_gpu():
    L = rbc.lib()
    # d = 0 and k > KMAX are rejected before any HIP call.
    st = L.rbc_brute_k(None, 0, 0, None, 0, 0, 0, 0, 1, None, None, 1, 0, None)
    assert st == 1
    st = L.rbc_brute_k(None, 0, 4, None, 0, 4, 4, 0, 33, None, None, 33, 0, None)
    assert st == 1
    # Null outputs of the peak probe.
    assert L.rbc_measure_peaks(None, None, None) == 1
    assert b"null" in L.rbc_last_error()


def test_internals_reject_bad_arguments_before_the_gpu():
    L = rbc.lib()
    # rbc_dist_sub: d = 0; rbc_compute_reps: no reps; rbc_plan_init: NULL plan;
    # rbc_plan_query: K above KMAX; rbc_setup_reps: NULL rows.
    assert L.rbc_dist_sub(None, 4, 0, 16, None, 16, 4, 0, None, 16, None) == 1
    assert L.rbc_compute_reps(None, 1, 4, None, 0, 4, 4, None, None, 0, 0, None) == 1
    assert L.rbc_plan_init(None, None, 0, None, None, 0) == 1
    assert L.rbc_plan_free(None) == 1
    assert L.rbc_plan_query(None, 16, 16, None, 1, 16, None, 1, 16, 4, None, 16, None, 1, None, None, 1, 0,
                            None) == 1
    assert b"compPlan" in L.rbc_last_error()
    assert L.rbc_setup_reps(None, 10, 4, 4, None, 2, None) == 1
