"""Native boundary tests: a C++ program that includes the reference's headers
(rbc_include.h) and links librbc_hip calls the whole reference surface (rbc.h:11-26,
brute.h:10-13, utils.h:10-28, utilsGPU.h:12-17); AddressSanitizer + UndefinedBehavior
Sanitizer builds of the oracle and of the library's host code run their self-checks."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
LIB = os.path.join(ROOT, "randomballcover-gpu_amd", "librbc_hip.so")
EXE = os.path.join(NATIVE, "_build", "link_surface")


def _run(args, timeout):
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, f"{args} exited {r.returncode}:\n{r.stdout[-2000:]}\n{r.stderr[-4000:]}"
    return r.stdout


@pytest.fixture(scope="module")
def link_exe():
    if not os.path.exists(LIB):
        pytest.skip("librbc_hip.so is not built")
    _run(["make", "-s", "-C", NATIVE], 300)
    return EXE


def test_reference_surface_links_and_host_helpers_hold(link_exe):
    out = _run([link_exe], 60)
    assert out.strip().endswith("ok")


def test_sanitized_oracle_and_host_code():
    if not os.path.exists(LIB):
        pytest.skip("librbc_hip.so is not built")
    out = _run(["make", "-s", "-j8", "-C", NATIVE, "sanitize"], 900)
    assert out.count("ok") >= 2


@pytest.mark.gpu
def test_reference_internals_on_gpu():
    # kqueryRBC / queryRBC assembled from computeReps, computeCounts, buildQMap,
    # idIntersection, initCompPlan, computeKNNs / computeNNs equal the one-call results;
    # distSubMat matches distVec; setupReps copies distinct rows.
    assert os.path.exists(EXE), "tests/native/_build/link_surface is not built"
    print("link_surface gpu", flush=True)
    out = _run([EXE, "gpu"], 120)
    assert out.strip().endswith("ok")
