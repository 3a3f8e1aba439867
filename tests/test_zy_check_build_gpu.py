"""The device bounds-checking build (librbc_hip_check.so, -DRBC_CHECK): the main entry
points run through it without a check firing and with the oracle's results, and the
checker catches a planted out-of-range index (clamped, so nothing is read out of bounds).
One child process, since the library is chosen when it is first loaded."""
import os
import subprocess
import sys

import pytest

from conftest import PKG, ROOT

CHECK_LIB = os.path.join(PKG, "librbc_hip_check.so")


@pytest.mark.gpu
def test_entry_points_pass_their_bounds_checks():
    assert os.path.exists(CHECK_LIB), "librbc_hip_check.so is not built"
    env = dict(os.environ, RBC_LIB=CHECK_LIB)
    print("check build: tests/check_build_run.py", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "tests", "check_build_run.py")], env=env,
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "check build ok" in r.stdout


def test_check_build_exports_what_the_library_exports():
    import ctypes

    from conftest import rbc
    if not os.path.exists(CHECK_LIB):
        pytest.skip("librbc_hip_check.so is not built")
    main, chk = rbc.lib(), ctypes.CDLL(CHECK_LIB)
    from test_abi import declared_c_symbols
    missing = [n for n in declared_c_symbols() if not hasattr(chk, n)]
    assert not missing
    chk.rbc_check_enabled.restype = ctypes.c_int
    assert chk.rbc_check_enabled() == 1 and main.rbc_check_enabled() == 0
    # The plain build refuses an out-of-range self-test before any launch.
    v = ctypes.c_uint32()
    assert main.rbc_check_selftest(8, 8, ctypes.byref(v)) == 1
