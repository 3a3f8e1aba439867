"""testRBC argument handling (driver.cu:125-198): usage text, and the reference's checks
that end the program before any device work."""
import os
import subprocess

import pytest

from conftest import PKG

EXE = os.path.join(PKG, "testRBC")


def run(*args):
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=60)


def test_usage_without_arguments():
    r = run()
    assert r.returncode == 0
    assert "usage:" in r.stdout and "-x datafileX" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["-x", "a", "-q", "b", "-n", "100", "-m", "10", "-d", "4"], "more arguments needed"),
    (["-x", "a", "-q", "b", "-n", "100", "-m", "10", "-d", "4", "-r", "16"], "at least 32"),
    (["-x", "a", "-q", "b", "-n", "100", "-m", "10", "-d", "4", "-r", "200"], "more representatives than points"),
    (["-x", "a", "-X", "c", "-q", "b", "-n", "100", "-m", "10", "-d", "4", "-r", "32"], "only give one database"),
    (["-z"], "unrecognized option"),
])
def test_argument_errors(args, msg):
    r = run(*args)
    assert r.returncode == 1
    assert msg in r.stderr
