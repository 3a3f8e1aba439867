"""Shared fixtures: loads the engine's ctypes binding and the oracle by path (the
package directory name has a hyphen, so it is not importable by name)."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "randomballcover-gpu_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_module(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


rbc = load_module("rbc_ctypes", os.path.join(PKG, "rbc_ctypes.py"))
datagen = load_module("rbc_datagen", os.path.join(PKG, "datagen.py"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def lib():
    return rbc.lib()
