"""Writes the golden fixtures under tests/golden/ from the FP64 oracle.

sample_shaped_fp64_knn.npz: a database of 1024 x 16 and 128 x 16 queries (the shapes of
the reference's sample_input/, readme.txt:75-84; the values are this project's own
seeded Gaussian mixture, not the reference's files) with the FP64 32-NN of every query,
ties to the lowest index.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

spec = importlib.util.spec_from_file_location("dg", os.path.join(ROOT, "randomballcover-gpu_amd", "datagen.py"))
dg = importlib.util.module_from_spec(spec)
spec.loader.exec_module(dg)


def main():
    x = dg.gmm(1024, 16, 0, components=8)
    q = dg.gmm(128, 16, 1, components=8)
    ids, dists = oracle.brute_k_f64(x, q, 32)
    ni, nd = oracle.knn_f64_numpy(x, q, 32)
    assert (ids == ni).all() and np.allclose(dists, nd, rtol=1e-13)
    np.savez_compressed(os.path.join(HERE, "sample_shaped_fp64_knn.npz"), x=x, q=q, ids=ids,
                        dists=dists)


if __name__ == "__main__":
    main()
