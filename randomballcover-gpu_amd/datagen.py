"""Seeded synthetic data sets for the five workloads the engine is measured on.

Every generator is deterministic and chunked: point i is drawn from a PCG64 stream
seeded by (seed, tag, i // CHUNK), so the first n rows of a large set are the same
rows a smaller call returns.  All outputs are C-contiguous float32 arrays (rows =
points).  The workloads:

* ``uniform``   -- i.i.d. U[0,1) coordinates (database seed 0, queries seed 1).
* ``gmm``       -- equal-weight Gaussian mixture, centres U[-10,10]^d, sigma 1.
* ``lowrank``   -- x = A z + 0.01 e with z ~ N(0, I_16), A a fixed d x 16 matrix of
                  N(0, 1/16) entries (drawn from seed 12345), e ~ N(0, I_d).
* ``gmm_dirichlet`` -- 1,024-component mixture with centres N(0, 25 I),
                  per-component sigma U[0.5, 2] and Dirichlet(1) weights.
* ``batched``   -- 256 independent U[0,1) databases, n log-uniform in [2^10, 2^18],
                  d in {32, 96}, query batches log-uniform in [1, 4096].
"""
from __future__ import annotations

import numpy as np

CHUNK = 1 << 16


def _rng(seed: int, tag: int, chunk: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64([seed, tag, chunk]))


def _chunked(n: int, d: int, seed: int, tag: int, fill) -> np.ndarray:
    out = np.empty((n, d), dtype=np.float32)
    for c0 in range(0, n, CHUNK):
        c1 = min(n, c0 + CHUNK)
        fill(out[c0:c1], _rng(seed, tag, c0 // CHUNK), c0)
    return out


def uniform(n: int, d: int, seed: int) -> np.ndarray:
    def fill(o, rng, _):
        full = rng.random((CHUNK, d), dtype=np.float32)
        o[:] = full[: o.shape[0]]
    return _chunked(n, d, seed, 1, fill)


def gmm(n: int, d: int, seed: int, components: int = 64, center_seed: int = 777) -> np.ndarray:
    crng = _rng(center_seed, 2, 0)
    centers = crng.uniform(-10.0, 10.0, size=(components, d)).astype(np.float32)

    def fill(o, rng, _):
        lab = rng.integers(0, components, size=CHUNK)[: o.shape[0]]
        noise = rng.standard_normal((CHUNK, d), dtype=np.float32)[: o.shape[0]]
        o[:] = centers[lab] + noise
    return _chunked(n, d, seed, 3, fill)


def lowrank_matrix(d: int = 128, r: int = 16, seed: int = 12345) -> np.ndarray:
    return (_rng(seed, 4, 0).standard_normal((d, r)) / np.sqrt(r)).astype(np.float32)


def lowrank(n: int, d: int, seed: int, r: int = 16) -> np.ndarray:
    A = lowrank_matrix(d, r)

    def fill(o, rng, _):
        z = rng.standard_normal((CHUNK, r), dtype=np.float32)[: o.shape[0]]
        e = rng.standard_normal((CHUNK, d), dtype=np.float32)[: o.shape[0]]
        o[:] = z @ A.T + np.float32(0.01) * e
    return _chunked(n, d, seed, 5, fill)


def gmm_dirichlet_params(d: int = 96, components: int = 1024, seed: int = 999):
    rng = _rng(seed, 6, 0)
    centers = (5.0 * rng.standard_normal((components, d))).astype(np.float32)
    sigma = rng.uniform(0.5, 2.0, size=components).astype(np.float32)
    weights = rng.dirichlet(np.ones(components))
    return centers, sigma, weights


def gmm_dirichlet(n: int, d: int, seed: int, components: int = 1024) -> np.ndarray:
    centers, sigma, weights = gmm_dirichlet_params(d, components)
    cdf = np.cumsum(weights)
    cdf[-1] = 1.0

    def fill(o, rng, _):
        u = rng.random(CHUNK)[: o.shape[0]]
        lab = np.searchsorted(cdf, u, side="right")
        noise = rng.standard_normal((CHUNK, d), dtype=np.float32)[: o.shape[0]]
        o[:] = centers[lab] + sigma[lab, None] * noise
    return _chunked(n, d, seed, 7, fill)


def batched_shapes(instances: int = 256, seed: int = 4242):
    """(n_i, d_i, m_i) for the batched workload."""
    rng = _rng(seed, 8, 0)
    n = np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 18), size=instances)).astype(np.int64)
    d = rng.choice([32, 96], size=instances)
    m = np.exp(rng.uniform(0.0, np.log(4096), size=instances)).astype(np.int64)
    return n, d, np.maximum(m, 1)


WORKLOADS = {
    # name: (generator, d, n, m, numReps, k)
    "uniform_n65536_d16": (uniform, 16, 65536, 8192, 256, 1),
    "gmm64_n1M_d64": (gmm, 64, 1 << 20, 65536, 1024, 32),
    "lowrank16_n4M_d128": (lowrank, 128, 1 << 22, 131072, 2048, 10),
    "gmm1024_n4M_d96": (gmm_dirichlet, 96, 1 << 22, 262144, 2048, 10),
}


def workload(name: str, n: int | None = None, m: int | None = None):
    """Database (seed 0) and queries (seed 1) of a named workload, optionally truncated."""
    gen, d, n0, m0, reps, k = WORKLOADS[name]
    n = n0 if n is None else n
    m = m0 if m is None else m
    return gen(n, d, 0), gen(m, d, 1), reps, k


def training_queries(name: str, count: int):
    """Training queries of a named workload (seed 2): same model, disjoint from the
    database (seed 0) and the measured queries (seed 1)."""
    gen, d = WORKLOADS[name][:2]
    return gen(count, d, 2)
