"""Seeded synthetic inputs for the five benchmark workloads.

Each generator returns C-contiguous float32 rows in a random order (rows are
shuffled, so no workload is stored in cluster order). Sizes default to the
benchmark sizes and can be reduced for tests.

* ``clusters``   5 isotropic Gaussian clusters (sigma 1, centres U[-10, 10]^d,
                 equal sizes) plus 1% uniform outliers in [-15, 15]^d.
                 n=20,000 / 200,000 with d=8, k=20 are the plain workloads.
* ``lowrank``    16 clusters, each x = c + B (s * z) + 0.1 eps with B a random
                 orthonormal d x 32 basis, s ~ U[1, 5]^32, centres N(0, 4^2);
                 1% outliers N(0, 8^2). n=100,000, d=128, k=50.
* ``grad``       8 Gaussian clusters (sigma 1, centres U[-10, 10]^d) plus 1%
                 uniform outliers in [-15, 15]^d, with upstream weights
                 w ~ N(0, 1) per point. n=50,000, d=32, k=20.
* ``ties``       32 Gaussian clusters (sigma 1, centres U[-20, 20]^d), all
                 coordinates rounded to a 0.05 grid, 0.5% exact duplicate rows
                 and 1% uniform outliers in [-30, 30]^d. n=500,000, d=64, k=32.
"""

from __future__ import annotations

import numpy as np


def _split(total: int, parts: int) -> list[int]:
    base, extra = divmod(total, parts)
    return [base + (1 if i < extra else 0) for i in range(parts)]


def clusters(n: int = 200_000, d: int = 8, n_clusters: int = 5, spread: float = 10.0,
             outlier_box: float = 15.0, outlier_frac: float = 0.01, seed: int = 0,
             return_outliers: bool = False):
    """X, or (X, mask of the planted uniform outlier rows) with ``return_outliers``."""
    rng = np.random.default_rng(seed)
    n_out = int(round(n * outlier_frac))
    centres = rng.uniform(-spread, spread, size=(n_clusters, d))
    blocks = [centres[c] + rng.standard_normal((m, d)) for c, m in enumerate(_split(n - n_out, n_clusters))]
    blocks.append(rng.uniform(-outlier_box, outlier_box, size=(n_out, d)))
    X = np.concatenate(blocks).astype(np.float32)
    perm = rng.permutation(n)
    X = np.ascontiguousarray(X[perm])
    if return_outliers:
        return X, perm >= n - n_out
    return X


def lowrank(n: int = 100_000, d: int = 128, rank: int = 32, n_clusters: int = 16,
            seed: int = 0, return_outliers: bool = False):
    """X, or (X, mask of the planted N(0, 8^2) outlier rows) with ``return_outliers``."""
    rng = np.random.default_rng(seed)
    n_out = int(round(n * 0.01))
    blocks = []
    for m in _split(n - n_out, n_clusters):
        basis, _ = np.linalg.qr(rng.standard_normal((d, rank)))
        s = rng.uniform(1.0, 5.0, size=rank)
        c = rng.normal(0.0, 4.0, size=d)
        z = rng.standard_normal((m, rank))
        blocks.append(c + (z * s) @ basis.T + 0.1 * rng.standard_normal((m, d)))
    blocks.append(rng.normal(0.0, 8.0, size=(n_out, d)))
    X = np.concatenate(blocks).astype(np.float32)
    perm = rng.permutation(n)
    X = np.ascontiguousarray(X[perm])
    if return_outliers:
        return X, perm >= n - n_out
    return X


def grad(n: int = 50_000, d: int = 32, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    X = clusters(n=n, d=d, n_clusters=8, spread=10.0, outlier_box=15.0, seed=seed)
    w = np.random.default_rng(seed + 1).standard_normal(n).astype(np.float32)
    return X, w


def ties(n: int = 500_000, d: int = 64, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n_out = int(round(n * 0.01))
    n_dup = int(round(n * 0.005))
    n_in = n - n_out - n_dup
    centres = rng.uniform(-20.0, 20.0, size=(32, d))
    inl = np.concatenate([centres[c] + rng.standard_normal((m, d)) for c, m in enumerate(_split(n_in, 32))])
    out = rng.uniform(-30.0, 30.0, size=(n_out, d))
    X = np.concatenate([inl, out])
    X = np.round(X / 0.05) * 0.05  # quantise in float64, then cast
    dup = X[rng.integers(0, n_in, size=n_dup)]
    X = np.concatenate([X, dup]).astype(np.float32)
    return np.ascontiguousarray(X[rng.permutation(n)])


WORKLOADS = {
    "plain-20k": dict(gen="clusters", n=20_000, d=8, k=20),
    "plain-200k": dict(gen="clusters", n=200_000, d=8, k=20),
    "bf16-100k": dict(gen="lowrank", n=100_000, d=128, k=50),
    "grad-50k": dict(gen="grad", n=50_000, d=32, k=20),
    "ties-500k": dict(gen="ties", n=500_000, d=64, k=32),
    # shapes past the list kernels (k > 64, d > 128)
    "k128-200k": dict(gen="clusters", n=200_000, d=8, k=128),
    "d256-200k": dict(gen="lowrank", n=200_000, d=256, k=20),
}


def make(name: str, seed: int = 0):
    """Input of a named workload: X, or (X, w) for the gradient workload."""
    spec = WORKLOADS[name]
    gen = spec["gen"]
    if gen == "clusters":
        return clusters(n=spec["n"], d=spec["d"], seed=seed)
    if gen == "lowrank":
        return lowrank(n=spec["n"], d=spec["d"], seed=seed)
    if gen == "grad":
        return grad(n=spec["n"], d=spec["d"], seed=seed)
    return ties(n=spec["n"], d=spec["d"], seed=seed)
