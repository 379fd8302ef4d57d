"""Writes the golden LOF vectors in this directory.

Each ``*.npz`` holds the float32 input ``X``, ``k``, and scikit-learn
LocalOutlierFactor's FP64 answer on ``X`` cast to float64: ``lof`` (positive
LOF, i.e. -negative_outlier_factor_) plus the (k+1)-neighbour distances and
indices from NearestNeighbors, used to locate near-ties. Generated with
scikit-learn 1.7.2, numpy 2.2; rerun this script to regenerate.

The shapes follow the reference's own test cases (tests/cpp/test_culof.cu:
102-149, tests/python/test_parity.py:167-188) plus a small copy of the plain
benchmark workload.
"""

from __future__ import annotations

import os
import sys

import numpy as np
from sklearn.neighbors import LocalOutlierFactor, NearestNeighbors

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from culof_amd import workloads  # noqa: E402


def gauss(n, d, seed, n_out):
    """N(0, 1) rows with the first n_out rows uniform in [-15, 15]^d."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d)).astype(np.float32)
    if n_out:
        X[:n_out] = rng.uniform(-15, 15, size=(n_out, d)).astype(np.float32)
    return X


def duplicates():
    """300 points on 15 distinct positions, five of them pushed away."""
    n, d = 300, 2
    X = np.zeros((n, d), np.float32)
    base = (np.arange(n) % 15).astype(np.float32)
    X[:, 0] = base
    X[:, 1] = base * 2.0
    X[:5, 0] += 40.0 * np.arange(1, 6, dtype=np.float32)
    return X


CASES = {
    "kat_docstring": (np.array([[-1.1], [0.2], [101.1], [0.3]], np.float32), 2),
    "gauss_n400_d3_k20": (gauss(400, 3, 11, 8), 20),
    "gauss_n512_d32_k15": (gauss(512, 32, 3, 10), 15),
    "gauss_n600_d4_k64": (gauss(600, 4, 5, 12), 64),
    "gauss_n128_d2_k127": (gauss(128, 2, 7, 4), 127),
    "gauss_n1000_d3_k1": (gauss(1000, 3, 1004, 10), 1),
    "dups_n300_d2_k10": (duplicates(), 10),
    "clusters_n2000_d8_k20": (workloads.clusters(n=2000, d=8, seed=7), 20),
    "lowrank_n1500_d128_k50": (workloads.lowrank(n=1500, d=128, seed=3), 50),
}


def main() -> None:
    for name, (X, k) in CASES.items():
        X64 = X.astype(np.float64)
        lof = -LocalOutlierFactor(n_neighbors=k, algorithm="brute").fit(X64).negative_outlier_factor_
        kk = min(k + 2, len(X))
        dist, idx = NearestNeighbors(n_neighbors=kk, algorithm="brute").fit(X64).kneighbors(X64)
        np.savez_compressed(
            os.path.join(HERE, f"{name}.npz"), X=X, k=np.int64(k), lof=lof,
            nn_dist=dist, nn_idx=idx.astype(np.int32),
        )
        print(f"{name}: n={len(X)} d={X.shape[1]} k={k} max lof {lof.max():.4f}")


if __name__ == "__main__":
    main()
