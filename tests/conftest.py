"""Shared fixtures for the culof_amd test-suite.

Markers: ``gpu`` tests need an MI355X and call the HIP path through the C ABI
or the Python package; everything else runs on the CPU (oracle, host logic,
library loading and exports).
"""

from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU; runs the HIP kernels")


def pytest_collection_modifyitems(config, items):
    if any("gpu" in item.keywords for item in items):
        try:
            import culof_amd

            have_gpu = culof_amd.cuda_available()
        except ImportError:
            have_gpu = False
        if not have_gpu:
            skip = pytest.mark.skip(reason="no GPU in this environment")
            for item in items:
                if "gpu" in item.keywords:
                    item.add_marker(skip)


def load_golden(name: str) -> dict:
    with np.load(os.path.join(GOLDEN, f"{name}.npz")) as f:
        return {key: f[key] for key in f.files}


GOLDEN_CASES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def rel_err(got, want) -> np.ndarray:
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    return np.abs(got - want) / np.maximum(np.abs(want), 1e-9)


def make_data(n: int, d: int, seed: int, n_outliers: int = 0) -> np.ndarray:
    """N(0, 1) rows with the first ``n_outliers`` rows uniform in [-15, 15]^d
    (the reference's test generator, tests/python/conftest.py:19-30)."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d)).astype(np.float32)
    if n_outliers:
        X[:n_outliers] = rng.uniform(-15, 15, size=(n_outliers, d)).astype(np.float32)
    return X


def make_tied(n: int, d: int, seed: int, n_clusters: int = 6, grid: float = 0.05,
              dup_frac: float = 0.02) -> np.ndarray:
    """Clustered rows quantised to a grid (in float64, then cast) plus exact
    duplicate rows: many equal and nearly equal FP32 distances, as in the
    500k x 64 ties workload, at a size the oracles finish quickly."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-3.0, 3.0, size=(n_clusters, d))
    X = centres[rng.integers(0, n_clusters, size=n)] + rng.standard_normal((n, d))
    X = np.round(X / grid) * grid
    n_dup = int(n * dup_frac)
    X[rng.choice(n, n_dup, replace=False)] = X[rng.integers(0, n, size=n_dup)]
    return np.ascontiguousarray(X.astype(np.float32))


@pytest.fixture
def data() -> np.ndarray:
    """1,000 points in 3-D with 25 planted outliers in the first rows."""
    return make_data(1000, 3, seed=42, n_outliers=25)
