"""k beyond the list kernels (64 < k <= 256) on the matrix-core collect pipeline.

The reference's selection is independent of k (include/culof.h:39-41,
src/cuda/knn.cu:119-232 serve any k). Here the candidate order, the seeds,
the collect passes and the selects serve k up to 256; rows the list pass
would take (heavy rows, second-pass overflow) get their exact keys stored and
the stored radix select, and every row of a run without the candidate order
takes the exact radix select. Every comparison is bit for
bit against the oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import make_data, make_tied

pytestmark = pytest.mark.gpu

TIES = 256


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _profiled(fn):
    """fn() with the kernel timers on; returns (result, {scope: launches})."""
    from culof_amd import _native

    _native.profile_reset()
    _native.profile_enable(True)
    try:
        out = fn()
    finally:
        _native.profile_enable(False)
    scopes = ("mfma_collect", "collect_select", "big_select", "radix_knn", "mfma_sweep", "brute_rows")
    return out, {s: _native.profile_read(s)[1] for s in scopes}


@pytest.mark.parametrize("form", [0, 1], ids=["fp32", "bf16"])
@pytest.mark.parametrize("n,d,k,seed", [(1500, 8, 65, 1), (2000, 4, 128, 2), (1800, 32, 100, 3),
                                        (1300, 64, 200, 4), (1600, 128, 256, 5), (900, 3, 256, 6)])
def test_forward_bit_exact_large_k(n, d, k, seed, form):
    from _gpu import forward

    X = make_data(n, d, seed, n // 100)
    got, launches = _profiled(lambda: forward(X, k, form=form))
    assert launches["mfma_collect"] > 0 and launches["mfma_sweep"] == 0
    want = oracle.lof(X, k, full=True)
    np.testing.assert_array_equal(got["idx"].T, want["idx"])
    for key, name in (("d2", "d2"), ("kdist2", "kdist2"), ("lrd", "lrd"), ("scores", "scores")):
        g = got[key].T if key == "d2" else got[key]
        np.testing.assert_array_equal(_bits(g), _bits(want[name]), err_msg=key)


@pytest.mark.parametrize("normalize", [False, True])
def test_large_k_every_path_matches_oracle(normalize):
    """Collect pipeline (both screens), CULOF_ORDER=0 (no order: radix for
    k > 64) and the radix path all return the oracle's bits."""
    from _gpu import knn

    X = make_data(2500, 16, 11, 25)
    k = 150
    oi, od, ok = oracle.knn_rows(oracle.key_rows_of(X, normalize), k)
    runs = {"mfma-fp32": knn(X, k, normalize=normalize, form=0),
            "mfma-bf16": knn(X, k, normalize=normalize, form=1),
            "radix": knn(X, k, normalize=normalize, method=1)}
    for label, got in runs.items():
        np.testing.assert_array_equal(got[0].T, oi, err_msg=f"{label} nbr_idx")
        np.testing.assert_array_equal(_bits(got[1].T), _bits(od), err_msg=f"{label} nbr_d2")
        np.testing.assert_array_equal(_bits(got[2]), _bits(ok), err_msg=f"{label} kdist2")


def test_large_k_without_order_takes_radix(monkeypatch):
    from _gpu import knn

    X = make_data(1200, 8, 12, 12)
    k = 96
    monkeypatch.setenv("CULOF_ORDER", "0")
    got, launches = _profiled(lambda: knn(X, k))
    assert launches["radix_knn"] > 0 and launches["mfma_sweep"] == 0
    oi, od, ok = oracle.knn_rows(X, k)
    np.testing.assert_array_equal(got[0].T, oi)
    np.testing.assert_array_equal(_bits(got[2]), _bits(ok))


@pytest.mark.parametrize("n,d,k,form", [(40_000, 8, 128, 0), (36_000, 32, 256, 1), (34_000, 64, 100, 0)])
def test_symmetric_pass_large_k(monkeypatch, n, d, k, form):
    """Full runs past 32,768 rows take the symmetric pass; CULOF_SYM=0 gives
    the same bits, and sampled rows (far outliers among them) match the oracle."""
    from _gpu import knn

    rng = np.random.default_rng(n + k)
    X = make_data(n, d, n % 89, max(1, n // 200))
    far = rng.choice(n, max(1, n // 100), replace=False)
    X[far] = X[far] * 40.0 + 25.0
    sym = knn(X, k, form=form)
    monkeypatch.setenv("CULOF_SYM", "0")
    plain = knn(X, k, form=form)
    for a, b, name in zip(sym, plain, ("idx", "d2", "kdist2")):
        np.testing.assert_array_equal(a, b, err_msg=name)
    for r in np.r_[far[:3], rng.choice(n, 5, replace=False)]:
        oi, od, ok = oracle.knn_rows(X, k, int(r), 1)
        np.testing.assert_array_equal(sym[0][:, r], oi[0], err_msg=f"row {r}")
        np.testing.assert_array_equal(_bits(sym[1][:, r]), _bits(od[0]), err_msg=f"row {r} d2")
        np.testing.assert_array_equal(_bits(sym[2][r]), _bits(ok[0]), err_msg=f"row {r} kdist2")


def test_heavy_rows_large_k_take_stored_keys():
    """At n > 65,536 wide outliers are heavy rows; with k > 64 they get stored
    exact keys and the stored radix select instead of the list pass. Those
    rows and sampled inliers match the oracle."""
    from _gpu import knn

    rng = np.random.default_rng(5)
    n, d, k = 70_000, 8, 128
    X = make_data(n, d, 31, 0)
    far = rng.choice(n, 300, replace=False)
    X[far] = rng.uniform(-60, 60, (far.size, d)).astype(np.float32)
    got, launches = _profiled(lambda: knn(X, k))
    assert launches["brute_rows"] > 0 and launches["mfma_sweep"] == 0
    for r in np.r_[far[:6], rng.choice(n, 6, replace=False)]:
        oi, od, ok = oracle.knn_rows(X, k, int(r), 1)
        np.testing.assert_array_equal(got[0][:, r], oi[0], err_msg=f"row {r}")
        np.testing.assert_array_equal(_bits(got[2][r]), _bits(ok[0]), err_msg=f"row {r} kdist2")


@pytest.mark.parametrize("form", [0, 1], ids=["fp32", "bf16"])
def test_large_k_ties_bit_exact(form):
    """The FP64 tie-break at k = 100 (the forward selects k + 1 = 101)."""
    from _gpu import knn

    X = make_tied(1500, 8, 7)
    k = 100
    oi, od, ok, tie = oracle.knn_ties(X, k)
    assert tie.sum() > 0
    gi, gd, gk = knn(X, k, form=form | TIES)
    np.testing.assert_array_equal(gi.T, oi)
    np.testing.assert_array_equal(_bits(gd.T), _bits(od))
    np.testing.assert_array_equal(_bits(gk), _bits(ok))


def test_large_k_at_scale():
    """n = 200,000, d = 8, k = 128 (the clustered generator): the collect
    pipeline serves it; 300 sampled rows (outliers among them) are bit-exact
    with the oracle, and a repeat gives the same bits."""
    from _gpu import knn
    from culof_amd import workloads

    X, outlier = workloads.clusters(n=200_000, d=8, seed=3, return_outliers=True)
    k = 128
    got, launches = _profiled(lambda: knn(X, k))
    assert launches["mfma_collect"] > 0 and launches["mfma_sweep"] == 0
    again = knn(X, k)
    for a, b in zip(got, again):
        np.testing.assert_array_equal(a, b)
    rng = np.random.default_rng(0)
    rows = np.r_[rng.choice(np.flatnonzero(outlier), 100, replace=False),
                 rng.choice(X.shape[0], 200, replace=False)]
    for r in rows:
        oi, od, ok = oracle.knn_rows(X, k, int(r), 1)
        np.testing.assert_array_equal(got[0][:, r], oi[0], err_msg=f"row {r}")
        np.testing.assert_array_equal(_bits(got[1][:, r]), _bits(od[0]), err_msg=f"row {r} d2")
        np.testing.assert_array_equal(_bits(got[2][r]), _bits(ok[0]), err_msg=f"row {r} kdist2")
