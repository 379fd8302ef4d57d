"""Rows wider than 128 features (128 < d <= 512) on the matrix cores.

The reference serves any d through its SGEMM Gram tile
(src/cuda/pipeline.cu:178-188). Here widths up to 512 are padded to 256 or
512 and take the candidate order (cells on the first 128 features), seeds
keyed over the whole row in 128-feature pieces, the K-chunked collect pass
(64 features per chunk, both screen forms) and the selects; rows the list
pass would take get stored exact keys and the stored radix select, and every
row of a run without the candidate order or wider than 512 takes the exact
radix select. Every comparison is bit for
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
    scopes = ("mfma_collect", "seed", "order", "radix_knn", "mfma_sweep", "sweep_knn", "brute_rows")
    return out, {s: _native.profile_read(s)[1] for s in scopes}


@pytest.mark.parametrize("form", [0, 1], ids=["fp32", "bf16"])
@pytest.mark.parametrize("n,d,k,seed", [(1500, 130, 20, 1), (1200, 200, 50, 2), (1000, 256, 10, 3),
                                        (900, 300, 15, 4), (700, 512, 20, 5), (1300, 256, 128, 6)])
def test_forward_bit_exact_wide(n, d, k, seed, form):
    from _gpu import forward

    X = make_data(n, d, seed, n // 100)
    got, launches = _profiled(lambda: forward(X, k, form=form))
    assert launches["mfma_collect"] > 0 and launches["mfma_sweep"] == 0 and launches["sweep_knn"] == 0
    want = oracle.lof(X, k, full=True)
    np.testing.assert_array_equal(got["idx"].T, want["idx"])
    for key in ("kdist2", "lrd", "scores"):
        np.testing.assert_array_equal(_bits(got[key]), _bits(want[key]), err_msg=key)
    np.testing.assert_array_equal(_bits(got["d2"].T), _bits(want["d2"]), err_msg="d2")


def test_beyond_512_takes_radix():
    from _gpu import forward

    X = make_data(500, 520, 7, 5)
    got, launches = _profiled(lambda: forward(X, 12))
    assert launches["mfma_collect"] == 0 and launches["radix_knn"] > 0
    want = oracle.lof(X, 12, full=True)
    np.testing.assert_array_equal(got["idx"].T, want["idx"])
    np.testing.assert_array_equal(_bits(got["scores"]), _bits(want["scores"]))


@pytest.mark.parametrize("normalize", [False, True])
def test_wide_every_path_matches_oracle(normalize):
    """Collect pipeline (both screens), CULOF_ORDER=0 (radix for wide rows)
    and the radix path all return the oracle's bits."""
    from _gpu import knn

    X = make_data(2000, 192, 11, 20) * np.float32(2.5) + np.float32(3.0)
    k = 24
    oi, od, ok = oracle.knn_rows(oracle.key_rows_of(X, normalize), k)
    runs = {"mfma-fp32": knn(X, k, normalize=normalize, form=0),
            "mfma-bf16": knn(X, k, normalize=normalize, form=1),
            "radix": knn(X, k, normalize=normalize, method=1)}
    for label, got in runs.items():
        np.testing.assert_array_equal(got[0].T, oi, err_msg=f"{label} nbr_idx")
        np.testing.assert_array_equal(_bits(got[1].T), _bits(od), err_msg=f"{label} nbr_d2")
        np.testing.assert_array_equal(_bits(got[2]), _bits(ok), err_msg=f"{label} kdist2")


def test_wide_switches_give_same_bits(monkeypatch):
    """CULOF_ORDER=0 (no order: radix) and CULOF_TILE_ROWS (row tiles) agree
    with the default run bit for bit."""
    from _gpu import knn

    X = make_data(3000, 256, 12, 30)
    base = knn(X, 20)
    monkeypatch.setenv("CULOF_TILE_ROWS", "700")
    tiled = knn(X, 20)
    monkeypatch.delenv("CULOF_TILE_ROWS")
    monkeypatch.setenv("CULOF_ORDER", "0")
    plain, launches = _profiled(lambda: knn(X, 20))
    assert launches["mfma_collect"] == 0 and launches["radix_knn"] > 0
    for other in (tiled, plain):
        for a, b in zip(base, other):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("form", [0, 1], ids=["fp32", "bf16"])
def test_wide_ties_bit_exact(form):
    """The FP64 tie-break on grid-quantised wide rows with duplicates."""
    from _gpu import knn

    X = make_tied(1500, 160, 8)
    k = 15
    oi, od, ok, tie = oracle.knn_ties(X, k)
    gi, gd, gk = knn(X, k, form=form | TIES)
    np.testing.assert_array_equal(gi.T, oi)
    np.testing.assert_array_equal(_bits(gd.T), _bits(od))
    np.testing.assert_array_equal(_bits(gk), _bits(ok))


def test_wide_overflowing_rows():
    """Wide uniform outliers collect thousands of candidates: they go through
    the second pass and the stored-key select and still match the oracle."""
    from _gpu import knn

    rng = np.random.default_rng(9)
    n, d, k = 6000, 256, 20
    X = make_data(n, d, 13, 0)
    far = rng.choice(n, 60, replace=False)
    X[far] = rng.uniform(-30, 30, (far.size, d)).astype(np.float32)
    got, launches = _profiled(lambda: knn(X, k))
    assert launches["brute_rows"] > 0
    for r in np.r_[far[:10], rng.choice(n, 10, replace=False)]:
        oi, od, ok = oracle.knn_rows(X, k, int(r), 1)
        np.testing.assert_array_equal(got[0][:, r], oi[0], err_msg=f"row {r}")
        np.testing.assert_array_equal(_bits(got[2][r]), _bits(ok[0]), err_msg=f"row {r} kdist2")


@pytest.mark.parametrize("form", [0, 1], ids=["fp32", "bf16"])
def test_wide_rows_at_scale(form):
    """n = 200,000, d = 256, k = 20 (16 low-rank clusters, 1 % outliers): the
    matrix-core pipeline serves it; both forms give the same bits and 200
    sampled rows (outliers among them) are bit-exact with the oracle."""
    from _gpu import knn
    from culof_amd import workloads

    X, outlier = workloads.lowrank(n=200_000, d=256, seed=4, return_outliers=True)
    k = 20
    got, launches = _profiled(lambda: knn(X, k, form=form))
    assert launches["mfma_collect"] > 0 and launches["mfma_sweep"] == 0
    rng = np.random.default_rng(1)
    rows = np.r_[rng.choice(np.flatnonzero(outlier), 40, replace=False),
                 rng.choice(X.shape[0], 160, replace=False)]
    for r in rows:
        oi, od, ok = oracle.knn_rows(X, k, int(r), 1)
        np.testing.assert_array_equal(got[0][:, r], oi[0], err_msg=f"row {r}")
        np.testing.assert_array_equal(_bits(got[1][:, r]), _bits(od[0]), err_msg=f"row {r} d2")
        np.testing.assert_array_equal(_bits(got[2][r]), _bits(ok[0]), err_msg=f"row {r} kdist2")
