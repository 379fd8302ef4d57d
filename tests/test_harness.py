"""Host logic of the sweep benchmark and the accuracy report (CPU only).

The harness enforces agreement with scikit-learn the way the reference's does
(benchmarks/benchmark.py:94-112); these tests drive it with scikit-learn on
both sides, or with planted disagreements, so no GPU is involved.
"""

from __future__ import annotations

import json

import numpy as np
import pytest
from sklearn.neighbors import LocalOutlierFactor

from culof_amd import accuracy as acc
from culof_amd import benchmark as bm


def sk_scores(X, k):
    return -LocalOutlierFactor(n_neighbors=k).fit(X).negative_outlier_factor_


def test_blobs_with_outliers_layout():
    X = bm.blobs_with_outliers(1000, 3, seed=0)
    assert X.shape == (1000, 3) and X.dtype == np.float32 and X.flags["C_CONTIGUOUS"]
    np.testing.assert_array_equal(X, bm.blobs_with_outliers(1000, 3, seed=0))
    assert not np.array_equal(X, bm.blobs_with_outliers(1000, 3, seed=1))
    # 10 outliers first, drawn after the inliers from the same generator
    rng = np.random.default_rng(0)
    centres = rng.uniform(-8, 8, size=(4, 3))
    labels = rng.integers(0, 4, size=990)
    inl = (centres[labels] + rng.standard_normal((990, 3))).astype(np.float32)
    out = rng.uniform(-20, 20, size=(10, 3)).astype(np.float32)
    np.testing.assert_array_equal(X[:10], out)
    np.testing.assert_array_equal(X[10:], inl)
    assert bm.blobs_with_outliers(50, 2).shape == (50, 2)  # max(1, n // 100) outliers


def test_relative_errors_and_stats():
    want = np.linspace(1.0, 3.0, 1000)
    eps = np.zeros(1000)
    eps[::100] = 1e-3          # 10 points at 1e-3
    eps[5] = 2e-2              # one at 2e-2
    got = want * (1 + eps)
    rel = acc.relative_errors(got, want)
    np.testing.assert_allclose(rel, eps, rtol=1e-9, atol=1e-15)
    s = acc.score_stats(got, want, threshold=1e-4)
    assert s["n"] == 1000
    assert s["median_rel"] == 0.0
    assert s["max_rel"] == pytest.approx(2e-2)
    assert s["p99_rel"] == pytest.approx(np.percentile(eps, 99))
    assert s["frac_above"] == pytest.approx(11 / 1000)
    assert 0.99 < s["pearson_r"] <= 1.0
    # the near-zero guard: |want| below 1e-9 divides by 1e-9
    assert acc.relative_errors(np.array([1e-9]), np.array([0.0]))[0] == pytest.approx(1.0)


def test_top_agreement():
    want = np.arange(1000, dtype=np.float64)
    assert acc.top_agreement(want, want) == 1.0
    got = want.copy()
    got[999], got[0] = got[0], got[999]   # the top point falls to the bottom
    assert acc.top_agreement(got, want) == pytest.approx(0.9)
    assert acc.top_agreement(want[:50], want[:50]) == 1.0  # top = max(1, n // 100)


def test_check_agreement_enforces_limits():
    X = bm.blobs_with_outliers(600, 4)
    want = sk_scores(X, 10)
    worst, agree = bm.check_agreement(want * (1 + 1e-6), want, 600, 4, 10)
    assert worst == pytest.approx(1e-6, rel=0.05) and agree == 1.0  # float32 scores
    with pytest.raises(bm.DisagreementError, match="max relative error"):
        bm.check_agreement(want * 1.06, want, 600, 4, 10)
    swapped = want.copy()
    top = int(np.argmax(want))
    swapped[top] = np.median(want)
    with pytest.raises(bm.DisagreementError, match="rank agreement"):
        bm.check_agreement(swapped, want, 600, 4, 10)
    bad = want.copy()
    bad[3] = np.nan
    with pytest.raises(bm.DisagreementError):
        bm.check_agreement(bad, want, 600, 4, 10)


def test_bench_one_and_sweeps_with_stand_in_scorers():
    ref = lambda X: sk_scores(X, 10)  # noqa: E731
    row = bm.bench_one("samples", 400, 3, 10, repeats=1, scorer=ref, reference=ref)
    assert (row.sweep, row.n_samples, row.n_features, row.k) == ("samples", 400, 3, 10)
    assert row.max_rel_err == 0.0 and row.rank_agreement == 1.0
    assert row.points_per_s > 0 and row.speedup > 0
    # a scorer that inverts the ranking is refused, not reported
    with pytest.raises(bm.DisagreementError):
        bm.bench_one("samples", 400, 3, 10, repeats=1, scorer=lambda X: 1.0 / ref(X), reference=ref)
    rows = bm.run_sweeps(k=10, repeats=1, max_n=500, sizes=(200, 500, 1000), dims=(2, 5),
                         n_dims=300, scorer=ref, reference=ref)
    assert [(r.sweep, r.n_samples, r.n_features) for r in rows] == [
        ("samples", 200, 8), ("samples", 500, 8), ("features", 300, 2), ("features", 300, 5)]
    doc = json.loads(json.dumps(bm.payload(rows, {"k": 10})))
    assert doc["k"] == 10 and len(doc["rows"]) == 4
    assert set(doc["rows"][0]) >= {"sweep", "n_samples", "n_features", "k", "sklearn_s", "culof_s",
                                   "speedup", "max_rel_err", "rank_agreement", "points_per_s"}
    assert bm.format_row(rows[0]).split()[0] == "200"


def test_near_tie_rows_and_spread():
    d2 = np.array([[1.0, 2.0, 2.0, 5.0],            # k=2: 2nd == 3rd -> tie
                   [1.0, 2.0, 2.0 * (1 + 2.0 ** -22), 5.0],  # inside the tolerance
                   [1.0, 2.0, 2.0 * (1 + 2.0 ** -18), 5.0],  # outside
                   [1.0, 2.0, 3.0, 5.0]])
    np.testing.assert_array_equal(acc.near_tie_rows(d2, 2), [True, True, False, False])
    # chain 0 -> 1 -> 2 -> 3 -> 4 -> 0 (row i lists i + 1): a tie at row 3
    # reaches rows 2 (one hop) and 1 (two hops) and no further
    nbrs = np.array([[1], [2], [3], [4], [0]])
    flags = np.array([False, False, False, True, False])
    np.testing.assert_array_equal(acc.spread(flags, nbrs, hops=2), [False, True, True, True, False])
    np.testing.assert_array_equal(acc.spread(flags, nbrs, hops=0), flags)


def test_fp64_neighbours_exclude_self_with_duplicates():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((200, 3)).astype(np.float32)
    X[100:110] = X[0]                                  # 11 copies of one point
    d2, idx = acc.fp64_neighbours(X, 5)
    assert d2.shape == (200, 6) and idx.shape == (200, 6)
    assert not (idx == np.arange(200)[:, None]).any()
    assert (np.diff(d2, axis=1) >= 0).all()
    assert (d2[0, :5] == 0).all()                      # its copies
    brute = ((X[:, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(brute, np.inf)
    np.testing.assert_allclose(d2, np.sort(brute, axis=1)[:, :6], rtol=1e-9, atol=1e-12)


def test_report_attributes_grid_ties():
    # An unjittered grid: every point has many exactly equidistant
    # neighbours, so every deviation is within two hops of a tie.
    g = np.arange(12, dtype=np.float32)
    X = np.stack(np.meshgrid(g, g), -1).reshape(-1, 2)
    want = sk_scores(X, 6)
    got = want.copy()
    got[[7, 50]] *= 1 + 1e-3
    rep = acc.report(X, 6, scores=got, reference=want)
    assert rep.n_above == 2 and rep.n_attributed == 2 and rep.unattributed == []
    assert rep.near_tie_rows > 0
    text = rep.text()
    assert "median" in text and "top-1% agreement" in text
    # generic positions: no near ties, so a planted deviation stays unattributed
    X2 = np.random.default_rng(5).standard_normal((300, 3)).astype(np.float32)
    want2 = sk_scores(X2, 6)
    got2 = want2.copy()
    got2[11] *= 1 + 1e-3
    rep2 = acc.report(X2, 6, scores=got2, reference=want2)
    assert rep2.near_tie_rows == 0 and rep2.unattributed == [11]
