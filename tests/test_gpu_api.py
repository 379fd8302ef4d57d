"""The Python surface on the GPU, against scikit-learn.

The behaviours the reference asserts through its package
(tests/python/test_api.py, test_parity.py, test_regressions.py), checked here
through ``culof`` -- the alias package a reference user imports -- on the HIP
path. Tolerances are tighter than the reference's 2e-4 where the difference-form
FP32 keys allow it; near ties are the only expected disagreement.
"""

from __future__ import annotations

import os

import numpy as np
import pytest
from sklearn.neighbors import LocalOutlierFactor, NearestNeighbors

import culof
from culof import LOF, lof

pytestmark = pytest.mark.gpu

REL_TOL = 2e-5


def points(n: int, d: int, seed: int, n_outliers: int = 0) -> np.ndarray:
    """Unit Gaussian float32 rows, the first ``n_outliers`` spread over [-15, 15]^d."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d)).astype(np.float32)
    if n_outliers:
        X[:n_outliers] = rng.uniform(-15, 15, size=(n_outliers, d)).astype(np.float32)
    return X


def sk(X, k):
    """scikit-learn's LOF fitted on float64 input (its FP64 path)."""
    return -LocalOutlierFactor(n_neighbors=k).fit(np.asarray(X, np.float64)).negative_outlier_factor_


def max_rel(got, want):
    return float(np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-9)))


def test_device_is_reported():
    assert culof.cuda_available()
    info = culof.device_info()
    assert "gfx950" in info and "CUs" in info


def test_scores_separate_planted_outliers():
    X = points(2000, 5, seed=1, n_outliers=20)
    s = lof(X, 20)
    assert s.dtype == np.float32 and s.shape == (2000,)
    assert set(np.argsort(-s)[:20]) == set(range(20))
    assert np.median(s[20:]) == pytest.approx(1.0, abs=0.1)


def test_estimator_attributes_and_labels():
    X = points(1500, 4, seed=2, n_outliers=15)
    m = LOF(n_neighbors=20, contamination=0.01)
    labels = m.fit_predict(X)
    assert labels.shape == (1500,) and set(np.unique(labels)) <= {-1, 1}
    np.testing.assert_array_equal(m.negative_outlier_factor_, -lof(X, 20))
    np.testing.assert_array_equal(m.score_samples(), m.negative_outlier_factor_)
    # predict flags exactly the scores below offset_, the contamination percentile
    np.testing.assert_array_equal(labels == -1, m.negative_outlier_factor_ < m.offset_)
    assert m.offset_ == pytest.approx(float(np.percentile(m.negative_outlier_factor_, 1.0)))
    assert m.n_neighbors_ == 20


@pytest.mark.parametrize("n,d,k", [(500, 2, 20), (2000, 8, 20), (3000, 16, 10), (1500, 64, 35),
                                   (1000, 3, 1), (1000, 3, 100)])
def test_values_match_sklearn(n, d, k):
    X = points(n, d, seed=n + d + k, n_outliers=max(1, n // 100))
    assert max_rel(lof(X, k), sk(X, k)) < REL_TOL


@pytest.mark.parametrize("k", [33, 64, 128, 256])
def test_large_k_match_sklearn(k):
    X = points(1000, 4, seed=k)
    assert max_rel(lof(X, k), sk(X, k)) < REL_TOL


def test_drop_in_for_sklearn_estimator():
    X = points(3000, 6, seed=7, n_outliers=30)
    theirs = LocalOutlierFactor(n_neighbors=20, contamination=0.01)
    ours = LOF(n_neighbors=20, contamination=0.01)
    np.testing.assert_array_equal(ours.fit_predict(X), theirs.fit_predict(X))
    assert max_rel(-ours.negative_outlier_factor_, -theirs.negative_outlier_factor_) < REL_TOL
    assert ours.offset_ == pytest.approx(theirs.offset_, rel=1e-5)
    assert LOF().fit(X).offset_ == LocalOutlierFactor().fit(X).offset_ == -1.5


def test_ranking_and_near_ties():
    """The bulk agrees tightly; any point above 1e-4 sits on a near tie (its
    own or a neighbour's k-th / (k+1)-th FP64 distances within 1e-3 relative);
    the top 1 % is the same set."""
    k, n = 20, 20_000
    X = points(n, 8, seed=99, n_outliers=n // 100)
    got, want = lof(X, k), sk(X, k)
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-9)
    assert np.median(rel) < 1e-6 and np.percentile(rel, 99) < 5e-6
    bad = np.flatnonzero(rel > 1e-4)
    assert len(bad) < n // 1000
    if len(bad):
        dist, idx = NearestNeighbors(n_neighbors=k + 2).fit(X).kneighbors(X)
        gap = (dist[:, k + 1] - dist[:, k]) / np.maximum(dist[:, k], 1e-12)
        near = gap < 1e-3
        near_any = near | near[idx[:, 1:k + 1]].any(axis=1)
        assert near_any[bad].all(), "disagreement away from a near tie"
    assert set(np.argsort(-got)[: n // 100]) == set(np.argsort(-want)[: n // 100])
    assert np.corrcoef(got, want)[0, 1] > 0.999999


def test_input_layouts_give_the_same_bits():
    X = points(1200, 6, seed=3, n_outliers=12)
    base = lof(X, 15)
    wide = np.repeat(X, 2, axis=1)                      # a strided view of the same values
    np.testing.assert_array_equal(lof(wide[:, ::2], 15).view(np.uint32), base.view(np.uint32))
    np.testing.assert_array_equal(lof(np.asfortranarray(X), 15).view(np.uint32), base.view(np.uint32))
    np.testing.assert_array_equal(lof(X.astype(np.float64), 15).view(np.uint32), base.view(np.uint32))
    np.testing.assert_array_equal(lof(X.tolist(), 15).view(np.uint32), base.view(np.uint32))
    np.testing.assert_array_equal(lof(X, 15).view(np.uint32), base.view(np.uint32))  # repeatable


def test_tile_rows_do_not_change_results(monkeypatch):
    X = points(3000, 8, seed=4, n_outliers=30)
    base = lof(X, 20)
    for rows in ("1", "700", "3000"):
        monkeypatch.setenv("CULOF_TILE_ROWS", rows)
        np.testing.assert_array_equal(lof(X, 20).view(np.uint32), base.view(np.uint32))
    monkeypatch.delenv("CULOF_TILE_ROWS")


@pytest.mark.filterwarnings("ignore:Duplicate values")  # scikit-learn's own note on planted copies
def test_duplicates_do_not_collapse():
    X = points(1000, 3, seed=5)
    X[500:530] = X[0]                                   # 31 copies of one point
    s = lof(X, 20)
    assert np.isfinite(s).all() and s.std() > 0
    assert max_rel(s, sk(X, 20)) < REL_TOL


def test_normalize_matches_zscore_and_keeps_constant_features():
    X = points(1500, 5, seed=6, n_outliers=15) * np.array([1, 10, 100, 0.1, 3], np.float32)
    Z = ((X - X.mean(0, dtype=np.float64)) / X.std(0, dtype=np.float64)).astype(np.float32)
    np.testing.assert_allclose(lof(X, 20, normalize=True), sk(Z, 20), rtol=1e-4)
    Xc = X.copy()
    Xc[:, 2] = 7.0                                      # a constant column
    s = lof(Xc, 20, normalize=True)
    assert np.isfinite(s).all()
    np.testing.assert_allclose(s, lof(np.delete(Xc, 2, axis=1), 20, normalize=True), rtol=1e-5)


def test_far_from_origin_is_stable():
    """Centring keeps a large common offset harmless; the comparison is on the
    same shifted float32 array, so its quantisation hits both sides."""
    X = points(1500, 3, seed=8, n_outliers=15)
    shifted = (X + np.float32(5000.0)).astype(np.float32)
    assert max_rel(lof(shifted, 20), sk(shifted, 20)) < REL_TOL


def test_overflow_raises_instead_of_non_finite():
    X = points(200, 4, seed=9)
    with pytest.raises(ValueError, match="overflow"):
        lof(X * np.float32(1e19), 10)
    s = lof(X * np.float32(1e6), 10)                    # large but representable
    assert np.isfinite(s).all()


def test_env_path_switches_give_the_same_bits(monkeypatch):
    X = points(2500, 8, seed=10, n_outliers=25)
    base = lof(X, 20)
    for var, val in (("CULOF_KNN_PATH", "valu"), ("CULOF_KNN_PATH", "radix"), ("CULOF_ORDER", "0")):
        monkeypatch.setenv(var, val)
        np.testing.assert_array_equal(lof(X, 20).view(np.uint32), base.view(np.uint32))
        monkeypatch.delenv(var)
    assert os.environ.get("CULOF_KNN_PATH") is None


@pytest.mark.parametrize("scale", [1e-6, 1.0, 1e3, 1e9])
def test_scale_regression_through_the_hip_path(scale):
    """The reference's overflow-guard regression (tests/python/test_regressions.py:140-146)
    on the HIP path: 64 x 4 points at scales 1e-6 ... 1e9 give finite scores,
    bit-exact with the oracle; and the same scales at n = 40,000, d = 8 (the
    symmetric matrix-core pass, whose screen margins scale with the norms)
    keep sampled rows bit-exact with the oracle in both screen forms."""
    import oracle
    from _gpu import forward

    from culof_amd import _native

    rng = np.random.default_rng(0)
    X = (rng.standard_normal((64, 4)) * scale).astype(np.float32)
    s = lof(X, 5)
    assert np.isfinite(s).all(), f"scale {scale} produced non-finite scores"
    np.testing.assert_array_equal(s.view(np.uint32), oracle.lof(X, 5).view(np.uint32))

    big = (points(40_000, 8, seed=12, n_outliers=400) * np.float32(scale)).astype(np.float32)
    print(f"scale {scale}: n=40000 d=8 k=20", flush=True)
    rows = [0, 1, 399, *np.random.default_rng(1).choice(40_000, 8, replace=False).tolist()]
    want = {r: oracle.knn_rows(big, 20, r, 1) for r in rows}
    for form in (_native.FORM_PLAIN, _native.FORM_BF16):
        got = forward(big, 20, form=form)
        assert np.isfinite(got["scores"]).all()
        for r in rows:
            oi, od, ok = want[r]
            np.testing.assert_array_equal(got["idx"][:, r], oi[0], err_msg=f"scale {scale} row {r}")
            assert got["kdist2"][r].view(np.uint32) == ok[0].view(np.uint32), f"scale {scale} row {r}"
