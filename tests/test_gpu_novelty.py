"""Novelty mode on the GPU: new points scored against a fitted set.

culof_lof_novelty_device against the oracle's restatement (bit-exact: the
query keys, the reference's tie rule and slot order, the FP64 density and
ratio sums) and against scikit-learn's LocalOutlierFactor(novelty=True).
"""

from __future__ import annotations

import numpy as np
import pytest
from sklearn.neighbors import LocalOutlierFactor

import oracle
from culof_amd import LOF, _native

pytestmark = pytest.mark.gpu


def data(n, m, d, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d)).astype(np.float32) * np.float32(2.0)
    Q = np.concatenate([rng.standard_normal((m - m // 10, d)) * 2.0,
                        rng.uniform(-15, 15, (m // 10, d))]).astype(np.float32)
    return X, Q


@pytest.mark.parametrize("n,m,d,k,normalize,form", [
    (3000, 500, 8, 20, False, _native.FORM_PLAIN),
    (3000, 500, 8, 20, True, _native.FORM_PLAIN),
    (8000, 700, 16, 25, False, _native.FORM_BF16),
    (2500, 300, 128, 50, False, _native.FORM_PLAIN),
    (1500, 200, 3, 1, False, _native.FORM_PLAIN),
    # more queries than fitted rows: every query block must still visit each
    # candidate chunk once (its start chunk is taken modulo the chunk count)
    (1000, 5000, 8, 20, False, _native.FORM_PLAIN),
    (600, 4100, 32, 10, True, _native.FORM_PLAIN),
])
def test_novelty_bit_exact_with_oracle(n, m, d, k, normalize, form):
    X, Q = data(n, m, d, seed=n + m + d)
    from _gpu import novelty

    print(f"novelty n={n} m={m} d={d} k={k} normalize={normalize} form={form}", flush=True)
    new, fit = novelty(X, Q, k, normalize, form)
    want = oracle.lof_novelty(X, Q, k, normalize=normalize)
    np.testing.assert_array_equal(new.view(np.uint32), want.view(np.uint32))
    np.testing.assert_array_equal(fit.view(np.uint32), oracle.lof(X, k, normalize=normalize).view(np.uint32))


def test_queries_equal_to_fitted_rows_and_no_queries():
    from _gpu import novelty

    X, _ = data(2000, 10, 5, seed=3)
    Q = np.concatenate([X[:50], X[:50] + np.float32(1e-3)])
    new, _ = novelty(X, Q, 10)
    np.testing.assert_array_equal(new.view(np.uint32), oracle.lof_novelty(X, Q, 10).view(np.uint32))
    empty, fit = novelty(X, np.zeros((0, 5), np.float32), 10)
    assert empty.shape == (0,) and np.isfinite(fit).all()


def test_estimator_novelty_matches_sklearn():
    X, Q = data(5000, 1000, 6, seed=9)
    ours = LOF(n_neighbors=20, contamination=0.05, novelty=True).fit(X)
    theirs = LocalOutlierFactor(n_neighbors=20, contamination=0.05, novelty=True).fit(X)
    got, want = ours.score_samples(Q), theirs.score_samples(Q)
    assert np.max(np.abs(got - want) / np.abs(want)) < 2e-5
    np.testing.assert_array_equal(got.view(np.uint32), (-oracle.lof_novelty(X, Q, 20)).view(np.uint32))
    np.testing.assert_array_equal(ours.predict(Q), theirs.predict(Q))
    np.testing.assert_allclose(ours.decision_function(Q), theirs.decision_function(Q), rtol=1e-4, atol=1e-5)
    assert ours.offset_ == pytest.approx(theirs.offset_, rel=1e-5)
    assert (ours.predict(Q)[-100:] == -1).mean() > 0.9       # the planted far queries
    with pytest.raises(AttributeError):
        ours.fit_predict(X)


def test_novelty_guard_routes_queries_to_radix():
    """A fitted norm past FLT_MAX / 16 sets the device guard: the sweep stands
    down and every query takes the exact radix select from its row list (keys
    stay finite at this scale); the scores still match the oracle bit for bit."""
    from _gpu import novelty

    rng = np.random.default_rng(92)
    X = (rng.standard_normal((3000, 4)) * 9.5e17).astype(np.float32)
    X[0] = np.array([5e18, 0.0, 0.0, 0.0], np.float32)      # norm 2.5e37 > FLT_MAX / 16
    Q = (rng.standard_normal((200, 4)) * 9.5e17).astype(np.float32)
    _native.profile_reset()
    _native.profile_enable(True)
    try:
        new, _ = novelty(X, Q, 10)
    finally:
        _native.profile_enable(False)
    assert _native.profile_read("radix_knn")[1] >= 2          # the fit's redo and the queries'
    np.testing.assert_array_equal(new.view(np.uint32), oracle.lof_novelty(X, Q, 10).view(np.uint32))
