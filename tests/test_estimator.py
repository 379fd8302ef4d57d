"""scikit-learn estimator protocol of culof_amd.LOF (beyond the reference's
surface, src/culof/__init__.py:116-227): get_params / set_params (so
sklearn.base.clone works) and kneighbors of the fitted samples."""

from __future__ import annotations

import numpy as np
import pytest
from sklearn.base import clone
from sklearn.neighbors import NearestNeighbors

from culof_amd import LOF


def test_params_round_trip_and_clone():
    m = LOF(n_neighbors=7, contamination=0.1, normalize=True, form="bf16", ties="fp64")
    p = m.get_params()
    assert p == {"n_neighbors": 7, "contamination": 0.1, "normalize": True, "form": "bf16", "ties": "fp64",
                 "novelty": False}
    c = clone(m)
    assert c is not m and c.get_params() == p
    assert m.set_params(n_neighbors=9, contamination="auto") is m
    assert (m.n_neighbors, m.contamination) == (9, "auto")


def test_set_params_validates():
    m = LOF()
    with pytest.raises(ValueError, match="invalid parameter"):
        m.set_params(metric="cosine")
    with pytest.raises(ValueError, match="n_neighbors"):
        m.set_params(n_neighbors=0)
    with pytest.raises(ValueError, match="contamination"):
        m.set_params(contamination=0.7)
    assert m.get_params()["n_neighbors"] == 20  # unchanged after a refused call


def test_kneighbors_needs_a_fit():
    with pytest.raises(ValueError, match="not fitted"):
        LOF().kneighbors()


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp32", "bf16"])
def test_kneighbors_matches_exact_search(form):
    rng = np.random.default_rng(11)
    X = rng.standard_normal((3000, 16)).astype(np.float32)
    m = LOF(n_neighbors=12, form=form).fit(X)
    assert (m.n_features_in_, m.n_samples_fit_) == (16, 3000)
    dist, idx = m.kneighbors()
    assert dist.shape == idx.shape == (3000, 12) and dist.dtype == np.float32
    assert (np.diff(dist, axis=1) >= 0).all()
    ref_d, ref_i = NearestNeighbors(n_neighbors=13).fit(X.astype(np.float64)).kneighbors()
    np.testing.assert_array_equal(idx, ref_i[:, :12])          # generic data: no ties
    np.testing.assert_allclose(dist, ref_d[:, :12], rtol=2e-6)
    assert m.kneighbors(n_neighbors=3, return_distance=False).shape == (3000, 3)
    # new rows: nothing excluded (a fitted row finds itself at distance 0)
    Q = np.concatenate([rng.standard_normal((400, 16)).astype(np.float32), X[:5]])
    qd, qi = m.kneighbors(Q, n_neighbors=7)
    rd, ri = NearestNeighbors(n_neighbors=7).fit(X.astype(np.float64)).kneighbors(Q.astype(np.float64))
    np.testing.assert_array_equal(qi, ri)
    np.testing.assert_allclose(qd, rd, rtol=2e-6, atol=1e-6)
    assert (qi[-5:, 0] == np.arange(5)).all() and (qd[-5:, 0] == 0).all()
    with pytest.raises(ValueError, match="n_neighbors"):
        m.kneighbors(n_neighbors=3000)


def test_novelty_argument_contract():
    m = LOF(novelty=True)
    assert m.get_params()["novelty"] is True and "novelty=True" in repr(m)
    with pytest.raises(AttributeError, match="novelty"):
        m.fit_predict(np.zeros((10, 2), np.float32))
    with pytest.raises(TypeError, match="new points"):
        m.score_samples()
    with pytest.raises(AttributeError, match="novelty"):
        LOF().decision_function(np.zeros((3, 2), np.float32))
    with pytest.raises(TypeError):
        LOF().score_samples(np.zeros((3, 2), np.float32))
