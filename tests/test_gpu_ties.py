"""The FP64 tie-break form (CULOF_FORM_FP64_TIES) on the GPU.

* Bit-exact against the oracle's restatement (oracle.knn_ties) through every
  kNN path and both screens: neighbour lists, keys, k-distances, and the
  scores built on them.
* Neighbour sets equal to the FP64 brute-force oracle's
  (tests/cpp/test_culof.cu:39-80 restated) on every row, at the reference's
  near-tie test shape (n = 20,000, d = 8, k = 20; tests/python/test_parity.py:70-100)
  and on grid-quantised data with duplicates, where the plain FP32 rule is
  shown to disagree on some rows.
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


@pytest.mark.parametrize("path", ["mfma-fp32", "mfma-bf16", "valu", "radix"])
@pytest.mark.parametrize("n,d,k,seed", [(1500, 8, 10, 0), (1200, 3, 20, 1), (900, 32, 15, 2),
                                        (700, 64, 31, 3)])
def test_ties_bit_exact_vs_oracle(path, n, d, k, seed):
    from _gpu import knn

    X = make_tied(n, d, seed)
    oi, od, ok, tie = oracle.knn_ties(X, k)
    assert tie.sum() > 0
    method, form = {"mfma-fp32": (0, 0), "mfma-bf16": (0, 1), "valu": (2, 0), "radix": (1, 0)}[path]
    gi, gd, gk = knn(X, k, method=method, form=form | TIES)
    np.testing.assert_array_equal(gi.T, oi)
    np.testing.assert_array_equal(_bits(gd.T), _bits(od))
    np.testing.assert_array_equal(_bits(gk), _bits(ok))


def test_ties_forward_scores_bit_exact():
    from _gpu import forward

    X = make_tied(2000, 8, 5)
    oi, od, ok, _ = oracle.knn_ties(X, 20)
    got = forward(X, 20, form=TIES)
    want_lrd = oracle.lrd(oi, od, ok)
    np.testing.assert_array_equal(got["idx"].T, oi)
    np.testing.assert_array_equal(_bits(got["lrd"]), _bits(want_lrd))
    np.testing.assert_array_equal(_bits(got["scores"]), _bits(oracle.scores(oi, want_lrd)))


def _sets(idx_rows):
    return [frozenset(r) for r in idx_rows.tolist()]


@pytest.mark.parametrize("shape", [(20000, 8, 20, 0), (6000, 64, 32, 1)], ids=["n20k-d8", "n6k-d64"])
def test_ties_sets_equal_fp64_oracle(shape):
    """Every row: the FP64 oracle's neighbour set (the near-tie contract)."""
    from _gpu import knn

    n, d, k, seed = shape
    X = make_data(n, d, seed, n // 100) if d == 8 else make_tied(n, d, seed, n_clusters=32, dup_frac=0.005)
    fp64 = _sets(oracle.lof_fp64(X, k, full=True)["idx"])
    gi, _, _ = knn(X, k, form=TIES)
    got = _sets(gi.T)
    bad = [r for r in range(n) if got[r] != fp64[r]]
    assert not bad, f"{len(bad)} rows differ from the FP64 oracle, first {bad[:5]}"


def test_plain_rule_disagrees_where_ties_fix_it():
    """On quantised data the reference's FP32 rule picks other sets on some
    rows; the tie-break form picks the FP64 ones on all rows."""
    from _gpu import knn

    X = make_tied(3000, 8, 7)
    fp64 = _sets(oracle.lof_fp64(X, 10, full=True)["idx"])
    plain = _sets(knn(X, 10)[0].T)
    tied = _sets(knn(X, 10, form=TIES)[0].T)
    assert sum(p != f for p, f in zip(plain, fp64)) > 0
    assert all(t == f for t, f in zip(tied, fp64))


def test_ties_python_api():
    import culof_amd

    X = make_tied(1000, 4, 11)
    oi, od, ok, _ = oracle.knn_ties(X, 15)
    want = oracle.scores(oi, oracle.lrd(oi, od, ok))
    np.testing.assert_array_equal(_bits(culof_amd.lof(X, 15, ties="fp64")), _bits(want))
    np.testing.assert_array_equal(_bits(culof_amd.lof(X, 15, form="bf16", ties="fp64")), _bits(want))
    est = culof_amd.LOF(n_neighbors=15, ties="fp64").fit(X)
    np.testing.assert_array_equal(_bits(-est.negative_outlier_factor_), _bits(want))
    with pytest.raises(ValueError, match="ties"):
        culof_amd.lof(X, 15, ties="nearest")


def test_ties_with_k_equal_n_minus_1_is_plain():
    from _gpu import knn

    X = make_tied(64, 3, 12)
    plain = knn(X, 63)
    tied = knn(X, 63, form=TIES)
    for a, b in zip(plain, tied):
        np.testing.assert_array_equal(a, b)


def test_many_exact_duplicates():
    """A row with thousands of zero-distance candidates: the band outgrows the
    LDS buffer and is compacted on the fly."""
    from _gpu import knn

    rng = np.random.default_rng(13)
    X = rng.standard_normal((6000, 5)).astype(np.float32)
    X[:4000] = X[0]
    oi, od, ok, tie = oracle.knn_ties(X, 20, 0, 50)
    gi, gd, gk = knn(X, 20, form=TIES)
    np.testing.assert_array_equal(gi.T[:50], oi)
    np.testing.assert_array_equal(_bits(gk[:50]), _bits(ok))
    assert tie.all()


@pytest.mark.parametrize("form", [0, 1], ids=["fp32", "bf16"])
def test_band_pass_matches_brute_force_scan(form):
    """Near-tie rows settled through the collect-based band pass (the matrix-core
    path) against the brute-force scan (the radix path), every row bit for bit.
    Integer-grid data makes almost every row a near-tie: more rows than the band
    queue holds (the excess is scanned), bands of a few hundred entries (600
    scattered copies of one point) and bands larger than one block holds (3,000
    scattered copies of another)."""
    from _gpu import knn
    from culof_amd import _native

    rng = np.random.default_rng(21)
    n, d, k = 40_000, 4, 16
    X = np.round(rng.standard_normal((n, d)) * 4).astype(np.float32)
    perm = rng.permutation(n)
    X[perm[:600]] = X[perm[0]]
    X[perm[600:3600]] = X[perm[600]]
    _native.profile_reset()
    _native.profile_enable(True)
    try:
        bi, bd, bk = knn(X, k, form=form | TIES)
    finally:
        _native.profile_enable(False)
    assert _native.profile_read("tie_band")[1] > 0
    ri, rd, rk = knn(X, k, method=1, form=TIES)
    np.testing.assert_array_equal(bi, ri)
    np.testing.assert_array_equal(_bits(bd), _bits(rd))
    np.testing.assert_array_equal(_bits(bk), _bits(rk))
    ties = 0
    for r in np.r_[perm[:3], perm[600:603], rng.choice(n, 6, replace=False)]:
        oi, od, ok, tie = oracle.knn_ties(X, k, int(r), 1)
        np.testing.assert_array_equal(bi[:, r], oi[0], err_msg=f"row {r}")
        np.testing.assert_array_equal(_bits(bk[r]), _bits(ok[0]), err_msg=f"row {r}")
        ties += int(tie[0])
    assert ties >= 6
