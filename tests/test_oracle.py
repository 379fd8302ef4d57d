"""Pins the CPU oracle before anything is compared against it.

Each check is chosen so that a plausible mistake in the oracle fails it:
* scikit-learn 1.7.2 FP64 golden vectors (tests/golden/) and the
  scikit-learn docstring known-answer test;
* closed forms: points equally spaced on a circle have LOF exactly 1 by
  symmetry, and so do the two points of n = 2;
* the selection rule against an independent brute-force lexicographic sort;
* centring against exact sums, invariants of the z-score, and norms;
* the Gram-screen error bound the fused sweep relies on, on adversarial data;
* planted mutations (the oracle's ``mutation`` argument: off-by-one
  k-distance, missing max() in the reach distance, inverted density ratio)
  that every score check must catch;
* the FP64 tie-break restatement against the FP64 brute-force oracle's
  neighbour sets on data full of near ties (where the plain FP32 rule is
  shown to disagree, so the check has teeth).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from conftest import GOLDEN_CASES, load_golden, make_data, make_tied, rel_err

# scikit-learn's LocalOutlierFactor docstring example (sklearn/neighbors/_lof.py:175-182).
KAT_X = np.array([[-1.1], [0.2], [101.1], [0.3]], np.float32)
KAT_LOF = np.array([0.98214286, 1.03703704, 73.36970899, 0.98214286])


def test_kat_docstring_fp64():
    got = oracle.lof_fp64(KAT_X, 2)
    np.testing.assert_allclose(got, KAT_LOF, rtol=1e-6)


def test_kat_docstring_fp32():
    got = oracle.lof(KAT_X, 2)
    np.testing.assert_allclose(got, KAT_LOF, rtol=1e-6)


@pytest.mark.parametrize("m,k", [(3, 1), (3, 4), (3, 7), (4, 5), (4, 15)])
def test_hypercube_is_exactly_one(m, k):
    """Vertices of {0, 3}^m: every point sees the same distances, so every
    k-distance, every density and every LOF is equal -- exactly 1 -- whatever
    the tie-breaking picks."""
    X = (np.array(np.meshgrid(*[[0, 1]] * m)).reshape(m, -1).T * 3).astype(np.float32)
    assert np.all(oracle.lof_fp64(X, k) == 1.0)
    assert np.all(oracle.lof(X, k) == 1.0)


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8])
def test_circle_is_one(k):
    """Equally spaced points on a circle: LOF 1 up to the float32 rounding of
    the positions themselves."""
    n = 64
    ang = 2 * np.pi * np.arange(n) / n
    X = np.stack([np.cos(ang), np.sin(ang)], 1).astype(np.float32) * 10
    np.testing.assert_allclose(oracle.lof_fp64(X, k), 1.0, rtol=1e-6)
    np.testing.assert_allclose(oracle.lof(X, k), 1.0, rtol=2e-6)


def test_two_points():
    X = np.array([[0.0, 0.0], [3.0, 4.0]], np.float32)
    assert np.all(oracle.lof_fp64(X, 1) == 1.0)
    assert np.all(oracle.lof(X, 1) == 1.0)
    full = oracle.lof(X, 1, full=True)
    assert full["kdist2"].tolist() == [25.0, 25.0]
    assert full["idx"].ravel().tolist() == [1, 0]


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_fp64_oracle_matches_sklearn(case):
    g = load_golden(case)
    k = int(g["k"])
    got = oracle.lof_fp64(g["X"], k)
    assert rel_err(got, g["lof"]).max() < 1e-9


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_fp32_oracle_within_accuracy_targets(case):
    """Median 1e-6, p99 5e-6 against scikit-learn FP64, identical top-1% sets."""
    g = load_golden(case)
    k = int(g["k"])
    got = oracle.lof(g["X"], k)
    r = rel_err(got, g["lof"])
    assert np.median(r) <= 1e-6 and np.percentile(r, 99) <= 5e-6, (np.median(r), np.percentile(r, 99))
    top = max(1, len(r) // 100)
    assert set(np.argsort(-got, kind="stable")[:top]) == set(np.argsort(-g["lof"], kind="stable")[:top])


@pytest.mark.parametrize("mutation", [1, 2, 3])
def test_planted_mutations_are_caught(mutation):
    """Each planted mistake moves scores far outside every tolerance above."""
    caught = []
    for case in ("clusters_n2000_d8_k20", "gauss_n400_d3_k20", "kat_docstring"):
        g = load_golden(case)
        k = int(g["k"])
        r32 = rel_err(oracle.lof(g["X"], k, mutation=mutation), g["lof"])
        r64 = rel_err(oracle.lof_fp64(g["X"], k, mutation=mutation), g["lof"])
        caught.append(r32.max() > 1e-3 and r64.max() > 1e-3)
    assert any(caught)
    # the clustered case alone catches every mutation
    assert caught[0]


def _brute_select(keys, k):
    """Independent statement of the selection rule: sort by (key, index)."""
    order = np.lexsort((np.arange(len(keys)), keys.view(np.uint32)))
    chosen = order[:k]
    kth = keys.view(np.uint32)[chosen[-1]]
    lt = np.sort(chosen[keys.view(np.uint32)[chosen] < kth])
    eq = np.sort(chosen[keys.view(np.uint32)[chosen] == kth])
    return np.concatenate([lt, eq]), keys[chosen[-1]]


@pytest.mark.parametrize("seed", range(6))
def test_select_row_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 400))
    keys = rng.uniform(0, 5, n).astype(np.float32)
    # heavy ties, zeros and +inf
    m = rng.random(n) < 0.5
    keys[m] = rng.choice(np.array([0.0, 1.0, 2.5, np.inf], np.float32), size=int(m.sum()))
    for k in sorted({1, 2, n // 2 or 1, n - 1 or 1, n}):
        idx, d2, kth = oracle.select_row(keys, k)
        want_idx, want_kth = _brute_select(keys, k)
        np.testing.assert_array_equal(idx, want_idx)
        np.testing.assert_array_equal(d2, keys[want_idx])
        assert np.float32(kth).view(np.uint32) == np.float32(want_kth).view(np.uint32)


def test_centring_exact_on_integer_data():
    rng = np.random.default_rng(0)
    X = rng.integers(-50, 50, size=(1000, 4)).astype(np.float32)
    xc, norms, mean, inv = oracle.center(X)
    np.testing.assert_array_equal(mean, X.astype(np.float64).sum(0) / 1000)  # sums are exact
    np.testing.assert_array_equal(inv, 1.0)
    np.testing.assert_array_equal(xc, (X.astype(np.float64) - mean).astype(np.float32))


def test_centring_statistics():
    X = make_data(3000, 5, 3, 30) * np.float32(4.0) + np.float32(1000.0)
    X[:, 2] = 7.0
    xc, norms, mean, inv = oracle.center(X, normalize=False)
    exact = np.array([math.fsum(X[:, f].astype(np.float64)) / len(X) for f in range(5)])
    np.testing.assert_allclose(mean, exact, rtol=1e-15)
    assert np.all(np.abs(xc.astype(np.float64).mean(0)) < 1e-4)
    z, _, _, inv_z = oracle.center(X, normalize=True)
    sd = np.delete(X.astype(np.float64).std(0), 2)
    np.testing.assert_allclose(np.delete(z.std(0), 2), 1.0, rtol=1e-5)
    np.testing.assert_allclose(np.delete(inv_z, 2), 1 / sd, rtol=1e-12)
    assert inv_z[2] == 1.0 and np.all(z[:, 2] == 0.0)  # constant feature stays centred at zero
    # norms: FP64 sum of squares of the centred float rows, rounded once
    want = (xc.astype(np.float64) ** 2).sum(1).astype(np.float32)
    assert np.max(np.abs(norms - want) / want) <= 6e-8


def test_keys_are_squared_distances():
    X = make_data(400, 6, 5, 4)
    for i in (0, 17, 399):
        keys = oracle.key_row(X, i)
        exact = ((X.astype(np.float64) - X[i].astype(np.float64)) ** 2).sum(1)
        assert keys[i] == np.inf
        mask = np.arange(len(X)) != i
        assert np.max(np.abs(keys[mask] - exact[mask]) / exact[mask]) <= 8 * 2.0 ** -24
    # symmetric bit for bit
    assert oracle.key_row(X, 3)[9] == oracle.key_row(X, 9)[3]


def _adversarial_sets():
    rng = np.random.default_rng(11)
    far = make_data(300, 8, 1, 3) * np.float32(0.01) + np.float32(40.0)  # tiny spread, large offset
    clus = np.concatenate([rng.normal(c, 0.05, size=(60, 16)) for c in rng.uniform(-30, 30, (5, 16))]).astype(np.float32)
    dups = np.repeat(make_data(40, 4, 2), 5, axis=0)
    wide = make_data(200, 128, 3, 2)
    return {"far": far, "clusters": clus, "duplicates": dups, "wide": wide}


@pytest.mark.parametrize("name", ["far", "clusters", "duplicates", "wide"])
def test_gram_screen_bound_holds(name):
    """|gram - e| <= c (|x_i|^2 + |x_j|^2 + e), c = (d + 16) 2^-24: the bound
    the fused sweep screens with (centred rows, FP64-accumulated norms)."""
    X = _adversarial_sets()[name]
    n, d = X.shape
    xc, norms, _, _ = oracle.center(X)
    c = (d + 16) * 2.0 ** -24
    worst = 0.0
    for i in range(0, n, max(1, n // 40)):
        e = oracle.key_row(X, i).astype(np.float64)
        raw = oracle.gram_key_row(xc, norms, i).astype(np.float64)
        mask = np.arange(n) != i
        # gram keys clamp at 0, which only moves them toward e
        slack = c * (norms[i] + norms[mask].astype(np.float64) + e[mask])
        ratio = np.abs(raw[mask] - e[mask]) / slack
        worst = max(worst, float(ratio.max()))
    assert worst <= 1.0, worst


def test_translation_and_permutation_invariance():
    # coordinates on a 1/64 grid, so the shifted copy is exact in float32
    X = np.round(make_data(600, 3, 8, 6) * 64) / 64
    base = oracle.lof_fp64(X, 15)
    shifted = (X + 512.0).astype(np.float32)
    np.testing.assert_allclose(oracle.lof_fp64(shifted, 15), base, rtol=1e-12)
    # far from the origin the FP32 keys keep their accuracy (no Gram cancellation)
    far = (make_data(1500, 3, 7, 15) + np.float32(5000.0)).astype(np.float32)
    assert rel_err(oracle.lof(far, 20), oracle.lof_fp64(far, 20)).max() < 2e-4
    # permutation equivariance, away from ties (unquantised data)
    Y = make_data(600, 3, 8, 6)
    perm = np.random.default_rng(0).permutation(len(Y))
    np.testing.assert_allclose(oracle.lof_fp64(Y[perm], 15), oracle.lof_fp64(Y, 15)[perm], rtol=1e-12)
    np.testing.assert_allclose(oracle.lof(Y[perm], 15), oracle.lof(Y, 15)[perm], rtol=1e-6)


def test_rejects_bad_arguments():
    X = make_data(50, 2, 1)
    for k in (0, 50):
        with pytest.raises(ValueError):
            oracle.lof(X, k)


@pytest.mark.parametrize("n,d,k,seed", [(1500, 8, 10, 0), (1200, 3, 20, 1), (900, 32, 15, 2),
                                        (600, 2, 40, 3)])
def test_tie_break_matches_fp64_oracle_sets(n, d, k, seed):
    X = make_tied(n, d, seed)
    idx, d2, kd, tie = oracle.knn_ties(X, k)
    pi, pd, pk = oracle.knn_rows(X, k)
    fp64 = [set(r) for r in oracle.lof_fp64(X, k, full=True)["idx"].tolist()]
    tied = [set(r) == f for r, f in zip(idx.tolist(), fp64)]
    plain = [set(r) == f for r, f in zip(pi.tolist(), fp64)]
    assert all(tied), f"{len(tied) - sum(tied)} rows differ from the FP64 oracle"
    # rows with a clear boundary keep the FP32 result bit for bit
    keep = tie == 0
    np.testing.assert_array_equal(idx[keep], pi[keep])
    np.testing.assert_array_equal(d2[keep].view(np.uint32), pd[keep].view(np.uint32))
    np.testing.assert_array_equal(kd[keep].view(np.uint32), pk[keep].view(np.uint32))
    # every row's slot k-1 holds a neighbour at the k-distance
    np.testing.assert_array_equal(d2[:, -1].view(np.uint32), kd.view(np.uint32))
    assert tie.sum() > 0
    if seed == 0:
        assert not all(plain), "data chosen to make the plain FP32 rule disagree"


def test_tie_break_on_well_separated_data_changes_nothing():
    X = make_data(800, 6, 9, 8)
    idx, d2, kd, tie = oracle.knn_ties(X, 12)
    pi, pd, pk = oracle.knn_rows(X, 12)
    np.testing.assert_array_equal(idx[tie == 0], pi[tie == 0])
    assert tie.sum() <= 2
    fp64 = oracle.lof_fp64(X, 12, full=True)["idx"]
    assert all(set(a) == set(b) for a, b in zip(idx.tolist(), fp64.tolist()))


def test_tie_band_ratio_covers_fp32_rounding():
    """The band must cover the relative error of the FP32 difference-form key."""
    for d in (1, 8, 64, 128, 1000):
        rho = oracle._load().oracle_tie_band_ratio(d)
        g = (d + 2) * 2.0**-24
        assert rho >= (1 + g) / (1 - g)
        assert rho < 1 + 4 * (d + 4) * 2.0**-24
    rng = np.random.default_rng(4)
    for d in (3, 64):
        A = rng.standard_normal((200, d)).astype(np.float32)
        keys = np.stack([oracle.key_row(A, i) for i in range(20)])
        exact = ((A[:20, None, :].astype(np.float64) - A[None, :, :]) ** 2).sum(-1)
        ok = np.isfinite(keys) & (exact > 0)
        rel = np.abs(keys[ok] - exact[ok]) / exact[ok]
        assert rel.max() <= (d + 2) * 2.0**-24


def _bf16(v):
    """Round-to-nearest-even float32 -> bfloat16, returned as float32."""
    b = np.ascontiguousarray(v, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("name", ["far", "clusters", "duplicates", "wide", "boundary"])
def test_split_bf16_screen_bound_holds(name):
    """The split BF16 screen (a_hi b_hi + a_hi b_lo + a_lo b_hi, bf16 products
    exact) drops at most 3.1 2^-16 sum_t |a_t b_t|, and with FP32 accumulation
    of the 3 d products (any order) stays within c (|a|^2 + |b|^2),
    c = 3.1 2^-16 + (12 d + 64) 2^-24, for b = -2 q."""
    sets = dict(_adversarial_sets())
    rng = np.random.default_rng(5)
    # values just off the bf16 rounding midpoints, magnitudes spread over 2^40
    mant = (1.0 + (rng.integers(0, 128, size=(64, 48)) + 0.5 + rng.uniform(-1e-3, 1e-3, (64, 48))) / 128)
    sets["boundary"] = (mant * 2.0 ** rng.integers(-20, 20, size=(64, 48)) *
                        rng.choice([-1.0, 1.0], size=(64, 48))).astype(np.float32)
    X = sets[name]
    xc, norms, _, _ = oracle.center(X) if name != "boundary" else (X, (X.astype(np.float64) ** 2).sum(1), 0, 0)
    xc = np.asarray(xc, np.float32)
    d = xc.shape[1]
    hi = _bf16(xc)
    lo = _bf16((xc - hi).astype(np.float32))
    c = 3.1 * 2.0 ** -16 + (12 * d + 64) * 2.0 ** -24
    worst_drop = worst_total = 0.0
    for i in range(0, xc.shape[0], max(1, xc.shape[0] // 24)):
        b = (-2.0 * xc[i]).astype(np.float32)          # exact scaling
        bh = _bf16(b)
        bl = _bf16((b - bh).astype(np.float32))
        exact = xc.astype(np.float64) @ b.astype(np.float64)
        prods = np.concatenate([hi.astype(np.float64) * bh, hi.astype(np.float64) * bl,
                                lo.astype(np.float64) * bh], axis=1)   # exact in FP64
        drop = np.abs(prods.sum(1) - exact)
        mag = np.abs(xc.astype(np.float64)) @ np.abs(b.astype(np.float64))
        ok = mag > 0
        worst_drop = max(worst_drop, float((drop[ok] / (3.1 * 2.0 ** -16 * mag[ok])).max(initial=0)))
        # FP32 accumulation in two different orders
        nn = np.asarray(norms, np.float64)
        scale = nn + float(nn[i]) if name != "boundary" else nn + nn[i]
        for order in (np.arange(prods.shape[1]), rng.permutation(prods.shape[1])):
            acc = np.zeros(xc.shape[0], np.float32)
            for t in order:
                acc = (acc + prods[:, t].astype(np.float32)).astype(np.float32)
            err = np.abs(acc.astype(np.float64) - exact)
            sel = scale > 0
            worst_total = max(worst_total, float((err[sel] / (c * scale[sel])).max(initial=0)))
    assert worst_drop <= 1.0, worst_drop
    assert worst_total <= 1.0, worst_total


def _coarse_slack(d):
    """The two-level BF16 screen's slack, as the kernels state it: the lo x hi
    and hi x lo corrections are at most 2^-7 (1 + 2^-8) sum |a_t b_t|, their
    2 d FP32 additions at most 2 d x 2u x 3.1 (N_q + N_c); x 1.01 + 16u."""
    u = 2.0 ** -24
    return 1.01 * (2.0 ** -7 * (1 + 2.0 ** -8) + 12.4 * d * u) + 16 * u


@pytest.mark.parametrize("name", ["far", "clusters", "duplicates", "wide", "boundary"])
def test_two_level_bf16_coarse_slack_holds(name):
    """The coarse value of the two-level BF16 screen (m_c plus the hi x hi
    products, FP32 accumulation) stays within g (N_q + N_c) of the full value
    (the same accumulator continued with the lo x hi and hi x lo products), so
    a tile whose coarse values all exceed a threshold widened by that margin
    holds no pair the full screen passes. Values just off the bf16 rounding
    midpoints make the residuals as large as they get (|lo(x)| near 2^-8
    |x|); the same check with the residual bound taken as 2^-9 must fail."""
    sets = dict(_adversarial_sets())
    rng = np.random.default_rng(9)
    mant = (1.0 + (rng.integers(0, 128, size=(64, 48)) + 0.5 + rng.uniform(-1e-3, 1e-3, (64, 48))) / 128)
    sets["boundary"] = (mant * 2.0 ** rng.integers(-10, 10, size=(64, 48)) *
                        rng.choice([-1.0, 1.0], size=(64, 48))).astype(np.float32)
    X = sets[name]
    if name != "boundary":
        xc, norms, _, _ = oracle.center(X)
    else:
        xc, norms = X, (X.astype(np.float64) ** 2).sum(1)
    xc = np.asarray(xc, np.float32)
    nn = np.asarray(norms, np.float64)
    d = xc.shape[1]
    hi = _bf16(xc)
    lo = _bf16((xc - hi).astype(np.float32))
    c = 3.1 * 2.0 ** -16 + (12 * d + 64) * 2.0 ** -24
    madj = (nn * (1 - c)).astype(np.float32)
    worst = worst_tight = 0.0
    for i in range(0, xc.shape[0], max(1, xc.shape[0] // 24)):
        b = (-2.0 * xc[i]).astype(np.float32)
        bh = _bf16(b)
        bl = _bf16((b - bh).astype(np.float32))
        p_hh = hi.astype(np.float64) * bh                       # exact in FP64 (and in FP32)
        p_cor = np.concatenate([lo.astype(np.float64) * bh, hi.astype(np.float64) * bl], axis=1)
        scale = nn + nn[i]
        sel = scale > 0
        for order in range(2):
            acc = madj.copy()
            cols = np.arange(d) if order == 0 else rng.permutation(d)
            for t in cols:
                acc = (acc + p_hh[:, t].astype(np.float32)).astype(np.float32)
            coarse = acc.astype(np.float64)
            cols = np.arange(2 * d) if order == 0 else rng.permutation(2 * d)
            for t in cols:
                acc = (acc + p_cor[:, t].astype(np.float32)).astype(np.float32)
            gap = np.abs(acc.astype(np.float64) - coarse)
            worst = max(worst, float((gap[sel] / (_coarse_slack(d) * scale[sel])).max(initial=0)))
            # a residual bound of 2^-9 (half the true one) is too small
            tight = 2.0 ** -8 * (1 + 2.0 ** -8) + 12.4 * d * 2.0 ** -24
            worst_tight = max(worst_tight, float((gap[sel] / (tight * scale[sel])).max(initial=0)))
    assert worst <= 1.0, worst
    if name == "boundary":
        assert worst_tight > 1.0, worst_tight


@pytest.mark.parametrize("normalize", [False, True])
def test_novelty_oracle_matches_sklearn(normalize):
    """oracle_lof_novelty against scikit-learn's LocalOutlierFactor(novelty=True)
    in FP64 (same z-scoring when normalising): the FP32 keys leave ~1e-7."""
    from sklearn.neighbors import LocalOutlierFactor

    rng = np.random.default_rng(21)
    X = rng.standard_normal((900, 6)).astype(np.float32) * np.float32(3.0)
    Q = np.concatenate([rng.standard_normal((150, 6)) * 3.0, rng.uniform(-20, 20, (10, 6))]).astype(np.float32)
    got = oracle.lof_novelty(X, Q, 15, normalize=normalize)
    Xs, Qs = X.astype(np.float64), Q.astype(np.float64)
    if normalize:
        mu, sd = Xs.mean(0), Xs.std(0)
        Xs, Qs = (Xs - mu) / sd, (Qs - mu) / sd
    want = -LocalOutlierFactor(n_neighbors=15, novelty=True).fit(Xs).score_samples(Qs)
    assert np.max(np.abs(got - want) / want) < 2e-6
    assert set(np.argsort(-got)[:10]) == set(range(150, 160))  # the far queries score highest


def test_novelty_oracle_of_fitted_rows_and_identities():
    """Closed forms on a unit grid (k = 4). The centre of an interior cell has
    its four corners at squared distance 0.5, each of whose k-distance is 1,
    as is every reach distance of their own neighbours: both densities are
    1 / (1 + 1e-10) and the LOF is exactly 1. A query equal to a fitted row
    finds that row at distance 0 (nothing is excluded). A query far outside
    scores far above 1."""
    g = np.arange(6, dtype=np.float32)
    X = np.stack(np.meshgrid(g, g), -1).reshape(-1, 2)
    Q = np.array([[2.0, 2.0], [2.5, 2.5], [30.0, 30.0]], np.float32)
    res = oracle.lof_novelty(X, Q, 4, full=True)
    assert res["scores"][1] == np.float32(1.0)
    assert sorted(res["idx"][1]) == [14, 15, 20, 21] and (res["d2"][1] == 0.5).all()
    assert res["d2"][0].min() == 0.0 and 14 in res["idx"][0]
    assert res["scores"][2] > 5.0
