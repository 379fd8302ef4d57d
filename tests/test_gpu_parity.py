"""Bit-exact parity of the HIP path with the CPU oracle, stage by stage.

Every comparison here is exact (bit patterns, not tolerances): the oracle
restates the same FP32/FP64 arithmetic the kernels perform, so any difference
is a bug. Inputs are seeded; shapes cover every sweep instantiation
(padded d in {4, 8, 16, 32, 64, 128}, k buckets 8/16/32/64) and the radix
path (k > 64, d > 128), plus k = 1, 2 and n - 1.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import make_data

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _assert_same_bits(got, want, what):
    got_b, want_b = _bits(got), _bits(want)
    bad = np.flatnonzero(got_b.ravel() != want_b.ravel())
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:5]}: " \
        f"got {np.asarray(got).ravel()[bad[:5]]} want {np.asarray(want).ravel()[bad[:5]]}"


SHAPES = [
    # (n, d, k, n_outliers, seed)
    (300, 1, 5, 3, 1),
    (500, 3, 20, 8, 2),
    (777, 8, 20, 9, 3),
    (1000, 8, 1, 10, 4),
    (640, 8, 2, 6, 5),
    (900, 16, 33, 9, 6),
    (700, 32, 50, 7, 7),
    (600, 64, 64, 6, 8),
    (400, 100, 10, 4, 9),
    (520, 128, 32, 5, 10),
    (300, 130, 7, 3, 11),    # d beyond the sweep: radix for every row
    (600, 4, 65, 6, 12),     # k beyond the sweep
    (128, 2, 127, 4, 13),    # k = n - 1
    (2000, 8, 20, 20, 14),
]


def _shape_id(s):
    return f"n{s[0]}-d{s[1]}-k{s[2]}"


@pytest.mark.parametrize("shape", SHAPES, ids=_shape_id)
def test_forward_bit_exact(shape):
    from _gpu import forward

    n, d, k, n_out, seed = shape
    X = make_data(n, d, seed, n_out)
    got = forward(X, k)
    want = oracle.lof(X, k, full=True)
    np.testing.assert_array_equal(got["idx"].T, want["idx"])
    _assert_same_bits(got["d2"].T, want["d2"], "nbr_d2")
    _assert_same_bits(got["kdist2"], want["kdist2"], "kdist2")
    _assert_same_bits(got["lrd"], want["lrd"], "lrd")
    _assert_same_bits(got["scores"], want["scores"], "scores")


@pytest.mark.parametrize("normalize", [False, True])
@pytest.mark.parametrize("shape", [(1000, 3, 10, 5, 21), (513, 7, 16, 5, 22), (300, 200, 9, 3, 23)],
                         ids=_shape_id)
def test_center_bit_exact(shape, normalize):
    from _gpu import center

    n, d, _, n_out, seed = shape
    X = make_data(n, d, seed, n_out) * np.float32(3.0) + np.float32(7.0)
    X[:, 0] = 4.0  # a constant feature
    xc, norms = center(X, normalize)
    want_xc, want_norms, _, _ = oracle.center(X, normalize)
    _assert_same_bits(xc, want_xc, "centred rows")
    _assert_same_bits(norms, want_norms, "norms")


@pytest.mark.parametrize("normalize", [False, True])
@pytest.mark.parametrize("shape", [s for s in SHAPES if s[1] <= 128 and s[2] <= 64], ids=_shape_id)
def test_every_path_matches_oracle(shape, normalize):
    """The matrix-core sweep (both screen forms), the VALU sweep and the
    radix path all return the oracle's bits."""
    from _gpu import knn

    n, d, k, n_out, seed = shape
    X = make_data(n, d, seed, n_out)
    oi, od, ok = oracle.knn_rows(oracle.key_rows_of(X, normalize), k)
    runs = {
        "mfma-fp32": knn(X, k, normalize=normalize, method=0, form=0),
        "mfma-bf16": knn(X, k, normalize=normalize, method=0, form=1),
        "valu": knn(X, k, normalize=normalize, method=2),
        "radix": knn(X, k, normalize=normalize, method=1),
    }
    for label, got in runs.items():
        np.testing.assert_array_equal(got[0].T, oi, err_msg=f"{label} nbr_idx")
        _assert_same_bits(got[1].T, od, f"{label} nbr_d2")
        _assert_same_bits(got[2], ok, f"{label} kdist2")


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("shape", [(5000, 8, 20, 50, 51), (6000, 32, 50, 60, 52), (4100, 128, 20, 41, 53),
                                   (7000, 3, 8, 70, 54)], ids=_shape_id)
def test_sliced_sweep_matches_oracle(shape, form):
    """Sizes where the candidate range is split into slices and merged."""
    from _gpu import knn

    n, d, k, n_out, seed = shape
    X = make_data(n, d, seed, n_out)
    got = knn(X, k, method=0, form=form)
    rows = np.r_[0:200, n // 2:n // 2 + 100, n - 200:n]
    oi, od, ok = oracle.knn_rows(X, k)
    np.testing.assert_array_equal(got[0].T, oi)
    _assert_same_bits(got[2], ok, "kdist2")


def test_forms_give_identical_forward():
    from _gpu import forward
    from culof_amd import workloads

    X = workloads.lowrank(n=3000, d=128, seed=2)
    plain = forward(X, 50, form=0)
    screened = forward(X, 50, form=1)
    for key in plain:
        np.testing.assert_array_equal(plain[key], screened[key], err_msg=key)


def test_partial_row_ranges():
    """A launch over a sub-range of query rows computes exactly those rows."""
    from _gpu import knn

    X = make_data(1500, 8, 31, 15)
    oi, od, ok = oracle.knn_rows(X, 20)
    for row0, rows in ((0, 1), (1, 63), (64, 65), (700, 800), (1499, 1)):
        gi, gd, gk = knn(X, 20, row0=row0, rows=rows)
        np.testing.assert_array_equal(gi.T, oi[row0:row0 + rows])
        _assert_same_bits(gk, ok[row0:row0 + rows], "kdist2")


def _random_key_rows(rng, rows, cols):
    """Non-negative keys with heavy ties: zeros, +inf, a few repeated values."""
    pool = np.concatenate([
        np.zeros(3, np.float32), np.full(2, np.inf, np.float32),
        rng.uniform(0, 4, size=12).astype(np.float32),
    ])
    keys = rng.uniform(0, 100, size=(rows, cols)).astype(np.float32)
    mask = rng.random((rows, cols)) < 0.4
    keys[mask] = rng.choice(pool, size=int(mask.sum()))
    # denormals and huge finite values exercise every radix digit
    keys[:, 0] = np.float32(1e-40)
    keys[:, 1] = np.float32(3e38)
    return keys


@pytest.mark.parametrize("cols,k", [(1, 1), (7, 3), (300, 1), (300, 37), (300, 299), (300, 300),
                                    (5000, 64), (5000, 1000), (70000, 20)])
def test_stored_select_bit_exact(cols, k):
    from _gpu import select_rows

    rng = np.random.default_rng(cols * 131 + k)
    rows = 17
    keys = _random_key_rows(rng, rows, cols) if cols >= 2 else np.zeros((rows, cols), np.float32)
    gi, gd, gk = select_rows(keys, k)
    for r in range(rows):
        wi, wd, wk = oracle.select_row(keys[r], k)
        np.testing.assert_array_equal(gi[:, r], wi, err_msg=f"row {r}")
        _assert_same_bits(gd[:, r], wd, f"row {r} keys")
        _assert_same_bits(gk[r], np.float32(wk), f"row {r} kth")


def test_density_stages_bit_exact():
    from _gpu import lrd, scores

    X = make_data(900, 5, 41, 9)
    want = oracle.lof(X, 15, full=True)
    got_lrd = lrd(want["idx"].T.copy(), want["d2"].T.copy(), want["kdist2"])
    _assert_same_bits(got_lrd, want["lrd"], "lrd")
    got_scores = scores(want["idx"].T.copy(), want["lrd"])
    _assert_same_bits(got_scores, want["scores"], "scores")


def test_tiling_is_bit_exact(monkeypatch):
    """CULOF_TILE_ROWS splits the sweep into launches; every row is independent."""
    from _gpu import forward

    X = make_data(700, 5, 13, 9)
    whole = forward(X, 25)["scores"]
    for rows in ("1", "7", "64", "699"):
        monkeypatch.setenv("CULOF_TILE_ROWS", rows)
        _assert_same_bits(forward(X, 25)["scores"], whole, f"tile_rows={rows}")


def test_repeated_calls_bit_identical():
    from _gpu import forward

    X = make_data(512, 6, 17, 7)
    first = forward(X, 20)
    for _ in range(3):
        again = forward(X, 20)
        for key in first:
            np.testing.assert_array_equal(first[key], again[key])


def test_large_collected_sets():
    """Rows whose candidates under the seed overflow the wave select: many
    exact duplicates (second collect pass, block select) and distances
    concentrated on a thin shell (block select), bit-exact with the oracle."""
    from _gpu import knn

    rng = np.random.default_rng(61)
    dup = rng.standard_normal((6000, 5)).astype(np.float32)
    dup[:3000] = dup[7]
    shell = rng.standard_normal((6000, 64))
    shell = (shell / np.linalg.norm(shell, axis=1, keepdims=True)).astype(np.float32)
    for X, k in ((dup, 20), (shell, 32)):
        oi, od, ok = oracle.knn_rows(X, k)
        for form in (0, 1):
            gi, gd, gk = knn(X, k, form=form)
            np.testing.assert_array_equal(gi.T, oi)
            _assert_same_bits(gd.T, od, "nbr_d2")
            _assert_same_bits(gk, ok, "kdist2")


def test_order_switch_gives_same_bits(monkeypatch):
    """CULOF_ORDER=0 (input order, unseeded lists) and the default agree."""
    from _gpu import forward

    X = make_data(3000, 32, 62, 30)
    default = forward(X, 20)
    monkeypatch.setenv("CULOF_ORDER", "0")
    plain = forward(X, 20)
    for key in default:
        np.testing.assert_array_equal(default[key], plain[key], err_msg=key)


@pytest.mark.parametrize("n,d,k,form", [(40_000, 8, 20, 0), (36_000, 16, 10, 1), (33_000, 64, 31, 0),
                                         (35_840, 64, 16, 1), (34_000, 128, 20, 1), (35_000, 32, 15, 1),
                                         (50_000, 4, 7, 0)])
def test_symmetric_pass_switch_gives_same_bits(monkeypatch, n, d, k, form):
    """The symmetric first pass (each pair of tiles computed once, the screen
    value tested against both rows' thresholds; full runs of at least 32,768
    rows) and CULOF_SYM=0 (every query row against every candidate tile) agree
    bit for bit, on clustered data with far outliers (large norms next to small
    thresholds), for odd and even block counts and both query-tile widths
    (two per wave for the wide BF16 screen); sampled rows also match the
    oracle."""
    from _gpu import knn

    rng = np.random.default_rng(n + d)
    X = make_data(n, d, n % 97, max(1, n // 200))
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
        _assert_same_bits(sym[2][r], ok[0], f"row {r} kdist2")


@pytest.mark.parametrize("n,d,k,form", [(40_000, 8, 20, 0), (36_000, 16, 10, 1), (33_000, 64, 31, 0),
                                         (35_840, 64, 16, 1), (34_000, 128, 50, 1), (35_000, 32, 15, 0),
                                         (50_000, 4, 7, 0), (45_000, 8, 100, 0)])
def test_pruned_symmetric_pass_gives_same_bits(monkeypatch, n, d, k, form):
    """The symmetric pass skips candidate chunks whose tile boxes lie farther
    from the block's than every seed that could use them (rows with the
    largest seeds ordered last). CULOF_PRUNE=0 (every chunk) gives the same
    bits, on clustered data with far outliers, and sampled rows match the
    oracle."""
    from _gpu import knn
    from culof_amd import workloads

    rng = np.random.default_rng(n + d + k)
    X = workloads.clusters(n=n, d=d, seed=n % 13)
    far = rng.choice(n, max(1, n // 200), replace=False)
    X[far] = X[far] * np.float32(6.0)
    pruned = knn(X, k, form=form)
    monkeypatch.setenv("CULOF_PRUNE", "0")
    full = knn(X, k, form=form)
    for a, b, name in zip(pruned, full, ("idx", "d2", "kdist2")):
        np.testing.assert_array_equal(a, b, err_msg=name)
    for r in np.r_[far[:3], rng.choice(n, 5, replace=False)]:
        oi, od, ok = oracle.knn_rows(X, k, int(r), 1)
        np.testing.assert_array_equal(pruned[0][:, r], oi[0], err_msg=f"row {r}")
        _assert_same_bits(pruned[2][r], ok[0], f"row {r} kdist2")


def test_two_level_screen_on_normalized_rows():
    """z-scored rows (normalize) of 128 features through the symmetric pass:
    the BF16 form's two-level screen gives the FP32 form's bits, on low-rank
    clusters scaled and shifted far from unit variance (so the z-scoring
    changes the norms the margins scale with)."""
    from _gpu import knn
    from culof_amd import workloads

    X = (workloads.lowrank(n=34_000, d=128, seed=4) * np.float32(30.0) + np.float32(7.0)).astype(np.float32)
    fp32 = knn(X, 20, normalize=True, form=0)
    bf16 = knn(X, 20, normalize=True, form=1)
    for a, b, name in zip(bf16, fp32, ("idx", "d2", "kdist2")):
        np.testing.assert_array_equal(_bits(a) if a.dtype == np.float32 else a,
                                      _bits(b) if b.dtype == np.float32 else b, err_msg=name)


@pytest.mark.parametrize("form", [0, 1], ids=["fp32", "bf16"])
def test_overflow_guard_in_symmetric_pass(form):
    """Norms beyond FLT_MAX / 16 set the device guard: the symmetric pass and
    its scatter and selects stand down and every row takes the radix path
    (keys stay finite at this scale); sampled rows match the oracle."""
    from _gpu import knn

    rng = np.random.default_rng(91)
    n, d, k = 36_000, 4, 10
    X = (rng.standard_normal((n, d)) * 9.5e17).astype(np.float32)  # one norm past the guard; keys < FLT_MAX
    from culof_amd import _native

    _native.profile_reset()
    _native.profile_enable(True)
    try:
        gi, gd, gk = knn(X, k, form=form)
    finally:
        _native.profile_enable(False)
    # every row went through the radix redo, which costs far more than the
    # stood-down collect pass
    assert _native.profile_read("radix_knn")[0] > 5 * _native.profile_read("mfma_collect")[0]
    for r in rng.choice(n, 6, replace=False):
        oi, od, ok = oracle.knn_rows(X, k, int(r), 1)
        np.testing.assert_array_equal(gi[:, r], oi[0], err_msg=f"row {r}")
        _assert_same_bits(gk[r], ok[0], f"row {r} kdist2")


def test_overflow_guard_with_wide_outliers_at_scale():
    """The guard at n > 65,536 on data with wide outliers (the reference's
    benchmark distribution, whose outliers would otherwise be queued for the
    list pass as heavy rows): every row goes through the radix redo exactly
    once, and sampled rows, outliers included, match the oracle."""
    from _gpu import knn

    from culof_amd import _native
    from culof_amd.benchmark import blobs_with_outliers

    n, d, k = 100_000, 8, 20
    X = blobs_with_outliers(n, d, seed=17) * np.float32(1.2e17)  # keys stay below FLT_MAX
    X[5] = np.float32(20 * 1.2e17)                                 # a norm past FLT_MAX / 16
    print(f"guard at scale: n={n} d={d} k={k}", flush=True)
    _native.profile_reset()
    _native.profile_enable(True)
    try:
        gi, gd, gk = knn(X, k)
    finally:
        _native.profile_enable(False)
    assert _native.profile_read("radix_knn")[0] > 5 * _native.profile_read("mfma_collect")[0]
    rng = np.random.default_rng(5)
    for r in [5, 0, 1, 2, *rng.choice(n, 6, replace=False).tolist()]:
        oi, od, ok = oracle.knn_rows(X, k, int(r), 1)
        np.testing.assert_array_equal(gi[:, r], oi[0], err_msg=f"row {r}")
        _assert_same_bits(gd[:, r], od[0], f"row {r} nbr_d2")
        _assert_same_bits(gk[r], ok[0], f"row {r} kdist2")


@pytest.mark.parametrize("d", [8, 32])
def test_scratch_reuse_after_a_larger_call(d):
    """A call reuses pool memory that a larger earlier call left behind
    (positions, counts, candidate buffers); nothing of it may reach the
    smaller call's result, which is bit-exact with the oracle."""
    from _gpu import forward

    print(f"scratch reuse after a larger call: d={d}", flush=True)
    big = make_data(40_000, d, seed=31, n_outliers=400)
    forward(big, 20)
    small = make_data(2_000, d, seed=32, n_outliers=40)
    got = forward(small, 20)
    want = oracle.lof(small, 20, full=True)
    np.testing.assert_array_equal(got["idx"].T, want["idx"])
    _assert_same_bits(got["scores"], want["scores"], "scores")


def test_small_scratch_budget_tiles_rows(monkeypatch):
    """CULOF_SCRATCH_BUDGET below the full run's collect buffers splits the
    query rows into tiles sized to the budget (the reference's tile_rows_for,
    src/cuda/pipeline.cu:99-110); the result is the same bits."""
    from _gpu import forward

    from culof_amd import _native

    X = make_data(50_000, 32, seed=23, n_outliers=500)
    print("budget tiling: n=50000 d=32 k=20", flush=True)
    want = forward(X, 20)
    monkeypatch.setenv("CULOF_SCRATCH_BUDGET", str(1 << 30))
    _native.profile_reset()
    _native.profile_enable(True)
    try:
        got = forward(X, 20)
    finally:
        _native.profile_enable(False)
    launches = _native.profile_read("mfma_collect")[1]
    assert launches >= 2, f"expected several row tiles under a 1 GiB budget, got {launches}"
    for key in ("scores", "idx", "d2", "kdist2", "lrd"):
        _assert_same_bits(got[key], want[key], key)


def test_every_fallback_tier():
    """10,000 copies of one point among 150,000: their collected sets overflow
    the first pass, the second pass and the list queue, so the rows run
    through every tier (block select, second collect pass, seeded list pass,
    radix). Sampled rows, duplicates included, are bit-exact with the oracle."""
    from _gpu import knn

    rng = np.random.default_rng(71)
    X = rng.standard_normal((150_000, 4)).astype(np.float32)
    dup = rng.choice(150_000, 10_000, replace=False)
    X[dup] = X[dup[0]]
    k = 12
    gi, gd, gk = knn(X, k)
    rows = np.r_[dup[:12], rng.choice(150_000, 12, replace=False)]
    for r in rows:
        oi, od, ok = oracle.knn_rows(X, k, int(r), 1)
        np.testing.assert_array_equal(gi[:, r], oi[0], err_msg=f"row {r}")
        _assert_same_bits(gk[r], ok[0], f"row {r} kdist2")
