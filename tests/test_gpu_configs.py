"""The five benchmark workloads at full size through the HIP path.

Each workload is built by ``culof_amd.workloads`` at the size the benchmark
quotes and checked against the oracle: the whole forward where the oracle
finishes in seconds (n = 20k), sampled rows bit for bit elsewhere (the
oracle's per-row kNN over all n candidates), plus the properties that do not
depend on size (identical results across screen forms, bit-identical
repeats), agreement with scikit-learn's LocalOutlierFactor fitted on float64
input, and the gradient against a float64 autograd reference with the kNN
structure fixed from the forward.

    plain-20k   clusters, n=20,000,  d=8,   k=20, contamination 0.01
    plain-200k  clusters, n=200,000, d=8,   k=20
    bf16-100k   low-rank, n=100,000, d=128, k=50, BF16 screen vs FP32
                (and, for the two-level BF16 screen's margins, the same
                generator with far rows and the reference's benchmark
                distribution at d = 128)
    grad-50k    clusters, n=50,000,  d=32,  k=20, forward + backward
    ties-500k   quantised clusters with duplicates, n=500,000, d=64, k=32,
                both forms x both tie rules, forward + backward
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from culof_amd import LOF, workloads
from culof_amd.autograd import backward_points, forward_state

pytestmark = pytest.mark.gpu

EPS = 1e-10


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _same_bits(got, want, what):
    g, w = _bits(got).ravel(), _bits(want).ravel()
    bad = np.flatnonzero(g != w)
    assert bad.size == 0, f"{what}: {bad.size} differences, first at {bad[:5]}"


def _state(x, k, form="fp32", ties="index"):
    st = forward_state(x, k, form=form, ties=ties)
    torch.cuda.synchronize()
    return st


def _host(st):
    return {"scores": st.scores.cpu().numpy(), "idx": st.idx.cpu().numpy(),
            "d2": st.d2.cpu().numpy(), "kdist2": st.kdist2.cpu().numpy(), "lrd": st.lrd.cpu().numpy()}


def _rows_match_oracle(X, got, rows, k, what):
    """kNN lists (slot order), keys and k-distances of `rows` bit-exact with the oracle."""
    for r in rows:
        oi, od, ok = oracle.knn_rows(X, k, int(r), 1)
        np.testing.assert_array_equal(got["idx"][:, r], oi[0], err_msg=f"{what} row {r} nbr_idx")
        _same_bits(got["d2"][:, r], od[0], f"{what} row {r} nbr_d2")
        _same_bits(got["kdist2"][r], ok[0], f"{what} row {r} kdist2")


def _sklearn_fp64(X, k, contamination="auto"):
    from sklearn.neighbors import LocalOutlierFactor

    est = LocalOutlierFactor(n_neighbors=k, contamination=contamination, n_jobs=-1)
    labels = est.fit_predict(X.astype(np.float64))
    return -est.negative_outlier_factor_, labels


def _agreement(scores, want):
    rel = np.abs(scores.astype(np.float64) - want) / np.abs(want)
    top = max(1, len(want) // 100)
    agree = len(set(np.argsort(-scores)[:top]) & set(np.argsort(-want)[:top])) / top
    return float(np.median(rel)), float(np.percentile(rel, 99)), agree


def _dist(a, b):
    """FP64 Euclidean distance with a zero (not NaN) gradient at coincidence."""
    ss = ((a - b) ** 2).sum(-1)
    pos = ss > 0
    return torch.where(pos, torch.sqrt(torch.where(pos, ss, torch.ones_like(ss))), torch.zeros_like(ss))


def _fixed_reference_grad(X, st, w):
    """d(sum w LOF)/dX in FP64 on the device (elementwise products and gathers
    only) with the neighbour sets and reach-distance branches of the forward
    state `st` held fixed; slot k - 1 holds a neighbour at the k-distance."""
    idx = st.idx.T.long()
    k = idx.shape[1]
    branch = st.kdist2[idx] > st.d2.T
    X64 = torch.from_numpy(np.asarray(X, np.float64)).cuda().requires_grad_(True)
    d_ij = _dist(X64[:, None, :], X64[idx])
    kd = _dist(X64, X64[idx[:, k - 1]])
    reach = torch.where(branch, kd[idx], d_ij)
    m = reach.mean(1)
    lrd = 1.0 / (m + EPS)
    lof = (m + EPS) * lrd[idx].mean(1)
    (torch.from_numpy(np.asarray(w, np.float64)).cuda() * lof).sum().backward()
    return X64.grad, lof.detach()


def _assert_grad_close(got, want, rel, what):
    scale = float(want.abs().max())
    err = float((got.double() - want).abs().max())
    assert err <= rel * scale, f"{what}: max abs err {err:.3e} > {rel} x max|grad| {scale:.3e}"


# ---------------------------------------------------------------- plain-20k
def test_plain_20k_whole_forward_and_fp64_sklearn():
    name = "plain-20k"
    spec = workloads.WORKLOADS[name]
    X = workloads.make(name)
    k = spec["k"]
    print(f"{name}: n={X.shape[0]} d={X.shape[1]} k={k}", flush=True)
    got = _host(_state(torch.from_numpy(X).cuda(), k))
    want = oracle.lof(X, k, full=True)
    for key in ("scores", "kdist2", "lrd"):
        _same_bits(got[key], want[key], f"{name} {key}")
    np.testing.assert_array_equal(got["idx"].T, want["idx"])
    _same_bits(got["d2"].T, want["d2"], f"{name} nbr_d2")

    # scikit-learn fitted on float64 input: its FP64 path, the comparator of the accuracy targets
    sk, sk_labels = _sklearn_fp64(X, k, contamination=0.01)
    med, p99, agree = _agreement(got["scores"], sk)
    assert med <= 1e-6 and p99 <= 5e-6 and agree == 1.0, (med, p99, agree)
    labels = LOF(n_neighbors=k, contamination=0.01).fit_predict(X)
    np.testing.assert_array_equal(labels, sk_labels)


# --------------------------------------------------------------- plain-200k
def test_plain_200k_outlier_rows_and_fp64_sklearn():
    name = "plain-200k"
    spec = workloads.WORKLOADS[name]
    X, outliers = workloads.clusters(n=spec["n"], d=spec["d"], return_outliers=True)
    k = spec["k"]
    print(f"{name}: n={X.shape[0]} d={X.shape[1]} k={k}", flush=True)
    x = torch.from_numpy(X).cuda()
    got = _host(_state(x, k))
    again = _host(_state(x, k))
    for key in got:
        _same_bits(again[key], got[key], f"{name} repeat {key}")
    rng = np.random.default_rng(2)
    rows = np.concatenate([np.flatnonzero(outliers), rng.choice(np.flatnonzero(~outliers), 200, replace=False)])
    assert rows.size == 2_200
    _rows_match_oracle(X, got, rows, k, name)

    sk, _ = _sklearn_fp64(X, k)
    med, p99, agree = _agreement(got["scores"], sk)
    assert med <= 1e-6 and p99 <= 5e-6 and agree == 1.0, (med, p99, agree)


# ---------------------------------------------------------------- bf16-100k
def test_bf16_100k_forms_identical_and_rows_exact():
    name = "bf16-100k"
    X = workloads.make(name)
    k = workloads.WORKLOADS[name]["k"]
    print(f"{name}: n={X.shape[0]} d={X.shape[1]} k={k}", flush=True)
    x = torch.from_numpy(X).cuda()
    fp32 = _host(_state(x, k, "fp32"))
    bf16 = _host(_state(x, k, "bf16"))
    for key in fp32:
        _same_bits(bf16[key], fp32[key], f"{name} bf16 vs fp32 {key}")
    rows = np.random.default_rng(3).choice(X.shape[0], 200, replace=False)
    rows = np.concatenate([rows, np.argsort(-fp32["scores"])[:20]])
    _rows_match_oracle(X, fp32, rows, k, name)


@pytest.mark.parametrize("data", ["lowrank-far", "reference-blobs"])
def test_bf16_two_level_screen_far_rows(data):
    """The two-level BF16 screen of the symmetric pass at 128 features where
    row norms and seeds spread widely, so its coarse margins (which grow with
    the norms and the column thresholds) are tested at large values: the
    low-rank workload with 2 % of its rows replaced by points uniform in
    [-30, 30]^128, and the reference's benchmark distribution (4 unit Gaussian
    clusters, 1 % uniform outliers in [-20, 20]^d) at n = 100k, d = 128. Both
    forms give identical results; the far rows and sampled others are
    bit-exact with the oracle."""
    from culof_amd.benchmark import blobs_with_outliers

    rng = np.random.default_rng(31)
    if data == "lowrank-far":
        X = workloads.lowrank(n=100_000, d=128, seed=31)
        far = rng.choice(X.shape[0], 2_000, replace=False)
        X[far] = rng.uniform(-30.0, 30.0, size=(far.size, X.shape[1])).astype(np.float32)
    else:
        X = blobs_with_outliers(100_000, 128, seed=3)
        far = np.arange(1_000)  # the outliers come first
    k = 20
    print(f"two-level {data}: n={X.shape[0]} d={X.shape[1]} k={k}", flush=True)
    x = torch.from_numpy(X).cuda()
    fp32 = _host(_state(x, k, "fp32"))
    bf16 = _host(_state(x, k, "bf16"))
    for key in fp32:
        _same_bits(bf16[key], fp32[key], f"two-level {data} bf16 vs fp32 {key}")
    rows = np.concatenate([far[:20], rng.choice(X.shape[0], 40, replace=False)])
    _rows_match_oracle(X, bf16, rows, k, f"two-level {data}")


# ----------------------------------------------------------------- grad-50k
def test_grad_50k_forward_backward_full_size():
    name = "grad-50k"
    X, w = workloads.make(name)
    k = workloads.WORKLOADS[name]["k"]
    print(f"{name}: n={X.shape[0]} d={X.shape[1]} k={k}", flush=True)
    x = torch.from_numpy(X).cuda()
    wt = torch.from_numpy(w).cuda()
    st = _state(x, k)
    got = _host(st)
    rows = np.random.default_rng(4).choice(X.shape[0], 200, replace=False)
    _rows_match_oracle(X, got, rows, k, name)

    g = backward_points(x, st, wt)
    g2 = backward_points(x, st, wt)
    torch.cuda.synchronize()
    assert torch.equal(g, g2), "backward not bit-identical on repeat"
    assert bool(torch.isfinite(g).all())
    st_bf16 = _state(x, k, "bf16")
    assert torch.equal(backward_points(x, st_bf16, wt), g), "gradient depends on the screen form"

    want, lof64 = _fixed_reference_grad(X, st, w)
    np.testing.assert_allclose(got["scores"], lof64.cpu().numpy(), rtol=1e-5)
    _assert_grad_close(g, want, 2e-5, f"{name} gradient vs FP64 autograd (fixed structure)")


def test_grad_generator_n4000_finite_differences_and_dense_fp64():
    """The gradient workload's generator at n=4,000: central differences
    (h = 1e-3, FP64) on 256 coordinates of the fixed-structure function, and
    an independent dense FP64 LOF ((distance, index) order) under autograd."""
    X, w = workloads.grad(n=4_000, d=32, seed=5)
    n, d = X.shape
    k = 20
    x = torch.from_numpy(X).cuda()
    st = _state(x, k)
    g = backward_points(x, st, torch.from_numpy(w).cuda()).cpu().numpy().astype(np.float64)

    # central differences of the fixed-structure FP64 function (on the CPU)
    idx = st.idx.T.long().cpu()
    branch = (st.kdist2[st.idx.T.long()] > st.d2.T).cpu()
    w64 = torch.from_numpy(w.astype(np.float64))

    def f(x64):
        d_ij = _dist(x64[:, None, :], x64[idx])
        kd = _dist(x64, x64[idx[:, k - 1]])
        m = torch.where(branch, kd[idx], d_ij).mean(1)
        return float((w64 * (m + EPS) * (1.0 / (m + EPS))[idx].mean(1)).sum())

    base = torch.from_numpy(X.astype(np.float64))
    coords = np.random.default_rng(6).choice(n * d, size=256, replace=False)
    h = 1e-3
    fd = np.empty(coords.size)
    for t, c in enumerate(coords):
        i, j = divmod(int(c), d)
        xp, xm = base.clone(), base.clone()
        xp[i, j] += h
        xm[i, j] -= h
        fd[t] = (f(xp) - f(xm)) / (2 * h)
    scale = np.abs(g).max()
    assert np.abs(fd - g.reshape(-1)[coords]).max() <= 1e-4 * scale

    # dense FP64 LOF with its own neighbour sets (CPU, no shared code)
    X64 = torch.from_numpy(X.astype(np.float64)).requires_grad_(True)
    with torch.no_grad():
        sq = (X64 * X64).sum(1)
        D = sq[:, None] + sq[None, :] - 2.0 * (X64 @ X64.T)
        D.fill_diagonal_(float("inf"))
        order = torch.sort(D, dim=1, stable=True).indices[:, :k]
    dist = _dist(X64[:, None, :], X64[order])
    kd = dist[:, k - 1]
    m = torch.maximum(kd[order], dist).mean(1)
    lof64 = (m + EPS) * (1.0 / (m + EPS))[order].mean(1)
    (w64 * lof64).sum().backward()
    same = [set(a) == set(b) for a, b in zip(idx.numpy().tolist(), order.numpy().tolist())]
    assert all(same), f"{len(same) - sum(same)} rows choose different neighbour sets"
    assert np.abs(g - X64.grad.numpy()).max() <= 2e-5 * np.abs(X64.grad.numpy()).max()


# ---------------------------------------------------------------- ties-500k
def test_ties_500k_forms_rules_rows_and_backward():
    name = "ties-500k"
    X = workloads.make(name)
    k = workloads.WORKLOADS[name]["k"]
    n = X.shape[0]
    print(f"{name}: n={n} d={X.shape[1]} k={k}", flush=True)
    x = torch.from_numpy(X).cuda()
    runs = {}
    for ties in ("index", "fp64"):
        for form in ("fp32", "bf16"):
            runs[form, ties] = _state(x, k, form, ties)
    host = {key: _host(st) for key, st in runs.items()}
    for ties in ("index", "fp64"):
        for key in host["fp32", ties]:
            _same_bits(host["bf16", ties][key], host["fp32", ties][key], f"{name} {ties} bf16 vs fp32 {key}")

    rng = np.random.default_rng(7)
    index_rule = host["fp32", "index"]
    _rows_match_oracle(X, index_rule, rng.choice(n, 60, replace=False), k, f"{name} index rule")
    fp64_rule = host["fp32", "fp64"]
    changed = np.flatnonzero((fp64_rule["idx"] != index_rule["idx"]).any(0))
    assert changed.size > 0, "the tie rule should change some rows of this workload"
    rows = np.concatenate([rng.choice(changed, min(16, changed.size), replace=False),
                           rng.choice(n, 24 - min(16, changed.size), replace=False)])
    for r in rows:
        oi, od, ok, _ = oracle.knn_ties(X, k, int(r), 1)
        np.testing.assert_array_equal(fp64_rule["idx"][:, r], oi[0], err_msg=f"fp64 rule row {r}")
        _same_bits(fp64_rule["d2"][:, r], od[0], f"fp64 rule row {r} nbr_d2")
        _same_bits(fp64_rule["kdist2"][r], ok[0], f"fp64 rule row {r} kdist2")

    # full-size backward in both forms (index rule): deterministic, finite,
    # form-independent, and equal to the FP64 fixed-structure gradient
    w = np.random.default_rng(8).standard_normal(n).astype(np.float32)
    wt = torch.from_numpy(w).cuda()
    g = backward_points(x, runs["fp32", "index"], wt)
    g_bf16 = backward_points(x, runs["bf16", "index"], wt)
    torch.cuda.synchronize()
    assert torch.equal(g, g_bf16), "gradient depends on the screen form"
    assert torch.equal(backward_points(x, runs["fp32", "index"], wt), g), "backward not bit-identical"
    assert bool(torch.isfinite(g).all())
    want, _ = _fixed_reference_grad(X, runs["fp32", "index"], w)
    _assert_grad_close(g, want, 2e-5, f"{name} gradient vs FP64 autograd (fixed structure)")
