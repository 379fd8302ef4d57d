"""The hand-written backward d(sum w LOF)/dX against FP64 references.

Three independent checks of ``culof_lof_backward_device`` (through
``culof_amd.autograd``, i.e. the C ABI):

* FP64 torch autograd of LOF with the forward's neighbour sets and
  reach-distance branches held fixed (the function the backward differentiates);
* central finite differences (FP64, h = 1e-3) of that same function on 256
  random coordinates, which pins the autograd reference itself;
* FP64 autograd of a *dense* LOF at n = 4000: an n x n FP64 distance matrix,
  neighbours chosen by (distance, index) as in tests/cpp/test_culof.cu:59-61,
  no state taken from the GPU at all.

The library accumulates in FP64 from the FP32 points and casts the result to
FP32, so gradients agree to FP32 rounding of the distances; the tolerance
written in each test is relative to the largest gradient entry.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import make_data

pytestmark = pytest.mark.gpu

EPS = 1e-10


def _torch():
    import torch

    return torch


def _gpu_grad(X, k, w, normalize=False, form="fp32"):
    torch = _torch()
    from culof_amd.autograd import lof_torch

    x = torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda().requires_grad_(True)
    wt = torch.from_numpy(np.asarray(w, np.float32)).cuda()
    scores = lof_torch(x, k, normalize=normalize, form=form)
    (wt * scores).sum().backward()
    torch.cuda.synchronize()
    return scores.detach().cpu().numpy(), x.grad.cpu().numpy()


def _state(X, k, normalize=False):
    from culof_amd.autograd import forward_state

    torch = _torch()
    st = forward_state(torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda(), k, normalize)
    torch.cuda.synchronize()
    return (st.idx.cpu().numpy().T.astype(np.int64), st.d2.cpu().numpy().T, st.kdist2.cpu().numpy(),
            st.scores.cpu().numpy())


def _zscore(X):
    """(X - mean) / std with population std; features with std <= 1e-12 only centred."""
    torch = _torch()
    mean = X.mean(0)
    c = X - mean
    var = (c * c).mean(0)
    scaled = (var > 1e-24).detach()  # std > 1e-12
    sd = torch.sqrt(torch.where(scaled, var, torch.ones_like(var)))
    return c / sd


def _dist(a, b):
    """FP64 Euclidean distance with a zero (not NaN) gradient at coincidence."""
    torch = _torch()
    ss = ((a - b) ** 2).sum(-1)
    pos = ss > 0
    return torch.where(pos, torch.sqrt(torch.where(pos, ss, torch.ones_like(ss))), torch.zeros_like(ss))


def _fixed_lof64(X64, idx, branch_kd, normalize):
    """LOF in FP64 with neighbour sets and max() branches held fixed."""
    torch = _torch()
    Z = _zscore(X64) if normalize else X64
    n, k = idx.shape
    d_ij = _dist(Z[:, None, :], Z[idx])
    kd = _dist(Z, Z[idx[:, k - 1]])  # slot k-1 holds a neighbour at the k-distance
    reach = torch.where(branch_kd, kd[idx], d_ij)
    m = reach.mean(1)
    lrd = 1.0 / (m + EPS)
    return (m + EPS) * lrd[idx].mean(1)


def _fixed_reference(X, k, w, normalize=False):
    torch = _torch()
    idx, d2, kd2, _ = _state(X, k, normalize)
    branch = torch.from_numpy(kd2[idx] > d2)
    idx_t = torch.from_numpy(idx)
    X64 = torch.from_numpy(np.asarray(X, np.float64)).requires_grad_(True)
    w64 = torch.from_numpy(np.asarray(w, np.float64))
    (w64 * _fixed_lof64(X64, idx_t, branch, normalize)).sum().backward()

    def f(x64):
        with torch.no_grad():
            return float((w64 * _fixed_lof64(x64, idx_t, branch, normalize)).sum())

    return X64.grad.numpy(), f


def _assert_close(got, want, rel, what):
    scale = np.abs(want).max()
    err = np.abs(got.astype(np.float64) - want).max()
    assert err <= rel * scale, f"{what}: max abs err {err:.3e} > {rel} x max|grad| {scale:.3e}"


@pytest.mark.parametrize("normalize", [False, True])
@pytest.mark.parametrize("n,d,k,seed", [(2000, 8, 20, 1), (1500, 32, 20, 2), (800, 3, 5, 3),
                                        (600, 100, 12, 4), (700, 200, 9, 5)])
def test_backward_matches_fp64_autograd(n, d, k, seed, normalize):
    X = make_data(n, d, seed, n // 50)
    if normalize:
        X = X * np.linspace(0.5, 3.0, d, dtype=np.float32) + np.float32(2.0)
        X[:, 0] = 1.25  # constant feature: centred, never scaled
    w = np.random.default_rng(seed).standard_normal(n).astype(np.float32)
    _, g = _gpu_grad(X, k, w, normalize)
    want, _ = _fixed_reference(X, k, w, normalize)
    _assert_close(g, want, 2e-5, "gradient vs FP64 autograd (fixed structure)")


def test_finite_differences_pin_the_reference():
    """Central differences of the fixed-structure FP64 LOF on 256 coordinates."""
    torch = _torch()
    n, d, k = 1200, 8, 20
    X = make_data(n, d, 7, 24)
    w = np.random.default_rng(7).standard_normal(n).astype(np.float32)
    _, g = _gpu_grad(X, k, w)
    want, f = _fixed_reference(X, k, w)
    rng = np.random.default_rng(8)
    coords = rng.choice(n * d, size=256, replace=False)
    h = 1e-3
    base = torch.from_numpy(np.asarray(X, np.float64))
    fd = np.empty(coords.size)
    for t, c in enumerate(coords):
        i, j = divmod(int(c), d)
        xp, xm = base.clone(), base.clone()
        xp[i, j] += h
        xm[i, j] -= h
        fd[t] = (f(xp) - f(xm)) / (2 * h)
    got = g.reshape(-1)[coords].astype(np.float64)
    scale = np.abs(want).max()
    assert np.abs(fd - want.reshape(-1)[coords]).max() <= 1e-4 * scale
    assert np.abs(fd - got).max() <= 1e-4 * scale


def test_dense_fp64_autograd_n4000():
    """An independent FP64 LOF (dense n x n distances, (distance, index) order)."""
    torch = _torch()
    n, d, k = 4000, 32, 20
    X = make_data(n, d, 11, 40)
    w = np.random.default_rng(11).standard_normal(n).astype(np.float32)
    scores, g = _gpu_grad(X, k, w)

    X64 = torch.from_numpy(np.asarray(X, np.float64)).requires_grad_(True)
    with torch.no_grad():
        sq = (X64 * X64).sum(1)
        D = sq[:, None] + sq[None, :] - 2.0 * (X64 @ X64.T)
        D.fill_diagonal_(float("inf"))
        # (distance, index) order: a stable sort keeps ascending index among equal distances
        order = torch.sort(D, dim=1, stable=True).indices[:, :k]
    dist = _dist(X64[:, None, :], X64[order])
    kd = dist[:, k - 1]
    reach = torch.maximum(kd[order], dist)
    m = reach.mean(1)
    lrd = 1.0 / (m + EPS)
    lof64 = (m + EPS) * lrd[order].mean(1)
    (torch.from_numpy(w.astype(np.float64)) * lof64).sum().backward()

    idx, _, _, _ = _state(X, k)
    same = np.array([set(a) == set(b) for a, b in zip(idx.tolist(), order.numpy().tolist())])
    assert same.all(), f"{(~same).sum()} rows choose different neighbour sets"
    np.testing.assert_allclose(scores, lof64.detach().numpy(), rtol=1e-5)
    _assert_close(g, X64.grad.numpy(), 2e-5, "gradient vs dense FP64 autograd")


def test_backward_is_deterministic_and_form_independent():
    torch = _torch()
    from culof_amd.autograd import backward_points, forward_state

    X = make_data(3000, 32, 13, 30)
    x = torch.from_numpy(X).cuda()
    w = torch.from_numpy(np.random.default_rng(13).standard_normal(3000).astype(np.float32)).cuda()
    st = forward_state(x, 20)
    first = backward_points(x, st, w).cpu().numpy()
    for _ in range(3):
        np.testing.assert_array_equal(backward_points(x, st, w).cpu().numpy(), first)
    _, g_bf16 = _gpu_grad(X, 20, w.cpu().numpy(), form="bf16")
    np.testing.assert_array_equal(g_bf16, first)


@pytest.mark.parametrize("k", [40, 140])
def test_hub_points_and_long_reverse_lists(k):
    """Hubs (points in hundreds of kNN lists): reverse lists longer than 64
    entries are summed in several chunks. At k = 40 every 256-point block
    sorts its lists in one LDS window (at most 32,768 entries); at k = 140
    most blocks hold more and sort them in several windows. Both against
    FP64 autograd, and a repeat gives the same bits."""
    torch = _torch()
    from culof_amd.autograd import backward_points, forward_state

    n, d = 3000, 32
    lds_entries = 32768
    X = make_data(n, d, 21, n // 100)
    w = np.random.default_rng(21).standard_normal(n).astype(np.float32)
    idx, _, _, _ = _state(X, k)
    indeg = np.bincount(idx.ravel(), minlength=n)
    spans = np.add.reduceat(indeg, np.arange(0, n, 256))
    assert indeg.max() > 4 * 64, indeg.max()   # a list of five or more chunks
    assert (indeg > 64).sum() > 100
    if k == 40:
        assert (spans <= lds_entries).all(), spans
    else:
        assert (spans > lds_entries).any() and (spans <= lds_entries).any(), spans
    _, g = _gpu_grad(X, k, w)
    want, _ = _fixed_reference(X, k, w)
    _assert_close(g, want, 2e-5, f"gradient vs FP64 autograd (hub points, k={k})")
    x = torch.from_numpy(X).cuda()
    st = forward_state(x, k)
    wt = torch.from_numpy(w).cuda()
    first = backward_points(x, st, wt).cpu().numpy()
    np.testing.assert_array_equal(backward_points(x, st, wt).cpu().numpy(), first)
    np.testing.assert_array_equal(first, g)


def test_duplicates_give_finite_gradients():
    X = make_data(1000, 8, 17, 10)
    X[500:520] = X[100]  # 21 coincident points: zero distances inside every list
    w = np.ones(1000, np.float32)
    _, g = _gpu_grad(X, 10, w)
    assert np.isfinite(g).all()
    want, _ = _fixed_reference(X, 10, w)
    _assert_close(g, want, 2e-5, "gradient with duplicates")


def test_backward_rejects_bad_input():
    torch = _torch()
    from culof_amd.autograd import backward_points, forward_state

    X = torch.from_numpy(make_data(300, 300, 19, 3)).cuda()
    st = forward_state(X, 5)
    with pytest.raises(ValueError, match="n_dims <= 256"):
        backward_points(X, st, torch.ones(300, device="cuda"))
    with pytest.raises(ValueError, match="grad_scores"):
        backward_points(X[:, :8].contiguous(), forward_state(X[:, :8], 5), torch.ones(3, device="cuda"))
    with pytest.raises(ValueError, match="k must satisfy"):
        forward_state(X, 300)
