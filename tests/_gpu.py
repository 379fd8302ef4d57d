"""Thin wrappers that drive the C ABI with torch-allocated device memory.

Used only by the ``gpu`` tests. Every call goes through libculof_amd.so's
``extern "C"`` entry points; torch provides device buffers and the stream.
Results come back as numpy arrays; neighbour arrays keep the library's
transposed (k, n) layout.
"""

from __future__ import annotations

import numpy as np
import torch

from culof_amd import _native

DEV = "cuda:0"


def _dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(DEV)


def _empty(shape, dtype):
    return torch.empty(shape, dtype=dtype, device=DEV)


def _sync():
    torch.cuda.synchronize()


def forward(X, k, normalize=False, form=_native.FORM_PLAIN):
    X = np.ascontiguousarray(X, np.float32)
    n, d = X.shape
    x = _dev(X)
    scores = _empty(n, torch.float32)
    idx = _empty((k, n), torch.int32)
    d2 = _empty((k, n), torch.float32)
    kd = _empty(n, torch.float32)
    lrd = _empty(n, torch.float32)
    _native.check(_native.lib().culof_lof_device(
        x.data_ptr(), n, d, k, int(normalize), form, scores.data_ptr(), idx.data_ptr(),
        d2.data_ptr(), kd.data_ptr(), lrd.data_ptr(), _native.current_stream()))
    _sync()
    return {"scores": scores.cpu().numpy(), "idx": idx.cpu().numpy(), "d2": d2.cpu().numpy(),
            "kdist2": kd.cpu().numpy(), "lrd": lrd.cpu().numpy()}


def novelty(X, Q, k, normalize=False, form=_native.FORM_PLAIN):
    """(scores of the new rows Q, scores of the fitted rows X) via culof_lof_novelty_device."""
    X = np.ascontiguousarray(X, np.float32)
    Q = np.ascontiguousarray(Q, np.float32).reshape(-1, X.shape[1])
    n, d = X.shape
    m = Q.shape[0]
    x = _dev(X)
    q = _dev(Q) if m else None
    fit = _empty(n, torch.float32)
    new = _empty(max(m, 1), torch.float32)
    _native.check(_native.lib().culof_lof_novelty_device(
        x.data_ptr(), n, d, q.data_ptr() if m else None, m, k, int(normalize), form, fit.data_ptr(),
        new.data_ptr(), None, None, _native.current_stream()))
    _sync()
    return new.cpu().numpy()[:m], fit.cpu().numpy()


def center(X, normalize=False):
    X = np.ascontiguousarray(X, np.float32)
    n, d = X.shape
    x = _dev(X)
    xc = _empty((n, d), torch.float32)
    norms = _empty(n, torch.float32)
    _native.check(_native.lib().culof_stage_center(
        x.data_ptr(), n, d, int(normalize), xc.data_ptr(), norms.data_ptr(), _native.current_stream()))
    _sync()
    return xc.cpu().numpy(), norms.cpu().numpy()


def knn(X, k, normalize=False, row0=0, rows=None, method=0, form=_native.FORM_PLAIN):
    X = np.ascontiguousarray(X, np.float32)
    n, d = X.shape
    rows = n - row0 if rows is None else rows
    x = _dev(X)
    idx = torch.full((k, n), -1, dtype=torch.int32, device=DEV)
    d2 = torch.full((k, n), float("nan"), dtype=torch.float32, device=DEV)
    kd = torch.full((n,), float("nan"), dtype=torch.float32, device=DEV)
    _native.check(_native.lib().culof_stage_knn(
        x.data_ptr(), n, d, k, int(normalize), form, row0, rows, method, idx.data_ptr(), d2.data_ptr(),
        kd.data_ptr(), _native.current_stream()))
    _sync()
    sl = slice(row0, row0 + rows)
    return idx.cpu().numpy()[:, sl], d2.cpu().numpy()[:, sl], kd.cpu().numpy()[sl]


def select_rows(keys, k):
    keys = np.ascontiguousarray(keys, np.float32)
    rows, cols = keys.shape
    kt = _dev(keys)
    idx = _empty((k, rows), torch.int32)
    d2 = _empty((k, rows), torch.float32)
    kd = _empty(rows, torch.float32)
    _native.check(_native.lib().culof_stage_select_rows(
        kt.data_ptr(), rows, cols, k, idx.data_ptr(), d2.data_ptr(), kd.data_ptr(),
        _native.current_stream()))
    _sync()
    return idx.cpu().numpy(), d2.cpu().numpy(), kd.cpu().numpy()


def lrd(idx_kn, d2_kn, kdist2):
    k, n = idx_kn.shape
    i_t, d_t, kd_t = _dev(idx_kn.astype(np.int32)), _dev(d2_kn.astype(np.float32)), _dev(kdist2.astype(np.float32))
    out = _empty(n, torch.float32)
    _native.check(_native.lib().culof_stage_lrd(
        i_t.data_ptr(), d_t.data_ptr(), kd_t.data_ptr(), n, k, out.data_ptr(), _native.current_stream()))
    _sync()
    return out.cpu().numpy()


def scores(idx_kn, lrd_values):
    k, n = idx_kn.shape
    i_t, l_t = _dev(idx_kn.astype(np.int32)), _dev(lrd_values.astype(np.float32))
    out = _empty(n, torch.float32)
    _native.check(_native.lib().culof_stage_scores(
        i_t.data_ptr(), l_t.data_ptr(), n, k, out.data_ptr(), _native.current_stream()))
    _sync()
    return out.cpu().numpy()
