"""ctypes front end for the CPU oracle in ``culof_oracle.c``.

Test infrastructure only: imported by ``tests/``, by ``__graft_entry__.smoke``
and by the CPU-baseline leg of ``bench.py``, always to check or time the
GPU path against, never in its place. The product package ``culof_amd``
never imports it.

Every wrapper returns numpy arrays. Neighbour lists come back row-major,
shape ``(n, k)``; the GPU path stores them transposed and the tests transpose
before comparing.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile ``liboracle.so`` with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    c_int, c_void = ctypes.c_int, ctypes.c_void_p
    lib.oracle_center.argtypes = [_f32p, c_int, c_int, c_int, _f32p, _f32p, _f64p, _f64p]
    lib.oracle_center.restype = None
    lib.oracle_key_row.argtypes = [_f32p, c_int, c_int, c_int, _f32p]
    lib.oracle_key_row.restype = None
    lib.oracle_gram_key_row.argtypes = [_f32p, _f32p, c_int, c_int, c_int, _f32p]
    lib.oracle_gram_key_row.restype = None
    lib.oracle_select_row.argtypes = [_f32p, c_int, c_int, _u32p, _i32p, _f32p, _f32p]
    lib.oracle_select_row.restype = None
    lib.oracle_knn_rows.argtypes = [_f32p, c_int, c_int, c_int, c_int, c_int, _i32p, _f32p, _f32p]
    lib.oracle_knn_rows.restype = None
    lib.oracle_knn_ties.argtypes = [_f32p, c_int, c_int, c_int, c_int, c_int, _i32p, _f32p, _f32p, _i32p]
    lib.oracle_knn_ties.restype = None
    lib.oracle_tie_band_ratio.argtypes = [c_int]
    lib.oracle_tie_band_ratio.restype = ctypes.c_double
    lib.oracle_lrd.argtypes = [_i32p, _f32p, _f32p, c_int, c_int, c_int, _f32p]
    lib.oracle_lrd.restype = None
    lib.oracle_scores.argtypes = [_i32p, _f32p, c_int, c_int, c_int, _f32p]
    lib.oracle_scores.restype = None
    lib.oracle_lof.argtypes = [_f32p, c_int, c_int, c_int, c_int, c_int, _f32p, c_void, c_void, c_void, c_void]
    lib.oracle_lof.restype = c_int
    lib.oracle_lof_fp64.argtypes = [_f32p, c_int, c_int, c_int, c_int, _f64p, c_void, c_void]
    lib.oracle_lof_fp64.restype = c_int
    lib.oracle_lof_novelty.argtypes = [_f32p, c_int, c_int, _f32p, c_int, c_int, c_int, _f32p, c_void, c_void]
    lib.oracle_lof_novelty.restype = c_int
    _lib = lib
    return lib


def _as_f32(X) -> np.ndarray:
    return np.ascontiguousarray(X, dtype=np.float32)


def center(X, normalize: bool = False):
    """Centred (optionally z-scored) float32 data, norms, FP64 means, FP64 inverse std."""
    X = _as_f32(X)
    n, d = X.shape
    xc = np.empty_like(X)
    norms = np.empty(n, np.float32)
    mean = np.empty(d, np.float64)
    inv = np.empty(d, np.float64)
    _load().oracle_center(X, n, d, int(normalize), xc, norms, mean, inv)
    return xc, norms, mean, inv


def key_rows_of(X, normalize: bool = False) -> np.ndarray:
    """The rows keys are computed from: the input itself, or its z-scored rows."""
    return center(X, True)[0] if normalize else _as_f32(X)


def key_row(xs, i: int) -> np.ndarray:
    """Selection keys of row ``i``: difference-form squared distances, +inf on the diagonal."""
    xs = _as_f32(xs)
    n, d = xs.shape
    out = np.empty(n, np.float32)
    _load().oracle_key_row(xs, n, d, int(i), out)
    return out


def gram_key_row(xc, norms, i: int) -> np.ndarray:
    """The reference's Gram-expansion keys of row ``i`` over centred rows."""
    xc = _as_f32(xc)
    n, d = xc.shape
    out = np.empty(n, np.float32)
    _load().oracle_gram_key_row(xc, _as_f32(norms), n, d, int(i), out)
    return out


def select_row(keys, k: int):
    """(neighbour indices in slot order, their keys, k-th key) for one key row."""
    keys = _as_f32(keys)
    n = keys.shape[0]
    scratch = np.empty(n, np.uint32)
    idx = np.empty(k, np.int32)
    d2 = np.empty(k, np.float32)
    kth = np.empty(1, np.float32)
    _load().oracle_select_row(keys, n, int(k), scratch, idx, d2, kth)
    return idx, d2, float(kth[0])


def knn_rows(xs, k: int, row0: int = 0, rows: int | None = None):
    """Row-major (rows, k) neighbour indices and keys plus k-distances squared,
    from key rows ``xs`` (see :func:`key_rows_of`)."""
    xs = _as_f32(xs)
    n, d = xs.shape
    rows = n - row0 if rows is None else rows
    idx = np.empty((rows, k), np.int32)
    d2 = np.empty((rows, k), np.float32)
    kd = np.empty(rows, np.float32)
    _load().oracle_knn_rows(xs, n, d, int(k), int(row0), int(rows), idx, d2, kd)
    return idx, d2, kd


def knn_ties(xs, k: int, row0: int = 0, rows: int | None = None):
    """kNN with the FP64 tie-break of near-tie rows: (idx, d2, kd, tie) with
    ``tie[r] = 1`` for rows settled on FP64 distances."""
    xs = _as_f32(xs)
    n, d = xs.shape
    rows = n - row0 if rows is None else rows
    idx = np.empty((rows, k), np.int32)
    d2 = np.empty((rows, k), np.float32)
    kd = np.empty(rows, np.float32)
    tie = np.empty(rows, np.int32)
    _load().oracle_knn_ties(xs, n, d, int(k), int(row0), int(rows), idx, d2, kd, tie)
    return idx, d2, kd, tie


def lrd(idx, d2, kdist2, mutation: int = 0) -> np.ndarray:
    idx = np.ascontiguousarray(idx, np.int32)
    n, k = idx.shape
    out = np.empty(n, np.float32)
    _load().oracle_lrd(idx, _as_f32(d2), _as_f32(kdist2), n, k, int(mutation), out)
    return out


def scores(idx, lrd_values, mutation: int = 0) -> np.ndarray:
    idx = np.ascontiguousarray(idx, np.int32)
    n, k = idx.shape
    out = np.empty(n, np.float32)
    _load().oracle_scores(idx, _as_f32(lrd_values), n, k, int(mutation), out)
    return out


def lof(X, k: int = 20, normalize: bool = False, mutation: int = 0, full: bool = False):
    """FP32 forward exactly as the GPU path defines it.

    With ``full=True`` returns a dict with scores, idx (n, k), d2 (n, k),
    kdist2 and lrd.
    """
    X = _as_f32(X)
    n, d = X.shape
    out = np.empty(n, np.float32)
    idx = np.empty((n, k), np.int32)
    d2 = np.empty((n, k), np.float32)
    kd = np.empty(n, np.float32)
    lr = np.empty(n, np.float32)
    rc = _load().oracle_lof(X, n, d, int(k), int(normalize), int(mutation), out,
                            idx.ctypes.data, d2.ctypes.data, kd.ctypes.data, lr.ctypes.data)
    if rc != 0:
        raise ValueError("invalid arguments to oracle_lof")
    if full:
        return {"scores": out, "idx": idx, "d2": d2, "kdist2": kd, "lrd": lr}
    return out


def lof_fp64(X, k: int = 20, mutation: int = 0, full: bool = False):
    """FP64 brute-force LOF with (distance, index) tie order."""
    X = _as_f32(X)
    n, d = X.shape
    out = np.empty(n, np.float64)
    idx = np.empty((n, k), np.int32)
    kd = np.empty(n, np.float64)
    rc = _load().oracle_lof_fp64(X, n, d, int(k), int(mutation), out, idx.ctypes.data, kd.ctypes.data)
    if rc != 0:
        raise ValueError("invalid arguments to oracle_lof_fp64")
    if full:
        return {"scores": out, "idx": idx, "kdist": kd}
    return out


def lof_novelty(X, Q, k: int = 20, normalize: bool = False, full: bool = False):
    """LOF of new rows ``Q`` against the fitted rows ``X`` (scikit-learn's
    novelty score_samples with culof's sign), in the forward's FP32/FP64 mix.
    With ``full=True`` also returns each query's neighbour indices and keys
    (row-major (m, k), slot order)."""
    X = _as_f32(X)
    Q = _as_f32(Q)
    n, d = X.shape
    m = Q.shape[0]
    if Q.ndim != 2 or Q.shape[1] != d:
        raise ValueError("queries must have the fitted rows' feature count")
    out = np.empty(m, np.float32)
    idx = np.empty((m, k), np.int32)
    d2 = np.empty((m, k), np.float32)
    rc = _load().oracle_lof_novelty(X, n, d, Q, m, int(k), int(normalize), out, idx.ctypes.data,
                                    d2.ctypes.data)
    if rc != 0:
        raise ValueError("invalid arguments to oracle_lof_novelty")
    if full:
        return {"scores": out, "idx": idx, "d2": d2}
    return out
