"""culof_amd - Local Outlier Factor on an AMD Instinct MI355X.

Python surface of the reference package (src/culof/__init__.py:31): the
estimator :class:`LOF` (``fit``, ``fit_predict``, ``score_samples``,
``predict``, ``negative_outlier_factor_``, ``offset_``, ``n_neighbors_``, with
scikit-learn's sign conventions) and the function :func:`lof`. It covers the
transductive fit_predict workflow and is not a full scikit-learn estimator.

    >>> from culof_amd import LOF, lof
    >>> labels = LOF(n_neighbors=20).fit_predict(X)   # -1 outlier, 1 inlier
    >>> scores = lof(X, k=20)                         # ~1.0 normal, higher = more anomalous

Every call runs the HIP kernels of ``libculof_amd.so`` through the ``_culof``
extension; there is no CPU fallback. The extension loads on first use: any
call (and ``CudaError``) raises ImportError while the native library is not
built (``python -m culof_amd.build``), so the builder itself imports from an
unbuilt tree.
"""

from __future__ import annotations

import math
from typing import TYPE_CHECKING

import numpy as np

# One process holds one HIP runtime. PyTorch-ROCm ships its own copy of it under
# the same SONAME as /opt/rocm's; importing torch first makes libculof_amd bind
# to that copy rather than load a second runtime (which would then own the GPU
# alone), so torch tensors and this library share devices, memory and streams.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - numpy-only installs
    pass

from ._version import __version__

if TYPE_CHECKING:  # pragma: no cover
    from numpy.typing import ArrayLike, NDArray

__all__ = ["LOF", "CudaError", "__version__", "cuda_available", "device_info", "lof"]

_culof_module = None


def _ext():
    """The pybind11 extension ``_culof``, imported on first use."""
    global _culof_module
    if _culof_module is None:
        try:
            from . import _culof
        except ImportError as exc:
            raise ImportError(
                "culof_amd's native extension is not built; run `python -m culof_amd.build` "
                f"(original error: {exc})"
            ) from exc
        _culof_module = _culof
    return _culof_module


def cuda_available() -> bool:
    """True when a HIP device is present (reference: include/culof.h:28)."""
    return bool(_ext().cuda_available())


def device_info() -> str:
    """Name, architecture, CU count and memory of the current device
    (reference: include/culof.h:31)."""
    return str(_ext().device_info())


def _validate(X: ArrayLike, k: int) -> NDArray[np.float32]:
    """Coerce and check the input (reference: src/culof/__init__.py:37-55)."""
    arr = np.ascontiguousarray(X, dtype=np.float32)
    if arr.ndim != 2:
        raise ValueError(
            f"expected a 2-D array of shape (n_samples, n_features), got shape {arr.shape}"
        )
    n_samples, n_features = arr.shape
    if n_features < 1:
        raise ValueError("need at least one feature")
    if k < 1:
        raise ValueError(f"n_neighbors must be >= 1, got {k}")
    if k >= n_samples:
        raise ValueError(f"n_neighbors={k} requires at least {k + 1} samples, got {n_samples}")
    if not np.isfinite(arr).all():
        raise ValueError("input contains NaN or infinity")
    _check_representable(arr, n_features)
    return arr


def _check_representable(arr: NDArray[np.float32], n_features: int) -> None:
    """Reject input whose squared distances would overflow float32.

    After centring a coordinate is at most twice the input's largest magnitude,
    so a squared distance is at most ``n_features * (4 * max|x|)**2``. Beyond
    FLT_MAX every key would saturate to +inf and every score would be NaN,
    which the estimator would read as "inlier". The bound is compared in logs
    so the check cannot overflow itself (reference: src/culof/__init__.py:58-80).
    """
    if arr.size == 0:
        return
    max_abs = float(np.max(np.abs(arr)))
    if max_abs == 0.0:
        return
    log_bound = math.log(n_features) + 2.0 * math.log(4.0 * max_abs)
    if log_bound > math.log(float(np.finfo(np.float32).max)):
        raise ValueError(
            f"input magnitude {max_abs:.3g} overflows float32 when squared "
            f"({n_features} features): distances would saturate to infinity and "
            f"every score would be NaN. Rescale the data, or pass normalize=True "
            f"to z-score each feature first."
        )


def lof(X: ArrayLike, k: int = 20, normalize: bool = False, form: str = "fp32",
        ties: str = "index") -> NDArray[np.float32]:
    """Local Outlier Factor of every row of ``X`` (reference: src/culof/__init__.py:83-113).

    Parameters
    ----------
    X : array-like of shape (n_samples, n_features)
        Converted to C-contiguous float32 if it is not already.
    k : int, default=20
        Neighbours per point, excluding the point itself; ``1 <= k <= n_samples - 1``.
    normalize : bool, default=False
        Z-score each feature before computing distances. scikit-learn does not,
        so leave it off when comparing.
    form : {"fp32", "bf16"}, default="fp32"
        Precision form of the distance screen. ``"bf16"`` screens candidates
        with BF16 matrix-core products and admits them on exact FP32 distances;
        it returns the same scores as ``"fp32"``. (Not in the reference.)
    ties : {"index", "fp64"}, default="index"
        ``"index"`` keeps the reference's rule (equal FP32 distances go to the
        lower index). ``"fp64"`` re-selects the neighbours of rows whose k-th
        and (k+1)-th FP32 distances are closer than FP32 resolution using FP64
        distances, (distance, index) order. (Not in the reference.)

    Returns
    -------
    ndarray of shape (n_samples,), float32
        About 1.0 for a point as densely surrounded as its neighbours, larger
        for one in a comparatively sparse region.
    """
    arr = _validate(X, k)
    if form == "fp32" and ties == "index":
        scores = _ext().lof(arr, k, normalize)
    else:
        from .autograd import form_code, forward_state

        code = form_code(form, ties)
        scores = forward_state(torch.from_numpy(arr).cuda(), k, normalize, code).scores.cpu().numpy()
    # A non-finite score would read as "not an outlier" downstream.
    if not np.isfinite(scores).all():
        raise ValueError(
            "LOF produced non-finite scores; this indicates float32 overflow or "
            "underflow in the distance computation. Rescale the data, or pass "
            "normalize=True."
        )
    return scores


def __getattr__(name: str):
    # The extension's exception type and the autograd binding (which needs
    # PyTorch-ROCm) load on first use.
    if name == "CudaError":
        return _ext().CudaError
    if name in ("lof_torch", "LOFFunction"):
        from . import autograd

        return getattr(autograd, name)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")


class LOF:
    """Unsupervised outlier detection with Local Outlier Factor on the GPU.

    The transductive ``fit_predict`` workflow of
    :class:`sklearn.neighbors.LocalOutlierFactor` with the same sign
    conventions (reference: src/culof/__init__.py:116-227). Beyond the
    reference: ``get_params`` / ``set_params`` (so ``sklearn.base.clone``
    works), ``kneighbors`` of the fitted samples, and scikit-learn's
    ``novelty=True`` mode (``score_samples(X)``, ``decision_function(X)`` and
    ``predict(X)`` on new points; ``fit_predict`` is then unavailable, as in
    scikit-learn). No metric selection (Euclidean only).

    Parameters
    ----------
    n_neighbors : int, default=20
        Neighbours used to measure local density; ``1 <= n_neighbors <= n_samples - 1``.
    contamination : {'auto'} or float, default='auto'
        Expected proportion of outliers, used to set :attr:`offset_`. ``'auto'``
        uses scikit-learn's threshold of -1.5.
    normalize : bool, default=False
        Z-score each feature before computing distances.
    form : {"fp32", "bf16"}, default="fp32"
        Precision form of the distance screen; both give the same scores.
    ties : {"index", "fp64"}, default="index"
        Tie rule at the k-th neighbour; see :func:`lof`. New points in
        novelty mode always take the index rule.
    novelty : bool, default=False
        Score new points against the fitted samples instead of the fitted
        samples themselves (scikit-learn's semantics).

    Attributes
    ----------
    negative_outlier_factor_ : ndarray of shape (n_samples,)
        Negated LOF of the fitted samples; the more negative, the more abnormal.
    offset_ : float
        Threshold on :attr:`negative_outlier_factor_` used by :meth:`predict`.
    n_neighbors_ : int
        The value of ``n_neighbors`` actually used.
    n_features_in_ : int
        Number of features of the fitted data.
    n_samples_fit_ : int
        Number of fitted samples.
    """

    _PARAMS = ("n_neighbors", "contamination", "normalize", "form", "ties", "novelty")

    def __init__(
        self,
        n_neighbors: int = 20,
        contamination: str | float = "auto",
        normalize: bool = False,
        form: str = "fp32",
        ties: str = "index",
        novelty: bool = False,
    ) -> None:
        self._check_params(n_neighbors, contamination)
        self.n_neighbors = n_neighbors
        self.contamination = contamination
        self.normalize = normalize
        self.form = form
        self.ties = ties
        self.novelty = novelty

    @staticmethod
    def _check_params(n_neighbors: int, contamination: str | float) -> None:
        if n_neighbors < 1:
            raise ValueError(f"n_neighbors must be >= 1, got {n_neighbors}")
        if contamination != "auto" and not 0.0 < float(contamination) <= 0.5:
            raise ValueError(f"contamination must be 'auto' or in (0, 0.5], got {contamination!r}")

    def get_params(self, deep: bool = True) -> dict:
        """The constructor parameters (scikit-learn's estimator protocol)."""
        del deep  # no nested estimators
        return {name: getattr(self, name) for name in self._PARAMS}

    def set_params(self, **params: object) -> LOF:
        """Set constructor parameters by name; returns ``self``. A refit is
        needed for them to take effect on the fitted attributes."""
        unknown = sorted(set(params) - set(self._PARAMS))
        if unknown:
            raise ValueError(f"invalid parameter(s) {unknown} for LOF; valid: {list(self._PARAMS)}")
        merged = {**self.get_params(), **params}
        self._check_params(merged["n_neighbors"], merged["contamination"])
        for name, value in params.items():
            setattr(self, name, value)
        return self

    def fit(self, X: ArrayLike, y: object = None) -> LOF:
        """Compute the LOF of every sample in ``X``. Returns ``self``."""
        del y  # accepted for scikit-learn call compatibility
        arr = _validate(X, self.n_neighbors)
        scores = lof(arr, self.n_neighbors, self.normalize, self.form, self.ties)
        self._fit_X = arr
        self.n_features_in_ = int(arr.shape[1])
        self.n_samples_fit_ = int(arr.shape[0])
        self.n_neighbors_ = self.n_neighbors
        self.negative_outlier_factor_ = -scores
        if self.contamination == "auto":
            self.offset_ = -1.5
        else:
            self.offset_ = float(
                np.percentile(self.negative_outlier_factor_, 100.0 * float(self.contamination))
            )
        return self

    def fit_predict(self, X: ArrayLike, y: object = None) -> NDArray[np.intp]:
        """Fit and label each sample: ``-1`` for outliers, ``1`` for inliers."""
        del y
        if self.novelty:
            raise AttributeError("fit_predict is not available when novelty=True; "
                                 "use fit(X).predict(X_new) on new data")
        return self.fit(X).predict()

    def score_samples(self, X: ArrayLike | None = None) -> NDArray[np.float32]:
        """Negated LOF; the lower, the more abnormal.

        Without novelty: of the fitted data, and no ``X`` is taken (scoring new
        points against each other would not be scikit-learn's novelty mode).
        With ``novelty=True``: of the new points ``X`` against the fitted data.
        """
        if not self.novelty:
            if X is not None:
                raise TypeError("score_samples takes no X unless novelty=True "
                                "(it scores the fitted samples)")
            return self._check_fitted()
        if X is None:
            raise TypeError("score_samples(X) needs the new points when novelty=True")
        self._check_fitted()
        return -self._novelty(X, self.n_neighbors_)[0]

    def decision_function(self, X: ArrayLike) -> NDArray[np.float32]:
        """``score_samples(X) - offset_`` for new points (novelty mode only):
        negative for outliers."""
        if not self.novelty:
            raise AttributeError("decision_function is only available when novelty=True")
        return self.score_samples(X) - np.float32(self._check_offset())

    def predict(self, X: ArrayLike | None = None) -> NDArray[np.intp]:
        """Label ``-1`` (outlier) or ``1`` (inlier): the fitted samples, or
        with ``novelty=True`` the new points ``X``."""
        if self.novelty:
            return np.where(self.decision_function(X) < 0, -1, 1)
        return np.where(self.score_samples(X) < self._check_offset(), -1, 1)

    def _novelty(self, X: ArrayLike, k: int, neighbours: bool = False):
        """(LOF of the new rows X against the fitted rows, and with
        ``neighbours`` their (k, m) neighbour indices and squared keys)."""
        q = np.ascontiguousarray(X, dtype=np.float32)
        if q.ndim != 2 or q.shape[1] != self.n_features_in_:
            raise ValueError(f"X has shape {q.shape}; expected (n, {self.n_features_in_})")
        if not np.isfinite(q).all():
            raise ValueError("X contains NaN or infinity")
        _check_representable(np.concatenate([self._fit_X, q]), self.n_features_in_)
        from . import _native
        from .autograd import form_code

        fit = torch.from_numpy(self._fit_X).cuda()
        new = torch.from_numpy(q).cuda()
        m = q.shape[0]
        out = torch.empty(m, dtype=torch.float32, device=fit.device)
        idx = torch.empty((k, m), dtype=torch.int32, device=fit.device) if neighbours else None
        d2 = torch.empty((k, m), dtype=torch.float32, device=fit.device) if neighbours else None
        _native.check(_native.lib().culof_lof_novelty_device(
            fit.data_ptr(), fit.shape[0], fit.shape[1], new.data_ptr(), m, k,
            int(bool(self.normalize)), form_code(self.form, self.ties), None, out.data_ptr(),
            _native.ptr(idx), _native.ptr(d2), _native.current_stream()))
        scores = out.cpu().numpy()
        if not np.isfinite(scores).all():
            raise ValueError("LOF produced non-finite scores for the new points; rescale the data")
        if not neighbours:
            return scores, None, None
        return scores, idx.cpu().numpy(), d2.cpu().numpy()

    def kneighbors(self, X: ArrayLike | None = None, n_neighbors: int | None = None,
                   return_distance: bool = True):
        """Nearest fitted samples, nearest first.

        With ``X=None``: of each fitted sample, itself excluded, the sets the
        LOF used (same tie rule and form). With ``X``: of each new row, nothing
        excluded (the index tie rule). Returns ``(distances, indices)`` of shape
        ``(n_rows, n_neighbors)`` (``indices`` alone when ``return_distance`` is
        false); distances are float32 Euclidean distances in the space the LOF
        was computed in (after z-scoring when ``normalize``), equal distances
        ordered by index.
        """
        self._check_fitted()
        k = self.n_neighbors_ if n_neighbors is None else int(n_neighbors)
        if not 1 <= k <= self.n_samples_fit_ - 1:
            raise ValueError(f"n_neighbors must satisfy 1 <= n_neighbors <= n_samples_fit_ - 1 "
                             f"(got {k}, n_samples_fit_={self.n_samples_fit_})")
        if X is not None:
            _, idx_t, d2_t = self._novelty(X, k, neighbours=True)
            idx = idx_t.T.astype(np.intp)
            d2 = np.ascontiguousarray(d2_t.T)
        else:
            from .autograd import forward_state

            state = forward_state(torch.from_numpy(self._fit_X).cuda(), k, self.normalize,
                                  self.form, self.ties)
            idx = state.idx.cpu().numpy().T.astype(np.intp)
            d2 = state.d2.cpu().numpy().T
        # (distance, index) order: non-negative float32 bit patterns sort like
        # the values, so one 64-bit key per entry orders both.
        key = (d2.view(np.uint32).astype(np.uint64) << np.uint64(32)) | idx.astype(np.uint64)
        order = np.argsort(key, axis=1)
        idx = np.take_along_axis(idx, order, axis=1)
        if not return_distance:
            return idx
        dist = np.sqrt(np.take_along_axis(d2, order, axis=1))
        return dist, idx

    def _check_fitted(self) -> NDArray[np.float32]:
        if not hasattr(self, "negative_outlier_factor_"):
            raise ValueError(
                "this LOF instance is not fitted yet - call fit() or fit_predict() first"
            )
        return self.negative_outlier_factor_

    def _check_offset(self) -> float:
        self._check_fitted()
        return self.offset_

    def __repr__(self) -> str:
        form = "" if self.form == "fp32" else f", form={self.form!r}"
        form += "" if self.ties == "index" else f", ties={self.ties!r}"
        form += ", novelty=True" if self.novelty else ""
        return (
            f"LOF(n_neighbors={self.n_neighbors}, "
            f"contamination={self.contamination!r}, normalize={self.normalize}{form})"
        )
