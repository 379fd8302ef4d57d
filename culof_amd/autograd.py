"""Differentiable LOF on device tensors: ``torch.autograd.Function`` over the C ABI.

The forward is ``culof_lof_device`` (the same kernels and bits as :func:`culof_amd.lof`),
the backward is the hand-written ``culof_lof_backward_device``: the gradient of
``sum_i grad_i * LOF_i`` with respect to the points, with the kNN sets and the
reach-distance ``max`` branches held at what the forward chose. PyTorch provides
device memory and the stream only; there is no torch fallback for either pass.

    >>> X = torch.randn(50_000, 32, device="cuda", requires_grad=True)
    >>> scores = lof_torch(X, k=20)
    >>> (w * scores).sum().backward()          # X.grad = d(sum w LOF)/dX

The reference is forward-only (src/cuda/pipeline.cu:131-207); this module is an
addition beside its API.
"""

from __future__ import annotations

from typing import NamedTuple

import torch

from . import _native

FORMS = {"fp32": _native.FORM_PLAIN, "bf16": _native.FORM_BF16}
TIES = {"index": 0, "fp64": _native.FORM_FP64_TIES}


def form_code(form: str | int, ties: str = "index") -> int:
    """Form word of the C ABI.

    ``form`` is the distance screen: ``"fp32"`` (plain) or ``"bf16"`` (BF16
    matrix-core screen, FP32 admission); both give identical kNN sets.
    ``ties`` is ``"index"`` (the reference's rule: equal FP32 keys go to the
    lower index) or ``"fp64"`` (rows whose k/(k+1) boundary is below FP32
    resolution are re-selected on FP64 distances, (distance, index) order).
    An int that is already a valid form word passes through unchanged.
    """
    if isinstance(form, int):
        screen = form & 0xFF
        if form & ~(0xFF | _native.FORM_FP64_TIES) or screen not in FORMS.values():
            raise ValueError(f"invalid form word {form:#x}")
        code = form
    else:
        try:
            code = FORMS[form]
        except (KeyError, TypeError):
            raise ValueError(f"form must be one of {sorted(FORMS)}, got {form!r}") from None
    try:
        return code | TIES[ties]
    except (KeyError, TypeError):
        raise ValueError(f"ties must be one of {sorted(TIES)}, got {ties!r}") from None


class Neighbours(NamedTuple):
    """Forward state on the device, in the library's transposed layout."""

    scores: torch.Tensor  # (n,) float32
    idx: torch.Tensor  # (k, n) int32, slot t of point i at [t, i]
    d2: torch.Tensor  # (k, n) float32 squared distances
    kdist2: torch.Tensor  # (n,) float32 squared k-distance
    lrd: torch.Tensor  # (n,) float32


def _check_points(X: torch.Tensor, k: int) -> torch.Tensor:
    if not isinstance(X, torch.Tensor) or X.device.type != "cuda":
        raise ValueError("X must be a torch tensor on a GPU device")
    if X.dim() != 2:
        raise ValueError(f"expected a 2-D tensor of shape (n_samples, n_features), got {tuple(X.shape)}")
    n = X.shape[0]
    if k < 1 or k >= n:
        raise ValueError(f"k must satisfy 1 <= k <= n_points - 1 (got k={k}, n_points={n})")
    return X.detach().to(torch.float32).contiguous()


def forward_state(X: torch.Tensor, k: int = 20, normalize: bool = False,
                  form: str | int = "fp32", ties: str = "index") -> Neighbours:
    """Run the forward on device tensor ``X`` and return scores plus the kNN state."""
    x = _check_points(X, k)
    n, d = x.shape
    dev = x.device
    out = Neighbours(
        torch.empty(n, dtype=torch.float32, device=dev),
        torch.empty((k, n), dtype=torch.int32, device=dev),
        torch.empty((k, n), dtype=torch.float32, device=dev),
        torch.empty(n, dtype=torch.float32, device=dev),
        torch.empty(n, dtype=torch.float32, device=dev),
    )
    with torch.cuda.device(dev):
        _native.check(_native.lib().culof_lof_device(
            x.data_ptr(), n, d, k, int(bool(normalize)), form_code(form, ties), out.scores.data_ptr(),
            out.idx.data_ptr(), out.d2.data_ptr(), out.kdist2.data_ptr(), out.lrd.data_ptr(),
            _native.current_stream()))
    return out


def backward_points(X: torch.Tensor, state: Neighbours, grad_scores: torch.Tensor,
                    normalize: bool = False) -> torch.Tensor:
    """d(sum grad_scores * LOF)/dX for the kNN state of a forward on the same ``X``."""
    x = X.detach().to(torch.float32).contiguous()
    n, d = x.shape
    k = state.idx.shape[0]
    g = grad_scores.detach().to(device=x.device, dtype=torch.float32).contiguous()
    if g.shape != (n,):
        raise ValueError(f"grad_scores must have shape ({n},), got {tuple(g.shape)}")
    out = torch.empty_like(x)
    with torch.cuda.device(x.device):
        _native.check(_native.lib().culof_lof_backward_device(
            x.data_ptr(), n, d, k, int(bool(normalize)), state.idx.data_ptr(), state.d2.data_ptr(),
            state.kdist2.data_ptr(), g.data_ptr(), out.data_ptr(), _native.current_stream()))
    return out


class LOFFunction(torch.autograd.Function):
    """``scores = LOFFunction.apply(X, k, normalize, form)``; differentiable in ``X``."""

    @staticmethod
    def forward(ctx, X, k, normalize, form):  # noqa: D102
        state = forward_state(X, k, normalize, form)
        ctx.save_for_backward(X.detach(), state.idx, state.d2, state.kdist2)
        ctx.normalize = bool(normalize)
        return state.scores

    @staticmethod
    def backward(ctx, grad_scores):  # noqa: D102
        X, idx, d2, kdist2 = ctx.saved_tensors
        state = Neighbours(None, idx, d2, kdist2, None)
        grad = backward_points(X, state, grad_scores, ctx.normalize)
        return grad.to(X.dtype), None, None, None


def lof_torch(X: torch.Tensor, k: int = 20, normalize: bool = False,
              form: str | int = "fp32", ties: str = "index") -> torch.Tensor:
    """LOF scores of the rows of device tensor ``X`` as a differentiable tensor.

    Same scores as :func:`culof_amd.lof` on the same float32 data; ``form``
    picks the distance screen (``"fp32"`` or ``"bf16"``, identical results),
    ``ties`` the tie rule (see :func:`form_code`).
    """
    return LOFFunction.apply(X, int(k), bool(normalize), form_code(form, ties))
