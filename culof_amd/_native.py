"""ctypes view of the C ABI in ``libculof_amd.so`` (``include/culof_c.h``).

The pybind11 module covers the host-array entry point; this module exposes
the device-pointer entry points (used by the torch binding and the GPU tests)
and the kernel-timing hooks (used by the benchmark). Status codes become
exceptions: ``CULOF_INVALID_ARGUMENT`` -> ``ValueError``, anything else ->
``culof_amd.CudaError``.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libculof_amd.so")

CULOF_OK = 0
CULOF_INVALID_ARGUMENT = 1
CULOF_DEVICE_ERROR = 2
CULOF_NO_DEVICE = 3

FORM_PLAIN = 0
FORM_BF16 = 1
FORM_FP64_TIES = 256  # flag OR-ed into a form

# Every symbol include/culof_c.h declares, with (restype, argtypes).
_c_int, _c_long, _c_double = ctypes.c_int, ctypes.c_long, ctypes.c_double
_vp, _cp, _sz = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t
SIGNATURES = {
    "culof_last_error": (_cp, []),
    "culof_device_available": (_c_int, []),
    "culof_device_info": (_c_int, [_cp, _sz]),
    "culof_lof": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp]),
    "culof_lof_device": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "culof_lof_device_noalloc": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "culof_lof_novelty_device": (_c_int, [_vp, _c_int, _c_int, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp,
                                          _vp, _vp, _vp]),
    "culof_lof_backward_device": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "culof_stage_center": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "culof_stage_knn": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "culof_stage_select_rows": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "culof_stage_lrd": (_c_int, [_vp, _vp, _vp, _c_int, _c_int, _vp, _vp]),
    "culof_stage_scores": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _vp]),
    "culof_choose_tile_rows": (_c_int, [_c_int, _sz]),
    "culof_profile_enable": (_c_int, [_c_int]),
    "culof_profile_reset": (_c_int, []),
    "culof_profile_read": (_c_int, [_cp, ctypes.POINTER(_c_double), ctypes.POINTER(_c_long)]),
    "culof_profile_tile_pairs": (_c_int, [ctypes.POINTER(ctypes.c_longlong)]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library, with every declared signature attached."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing; build it with `python -m culof_amd.build`"
            )
        handle = ctypes.CDLL(LIB_PATH)
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = handle
    return _lib


def check(rc: int) -> None:
    """Raise the Python exception matching a C ABI status code."""
    if rc == CULOF_OK:
        return
    message = lib().culof_last_error().decode(errors="replace")
    if rc == CULOF_INVALID_ARGUMENT:
        raise ValueError(message)
    from . import CudaError

    raise CudaError(message)


def ptr(tensor) -> int | None:
    """Device address of a torch tensor (None for None)."""
    return None if tensor is None else int(tensor.data_ptr())


def current_stream() -> int:
    """The hipStream_t torch is currently issuing work on, as an integer."""
    import torch

    return int(torch.cuda.current_stream().cuda_stream)


def profile_enable(on: bool = True) -> None:
    check(lib().culof_profile_enable(1 if on else 0))


def profile_reset() -> None:
    check(lib().culof_profile_reset())


def profile_read(name: str) -> tuple[float, int]:
    """(summed milliseconds, launches) of the library kernels matching ``name``."""
    ms = ctypes.c_double()
    launches = ctypes.c_long()
    check(lib().culof_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(launches)))
    return ms.value, launches.value


def profile_tile_pairs() -> int:
    """Query x candidate tile pairs (32 x 32 rows) the symmetric passes screened
    while timing was on, since the last reset."""
    v = ctypes.c_longlong()
    check(lib().culof_profile_tile_pairs(ctypes.byref(v)))
    return v.value
