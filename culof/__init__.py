"""``import culof`` for code written against the reference package.

Re-exports the public surface of :mod:`culof_amd` (the MI355X build) under the
reference's package name, so ``from culof import LOF, lof`` keeps working.
"""

from culof_amd import LOF, CudaError, __version__, cuda_available, device_info, lof

__all__ = ["LOF", "CudaError", "__version__", "cuda_available", "device_info", "lof"]
