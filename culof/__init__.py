"""``import culof`` for code written against the reference package.

Re-exports the public surface of :mod:`culof_amd` (the MI355X build) under the
reference's package name, so ``from culof import LOF, lof`` keeps working.
``CudaError`` comes from the native extension and, as in :mod:`culof_amd`,
loads on first use.
"""

from culof_amd import LOF, __version__, cuda_available, device_info, lof

__all__ = ["LOF", "CudaError", "__version__", "cuda_available", "device_info", "lof"]


def __getattr__(name: str):
    if name == "CudaError":
        import culof_amd

        return culof_amd.CudaError
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
