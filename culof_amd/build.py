"""In-tree build of the native parts of culof_amd.

* ``libculof_amd.so``: every ``csrc/*.hip`` translation unit compiled by hipcc
  for gfx950 and linked into one shared library exporting the C ABI
  (``include/culof_c.h``) and the C++ API (``include/culof.h``).
* ``_culof*.so``: the pybind11 module (``csrc/binding.cpp``), host code only,
  built with g++ and linked against ``libculof_amd.so`` (rpath ``$ORIGIN``).

Run ``python -m culof_amd.build`` or call :func:`build`. Objects go to
``culof_amd/_build/`` (git-ignored); the two ``.so`` files land beside this
file so the package loads them in place.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libculof_amd.so")
ARCH = os.environ.get("CULOF_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    # Packed FP32 VALU ops stay off: plain float code is compiled to scalar ops.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Keep MFMA accumulators in VGPRs: the screen reads every result on the
    # VALU, and AGPR copies would cost two moves per element per tile.
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall",
    "-Wno-unused-result",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    # Surface warnings other than hipcc's host-pass note about the GPU-only feature flag.
    noise = ("packed-fp32-ops", "not a recognized feature", "warning generated", "warnings generated")
    for line in (proc.stdout + proc.stderr).splitlines():
        if line.strip() and not any(s in line for s in noise):
            sys.stderr.write(line + "\n")


def _compile_hip(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
    return obj


def build_library(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    sources = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile_hip, sources))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_binding() -> str:
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "_culof" + suffix)
    tmp = out + ".tmp"
    cmd = [
        "g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
        f"-I{pybind11.get_include()}",
        f"-I{sysconfig.get_paths()['include']}",
        f"-I{INCLUDE}",
        os.path.join(CSRC, "binding.cpp"),
        f"-L{HERE}", "-lculof_amd", "-Wl,-rpath,$ORIGIN",
        "-o", tmp,
    ]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build(jobs: int = 8) -> None:
    build_library(jobs)
    build_binding()


if __name__ == "__main__":
    build()
    print(LIB)
