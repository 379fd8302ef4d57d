"""The build bootstraps from a source-only tree.

A checkout holds no built library (``*.so`` files are not versioned), so the
package and its builder must import without one, and the product path must
refuse to run until the library is built. The tree is copied without any
built file and exercised in a fresh interpreter.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCE_DIRS = ("culof_amd", "culof", "include", "oracle")


def _source_tree(dst: str) -> str:
    ignore = shutil.ignore_patterns("*.so", "*.o", "_build", "__pycache__", "_ref", "*.tmp")
    for name in SOURCE_DIRS:
        shutil.copytree(os.path.join(ROOT, name), os.path.join(dst, name), ignore=ignore)
    shutil.copy(os.path.join(ROOT, "__graft_entry__.py"), dst)
    return dst


def _run(tree: str, code: str) -> subprocess.CompletedProcess:
    return subprocess.run([sys.executable, "-c", code], cwd=tree, capture_output=True, text=True,
                          timeout=600)


@pytest.fixture(scope="module")
def tree(tmp_path_factory):
    t = _source_tree(str(tmp_path_factory.mktemp("clean")))
    for dirpath, _, files in os.walk(t):
        assert not [f for f in files if f.endswith(".so")], dirpath
    return t


def test_package_and_builder_import_without_native_library(tree):
    proc = _run(tree, "import culof_amd, culof, culof_amd.build, __graft_entry__\n"
                      "assert callable(culof_amd.build.build) and callable(__graft_entry__.build)\n"
                      "print(culof_amd.__version__)")
    assert proc.returncode == 0, proc.stderr


def test_product_path_refuses_to_run_unbuilt(tree):
    proc = _run(tree, "import numpy as np, culof_amd\n"
                      "try:\n"
                      "    culof_amd.lof(np.zeros((4, 2), np.float32) + np.arange(4)[:, None], 2)\n"
                      "except ImportError as e:\n"
                      "    print('refused:', e)\n"
                      "else:\n"
                      "    raise SystemExit('ran without the native library')")
    assert proc.returncode == 0, proc.stderr
    assert "not built" in proc.stdout


def test_builder_compiles_a_translation_unit_from_the_clean_tree(tree):
    if not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not installed")
    proc = _run(tree, "import os\n"
                      "from culof_amd import build\n"
                      "os.makedirs(build.OBJ, exist_ok=True)\n"
                      "print(build._compile_hip(os.path.join(build.CSRC, 'density.hip')))")
    assert proc.returncode == 0, proc.stderr
    assert os.path.exists(proc.stdout.strip().splitlines()[-1])
