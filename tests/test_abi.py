"""The drop-in boundary, checked without a GPU.

* libculof_amd.so loads and exports every function include/culof_c.h declares
  (parsed from the header, so a new declaration without a definition fails)
  and the C++ API of include/culof.h.
* Argument checking and error reporting match the reference
  (src/cuda/pipeline.cu:131-142, tests/cpp/test_culof.cu:198-212).
* The Python surface validates before touching the device
  (src/culof/__init__.py:37-80) and keeps the estimator contract
  (tests/python/test_api.py:83-128, tests/python/test_regressions.py:157-189).
No call here launches a kernel.
"""

from __future__ import annotations

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import culof
import culof_amd
from culof_amd import LOF, _native, lof

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_c_functions() -> list[str]:
    text = open(os.path.join(ROOT, "include", "culof_c.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"\b(culof_\w+)\s*\(", text)


def test_header_declarations_are_exported():
    names = _declared_c_functions()
    assert len(names) >= 12
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in names:
        assert hasattr(lib, name), f"{name} declared in culof_c.h but not exported"
    # and the Python binding knows every one of them
    assert set(names) == set(_native.SIGNATURES)


def test_cpp_api_is_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", _native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for sym in ("culof::lof(float const*, int, int, int, bool)", "culof::cuda_available()",
                "culof::device_info[abi:cxx11]()"):
        assert sym in out, sym


def test_choose_tile_rows_matches_reference_budget_math():
    lib = _native.lib()
    assert lib.culof_choose_tile_rows(1000, 40000) == 10
    assert lib.culof_choose_tile_rows(1000, 0) == 1
    assert lib.culof_choose_tile_rows(1000, 1 << 40) == 1000


def _skip_if_gpu():
    if culof_amd.cuda_available():
        pytest.skip("checks the no-device path")


def test_no_device_reports_cleanly():
    _skip_if_gpu()
    lib = _native.lib()
    assert lib.culof_device_available() == 0
    assert culof_amd.cuda_available() is False
    x = np.zeros((10, 2), np.float32)
    out = np.zeros(10, np.float32)
    rc = lib.culof_lof(x.ctypes.data, 10, 2, 3, 0, out.ctypes.data)
    assert rc == _native.CULOF_NO_DEVICE
    assert "no HIP device available" in lib.culof_last_error().decode()
    with pytest.raises(culof_amd.CudaError, match="no HIP device"):
        lof(np.random.default_rng(0).standard_normal((50, 3)), 5)


CASES = [
    (False, 50, 2, 5, "points must not be null"),
    (True, 1, 2, 1, "n_points must be >= 2"),
    (True, 50, 0, 5, "n_dims must be >= 1"),
    (True, 50, 2, 0, "k must satisfy 1 <= k <= n_points - 1 (got k=0, n_points=50)"),
    (True, 50, 2, 50, "k must satisfy 1 <= k <= n_points - 1 (got k=50, n_points=50)"),
]


@pytest.mark.parametrize("has_points,n,d,k,msg", CASES)
def test_reference_argument_checks(has_points, n, d, k, msg):
    lib = _native.lib()
    x = np.zeros((50, 2), np.float32)
    out = np.zeros(50, np.float32)
    rc = lib.culof_lof(x.ctypes.data if has_points else None, n, d, k, 0, out.ctypes.data)
    assert rc == _native.CULOF_INVALID_ARGUMENT
    assert lib.culof_last_error().decode() == msg
    with pytest.raises(ValueError, match=re.escape(msg)):
        _native.check(rc)


def test_device_entry_points_validate_before_the_device():
    lib = _native.lib()
    rc = lib.culof_lof_device(None, 100, 3, 5, 0, 0, None, None, None, None, None, None)
    assert rc == _native.CULOF_INVALID_ARGUMENT
    rc = lib.culof_stage_select_rows(None, 1, 1, 1, None, None, None, None)
    assert rc == _native.CULOF_INVALID_ARGUMENT
    rc = lib.culof_stage_knn(None, 10, 2, 3, 0, 0, 0, 10, 0, None, None, None, None)
    assert rc == _native.CULOF_INVALID_ARGUMENT


# ---- Python surface, validation only (reference tests/python/test_api.py) ----

def test_public_names():
    assert set(culof_amd.__all__) >= {"LOF", "__version__", "cuda_available", "device_info", "lof"}
    assert culof.LOF is culof_amd.LOF and culof.lof is culof_amd.lof
    assert culof_amd.__version__ == "0.2.0"


def test_unfitted_access_raises():
    with pytest.raises(ValueError, match="not fitted"):
        LOF().predict()
    with pytest.raises(ValueError, match="not fitted"):
        LOF().score_samples()


@pytest.mark.parametrize("bad", [0, -1])
def test_rejects_non_positive_n_neighbors(bad):
    with pytest.raises(ValueError, match="n_neighbors must be >= 1"):
        LOF(n_neighbors=bad)


@pytest.mark.parametrize("bad", [0.0, 0.9, 1.0, -0.1])
def test_rejects_bad_contamination(bad):
    with pytest.raises(ValueError, match="contamination"):
        LOF(contamination=bad)


@pytest.mark.parametrize("bad", [100, 500])
def test_rejects_n_neighbors_above_sample_count(bad):
    X = np.random.default_rng(1).standard_normal((100, 2)).astype(np.float32)
    with pytest.raises(ValueError, match="requires at least"):
        LOF(n_neighbors=bad).fit(X)


def test_rejects_wrong_shape():
    with pytest.raises(ValueError, match="2-D array"):
        lof(np.zeros(10, dtype=np.float32), 3)


def test_rejects_non_finite_input():
    X = np.ones((100, 2), dtype=np.float32)
    X[3, 1] = np.nan
    with pytest.raises(ValueError, match="NaN or infinity"):
        lof(X, 10)


def test_float32_overflow_raises_before_the_device():
    X = (np.random.default_rng(0).standard_normal((64, 4)) * 1e20).astype(np.float32)
    with pytest.raises(ValueError, match="overflow"):
        lof(X, 5)


def test_overflow_guard_leaves_ordinary_scales_alone():
    rng = np.random.default_rng(0)
    for scale in (1e-6, 1.0, 1e3, 1e9):
        X = (rng.standard_normal((64, 4)) * scale).astype(np.float32)
        culof_amd._validate(X, 5)  # does not raise


def test_repr_is_informative():
    assert repr(LOF(n_neighbors=7)) == "LOF(n_neighbors=7, contamination='auto', normalize=False)"


def test_predict_and_score_samples_take_no_X():
    model = LOF(n_neighbors=10)
    with pytest.raises(TypeError):
        model.predict(np.zeros((8, 3), np.float32))
    with pytest.raises(TypeError):
        model.score_samples(np.zeros((8, 3), np.float32))


def test_readme_does_not_claim_drop_in_replacement():
    """tests/python/test_regressions.py:179-189: the estimator contract is
    narrower than scikit-learn's, and the README says so."""
    readme = open(os.path.join(ROOT, "README.md")).read().lower()
    assert "not a drop-in replacement" in readme
    assert "a drop-in replacement for scikit-learn" not in readme


def test_version_has_one_source():
    from culof_amd import _version

    assert culof_amd.__version__ == _version.__version__
