"""culof_lof_device_noalloc: the forward with every scratch buffer in a fixed
device array, no allocator called.

Checked bit for bit against culof_lof_device (whose own parity with the
oracle the other GPU tests pin) on two calls in a row, the second with a
smaller input, so stale contents of the shared array cannot leak into a
result; over the screen forms, the FP64 tie rule, normalising, shapes served
by the radix path, and omitted optional outputs. Oversized calls are refused
before anything runs, and the call records into a HIP graph whose replays
match the eager result.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import make_data, make_tied
from culof_amd import _native

pytestmark = pytest.mark.gpu

KEYS = ("scores", "idx", "d2", "kdist2", "lrd")


def _call(entry, x, n, d, k, form=_native.FORM_PLAIN, normalize=False, outputs=True):
    out = {
        "scores": torch.full((n,), float("nan"), device=x.device),
        "idx": torch.full((k, n), -7, dtype=torch.int32, device=x.device),
        "d2": torch.full((k, n), float("nan"), device=x.device),
        "kdist2": torch.full((n,), float("nan"), device=x.device),
        "lrd": torch.full((n,), float("nan"), device=x.device),
    }
    extra = [out[key].data_ptr() if outputs else None for key in KEYS[1:]]
    _native.check(getattr(_native.lib(), entry)(
        x.data_ptr(), n, d, k, int(normalize), form, out["scores"].data_ptr(), *extra,
        _native.current_stream()))
    if not outputs:
        for key in KEYS[1:]:
            del out[key]
    return out


def _host(out):
    torch.cuda.synchronize()
    return {key: t.cpu().numpy() for key, t in out.items()}


def _same(got, want, what):
    for key in got:
        g = np.ascontiguousarray(got[key]).view(np.uint32)
        w = np.ascontiguousarray(want[key]).view(np.uint32)
        bad = np.flatnonzero(g.ravel() != w.ravel())
        assert bad.size == 0, f"{what} {key}: {bad.size} differences, first at {bad[:5]}"


CASES = [
    # (name, big input, small input, k, form, normalize, outputs)
    ("plain-d8", (20_000, 8, 200), (3_000, 8, 30), 20, _native.FORM_PLAIN, False, True),
    ("bf16-d32", (12_000, 32, 120), (2_500, 32, 25), 20, _native.FORM_BF16, False, True),
    ("plain-d128-k50", (6_000, 128, 60), (1_500, 128, 15), 50, _native.FORM_PLAIN, False, True),
    ("normalize-d5", (9_000, 5, 90), (800, 5, 8), 10, _native.FORM_PLAIN, True, True),
    ("radix-k80", (4_000, 6, 40), (1_000, 6, 10), 80, _native.FORM_PLAIN, False, True),
    ("radix-d150", (3_000, 150, 30), (700, 150, 7), 12, _native.FORM_PLAIN, False, True),
    ("no-optional-outputs", (15_000, 16, 150), (2_000, 16, 20), 20, _native.FORM_PLAIN, False, False),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_two_calls_match_the_allocating_entry_point(case):
    name, big, small, k, form, normalize, outputs = case
    inputs = [make_data(n, d, seed=n + d, n_outliers=o) for n, d, o in (big, small)]
    xs = [torch.from_numpy(X).cuda() for X in inputs]
    got = []
    for x, X in zip(xs, inputs):  # two calls in a row, the second smaller
        print(f"no-alloc {name}: n={X.shape[0]} d={X.shape[1]} k={k} form={form}", flush=True)
        got.append(_host(_call("culof_lof_device_noalloc", x, *X.shape, k, form, normalize, outputs)))
    for x, X, g in zip(xs, inputs, got):
        want = _host(_call("culof_lof_device", x, *X.shape, k, form, normalize, outputs))
        _same(g, want, f"{name} n={X.shape[0]}")


def test_fp64_tie_rule_matches():
    X1 = make_tied(8_000, 8, seed=3)
    X2 = make_tied(1_200, 8, seed=4)
    form = _native.FORM_PLAIN | _native.FORM_FP64_TIES
    xs = [torch.from_numpy(X).cuda() for X in (X1, X2)]
    got = [_host(_call("culof_lof_device_noalloc", x, *X.shape, 20, form)) for x, X in zip(xs, (X1, X2))]
    for x, X, g in zip(xs, (X1, X2), got):
        _same(g, _host(_call("culof_lof_device", x, *X.shape, 20, form)), f"ties n={X.shape[0]}")


def test_oversized_call_is_refused_before_running():
    n, d, k = 400_000, 128, 20
    x = torch.zeros((n, d), dtype=torch.float32, device="cuda")
    scores = torch.full((n,), -1.0, device="cuda")
    rc = _native.lib().culof_lof_device_noalloc(x.data_ptr(), n, d, k, 0, _native.FORM_PLAIN,
                                                scores.data_ptr(), None, None, None, None,
                                                _native.current_stream())
    assert rc == _native.CULOF_INVALID_ARGUMENT
    msg = _native.lib().culof_last_error().decode()
    assert "no-allocation" in msg and "MiB" in msg, msg
    torch.cuda.synchronize()
    assert bool((scores == -1.0).all())  # nothing ran


def test_argument_errors_match_the_allocating_entry_point():
    x = torch.zeros((10, 3), dtype=torch.float32, device="cuda")
    s = torch.empty(10, device="cuda")
    for args in ((x.data_ptr(), 10, 3, 10), (None, 10, 3, 2), (x.data_ptr(), 1, 3, 1)):
        rcs = [getattr(_native.lib(), e)(*args, 0, 0, s.data_ptr(), None, None, None, None,
                                          _native.current_stream())
               for e in ("culof_lof_device", "culof_lof_device_noalloc")]
        assert rcs[0] == rcs[1] == _native.CULOF_INVALID_ARGUMENT


def test_records_into_a_graph():
    """The whole call captures into a HIP graph (no allocation, no memset
    node); every replay, with eager work and synchronisations between them,
    equals the eager result."""
    X = make_data(6_000, 8, seed=77, n_outliers=60)
    n, d = X.shape
    x = torch.from_numpy(X).cuda()
    want = _host(_call("culof_lof_device_noalloc", x, n, d, 20))
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        out = _call("culof_lof_device_noalloc", x, n, d, 20)
    torch.cuda.synchronize()
    for replay in range(4):
        for t in out.values():
            t.fill_(float("nan") if t.dtype == torch.float32 else -7)
        other = torch.from_numpy(make_data(1_000, 8, seed=replay)).cuda()
        _host(_call("culof_lof_device_noalloc", other, 1_000, 8, 20))  # eager work on the array
        graph.replay()
        _same(_host(out), want, f"replay {replay}")
