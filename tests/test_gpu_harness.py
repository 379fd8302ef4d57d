"""The sweep benchmark and the accuracy report on the GPU.

The harness runs culof_amd against scikit-learn with agreement enforced
(benchmarks/benchmark.py:94-112); the accuracy report checks the figures the
reference publishes (benchmarks/figures.py:218-267) against tighter
targets: median relative difference <= 1e-6, p99 <= 5e-6 and an
identical top 1%.
"""

from __future__ import annotations

import numpy as np
import pytest

from culof_amd import accuracy as acc
from culof_amd import benchmark as bm

pytestmark = pytest.mark.gpu


def test_sweeps_agree_with_sklearn():
    rows = bm.run_sweeps(k=20, repeats=1, max_n=5000, n_jobs=16, sizes=(1_000, 5_000),
                         dims=(2, 8, 32, 128), n_dims=5_000)
    print(bm.HEADER)
    for r in rows:
        print(bm.format_row(r))
        assert r.max_rel_err <= bm.MAX_REL_ERR and r.rank_agreement == 1.0
        assert r.culof_s > 0 and r.points_per_s > 0
    assert [r.n_features for r in rows] == [8, 8, 2, 8, 32, 128]


@pytest.mark.parametrize("form", ["fp32", "bf16"])
def test_accuracy_report_meets_targets(form):
    X = bm.blobs_with_outliers(20_000, 8, seed=7)
    rep = acc.report(X, 20, form=form)
    print(rep.text())
    s = rep.stats
    assert s["median_rel"] <= 1e-6
    assert s["p99_rel"] <= 5e-6
    assert s["top1pct_agreement"] == 1.0
    assert 0 <= rep.n_attributed <= rep.n_above
