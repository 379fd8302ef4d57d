"""Sweep benchmark of culof_amd against scikit-learn, agreement enforced.

The reference's harness (benchmarks/benchmark.py): a sample-count sweep at
d = 8 and a dimensionality sweep at n = 20,000, each configuration timed as
the best of ``repeats`` calls after one warm-up for both libraries, and each
checked against scikit-learn in the same run -- a configuration whose scores
disagree (max relative error above MAX_REL_ERR, or top-1% ranking agreement
below MIN_RANK_AGREEMENT) raises instead of printing a speedup
(benchmark.py:94-112). One-time runtime initialisation is measured once and
reported apart (benchmark.py:175-181).

Rows keep the reference's schema (benchmark.py:43-56) and add points/s, the
median and p99 relative errors and the screen form. The inputs are the
reference's benchmark distribution (4 unit Gaussian clusters, centres
U(-8, 8)^d, 1% uniform outliers in [-20, 20]^d, outliers first), drawn in the
same order from the same seeded generator, so rows are comparable with the
reference's published tables.

    python -m culof_amd.benchmark [--max-n 200000] [--form bf16] [--json out.json]
"""

from __future__ import annotations

import argparse
import json
import platform
import subprocess
import sys
import time
from dataclasses import asdict, dataclass
from typing import Callable

import numpy as np

from .accuracy import relative_errors, top_agreement

# The reference's agreement limits (benchmark.py:37-41).
MAX_REL_ERR = 0.05
MIN_RANK_AGREEMENT = 1.0

SAMPLE_SIZES = (1_000, 2_000, 5_000, 10_000, 20_000, 50_000, 100_000, 200_000)
DIMS = (2, 8, 32, 128)


class DisagreementError(RuntimeError):
    """Scores disagree with scikit-learn beyond the harness limits."""


@dataclass
class Row:
    sweep: str            # "samples" or "features" (both contain n=20k, d=8)
    n_samples: int
    n_features: int
    k: int
    sklearn_s: float
    culof_s: float
    speedup: float
    max_rel_err: float
    rank_agreement: float
    points_per_s: float
    median_rel_err: float
    p99_rel_err: float
    form: str


def blobs_with_outliers(n: int, d: int, seed: int = 0) -> np.ndarray:
    """float32 rows: max(1, n // 100) uniform outliers in [-20, 20]^d first,
    then points of 4 unit Gaussian clusters with centres U(-8, 8)^d and
    uniformly drawn labels (the reference's benchmark data, benchmark.py:59-66)."""
    rng = np.random.default_rng(seed)
    n_out = max(1, n // 100)
    n_in = n - n_out
    centres = rng.uniform(-8, 8, size=(4, d))
    labels = rng.integers(0, 4, size=n_in)
    inliers = (centres[labels] + rng.standard_normal((n_in, d))).astype(np.float32)
    outliers = rng.uniform(-20, 20, size=(n_out, d)).astype(np.float32)
    return np.ascontiguousarray(np.concatenate([outliers, inliers]), dtype=np.float32)


def best_of(fn: Callable[[], object], repeats: int) -> float:
    """Fastest of ``repeats`` timed calls after one untimed warm-up (the
    minimum, as the reference takes it: benchmark.py:69-79)."""
    fn()
    best = float("inf")
    for _ in range(max(1, repeats)):
        t0 = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t0)
    return best


def sklearn_scorer(k: int, n_jobs: int) -> Callable[[np.ndarray], np.ndarray]:
    from sklearn.neighbors import LocalOutlierFactor

    def run(X: np.ndarray) -> np.ndarray:
        return -LocalOutlierFactor(n_neighbors=k, n_jobs=n_jobs).fit(X).negative_outlier_factor_

    return run


def culof_scorer(k: int, form: str) -> Callable[[np.ndarray], np.ndarray]:
    from . import lof

    def run(X: np.ndarray) -> np.ndarray:
        return lof(X, k, form=form)

    return run


def check_agreement(got: np.ndarray, want: np.ndarray, n: int, d: int, k: int) -> tuple[float, float]:
    """(max relative error, top-1% agreement); raises DisagreementError past
    the limits, or when a score is not finite."""
    rel = relative_errors(got, want)
    worst = float(rel.max()) if len(rel) else 0.0
    agree = top_agreement(got, want)
    if not np.isfinite(worst) or worst > MAX_REL_ERR or agree < MIN_RANK_AGREEMENT:
        raise DisagreementError(
            f"disagreement with scikit-learn at n={n}, d={d}, k={k}: max relative error "
            f"{worst:.3e} (limit {MAX_REL_ERR:.0e}), top-1% rank agreement {agree:.3f} "
            f"(limit {MIN_RANK_AGREEMENT:.2f})"
        )
    return worst, agree


def bench_one(sweep: str, n: int, d: int, k: int, repeats: int, n_jobs: int = -1,
              form: str = "fp32", scorer: Callable[[np.ndarray], np.ndarray] | None = None,
              reference: Callable[[np.ndarray], np.ndarray] | None = None) -> Row:
    """Time and check one configuration (``scorer`` / ``reference`` replace
    culof_amd / scikit-learn, e.g. in tests)."""
    X = blobs_with_outliers(n, d)
    ours = scorer or culof_scorer(k, form)
    theirs = reference or sklearn_scorer(k, n_jobs)
    t_ref = best_of(lambda: theirs(X), repeats)
    t_ours = best_of(lambda: ours(X), repeats)
    got, want = ours(X), theirs(X)
    worst, agree = check_agreement(got, want, n, d, k)
    rel = relative_errors(got, want)
    return Row(sweep, n, d, k, t_ref, t_ours, t_ref / t_ours, worst, agree, n / t_ours,
               float(np.median(rel)), float(np.percentile(rel, 99)), form)


def run_sweeps(k: int = 20, repeats: int = 7, max_n: int = 50_000, n_jobs: int = -1,
               form: str = "fp32", sizes: tuple[int, ...] = SAMPLE_SIZES,
               dims: tuple[int, ...] = DIMS, n_dims: int = 20_000,
               show: Callable[[Row], None] | None = None, **kw) -> list[Row]:
    """The sample sweep (d = 8, every size up to ``max_n``) then the feature
    sweep (n = min(n_dims, max_n)); every row passes check_agreement."""
    rows = []
    for n in (s for s in sizes if s <= max_n):
        rows.append(bench_one("samples", n, 8, k, repeats, n_jobs, form, **kw))
        if show:
            show(rows[-1])
    nd = min(n_dims, max_n)
    for d in dims:
        rows.append(bench_one("features", nd, d, k, repeats, n_jobs, form, **kw))
        if show:
            show(rows[-1])
    return rows


def payload(rows: list[Row], meta: dict) -> dict:
    """The JSON document: run metadata plus one object per row."""
    return {**meta, "rows": [asdict(r) for r in rows]}


def cpu_name() -> str:
    try:
        out = subprocess.check_output(["lscpu"], text=True)
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:  # pragma: no cover - lscpu missing
        pass
    return platform.processor() or "unknown"


HEADER = (f"{'n':>8} {'d':>4} {'sklearn (s)':>12} {'culof (s)':>11} {'speedup':>9} "
          f"{'pts/s':>10} {'max rel err':>12} {'top-1% agree':>13}")


def format_row(r: Row) -> str:
    return (f"{r.n_samples:>8} {r.n_features:>4} {r.sklearn_s:>12.4f} {r.culof_s:>11.5f} "
            f"{r.speedup:>8.1f}x {r.points_per_s:>10.3g} {r.max_rel_err:>12.2e} {r.rank_agreement:>12.0%}")


def main(argv: list[str] | None = None) -> int:
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--k", type=int, default=20)
    p.add_argument("--repeats", type=int, default=7)
    p.add_argument("--max-n", type=int, default=50_000)
    p.add_argument("--n-jobs", type=int, default=-1, help="cores for scikit-learn")
    p.add_argument("--form", choices=("fp32", "bf16"), default="fp32")
    p.add_argument("--json", type=str, default=None)
    args = p.parse_args(argv)

    from . import __version__, cuda_available, device_info, lof

    if not cuda_available():
        print("No HIP device available.", file=sys.stderr)
        return 1
    # One-time runtime and module initialisation, before anything else.
    warm = np.zeros((64, 2), dtype=np.float32)
    warm[:, 0] = np.arange(64)
    t0 = time.perf_counter()
    lof(warm, 5)
    init_s = time.perf_counter() - t0
    print(f"GPU : {device_info()}")
    print(f"CPU : {cpu_name()}")
    print(f"culof_amd {__version__}, k={args.k}, form={args.form}, best of {args.repeats} runs")
    print(f"one-time HIP initialisation: {init_s * 1e3:.0f} ms (excluded below)\n")
    print(HEADER)
    print("-" * len(HEADER))
    rows = run_sweeps(args.k, args.repeats, args.max_n, args.n_jobs, args.form,
                      show=lambda r: print(format_row(r), flush=True))
    if args.json:
        meta = {"gpu": device_info(), "cpu": cpu_name(), "culof_version": __version__, "k": args.k,
                "repeats": args.repeats, "form": args.form, "init_seconds": init_s,
                "limits": {"max_rel_err": MAX_REL_ERR, "min_rank_agreement": MIN_RANK_AGREEMENT}}
        with open(args.json, "w") as fh:
            json.dump(payload(rows, meta), fh, indent=2)
        print(f"\nwrote {args.json}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
