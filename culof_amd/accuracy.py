"""Agreement of culof_amd's scores with scikit-learn's FP64 LocalOutlierFactor.

The figures the reference publishes for its accuracy plot
(benchmarks/figures.py:218-267: median, p99 and max relative difference, and
the share of points above 1e-4), the top-1% ranking agreement its benchmark
enforces (benchmarks/benchmark.py:94-112), and what the reference only states
in prose (figures.py:285-292): that the tail comes from points whose k-th and
(k+1)-th neighbours are closer together than float32 resolves. Here that claim
is checked per point: every point above the threshold is attributed to a near
tie in its two-level neighbourhood or listed as unattributed.

A point's score depends on the k-distances of its neighbours (through their
reach-distances) and on their LRDs, which depend on the k-distances of *their*
neighbours; so a near tie at row j can move the score of i when j is i itself,
a neighbour of i, or a neighbour of one (two hops in the FP64 kNN graph).

    python -m culof_amd.accuracy --n 20000 --d 8 --k 20 [--form bf16] [--json out.json]
"""

from __future__ import annotations

import argparse
import json
import sys
from dataclasses import asdict, dataclass, field

import numpy as np

# Relative gap between the FP64 squared k-th and (k+1)-th neighbour distances
# below which FP32 keys can order the two either way: FP32 squared distances
# carry a relative error of a few units of 2^-24 per feature term, so 2^-20
# (about 1e-6) leaves margin for d up to a few hundred features.
NEAR_TIE_RTOL = 2.0 ** -20


def relative_errors(got: np.ndarray, want: np.ndarray) -> np.ndarray:
    """|got - want| / max(|want|, 1e-9), in float64 (figures.py:233)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    return np.abs(got - want) / np.maximum(np.abs(want), 1e-9)


def top_agreement(got: np.ndarray, want: np.ndarray, frac: float = 0.01) -> float:
    """Share of the top ``max(1, n * frac)`` points by ``want`` that are also
    the top points by ``got`` (benchmark.py:100-101 for frac = 0.01)."""
    n = len(want)
    top = max(1, int(n * frac))
    a = set(np.argsort(-np.asarray(got), kind="stable")[:top].tolist())
    b = set(np.argsort(-np.asarray(want), kind="stable")[:top].tolist())
    return len(a & b) / top


def score_stats(got: np.ndarray, want: np.ndarray, threshold: float = 1e-4) -> dict:
    """Summary of the relative differences of ``got`` against ``want``."""
    rel = relative_errors(got, want)
    g = np.asarray(got, dtype=np.float64)
    w = np.asarray(want, dtype=np.float64)
    pearson = float(np.corrcoef(g, w)[0, 1]) if len(g) > 1 and g.std() > 0 and w.std() > 0 else 1.0
    return {
        "n": int(len(rel)),
        "median_rel": float(np.median(rel)),
        "p99_rel": float(np.percentile(rel, 99)),
        "max_rel": float(rel.max()),
        "frac_above": float(np.mean(rel > threshold)),
        "threshold": threshold,
        "pearson_r": pearson,
        "top1pct_agreement": top_agreement(got, want),
    }


def fp64_neighbours(X: np.ndarray, k: int) -> tuple[np.ndarray, np.ndarray]:
    """FP64 squared distances and indices of each row's k + 1 nearest other
    rows (ascending), by scikit-learn's exact search."""
    from sklearn.neighbors import NearestNeighbors

    X64 = np.asarray(X, dtype=np.float64)
    nn = NearestNeighbors(n_neighbors=min(k + 2, len(X64))).fit(X64)
    dist, idx = nn.kneighbors(X64)
    # Drop each row's own entry (distance 0; with duplicates it may not come
    # first, so remove it by index, falling back to the last column).
    rows = np.arange(len(X64))[:, None]
    own = idx == rows
    keep = ~own
    missing = keep.all(axis=1)
    keep[missing, -1] = False
    idx = idx[keep].reshape(len(X64), -1)[:, : k + 1]
    dist = dist[keep].reshape(len(X64), -1)[:, : k + 1]
    return dist * dist, idx


def near_tie_rows(d2: np.ndarray, k: int, rtol: float = NEAR_TIE_RTOL) -> np.ndarray:
    """Rows whose k-th and (k+1)-th FP64 squared distances differ by at most
    ``rtol`` relative (``d2``: ascending, at least k + 1 columns)."""
    if d2.shape[1] <= k:
        return np.zeros(d2.shape[0], dtype=bool)
    kth = d2[:, k - 1]
    nxt = d2[:, k]
    return (nxt - kth) <= rtol * np.maximum(kth, np.finfo(np.float64).tiny)


def spread(flags: np.ndarray, nbrs: np.ndarray, hops: int = 2) -> np.ndarray:
    """Rows within ``hops`` steps of a flagged row along the kNN lists:
    row i is marked when it, or a row on its list, was marked one step before."""
    out = np.asarray(flags, dtype=bool).copy()
    for _ in range(hops):
        out = out | out[nbrs].any(axis=1)
    return out


@dataclass
class AccuracyReport:
    n: int
    d: int
    k: int
    form: str
    ties: str
    stats: dict
    n_above: int
    n_attributed: int
    near_tie_rows: int
    unattributed: list = field(default_factory=list)  # up to 20 row indices

    def text(self) -> str:
        s = self.stats
        lines = [
            f"n={self.n} d={self.d} k={self.k} form={self.form} ties={self.ties}",
            f"median     {s['median_rel']:.2e}",
            f"p99        {s['p99_rel']:.2e}",
            f"max        {s['max_rel']:.2e}",
            f"above {s['threshold']:.0e} {s['frac_above']:.3%} of points ({self.n_above})",
            f"pearson r  {s['pearson_r']:.8f}",
            f"top-1% agreement {s['top1pct_agreement']:.3f}",
            f"near-tie rows (FP64 k-th/(k+1)-th gap <= {NEAR_TIE_RTOL:.1e}): {self.near_tie_rows}",
            f"above-threshold points within two kNN hops of a near tie: "
            f"{self.n_attributed} of {self.n_above}",
        ]
        if self.unattributed:
            lines.append(f"unattributed (first {len(self.unattributed)}): {self.unattributed}")
        return "\n".join(lines)


def sklearn_scores(X: np.ndarray, k: int, dtype=np.float64) -> np.ndarray:
    """LOF of scikit-learn's LocalOutlierFactor fitted on ``X`` cast to ``dtype``.

    Fitted on float64 input, scikit-learn keeps its kNN distances and
    densities in FP64; fitted on float32 (as the reference's own comparisons
    do, tests/python/conftest.py:10-12) it casts the distances to float32
    first (sklearn/neighbors/_lof.py:294-298), a different comparator."""
    from sklearn.neighbors import LocalOutlierFactor

    est = LocalOutlierFactor(n_neighbors=k, n_jobs=-1)
    return -est.fit(np.asarray(X, dtype=dtype)).negative_outlier_factor_


def report(X: np.ndarray, k: int, scores: np.ndarray | None = None, form: str = "fp32",
           ties: str = "index", threshold: float = 1e-4,
           reference: np.ndarray | None = None, comparator: str = "fp64") -> AccuracyReport:
    """Accuracy report of culof_amd (or of ``scores``) on ``X`` against
    scikit-learn's LocalOutlierFactor fitted on float64 input (``comparator``
    "fp32": on float32 input, the reference's comparator), or against
    ``reference``."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    if scores is None:
        from . import lof

        scores = lof(X, k, form=form, ties=ties)
    if reference is None:
        if comparator not in ("fp64", "fp32"):
            raise ValueError(f"comparator must be 'fp64' or 'fp32', got {comparator!r}")
        reference = sklearn_scores(X, k, np.float64 if comparator == "fp64" else np.float32)
    stats = score_stats(scores, reference, threshold)
    rel = relative_errors(scores, reference)
    above = np.flatnonzero(rel > threshold)
    d2, nbrs = fp64_neighbours(X, k)
    ties64 = near_tie_rows(d2, k)
    affected = spread(ties64, nbrs[:, :k])
    unattributed = [int(i) for i in above if not affected[i]]
    return AccuracyReport(
        n=int(X.shape[0]), d=int(X.shape[1]), k=k, form=form, ties=ties, stats=stats,
        n_above=int(len(above)), n_attributed=int(len(above) - len(unattributed)),
        near_tie_rows=int(ties64.sum()), unattributed=unattributed[:20],
    )


def main(argv: list[str] | None = None) -> int:
    from .benchmark import blobs_with_outliers

    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--n", type=int, default=20_000)
    p.add_argument("--d", type=int, default=8)
    p.add_argument("--k", type=int, default=20)
    p.add_argument("--seed", type=int, default=7)
    p.add_argument("--form", choices=("fp32", "bf16"), default="fp32")
    p.add_argument("--ties", choices=("index", "fp64"), default="index")
    p.add_argument("--threshold", type=float, default=1e-4)
    p.add_argument("--comparator", choices=("fp64", "fp32"), default="fp64",
                   help="scikit-learn fitted on float64 input (default) or on float32 input "
                        "(the comparator of the reference's published accuracy figures)")
    p.add_argument("--json", type=str, default=None)
    args = p.parse_args(argv)
    X = blobs_with_outliers(args.n, args.d, seed=args.seed)
    rep = report(X, args.k, form=args.form, ties=args.ties, threshold=args.threshold,
                 comparator=args.comparator)
    print(rep.text())
    if args.json:
        with open(args.json, "w") as fh:
            json.dump(asdict(rep), fh, indent=2)
        print(f"wrote {args.json}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
