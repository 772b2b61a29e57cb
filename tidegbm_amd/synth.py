"""Seeded synthetic panels for the five benchmark workloads.

Every generator is deterministic in its seed (numpy PCG64) and returns rows ordered by
(date, entity), so the sort inside a fit is the identity.  Features are generated as float32
and handed to the C API as float64, matching how the reference's Python layer casts inputs
(reference src/tidegbm/data.py:36-38).

Workloads:
  * panel_small   — 500 dates x 30 entities = 15,000 rows, 20 AR(1) features (phi 0.9).
  * panel_large   — 2,500 dates x 400 entities = 1,000,000 rows, 128 AR(1) features with
                    phi ~ U(0.5, 0.99), 13 Student-t(3) features, 1% NaN, depth-3 tree term.
  * panel_ranking — 5,000 dates x U(200, 1000) entities (~3M rows), 64 AR(1) features.
  * walk_forward_tasks — 16 expanding folds x horizons (1, 3, 5, 10) over panel_large.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Panel:
    X: np.ndarray          # float64 (n, F), values representable in float32
    y: np.ndarray          # float64 (n,)
    ts: np.ndarray         # int64 (n,) date index
    entity: np.ndarray     # int64 (n,)
    y_raw: np.ndarray | None = None  # float64 (n,) target before within-date demeaning

    @property
    def n(self) -> int:
        return int(self.X.shape[0])

    @property
    def n_features(self) -> int:
        return int(self.X.shape[1])


def _demean_by_date(y: np.ndarray, ts: np.ndarray) -> np.ndarray:
    # ts is sorted ascending; subtract each date's mean.
    uniq, start, counts = np.unique(ts, return_index=True, return_counts=True)
    sums = np.add.reduceat(y, start)
    means = sums / counts
    return y - np.repeat(means, counts)


def _ar1_panel(rng, n_dates, n_ent, n_feat, phi, heavy_mask=None, df=3.0, chunk=256):
    """AR(1) per (entity, feature), stationary start; returns float32 (n_dates, n_ent, n_feat)."""
    phi = np.asarray(phi, dtype=np.float64).reshape(1, n_feat)
    out = np.empty((n_dates, n_ent, n_feat), dtype=np.float32)
    state = rng.standard_normal((n_ent, n_feat)) / np.sqrt(1.0 - phi ** 2)
    for d0 in range(0, n_dates, chunk):
        d1 = min(n_dates, d0 + chunk)
        eps = rng.standard_normal((d1 - d0, n_ent, n_feat))
        if heavy_mask is not None and heavy_mask.any():
            k = int(heavy_mask.sum())
            z = rng.standard_normal((d1 - d0, n_ent, k))
            chi = rng.chisquare(df, size=(d1 - d0, n_ent, k))
            eps[:, :, heavy_mask] = z / np.sqrt(chi / df)
        for d in range(d1 - d0):
            state = phi * state + eps[d]
            out[d0 + d] = state
    return out


def _tree_term(rng, X: np.ndarray, depth: int = 3) -> np.ndarray:
    """A random axis-aligned tree of the given depth over X's columns (leaf values N(0,1))."""
    n, F = X.shape
    node = np.zeros(n, dtype=np.int64)
    for _ in range(depth):
        n_nodes = int(node.max()) + 1
        feats = rng.integers(0, F, size=n_nodes)
        thr = rng.standard_normal(n_nodes) * 0.5
        xv = X[np.arange(n), feats[node]]
        node = node * 2 + (xv > thr[node]).astype(np.int64)
    leaf_vals = rng.standard_normal(1 << depth)
    return leaf_vals[node]


def panel_small(seed: int = 0, n_dates: int = 500, n_entities: int = 30, n_features: int = 20) -> Panel:
    rng = np.random.default_rng(seed)
    cube = _ar1_panel(rng, n_dates, n_entities, n_features, np.full(n_features, 0.9))
    X = cube.reshape(n_dates * n_entities, n_features).astype(np.float64)
    K = rng.choice(n_features, size=5, replace=False)
    beta = rng.standard_normal(5)
    y_raw = X[:, K] @ beta + rng.normal(0.0, 0.5, size=X.shape[0])
    ts = np.repeat(np.arange(n_dates, dtype=np.int64), n_entities)
    ent = np.tile(np.arange(n_entities, dtype=np.int64), n_dates)
    return Panel(X, _demean_by_date(y_raw, ts), ts, ent, y_raw)


def panel_large(seed: int = 1, n_dates: int = 2500, n_entities: int = 400, n_features: int = 128,
                nan_frac: float = 0.01) -> Panel:
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0.5, 0.99, size=n_features)
    n_heavy = int(round(0.1 * n_features))
    heavy = np.zeros(n_features, dtype=bool)
    heavy[rng.choice(n_features, size=n_heavy, replace=False)] = True
    cube = _ar1_panel(rng, n_dates, n_entities, n_features, phi, heavy_mask=heavy)
    X = cube.reshape(n_dates * n_entities, n_features)
    K = rng.choice(n_features, size=5, replace=False)
    beta = rng.standard_normal(5)
    Xk = X[:, K].astype(np.float64)
    y_raw = Xk @ beta + _tree_term(rng, X) + rng.normal(0.0, 0.5, size=X.shape[0])
    ts = np.repeat(np.arange(n_dates, dtype=np.int64), n_entities)
    ent = np.tile(np.arange(n_entities, dtype=np.int64), n_dates)
    X64 = X.astype(np.float64)
    if nan_frac > 0:
        m = rng.random(X64.shape) < nan_frac
        X64[m] = np.nan
    return Panel(X64, _demean_by_date(y_raw, ts), ts, ent, y_raw)


def panel_ranking(seed: int = 2, n_dates: int = 5000, lo: int = 200, hi: int = 1000,
                  n_features: int = 64) -> Panel:
    rng = np.random.default_rng(seed)
    counts = rng.integers(lo, hi + 1, size=n_dates)
    phi = np.full((1, n_features), 0.9)
    state = rng.standard_normal((hi, n_features)) / np.sqrt(1.0 - phi ** 2)
    K = rng.choice(n_features, size=5, replace=False)
    beta = rng.standard_normal(5)
    n = int(counts.sum())
    X = np.empty((n, n_features), dtype=np.float32)
    ent = np.empty(n, dtype=np.int64)
    ts = np.repeat(np.arange(n_dates, dtype=np.int64), counts)
    r = 0
    for d in range(n_dates):
        state = phi * state + rng.standard_normal((hi, n_features))
        m = int(counts[d])
        sel = np.sort(rng.choice(hi, size=m, replace=False))
        X[r:r + m] = state[sel]
        ent[r:r + m] = sel
        r += m
    X64 = X.astype(np.float64)
    y_raw = X64[:, K] @ beta + rng.normal(0.0, 0.5, size=n)
    return Panel(X64, _demean_by_date(y_raw, ts), ts, ent, y_raw)


@dataclass
class WalkForwardTask:
    fold: int
    horizon: int
    n_train: int        # rows [0, n_train) of the shared panel
    y: np.ndarray       # float64 (n_train,) target for this horizon, demeaned within date


def walk_forward_tasks(panel: Panel, n_folds: int = 16, horizons=(1, 3, 5, 10),
                       first_dates: int = 250) -> list[WalkForwardTask]:
    """Expanding folds whose train window ends `h` dates (purge gap) before the fold's test start.

    Fold k's test window starts at date first_dates + k*step; its train rows are every date
    < start - h (the purge gap of WalkForwardSplit, reference src/tidegbm/cv.py:45-56).
    The target for horizon h is the forward h-date sum of the one-date raw target, demeaned
    within date.
    """
    ts = panel.ts
    n_dates = int(ts.max()) + 1
    n_ent = int(np.sum(ts == 0))
    assert panel.n == n_dates * n_ent, "walk-forward tasks need a balanced panel"
    step = (n_dates - first_dates) // max(n_folds - 1, 1)
    y1 = panel.y_raw.reshape(n_dates, n_ent)
    tasks = []
    for h in horizons:
        csum = np.concatenate([np.zeros((1, n_ent)), np.cumsum(y1, axis=0)], axis=0)
        end = np.minimum(np.arange(n_dates) + h, n_dates)
        yh = (csum[end] - csum[np.arange(n_dates)]).reshape(-1)
        yh = _demean_by_date(yh, ts)
        for k in range(n_folds):
            start = min(first_dates + k * step, n_dates)
            train_dates = max(start - h, 1)
            n_train = train_dates * n_ent
            tasks.append(WalkForwardTask(k, h, n_train, np.ascontiguousarray(yh[:n_train])))
    return tasks
