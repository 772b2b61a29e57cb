"""Categorical-feature cases shared by the oracle pin tests and the GPU parity tests."""
import numpy as np

from tidegbm_amd import synth


def cat_panel(seed, n_dates=150, n_entities=40):
    """Panel with three categorical columns: 11 codes (some rare, NaN sprinkled), 40 codes with
    per-code effects on y, and 300 codes (more than max_bin - 1, so the vocabulary keeps the most
    frequent)."""
    p = synth.panel_small(seed, n_dates=n_dates, n_entities=n_entities)
    rng = np.random.default_rng(seed + 100)
    X = p.X.copy()
    n = p.n
    X[:, 3] = np.clip(np.round(X[:, 3] * 2.0), -5, 5)
    X[rng.random(n) < 0.03, 3] = np.nan
    codes = rng.integers(0, 40, size=n)
    effect = rng.normal(size=40)
    X[:, 7] = codes.astype(np.float64)
    X[:, 11] = rng.integers(0, 300, size=n).astype(np.float64)
    y = p.y + 0.8 * effect[codes] + 0.3 * np.nan_to_num(X[:, 3])
    is_cat = np.zeros(p.n_features, np.int32)
    is_cat[[3, 7, 11]] = 1
    return synth.Panel(X, y, p.ts, p.entity), is_cat


CAT_CASES = {
    "time_pair_default": dict(iid=False, objective=0, params=dict(n_estimators=12, max_depth=6)),
    "iid_device_like": dict(iid=True, objective=0, params=dict(n_estimators=10, max_depth=5, pair_splits=0)),
    "pairwise": dict(iid=False, objective=5, params=dict(n_estimators=8, max_depth=5, pair_splits=0)),
    "min_count_smooth": dict(iid=False, objective=0,
                             params=dict(n_estimators=8, max_depth=5, cat_min_count=40, cat_smooth=5.0)),
}
