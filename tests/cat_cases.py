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


# feature groups (Booster::fit's group_of), interaction groups (Config::interaction_group) and the
# regime feature (tree.cpp:135-176, 300-317, 436-445), on the categorical panel
GROUP_CASES = {
    # xor target: weak leaves take PAIR splits, whose top-k list the groups reshape
    "feature_groups_pair": dict(xor=True, group_of={0: 0, 2: 0, 1: 1, 4: 1, 9: 300, 10: 300},
                                params=dict(n_estimators=8)),
    # the strongest feature binds its subtree to group 0, keeping the next two out of it
    "interaction_groups": dict(interaction={19: 0, 15: 0, 18: 1, 13: 1}, params=dict(n_estimators=8, pair_splits=0)),
    "regime_numeric": dict(params=dict(n_estimators=8, regime_feature=18, pair_splits=0)),
    "regime_categorical": dict(params=dict(n_estimators=8, regime_feature=7)),
    "all_together": dict(group_of={0: 0, 1: 0, 5: 1, 6: 1}, interaction={2: 0, 3: 0, 4: 1, 8: 1},
                         params=dict(n_estimators=8, regime_feature=7, max_depth=5)),
}


def group_panel(case, seed=33):
    """The categorical panel, or for the cases that need PAIR splits a numeric panel with an XOR
    target of features 0 and 1."""
    if GROUP_CASES[case].get("xor"):   # numeric features only
        p = synth.panel_small(seed)
        rng = np.random.default_rng(seed + 7)
        y = np.sign(p.X[:, 0] * p.X[:, 1]) + 0.1 * rng.normal(size=p.n)
        return synth.Panel(p.X, y, p.ts, p.entity), np.zeros(p.n_features, np.int32)
    return cat_panel(seed)


def group_arrays(case, n_features):
    gc = GROUP_CASES[case]
    grp = ix = None
    if "group_of" in gc:
        grp = np.full(n_features, -1, np.int32)
        for f, g in gc["group_of"].items():
            grp[f] = g
    if "interaction" in gc:
        ix = np.full(n_features, -1, np.int32)
        for f, g in gc["interaction"].items():
            ix[f] = g
    return grp, ix
