"""Categorical features on the GPU path (reference Binner::fit's categorical branch,
binning.cpp:48-82; best_categorical, histogram.cpp:195-244; cat_goes_left, tree.hpp:27-30), and
the leaf-candidate rules that use them: feature groups, interaction groups and the regime feature
(tree.cpp:135-176, 293-317, 436-445).

Bar: category vocabularies equal; tree structure and the categorical bin masks bit-exact against
the CPU oracle and the reference core (a split decided by equal G/(H+cat_smooth) ratios may differ
in order, as the reference's std::sort is not stable; such a node is accepted only when both
splits cut the node's training rows into the same sets); leaf values within rtol 1e-7;
predictions within rtol 1e-5; a saved model predicts identically after loading.
"""
import sys

import numpy as np
import pytest

import checkers
from cat_cases import CAT_CASES, GROUP_CASES, cat_panel, group_arrays, group_panel
from tidegbm_amd import Booster, TideError, TideRegressor

pytestmark = pytest.mark.gpu

VALUE_RTOL = 1e-7
PRED_RTOL = 1e-5


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fit(case):
    cc = CAT_CASES[case]
    p, is_cat = cat_panel(31)
    params = dict(learning_rate=0.1, goss=0, objective=cc["objective"], **cc["params"])
    gpu = Booster(**params)
    gpu.fit(p.X, p.y, timestamps=None if cc["iid"] else p.ts, entity=None if cc["iid"] else p.entity,
            rank_groups=p.ts if cc["objective"] == 5 else None, iid=cc["iid"], horizon=1, is_cat=is_cat)
    cpu = []
    for lib in (checkers.oracle(), checkers.reference()):
        if lib is None:
            continue
        m = lib.model(**params)
        m.set_cat(is_cat)
        m.fit(p.X, p.y, ts=None if cc["iid"] else p.ts, entity=None if cc["iid"] else p.entity,
              rank_groups=p.ts if cc["objective"] == 5 else None, iid=cc["iid"], horizon=1)
        cpu.append(m)
    return p, is_cat, gpu, cpu


@pytest.mark.parametrize("case", sorted(CAT_CASES))
def test_categorical_fit_matches_oracle(case):
    _say(f"[case] categorical fit {case} {CAT_CASES[case]}")
    p, is_cat, gpu, cpu = _fit(case)
    F = p.n_features
    cats = [gpu.cat_values(f) if is_cat[f] else None for f in range(F)]
    xb = checkers.bins_of(p.X, [gpu.edges(f) for f in range(F)], cats)
    nb = gpu.model_info()["n_bins"]
    trees = gpu.trees()
    n_cat = sum(int(np.isin(t["ints"][:, 1], (2, 3)).sum()) for t in trees)
    _say(f"  categorical splits: {n_cat}")
    assert n_cat >= 5
    for m in cpu:
        for f in np.flatnonzero(is_cat):
            assert np.array_equal(gpu.cat_values(f), m.cat_values(f))
        diff = checkers.compare_trees(trees, m.trees(), rtol=VALUE_RTOL, atol=1e-9, xb=xb, n_bins=nb)
        assert diff is None, diff
        a, b = gpu.predict(p.X), m.predict(p.X)
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < PRED_RTOL


def test_categorical_save_load_and_unseen_values(tmp_path):
    p, is_cat, gpu, _ = _fit("iid_device_like")
    X = p.X[:3000].copy()
    X[:50, 7] = 1234.0    # a category the fit never saw bins to 0, like a missing value
    X[50:60, 3] = np.nan
    before = gpu.predict(X)
    path = tmp_path / "cat.tgbm"
    gpu.save(path)
    loaded = Booster.load(path)
    assert np.array_equal(loaded.predict(X), before)
    for f in np.flatnonzero(is_cat):
        assert np.array_equal(loaded.cat_values(f), gpu.cat_values(f))
    for a, b in zip(loaded.trees(), gpu.trees()):
        assert np.array_equal(a["cat"], b["cat"])


def test_estimator_categorical_feature_indices():
    p, is_cat, _, _ = _fit("iid_device_like")
    m = TideRegressor(n_estimators=5, iid_mode=True, pair_splits=False)
    m.fit(p.X, p.y, categorical_feature=list(np.flatnonzero(is_cat)))
    assert m.booster_.cat_values(7).size == 40
    with pytest.raises(ValueError, match="not in X"):
        TideRegressor(n_estimators=2, iid_mode=True).fit(p.X, p.y, categorical_feature=[p.n_features])


def test_categorical_ensembles_are_served(tmp_path):
    # The batched serving kernel routes categorical splits by their bin masks: served scores of
    # panel rows and of raw rows (unseen categories and NaN included) equal the one-call predict
    # bit for bit, and the reference core reading the same model files; refits of categorical
    # ensembles are refused (their splits are chosen by the host-driven grower).
    from tidegbm_amd import Server
    p, is_cat = cat_panel(12)
    models = []
    for k, y in enumerate((p.y, p.y + 0.5 * np.nan_to_num(p.X[:, 0]))):
        b = Booster(n_estimators=12, max_depth=6, learning_rate=0.1, goss=0, pair_splits=0)
        b.fit(p.X, y, timestamps=p.ts, entity=p.entity, horizon=1, is_cat=is_cat)
        models.append(b)
    assert any((t["ints"][:, 1] == 2).any() for t in models[0].trees()), "the case should have categorical splits"
    X = p.X[:2500].copy()
    X[:40, 7] = 1234.0
    X[40:60, 11] = np.nan
    srv = Server(models, p.X, predict_budget=64 << 20, refit_slots=2, refit_max_rows=2000)
    try:
        rows = np.arange(300, 2800, dtype=np.int32)
        for m, b in enumerate(models):
            a, _, _ = srv.wait(srv.predict(m, rows))
            c, _, _ = srv.wait(srv.predict_x(m, X))
            assert np.array_equal(a, b.predict(p.X[rows]))
            assert np.array_equal(c, b.predict(X))
            ref = checkers.reference()
            if ref is not None:
                b.save(tmp_path / f"m{m}.tgbm")
                r = checkers.CpuModel.load(ref, tmp_path / f"m{m}.tgbm")
                assert np.max(np.abs(c - r.predict(X)) / np.maximum(np.abs(r.predict(X)), 1e-12)) <= 1e-12
        with pytest.raises(TideError, match="categorical"):
            srv.refit(0, 0, p.y[:1000], extra_trees=2)
    finally:
        srv.close()


@pytest.mark.parametrize("case", sorted(GROUP_CASES))
def test_groups_interactions_regime_match_oracle(case):
    gc = GROUP_CASES[case]
    _say(f"[case] groups {case} {gc}")
    p, is_cat = group_panel(case)
    grp, ix = group_arrays(case, p.n_features)
    params = dict(learning_rate=0.1, goss=0, **gc["params"])
    extra = {}
    if grp is not None:
        extra["feature_groups"] = grp
    if ix is not None:
        extra["interaction_group"] = ix
    gpu = Booster(**params, **extra)
    gpu.fit(p.X, p.y, timestamps=p.ts, entity=p.entity, horizon=1, is_cat=is_cat if is_cat.any() else None)
    F = p.n_features
    cats = [gpu.cat_values(f) if is_cat[f] else None for f in range(F)]
    xb = checkers.bins_of(p.X, [gpu.edges(f) for f in range(F)], cats)
    nb = gpu.model_info()["n_bins"]
    for lib in (checkers.oracle(), checkers.reference()):
        if lib is None:
            continue
        m = lib.model(**params)
        m.set_cat(is_cat)
        m.set_groups(grp, ix)
        m.fit(p.X, p.y, ts=p.ts, entity=p.entity, horizon=1)
        diff = checkers.compare_trees(gpu.trees(), m.trees(), rtol=VALUE_RTOL, atol=1e-9, xb=xb, n_bins=nb)
        assert diff is None, diff
        a, b = gpu.predict(p.X), m.predict(p.X)
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < PRED_RTOL
    types = np.concatenate([t["ints"][t["ints"][:, 0] == 0, 1] for t in gpu.trees()])
    _say(f"  split types: {np.bincount(types, minlength=4).tolist()}")


def test_categorical_with_eval_set():
    # the eval rows are binned with the vocabularies and scored through the categorical masks
    p, is_cat = cat_panel(37)
    dates = np.unique(p.ts)
    cut = dates[int(len(dates) * 0.7)]
    tr, va = p.ts < cut, p.ts >= cut + 1
    params = dict(n_estimators=10, max_depth=5, learning_rate=0.1, goss=0, pair_splits=0, early_stopping_rounds=0)
    gpu = Booster(eval_metric="loss", **params)
    gpu.fit(p.X[tr], p.y[tr], timestamps=p.ts[tr], entity=p.entity[tr], horizon=1, is_cat=is_cat,
            eval_set=(p.X[va], p.y[va]), eval_timestamp=p.ts[va])
    m = checkers.oracle().model(eval_metric=1, **params)
    m.set_cat(is_cat)
    m.set_eval(p.X[va], p.y[va], p.ts[va])
    m.fit(p.X[tr], p.y[tr], ts=p.ts[tr], entity=p.entity[tr], horizon=1)
    ev, eo = gpu.evals_result()["valid"], m.evals_valid()
    assert len(ev) == len(eo) == 10
    assert np.allclose(ev, eo, rtol=1e-6, atol=0), (ev, eo)
    assert gpu.best_iteration == m.best_iteration() == 1
    a, b = gpu.predict(p.X[va]), m.predict(p.X[va])
    assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < PRED_RTOL
