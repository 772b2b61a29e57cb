"""Batched walk-forward fits on the GPU.

Each fold x horizon trains on a row prefix of one shared binned matrix.  Checked here:
* batched == sequential: the same fits grown together and one after another give identical
  trees (every integer field and every double bit-for-bit) and identical predictions;
* each batched fit == the oracle's fit of the same rows with the same (shared) bin edges:
  tree structure exact, leaf values within 1e-7 relative, predictions within 1e-5 relative;
"""
import numpy as np
import pytest

import checkers
import tidegbm_amd as t
from tidegbm_amd import synth

pytestmark = pytest.mark.gpu

PARAMS = dict(n_estimators=10, max_depth=5, learning_rate=0.1, pair_splits=0, goss=0, min_samples_leaf=10)


def _tasks():
    p = synth.panel_small(5, n_dates=300, n_entities=24, n_features=10)
    tasks = synth.walk_forward_tasks(p, n_folds=3, horizons=(1, 3, 10), first_dates=80)
    return p, tasks


def _fit(p, tasks, schedule, **extra):
    return t.fit_batch(p.X, p.ts, [k.n_train for k in tasks], [k.y for k in tasks], entity=p.entity,
                       horizons=[k.horizon for k in tasks], schedule=schedule, **{**PARAMS, **extra})


def test_batched_equals_sequential():
    p, tasks = _tasks()
    print(f"[batch] {len(tasks)} tasks, rows {[k.n_train for k in tasks]}", flush=True)
    a = _fit(p, tasks, "batched")
    b = _fit(p, tasks, "sequential")
    for k, (ma, mb) in enumerate(zip(a, b)):
        ta, tb = ma.trees(), mb.trees()
        assert len(ta) == PARAMS["n_estimators"]
        for x, y in zip(ta, tb):
            for key in ("ints", "dbl", "values", "cover", "prefix"):
                assert np.array_equal(x[key], y[key]), f"task {k}: {key} differs"
        pa, pb = ma.predict(p.X), mb.predict(p.X)
        assert np.array_equal(pa, pb), f"task {k}: predictions differ"


def test_batched_matches_oracle_with_shared_edges():
    p, tasks = _tasks()
    models = _fit(p, tasks, "batched")
    orc = checkers.oracle()
    F = p.n_features
    edges = [models[0].edges(f) for f in range(F)]
    for k, (task, m) in enumerate(zip(tasks, models)):
        for f in range(F):
            assert np.array_equal(m.edges(f), edges[f])
        n = task.n_train
        o = orc.model(**PARAMS).set_edges(edges)
        o.fit(p.X[:n], task.y, ts=p.ts[:n], entity=p.entity[:n], horizon=task.horizon)
        assert np.allclose(m.init_prefix(), o.init_prefix(), rtol=0, atol=0), f"task {k}: init prefix"
        xb = checkers.bins_of(p.X[:n], edges)
        err = checkers.compare_trees(m.trees(), o.trees(), rtol=1e-7, atol=1e-9, xb=xb,
                                     n_bins=m.model_info()["n_bins"])
        assert err is None, f"task {k} (rows {n}, h {task.horizon}): {err}"
        pg, po = m.predict(p.X), o.predict(p.X)
        assert np.allclose(pg, po, rtol=1e-5, atol=1e-9), f"task {k}: predictions"
        assert np.allclose(m.evals_result()["train"], o.evals_train(), rtol=1e-10, atol=0), f"task {k}: loss"
