"""Serving loop on the GPU: resident ensembles over a resident binned panel.

* predict results (panel row ids and raw feature rows) equal the plain predict path bit-for-bit;
* mixed load: predicts and refits submitted together; every predict equals the plain predict of
  the version it reports, and every refit's published trees equal Booster.update run alone on the
  parent version (all integer fields and doubles identical);
* the unbatched baseline mode gives the same per-request results;
* admission rejects a request larger than the predict memory budget.
"""
import numpy as np
import pytest

import tidegbm_amd as t
from tidegbm_amd import synth

pytestmark = pytest.mark.gpu

PARAMS = dict(n_estimators=30, max_depth=6, learning_rate=0.1, pair_splits=0, goss=0, min_samples_leaf=10)


def _models():
    p = synth.panel_small(7, n_dates=300, n_entities=40, n_features=12)
    tasks = synth.walk_forward_tasks(p, n_folds=2, horizons=(1, 5), first_dates=200)
    models = t.fit_batch(p.X, p.ts, [k.n_train for k in tasks], [k.y for k in tasks], entity=p.entity,
                         horizons=[k.horizon for k in tasks], **PARAMS)
    return p, models


def _same_trees(a, b):
    ta, tb = a.trees(), b.trees()
    assert len(ta) == len(tb)
    for k, (x, y) in enumerate(zip(ta, tb)):
        for key in ("ints", "dbl", "values", "cover", "prefix"):
            assert np.array_equal(x[key], y[key]), f"tree {k} {key}"


def test_server_predict_matches_plain():
    p, models = _models()
    srv = t.Server(models, p.X, predict_budget=64 << 20, refit_slots=4, refit_max_rows=4000)
    rng = np.random.default_rng(0)
    jobs = []
    for _ in range(24):
        m = int(rng.integers(len(models)))
        n = int(rng.integers(1, 3000))
        rows = np.sort(rng.choice(p.n, size=n, replace=False)).astype(np.int32)
        jobs.append((m, rows, srv.predict(m, rows), srv.predict_x(m, p.X[rows])))
    print(f"[serve] {len(jobs)} x 2 predicts submitted", flush=True)
    for m, rows, ta, tb in jobs:
        ref = models[m].predict(p.X[rows])
        a, va, _ = srv.wait(ta)
        b, vb, _ = srv.wait(tb)
        assert va == 0 and vb == 0
        assert np.array_equal(a, ref), "panel-row predict differs from the plain path"
        assert np.array_equal(b, ref), "raw-row predict differs from the plain path"
    st = srv.stats()
    assert st["predict_requests"] == 48
    srv.close()


def _mixed(batched):
    p, models = _models()
    srv = t.Server(models, p.X, predict_budget=64 << 20, refit_slots=3, refit_max_rows=4000, batched=batched)
    rng = np.random.default_rng(1)
    jobs = []
    for k in range(60):
        m = int(rng.integers(len(models)))
        if k % 8 == 3:
            n = int(rng.integers(500, 4000))
            row0 = int(rng.integers(0, p.n - n))
            y = p.y[row0:row0 + n]
            jobs.append(("refit", m, (row0, n, y), srv.refit(m, row0, y, extra_trees=5)))
        else:
            n = int(rng.integers(1, 2000))
            rows = rng.integers(0, p.n, size=n).astype(np.int32)
            jobs.append(("predict", m, rows, srv.predict(m, rows)))
    print(f"[serve] {len(jobs)} mixed requests submitted (batched={batched})", flush=True)
    results = [(j, srv.wait(j[3])) for j in jobs]
    n_refit = 0
    for (kind, m, arg, _), (out, ver, lat) in results:
        assert lat >= 0
        if kind == "predict":
            snap, _ = srv.model(m, ver)
            assert np.array_equal(out, snap.predict(p.X[arg])), f"predict on model {m} v{ver} differs"
        else:
            n_refit += 1
            row0, n, y = arg
            snap, parent = srv.model(m, ver)
            base, _ = srv.model(m, parent)
            base.update(p.X[row0:row0 + n], y, 5)
            _same_trees(snap, base)
    assert n_refit > 0
    st = srv.stats()
    srv.close()
    return st


def test_server_mixed_load_equals_alone():
    st = _mixed(True)
    assert st["refits"] == 8


def test_server_unbatched_mode_equals_alone():
    st = _mixed(False)
    assert st["predict_launches"] == st["predict_requests"]


def test_server_admission_budget():
    p, models = _models()
    srv = t.Server(models, p.X, predict_budget=1 << 20, refit_slots=2, refit_max_rows=2000)
    with pytest.raises(t.TideError, match="budget"):
        srv.predict_x(0, np.zeros((20000, p.n_features)))
    with pytest.raises(t.TideError, match="capacity"):
        srv.refit(0, 0, np.zeros(3000))
    out, _, _ = srv.wait(srv.predict(0, np.arange(10, dtype=np.int32)))
    assert np.array_equal(out, models[0].predict(p.X[:10]))
    srv.close()
