"""The benchmark configurations at their full sizes, on the GPU, against the reference core (and the
oracle where bins are shared), plus predictions on rows the fits never saw.

* 1,000,000 x 128 panel, all 200 trees (depth 8, lr 0.05, GOSS off, PAIR auto-off);
* ranking panel, 2,988,996 rows x 64, within-date pairwise, GOSS auto (on), 10 trees: the reference
  core takes its memory-mapped, row-major histogram path at this size (booster.cpp:201-229);
* walk-forward batch, 64 fits x 100 trees: batched == sequential for every tree, and two folds
  against the oracle with the shared bin edges;
* serving, 8 resident ensembles of 500 depth-8 trees: served predicts and a served refit against
  the reference core reading the served model versions from TGBM files.

Bar: tree structure exact up to the accepted, row-verified exceptions of checkers.compare_trees,
whose count is printed per test; leaf values within 1e-7 relative; predictions within 1e-5
relative.  On held-out rows (with extra NaNs), a tree's leaf value may differ from the
reference's only in a tree where an exception was accepted, and every other row's prediction
agrees within 1e-5.
"""
import sys

import numpy as np
import pytest

import checkers
import tidegbm_amd as t
from tidegbm_amd import synth

pytestmark = pytest.mark.gpu
REF = checkers.reference()
needs_ref = pytest.mark.skipif(REF is None, reason="reference core (oracle/_ref) not built")
VALUE_RTOL, PRED_RTOL = 1e-7, 1e-5
THREADS = 16


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bins(X, model):
    """uint8 bins of raw rows with a model's edges (the oracle's transform)."""
    F = X.shape[1]
    E = np.zeros((F, 256))
    ne = np.zeros(F, np.int32)
    for f in range(F):
        e = model.edges(f)
        E[f, :e.size] = e
        ne[f] = e.size
    X = np.ascontiguousarray(X)
    out = np.empty(X.shape, np.uint8)
    checkers.oracle().lib.orc_bin_transform(X.ctypes.data, X.shape[0], F, E.ctypes.data, ne.ctypes.data,
                                            out.ctypes.data)
    return out


def _rel(a, b):
    scale = np.maximum(np.abs(b), 1e-3 * np.max(np.abs(b)) + 1e-12)
    return np.abs(a - b) / scale


def _held_out(seed, n_dates, make, nan_frac=0.05):
    p = make(seed, n_dates)
    X = p.X.copy()
    X[np.random.default_rng(seed).random(X.shape) < nan_frac] = np.nan
    return X


def _compare(name, gpu, ref, xb, X_train_sample, X_held):
    stats = {}
    diff = checkers.compare_trees(gpu.trees(), ref.trees(), rtol=VALUE_RTOL, atol=1e-9, xb=xb,
                                  n_bins=gpu.model_info()["n_bins"], stats=stats)
    kinds = {k: v for k, v in stats.items() if k != "where"}
    _say(f"  {name}: accepted checker exceptions {kinds or 'none'} {stats.get('where', [])[:8]}")
    assert diff is None, f"{name}: {diff}"
    rel = _rel(gpu.predict(X_train_sample), ref.predict(X_train_sample))
    assert rel.max() < PRED_RTOL, f"{name}: training-row prediction rel err {rel.max()}"
    xh = _bins(X_held, gpu)
    unexplained, rows = checkers.held_out_differences(gpu.trees(), ref.trees(), xh, stats=stats,
                                                      n_bins=gpu.model_info()["n_bins"])
    assert not unexplained, f"{name}: held-out leaf values differ in trees without an exception: {unexplained}"
    rel = _rel(gpu.predict(X_held), ref.predict(X_held))
    _say(f"  {name}: held-out rows {X_held.shape[0]}, routed differently by an excepted node {int(rows.sum())}, "
         f"max rel err elsewhere {rel[~rows].max() if (~rows).any() else 0:.2e}")
    assert (~rows).sum() == 0 or rel[~rows].max() < PRED_RTOL
    return kinds


@needs_ref
@pytest.mark.timeout(900)
def test_full_panel_all_trees_against_reference():
    p = synth.panel_large(1)
    params = dict(n_estimators=200, max_depth=8, learning_rate=0.05, goss=0)
    _say(f"[case] 1M x 128, {params}")
    gpu = t.Booster(**params).fit(p.X, p.y, timestamps=p.ts, entity=p.entity, horizon=1)
    _say("  reference core fit (host threads)")
    ref = REF.model(n_jobs=THREADS, **params)
    ref.fit(p.X, p.y, ts=p.ts, entity=p.entity, horizon=1)
    for f in range(p.n_features):
        assert np.array_equal(gpu.edges(f), ref.edges(f)), f"edges of feature {f}"
    assert np.allclose(gpu.evals_result()["train"], ref.evals_train(), rtol=1e-10, atol=0)
    xb = _bins(p.X, gpu)
    X_held = _held_out(7, 60, lambda s, d: synth.panel_large(s, n_dates=d))
    _compare("1Mx128-200", gpu, ref, xb, p.X[::50], X_held)


@needs_ref
@pytest.mark.timeout(900)
def test_full_ranking_panel_against_reference():
    p = synth.panel_ranking(2)
    params = dict(n_estimators=10, max_depth=6, learning_rate=0.08, objective=5)
    _say(f"[case] ranking {p.n} x {p.n_features}, {params} (GOSS auto on, PAIR auto off)")
    gpu = t.Booster(**params).fit(p.X, p.y, timestamps=p.ts, entity=p.entity, rank_groups=p.ts, horizon=1)
    host_trees, _, _ = gpu.goss_stats()
    _say(f"  GOSS trees sampled on the host: {host_trees} of 10")
    ref = REF.model(n_jobs=THREADS, **params)
    ref.fit(p.X, p.y, ts=p.ts, entity=p.entity, rank_groups=p.ts, horizon=1)
    xb = _bins(p.X, gpu)
    X_held = _held_out(9, 100, lambda s, d: synth.panel_ranking(s, n_dates=d))
    _compare("ranking-full", gpu, ref, xb, p.X[::150], X_held)


@pytest.mark.timeout(900)
def test_full_walk_forward_batch():
    p = synth.panel_large(1)
    tasks = synth.walk_forward_tasks(p, n_folds=16, horizons=(1, 3, 5, 10), first_dates=250)
    params = dict(n_estimators=100, max_depth=6, learning_rate=0.05, max_bin=255, goss=0, pair_splits=0)
    kw = dict(entity=p.entity, horizons=[k.horizon for k in tasks], **params)
    rows = [k.n_train for k in tasks]
    _say(f"[case] walk-forward {len(tasks)} fits x 100 trees, rows {min(rows)}..{max(rows)}")
    a = t.fit_batch(p.X, p.ts, rows, [k.y for k in tasks], schedule="batched", **kw)
    b = t.fit_batch(p.X, p.ts, rows, [k.y for k in tasks], schedule="sequential", **kw)
    for i, (ma, mb) in enumerate(zip(a, b)):
        ta, tb = ma.trees(), mb.trees()
        assert len(ta) == len(tb) == 100
        for j, (x, y) in enumerate(zip(ta, tb)):
            for key in ("ints", "dbl", "values", "cover", "prefix"):
                assert np.array_equal(x[key], y[key]), f"fit {i} tree {j}: {key}"
    _say("  batched == sequential for all 64 x 100 trees")
    # two folds (the smallest of horizons 1 and 10) against the oracle with the shared edges,
    # over their first 30 trees
    F = p.n_features
    edges = [a[0].edges(f) for f in range(F)]
    for i in (0, 48):
        task, m = tasks[i], a[i]
        n = task.n_train
        o = checkers.oracle().model(**{**params, "n_estimators": 30}).set_edges(edges)
        o.fit(p.X[:n], task.y, ts=p.ts[:n], entity=p.entity[:n], horizon=task.horizon)
        stats = {}
        xb = _bins(p.X[:n], m)
        err = checkers.compare_trees(m.trees()[:30], o.trees(), rtol=VALUE_RTOL, atol=1e-9, xb=xb,
                                     n_bins=m.model_info()["n_bins"], stats=stats)
        _say(f"  fold {i} (rows {n}, horizon {task.horizon}): exceptions "
             f"{ {k: v for k, v in stats.items() if k != 'where'} or 'none'}")
        assert err is None, f"fold {i}: {err}"


@needs_ref
@pytest.mark.timeout(900)
def test_full_serving_against_reference(tmp_path):
    p = synth.panel_large(1)
    tasks = synth.walk_forward_tasks(p, n_folds=2, horizons=(1, 3, 5, 10), first_dates=2200)
    params = dict(n_estimators=500, max_depth=8, learning_rate=0.05, max_bin=255, goss=0, pair_splits=0)
    _say("[case] serving: 8 ensembles x 500 trees over the 1M x 128 panel")
    models = t.fit_batch(p.X, p.ts, [k.n_train for k in tasks], [k.y for k in tasks], entity=p.entity,
                         horizons=[k.horizon for k in tasks], **params)
    srv = t.Server(models, p.X, predict_budget=1 << 30, refit_slots=4, refit_max_rows=60000)
    try:
        rng = np.random.default_rng(11)
        jobs = []
        for k in range(24):
            m = int(rng.integers(len(models)))
            n = int(min(65536, max(1, np.floor(np.exp(rng.uniform(0.0, np.log(65536.0)))))))
            row0 = int(rng.integers(0, p.n - n + 1))
            jobs.append(("predict", m, row0, n, srv.predict(m, np.arange(row0, row0 + n, dtype=np.int32))))
        m_refit, row0_refit, n_refit = 2, 400000, 30000
        y_refit = p.y[row0_refit:row0_refit + n_refit]
        jobs.append(("refit", m_refit, row0_refit, n_refit, srv.refit(m_refit, row0_refit, y_refit, extra_trees=20)))
        results = [(j, srv.wait(j[4])) for j in jobs]
        _say(f"  {len(jobs)} requests served")
        ref_models = {}

        def ref_version(m, ver):
            if (m, ver) not in ref_models:
                snap, parent = srv.model(m, ver)
                path = tmp_path / f"m{m}_v{ver}.tgbm"
                snap.save(path)
                ref_models[(m, ver)] = (checkers.CpuModel.load(REF, path), snap, parent)
            return ref_models[(m, ver)]

        for (kind, m, row0, n, _), (out, ver, _) in results:
            X = p.X[row0:row0 + n]
            if kind == "predict":
                r, _, _ = ref_version(m, ver)
                rel = _rel(out, r.predict(X))
                assert rel.max() < PRED_RTOL, f"served predict model {m} v{ver} rows {n}: rel err {rel.max()}"
            else:
                r, snap, parent = ref_version(m, ver)
                base, _, _ = ref_version(m, parent)
                base.update(X, y_refit, 20)   # the reference's Booster::update on the parent version
                stats = {}
                xb = _bins(X, snap)
                err = checkers.compare_trees(snap.trees(), base.trees(), rtol=VALUE_RTOL, atol=1e-9, xb=xb,
                                             n_bins=snap.model_info()["n_bins"], stats=stats)
                _say(f"  served refit of model {m} (v{parent} -> v{ver}): exceptions "
                     f"{ {k: v for k, v in stats.items() if k != 'where'} or 'none'}")
                assert err is None, f"served refit: {err}"
                assert _rel(snap.predict(X[:5000]), base.predict(X[:5000])).max() < PRED_RTOL
        _say(f"  served predicts equal the reference core's predictions of the same versions")
    finally:
        srv.close()
