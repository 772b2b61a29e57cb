"""Eval sets and early stopping on the GPU path (reference Booster::fit with X_eval,
booster.cpp:459-478 and :707-732).

Bar: the eval kernels bit-exact against the reference's arithmetic restated in Python on the same
trees (predictions, mean squared error, per-date correlation); whole fits against the CPU oracle
and the reference core: same number of trees, same best_iteration, tree structure bit-exact, the
validation metric within rtol 1e-6 (the device leaf values differ from the reference's in the last
bits, within rtol 1e-7, because they come from exact fixed-point sums).
"""
import ctypes
import math
import sys
from pathlib import Path

import numpy as np
import pytest

import checkers
from eval_cases import EVAL_CASES, METRIC_NAMES, eval_target, split_by_date
from tidegbm_amd import Booster, TideError, TideRegressor, synth

pytestmark = pytest.mark.gpu

METRIC_RTOL = 1e-6
PRED_RTOL = 1e-5
HIP = Path(__file__).resolve().parent.parent / "tidegbm_amd" / "lib" / "libtgbm_hip.so"
_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _lib():
    L = ctypes.CDLL(str(HIP))
    L.tgbm_eval_create.argtypes = [_P, _I, _I, _P, _P, _P, _P, _I, _P, _P, _I]
    L.tgbm_eval_destroy.argtypes = [_P]
    L.tgbm_eval_destroy.restype = None
    L.tgbm_eval_reset.argtypes = [_P, _D]
    L.tgbm_eval_add_tree.argtypes = [_P, _I, _P, _P, _D, _P]
    L.tgbm_eval_metric.argtypes = [_P, _I, _P]
    L.tgbm_eval_predictions.argtypes = [_P, _P, _I]
    L.tgbm_last_error.restype = ctypes.c_char_p
    return L


def _walk(ints, row, nb):
    k = 0
    while not ints[k, 0]:
        b1 = row[ints[k, 2]]
        if ints[k, 1] == 1:
            b2 = row[ints[k, 3]]
            if b1 == 0 or b2 == 0:
                left = bool(ints[k, 7])
            else:
                left = ((b1 * 2654435761 + b2 * 40503) % (1 << 32)) % nb <= ints[k, 4]
        else:
            left = bool(ints[k, 7]) if b1 == 0 else b1 <= ints[k, 4]
        k = ints[k, 5] if left else ints[k, 6]
    return k


def _node(is_leaf, typ=0, f=0, f2=0, thr=0, left=0, right=0, mgl=1):
    return [is_leaf, typ, f, f2, thr, left, right, mgl, 0, 0]


def _pearson(y, p, ix):  # booster.cpp:132-153, same operation order
    m = len(ix)
    if m < 3:
        return 0.0
    my = mp = 0.0
    for k in ix:
        my += y[k]
        mp += p[k]
    my /= m
    mp /= m
    vy = vp = c = 0.0
    for k in ix:
        dy = y[k] - my
        dp = p[k] - mp
        vy += dy * dy
        vp += dp * dp
        c += dy * dp
    if vy < 1e-18 or vp < 1e-18:
        return 0.0
    return c / math.sqrt(vy * vp)


def test_eval_kernels_match_reference_arithmetic():
    rng = np.random.default_rng(7)
    n, F, nb = 3000, 4, 255
    X = rng.normal(size=(n, F))
    X[rng.random(size=(n, F)) < 0.05] = np.nan
    y = rng.normal(size=n)
    edges = [np.unique(np.quantile(X[~np.isnan(X[:, f]), f], np.linspace(0.02, 0.98, 40))) for f in range(F)]
    E = np.zeros((F, 256))
    ne = np.zeros(F, np.int32)
    for f, e in enumerate(edges):
        E[f, :len(e)] = e
        ne[f] = len(e)
    xb = checkers.bins_of(X, edges)
    # dates of 1..9 rows in a shuffled order; only dates of >= 3 rows are listed
    ts = rng.integers(0, 700, size=n)
    order, gstart = [], [0]
    for d in np.unique(ts):
        rows = np.flatnonzero(ts == d)
        rng.shuffle(rows)
        if rows.size >= 3:
            order.extend(rows.tolist())
            gstart.append(len(order))
    order = np.asarray(order, np.int32)
    gstart = np.asarray(gstart, np.int32)
    trees = [
        (np.array([_node(0, 0, 0, 0, 20, 1, 2, 1), _node(0, 0, 1, 0, 11, 3, 4, 0), _node(1), _node(1), _node(1)],
                  np.int32), np.array([0.0, 0.0, 0.25, -0.5, 1.0])),
        (np.array([_node(0, 1, 2, 3, 100, 1, 2, 0), _node(1), _node(1)], np.int32), np.array([0.0, -0.75, 0.4])),
        (np.array([_node(1)], np.int32), np.array([0.125])),
    ]
    init, lr = 0.3, 0.1
    L = _lib()
    h = _P()
    _say(f"[case] eval kernels n={n} F={F} dates={gstart.size - 1}")
    assert L.tgbm_eval_create(ctypes.byref(h), F, nb, E.ctypes.data, ne.ctypes.data, X.ctypes.data, y.ctypes.data,
                              n, order.ctypes.data, gstart.ctypes.data, gstart.size - 1) == 0, L.tgbm_last_error()
    try:
        assert L.tgbm_eval_reset(h, init) == 0
        expect = [init] * n
        for ints, vals in trees:
            assert L.tgbm_eval_add_tree(h, ints.shape[0], ints.ctypes.data, vals.ctypes.data, lr, None) == 0
            for i in range(n):
                expect[i] = expect[i] + lr * float(vals[_walk(ints, xb[i], nb)])
        got = np.empty(n)
        assert L.tgbm_eval_predictions(h, got.ctypes.data, n) == 0
        assert np.array_equal(got, np.asarray(expect))
        out = _D()
        assert L.tgbm_eval_metric(h, 0, ctypes.byref(out)) == 0
        s = 0.0
        for i in range(n):
            d = expect[i] - y[i]
            s += 1.0 * d * d
        assert out.value == s / n
        assert L.tgbm_eval_metric(h, 1, ctypes.byref(out)) == 0
        acc = 0.0
        for g in range(gstart.size - 1):
            acc += _pearson(y, expect, order[gstart[g]:gstart[g + 1]])
        assert out.value == acc / (gstart.size - 1)
    finally:
        L.tgbm_eval_destroy(h)


def _fit_pair(case):
    iid, objective, metric, esr, nan_y = EVAL_CASES[case]
    p = synth.panel_small(21, n_dates=120, n_entities=40)
    tr, va = split_by_date(p)
    ye = eval_target(p, va, nan_y)
    params = dict(n_estimators=15, max_depth=5, learning_rate=0.1, goss=0, objective=objective,
                  early_stopping_rounds=esr)
    gpu = Booster(eval_metric=METRIC_NAMES[metric], **params)
    gpu.fit(p.X[tr], p.y[tr], timestamps=None if iid else p.ts[tr], entity=None if iid else p.entity[tr],
            rank_groups=p.ts[tr] if objective == 5 else None, iid=iid, horizon=1,
            eval_set=(p.X[va], ye), eval_timestamp=None if iid else p.ts[va])
    cpu = []
    for lib in (checkers.oracle(), checkers.reference()):
        if lib is None:
            continue
        m = lib.model(eval_metric=metric, **params)
        m.set_eval(p.X[va], ye, None if iid else p.ts[va])
        m.fit(p.X[tr], p.y[tr], ts=None if iid else p.ts[tr], entity=None if iid else p.entity[tr],
              rank_groups=p.ts[tr] if objective == 5 else None, iid=iid, horizon=1)
        cpu.append(m)
    return p, tr, va, gpu, cpu


@pytest.mark.parametrize("case", sorted(EVAL_CASES))
def test_eval_set_fit_matches_oracle(case):
    _say(f"[case] eval fit {case} {EVAL_CASES[case]}")
    p, tr, va, gpu, cpu = _fit_pair(case)
    ev = gpu.evals_result()
    xb = checkers.bins_of(p.X[tr], [gpu.edges(f) for f in range(p.n_features)])
    nb = gpu.model_info()["n_bins"]
    for m in cpu:
        vo = m.evals_valid()
        assert len(ev["valid"]) == len(vo) >= 1, (len(ev["valid"]), len(vo))
        assert np.array_equal(np.isnan(ev["valid"]), np.isnan(vo))
        ok = ~np.isnan(vo)
        assert np.allclose(ev["valid"][ok], vo[ok], rtol=METRIC_RTOL, atol=1e-12), (ev["valid"], vo)
        assert len(ev["train"]) == len(m.evals_train())
        assert gpu.best_iteration == m.best_iteration(), (gpu.best_iteration, m.best_iteration())
        diff = checkers.compare_trees(gpu.trees(), m.trees(), rtol=1e-7, atol=1e-9, xb=xb, n_bins=nb)
        assert diff is None, diff
        a, b = gpu.predict(p.X[va]), m.predict(p.X[va])
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < PRED_RTOL
    _say(f"  trees kept {len(gpu.trees())}, best_iteration {gpu.best_iteration}, valid[0] {ev['valid'][0]}")


def test_estimator_eval_set_like_reference_tests():
    # reference tests/test_tidegbm.py:79-94 (future eval ok) and :468-491 (eval_metric="ic")
    p = synth.panel_small(40, n_dates=90, n_entities=6, n_features=5)
    tr, va = p.ts < 60, p.ts >= 70
    m = TideRegressor(n_estimators=25, random_state=0, horizon=1, pair_splits=False, early_stopping_rounds=8,
                      eval_metric="ic")
    m.fit(p.X[tr], p.y[tr], timestamp=p.ts[tr], eval_set=(p.X[va], p.y[va]), eval_timestamp=p.ts[va])
    v = np.asarray(m.evals_result_["valid"])
    assert len(v) >= 1
    assert np.all(np.abs(v) <= 1.0 + 1e-9)
    assert m.best_iteration_ is not None


def test_eval_in_the_past_raises():
    # reference tests/test_tidegbm.py:63-76
    p = synth.panel_small(3, n_dates=80, n_entities=4, n_features=5)
    tr, va = p.ts < 50, p.ts < 20
    m = TideRegressor(n_estimators=5, early_stopping_rounds=2, horizon=1)
    with pytest.raises(ValueError, match="overlap"):
        m.fit(p.X[tr], p.y[tr], timestamp=p.ts[tr], eval_set=(p.X[va], p.y[va]), eval_timestamp=p.ts[va])
    with pytest.raises(TideError, match="eval_timestamp"):
        Booster(n_estimators=3).fit(p.X[tr], p.y[tr], timestamps=p.ts[tr], horizon=1,
                                    eval_set=(p.X[va], p.y[va]))
