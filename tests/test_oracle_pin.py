"""Pin the CPU oracle (oracle/tide_oracle.cpp) to the reference core built from its own sources.

Every comparison is bit-exact: the oracle restates the reference's sequential FP64 arithmetic, so
trees (all node fields, values, TORB prefix values, cover) and predictions must be identical,
not merely close.  Inputs are seeded panels sized so the whole file runs in well under a minute.
Skipped when oracle/_ref was not built (no reference tree on this machine); the committed golden
fixtures (tests/test_golden.py) pin the oracle there.
"""
import numpy as np
import pytest

import checkers
from cat_cases import CAT_CASES, GROUP_CASES, cat_panel, group_arrays, group_panel
from eval_cases import EVAL_CASES, eval_target, split_by_date
from tidegbm_amd import synth

REF = checkers.reference()
pytestmark = pytest.mark.skipif(REF is None, reason="reference core (oracle/_ref) not built")


def _both(p, rg=None, iid=False, **params):
    out = []
    for lib in (checkers.oracle(), REF):
        m = lib.model(**params)
        m.fit(p.X, p.y, ts=None if iid else p.ts, entity=None if iid else p.entity, rank_groups=rg,
              iid=iid, horizon=1)
        out.append(m)
    return out


def _assert_identical(a, b, X):
    ta, tb = a.trees(), b.trees()
    assert len(ta) == len(tb)
    for t, (x, y) in enumerate(zip(ta, tb)):
        for k in ("ints", "dbl", "values", "cover", "prefix"):
            assert np.array_equal(x[k], y[k]), f"tree {t} field {k}"
    assert np.array_equal(a.predict(X), b.predict(X))


def xor_panel(seed):
    """Small panel whose target is an interaction, so weak leaves take PAIR splits."""
    p = synth.panel_small(seed)
    rng = np.random.default_rng(seed)
    y = np.sign(p.X[:, 0] * p.X[:, 1]) + 0.1 * rng.normal(size=p.n)
    return synth.Panel(p.X, y, p.ts, p.entity)


def test_mse_time_mode_default_config():
    p = synth.panel_small(0)
    o, r = _both(p, n_estimators=40, max_depth=6, learning_rate=0.08, goss=0)
    _assert_identical(o, r, p.X)


@pytest.mark.parametrize("seed", [0, 11])
def test_mse_time_mode_with_pair_splits(seed):
    p = xor_panel(seed)
    o, r = _both(p, n_estimators=20, max_depth=6, learning_rate=0.08, goss=0)
    _assert_identical(o, r, p.X)
    n_pair = sum(int((t["ints"][:, 1] == 1).sum()) for t in o.trees())
    assert n_pair >= 5, "case should exercise PAIR splits"


def test_mse_iid_mode():
    p = synth.panel_small(2)
    o, r = _both(p, iid=True, n_estimators=20, max_depth=5, learning_rate=0.1, goss=0)
    _assert_identical(o, r, p.X)


def test_pairwise_objective():
    p = synth.panel_small(3, n_dates=150, n_entities=70)
    o, r = _both(p, rg=p.ts, n_estimators=12, max_depth=6, learning_rate=0.08, goss=0, objective=5)
    _assert_identical(o, r, p.X)


def test_goss_sampling():
    p = synth.panel_small(4)
    o, r = _both(p, n_estimators=12, max_depth=5, learning_rate=0.1, goss=1)
    _assert_identical(o, r, p.X)


def test_bin_sampling_nan_and_long_horizon():
    # > bin_sample values per feature exercise the mt19937 + std::shuffle sub-sample; horizon 10
    # gives 64 TORB buckets; NaNs route through missing_go_left.
    p = synth.panel_large(5, n_dates=120, n_entities=60, n_features=12, nan_frac=0.05)
    params = dict(n_estimators=6, max_depth=6, learning_rate=0.1, goss=0, bin_sample=1000)
    ms = []
    for lib in (checkers.oracle(), REF):
        m = lib.model(**params)
        m.fit(p.X, p.y, ts=p.ts, entity=p.entity, horizon=10)
        ms.append(m)
    _assert_identical(ms[0], ms[1], p.X)
    for f in range(p.n_features):
        assert np.array_equal(ms[0].edges(f), ms[1].edges(f))
    assert np.array_equal(ms[0].init_prefix(), ms[1].init_prefix())


def test_unsorted_rows_and_update():
    p = synth.panel_small(6)
    perm = np.random.default_rng(0).permutation(p.n)
    q = synth.Panel(p.X[perm], p.y[perm], p.ts[perm], p.entity[perm])
    o, r = _both(q, n_estimators=8, max_depth=5, learning_rate=0.1, goss=0, pair_splits=0)
    _assert_identical(o, r, q.X)
    extra = synth.panel_small(7, n_dates=80)
    o.update(extra.X, extra.y, 4)
    r.update(extra.X, extra.y, 4)
    _assert_identical(o, r, extra.X)


@pytest.mark.parametrize("objective", [0, 5])
def test_training_loss_per_tree(objective):
    # evals.train (objective.cpp:262-329, booster.cpp:704-706): weighted MSE, or minus the mean
    # within-date Pearson correlation for the pairwise objective; bit-exact sequential sums
    p = synth.panel_small(12)
    w = np.random.default_rng(1).uniform(0.5, 2.0, p.n)
    rg = p.ts if objective == 5 else None
    out = []
    for lib in (checkers.oracle(), REF):
        m = lib.model(n_estimators=6, max_depth=5, learning_rate=0.1, goss=0, pair_splits=0, objective=objective)
        m.fit(p.X, p.y, ts=p.ts, w=w if objective == 0 else None, entity=p.entity, rank_groups=rg, horizon=1)
        out.append(m.evals_train())
    assert len(out[0]) == 6
    assert np.array_equal(out[0], out[1]), (out[0], out[1])
    if objective == 0:
        assert np.all(np.diff(out[0]) < 0)      # the training MSE falls tree by tree


@pytest.mark.parametrize("case", sorted(EVAL_CASES))
def test_eval_set_and_early_stopping(case):
    iid, objective, metric, esr, nan_y = EVAL_CASES[case]
    p = synth.panel_small(21, n_dates=120, n_entities=40)
    tr, va = split_by_date(p)
    ye = eval_target(p, va, nan_y)
    out = []
    for lib in (checkers.oracle(), REF):
        m = lib.model(n_estimators=15, max_depth=5, learning_rate=0.1, goss=0, objective=objective,
                      eval_metric=metric, early_stopping_rounds=esr)
        m.set_eval(p.X[va], ye, None if iid else p.ts[va])
        m.fit(p.X[tr], p.y[tr], ts=None if iid else p.ts[tr], entity=None if iid else p.entity[tr],
              rank_groups=p.ts[tr] if objective == 5 else None, iid=iid, horizon=1)
        out.append(m)
    o, r = out
    vo, vr = o.evals_valid(), r.evals_valid()
    assert len(vo) >= 1
    assert np.array_equal(vo, vr, equal_nan=True), (vo, vr)
    assert o.best_iteration() == r.best_iteration()
    assert np.array_equal(o.evals_train(), r.evals_train())
    _assert_identical(o, r, p.X)
    if nan_y:
        # a NaN metric never improves: every tree is scored and kept until the stall limit
        assert np.isnan(vo).all() and len(o.trees()) == esr and o.best_iteration() == 15
    elif esr:
        # the first iteration improves; the validation score then uses that one tree, so the
        # metric stays put and the fit stops after esr more trees, keeping one
        assert len(vo) == esr + 1 and len(o.trees()) == 1 and o.best_iteration() == 1
        assert np.all(vo == vo[0])


def test_eval_in_the_past_raises():
    p = synth.panel_small(22, n_dates=80, n_entities=20)
    tr = p.ts < 50
    va = p.ts < 20
    for lib in (checkers.oracle(), REF):
        m = lib.model(n_estimators=5, early_stopping_rounds=2, goss=0)
        m.set_eval(p.X[va], p.y[va], p.ts[va])
        with pytest.raises(ValueError, match="overlap"):
            m.fit(p.X[tr], p.y[tr], ts=p.ts[tr], entity=p.entity[tr], horizon=1)


@pytest.mark.parametrize("case", sorted(CAT_CASES))
def test_categorical_features(case):
    cc = CAT_CASES[case]
    p, is_cat = cat_panel(31)
    out = []
    for lib in (checkers.oracle(), REF):
        m = lib.model(learning_rate=0.1, goss=0, objective=cc["objective"], **cc["params"])
        m.set_cat(is_cat)
        m.fit(p.X, p.y, ts=None if cc["iid"] else p.ts, entity=None if cc["iid"] else p.entity,
              rank_groups=p.ts if cc["objective"] == 5 else None, iid=cc["iid"], horizon=1)
        out.append(m)
    o, r = out
    for f in (3, 7, 11):
        assert np.array_equal(o.cat_values(f), r.cat_values(f))
    # 300 codes of about 20 rows each: the 254 most frequent are kept; with cat_min_count 40
    # none qualifies and the single most frequent code is kept (binning.cpp:66-74)
    assert o.cat_values(11).size == (1 if cc["params"].get("cat_min_count") else 254)
    _assert_identical(o, r, p.X)
    for x, y in zip(o.trees(), r.trees()):
        assert np.array_equal(x["cat"], y["cat"])
    n_cat = sum(int(np.isin(t["ints"][:, 1], (2, 3)).sum()) for t in o.trees())
    assert n_cat >= 5, "case should exercise categorical splits"


@pytest.mark.parametrize("case", sorted(GROUP_CASES))
def test_feature_groups_interactions_regime(case):
    gc = GROUP_CASES[case]
    p, is_cat = group_panel(case)
    grp, ix = group_arrays(case, p.n_features)
    out = []
    for lib in (checkers.oracle(), REF):
        m = lib.model(learning_rate=0.1, goss=0, **gc["params"])
        m.set_cat(is_cat)
        m.set_groups(grp, ix)
        m.fit(p.X, p.y, ts=p.ts, entity=p.entity, horizon=1)
        out.append(m)
    o, r = out
    _assert_identical(o, r, p.X)
    for x, y in zip(o.trees(), r.trees()):
        assert np.array_equal(x["cat"], y["cat"])
    if gc["params"].get("regime_feature") == 7:
        assert sum(int((t["ints"][:, 1] == 3).sum()) for t in o.trees()) >= 3, "expected Regime nodes"


def test_histogram_and_scan_functions():
    # per-function pins: the restated histogram / split scan / gradients vs the reference's own
    rng = np.random.default_rng(3)
    n, F, nb = 3000, 7, 255
    xb = rng.integers(0, nb, size=(n, F)).astype(np.uint8)
    xb[rng.random((n, F)) < 0.05] = 0
    g = rng.normal(size=n)
    h = rng.uniform(0.5, 2.0, size=n)
    idx = np.sort(rng.choice(n, size=1700, replace=False)).astype(np.int32)
    O = checkers.oracle().lib
    G = np.zeros(F * nb); H = np.zeros(F * nb); C = np.zeros(F * nb, np.int64)
    O.orc_hist(xb.ctypes.data, F, idx.ctypes.data, idx.size, g.ctypes.data, h.ctypes.data, nb,
               G.ctypes.data, H.ctypes.data, C.ctypes.data)
    xcol = np.ascontiguousarray(xb.T.astype(np.uint16))
    G2 = np.zeros(F * nb); H2 = np.zeros(F * nb); C2 = np.zeros(F * nb, np.int64)
    REF.lib.tref_build_hist_col(xcol.ctypes.data, n, F, idx.ctypes.data, idx.size, g.ctypes.data,
                                h.ctypes.data, nb, G2.ctypes.data, H2.ctypes.data, C2.ctypes.data)
    assert np.array_equal(G, G2) and np.array_equal(H, H2) and np.array_equal(C, C2)
    gain_f = np.zeros(F); bin_f = np.zeros(F, np.int32); ml_f = np.zeros(F, np.int32)
    bf = np.zeros(1, np.int32)
    O.orc_scan(G.ctypes.data, H.ctypes.data, C.ctypes.data, F, nb, 1.0, 0.0, 20, 1e-3, 0.0,
               gain_f.ctypes.data, bin_f.ctypes.data, ml_f.ctypes.data, bf.ctypes.data)
    import ctypes
    for f in range(F):
        gg = ctypes.c_double(); bb = ctypes.c_int(); mm = ctypes.c_int()
        REF.lib.tref_best_numeric(G.ctypes.data, H.ctypes.data, C.ctypes.data, F, nb, f, 1.0, 0.0, 20, 1e-3,
                                  0, 0.0, ctypes.byref(gg), ctypes.byref(bb), ctypes.byref(mm))
        assert gg.value == gain_f[f] and bb.value == bin_f[f] and mm.value == ml_f[f]
    y = rng.normal(size=n)
    pred = rng.normal(size=n) * 0.1
    grp = np.repeat(np.arange(n // 100), 100).astype(np.int64)
    w = rng.uniform(0.5, 1.5, size=n)
    for kind in (0, 5):
        ga = np.zeros(n); ha = np.zeros(n); gb = np.zeros(n); hb = np.zeros(n)
        O.orc_grad(kind, y.ctypes.data, pred.ctypes.data, w.ctypes.data, grp.ctypes.data, n, ga.ctypes.data, ha.ctypes.data)
        REF.lib.tref_grad(kind, y.ctypes.data, pred.ctypes.data, w.ctypes.data, grp.ctypes.data, n, gb.ctypes.data, hb.ctypes.data)
        assert np.array_equal(ga, gb) and np.array_equal(ha, hb)
