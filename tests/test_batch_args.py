"""Argument checks of the batched walk-forward entry point (host side, no GPU needed: every
check runs before the device is touched)."""
import numpy as np
import pytest

import tidegbm_amd as t
from tidegbm_amd import synth

PARAMS = dict(n_estimators=3, max_depth=4, pair_splits=0, goss=0)


def _data():
    p = synth.panel_small(5, n_dates=120, n_entities=10, n_features=6)
    tasks = synth.walk_forward_tasks(p, n_folds=2, horizons=(1, 3), first_dates=40)
    return p, [k.n_train for k in tasks], [k.y for k in tasks], [k.horizon for k in tasks]


def test_pair_splits_rejected():
    p, rows, ys, hz = _data()
    with pytest.raises(t.TideError, match="pair_splits"):
        t.fit_batch(p.X, p.ts, rows, ys, entity=p.entity, horizons=hz, **{**PARAMS, "pair_splits": 1})


def test_goss_rejected():
    p, rows, ys, hz = _data()
    with pytest.raises(t.TideError, match="goss"):
        t.fit_batch(p.X, p.ts, rows, ys, entity=p.entity, horizons=hz, **{**PARAMS, "goss": 1})


def test_unsorted_rows_rejected():
    p, rows, ys, hz = _data()
    perm = np.random.default_rng(0).permutation(p.n)
    with pytest.raises(t.TideError, match="sorted"):
        t.fit_batch(p.X[perm], p.ts[perm], rows, ys, entity=p.entity[perm], horizons=hz, **PARAMS)


def test_missing_horizon_rejected():
    p, rows, ys, _ = _data()
    with pytest.raises(t.TideError, match="horizon"):
        t.fit_batch(p.X, p.ts, rows, ys, entity=p.entity, **PARAMS)


def test_shape_errors():
    p, rows, ys, hz = _data()
    with pytest.raises(t.TideError, match="target"):
        t.fit_batch(p.X, p.ts, rows, ys[:-1], entity=p.entity, horizons=hz, **PARAMS)
    with pytest.raises(t.TideError, match="rows"):
        t.fit_batch(p.X, p.ts, [p.n + 1] + rows[1:], [np.zeros(p.n + 1)] + ys[1:], entity=p.entity,
                    horizons=hz, **PARAMS)
    with pytest.raises(t.TideError, match="schedule"):
        t.fit_batch(p.X, p.ts, rows, ys, entity=p.entity, horizons=hz, schedule="bogus", **PARAMS)


def test_valid_call_reaches_the_device_check():
    p, rows, ys, hz = _data()
    if t.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(t.TideError, match="GPU"):
        t.fit_batch(p.X, p.ts, rows, ys, entity=p.entity, horizons=hz, **PARAMS)
