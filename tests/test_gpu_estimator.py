"""The estimator front end on the GPU against the oracle fed the same rows.

TideRegressor / TideRanker prepare data like the reference's Python layer (src/tidegbm/native.py:
81-102, 134-137; utils.py:47-70): entity codes from np.unique, rows ordered by (timestamp, entity
code) with ties in input order, horizon and embargo given in steps converted to timestamp units by
the median positive step, rank groups = timestamps.  Here the rows arrive shuffled, with
timestamps 1000 + 7 * date (step 7) and string entity labels; the oracle receives the same
shuffled rows with the reference's own conversions restated in closed form (its Booster::fit
sorts by (timestamp, entity, row) itself, booster.cpp:264-309).  Bar: tree structure exact (the
checker's row-verified exceptions aside), leaf values within 1e-7 relative, predictions within
1e-5 relative.  Also the reference's test_cuda_compiled_matches_device_gpu (tests/test_markets.py:
22-34) with a GPU present: device="gpu" fits and predicts.
"""
import numpy as np
import pytest

import checkers
import tidegbm_amd as t
from tidegbm_amd import synth

pytestmark = pytest.mark.gpu


def _shuffled_panel(seed, **kw):
    p = synth.panel_small(seed, **kw)
    rng = np.random.default_rng(seed + 1)
    perm = rng.permutation(p.n)
    ts = 1000 + 7 * p.ts[perm]
    names = np.array([f"asset-{e:03d}" for e in range(int(p.entity.max()) + 1)])
    ent = names[p.entity[perm]]
    return p.X[perm], p.y[perm], ts, ent


@pytest.mark.parametrize("ranker", [False, True])
def test_estimator_data_prep_matches_oracle(ranker):
    X, y, ts, ent = _shuffled_panel(31, n_dates=160, n_entities=50)
    params = dict(n_estimators=8, max_depth=5, learning_rate=0.1, goss=0, pair_splits=False)
    est = (t.TideRanker if ranker else t.TideRegressor)(horizon=2, embargo=1, **params)
    print(f"[case] estimator ranker={ranker} n={len(y)}", flush=True)
    est.fit(X, y, timestamp=ts, entity=ent)
    # the reference's conversions, restated: median positive step 7, entity codes by np.unique
    codes = np.unique(ent, return_inverse=True)[1].astype(np.int64)
    o = checkers.oracle().model(**params, embargo=7, objective=5 if ranker else 0)
    o.fit(X, y, ts=ts, entity=codes, rank_groups=ts if ranker else None, horizon=14)
    b = est.booster_
    order = np.lexsort((codes, ts))
    xb = checkers.bins_of(X[order], [b.edges(f) for f in range(X.shape[1])])
    diff = checkers.compare_trees(b.trees(), o.trees(), rtol=1e-7, atol=1e-9, xb=xb,
                                  n_bins=b.model_info()["n_bins"])
    assert diff is None, diff
    assert np.array_equal(b.init_prefix(), o.init_prefix())
    a, r = est.predict(X), o.predict(X)
    assert np.max(np.abs(a - r) / np.maximum(np.abs(r), 1e-3)) < 1e-5


def test_device_gpu_fits_when_compiled():
    # reference tests/test_markets.py:22-34, the branch taken when GPU support is compiled in
    assert t.gpu_compiled and t.device_count() > 0
    rng = np.random.default_rng(0)
    X = rng.normal(size=(40, 3))
    y = X[:, 0]
    m = t.TideRegressor(n_estimators=5, iid_mode=True, device="gpu", pair_splits=False)
    m.fit(X, y)
    assert m.predict(X).shape == (40,)
    with pytest.raises(ValueError, match="device=cpu"):
        t.TideRegressor(n_estimators=5, iid_mode=True, device="cpu").fit(X, y)
