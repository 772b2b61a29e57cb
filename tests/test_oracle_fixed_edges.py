"""The oracle's fixed-edge mode (used to check fits that share one binned matrix).

Supplying the edges a fit would compute itself must reproduce that fit bit-for-bit, so the mode
changes nothing but where the edges come from; edges from a larger window are kept verbatim and
the fit over a row prefix uses them.
"""
import numpy as np

import checkers
from tidegbm_amd import synth

PARAMS = dict(n_estimators=8, max_depth=5, learning_rate=0.1, pair_splits=0, goss=0, min_samples_leaf=10)


def test_own_edges_reproduce_the_fit():
    p = synth.panel_small(2, n_dates=200, n_entities=20, n_features=8)
    orc = checkers.oracle()
    a = orc.model(**PARAMS).fit(p.X, p.y, ts=p.ts, entity=p.entity, horizon=3)
    edges = [a.edges(f) for f in range(p.n_features)]
    b = orc.model(**PARAMS).set_edges(edges).fit(p.X, p.y, ts=p.ts, entity=p.entity, horizon=3)
    for x, y in zip(a.trees(), b.trees()):
        for k in ("ints", "dbl", "values", "cover", "prefix"):
            assert np.array_equal(x[k], y[k])
    assert np.array_equal(a.predict(p.X), b.predict(p.X))


def test_prefix_fit_keeps_supplied_edges():
    p = synth.panel_small(4, n_dates=200, n_entities=20, n_features=8)
    orc = checkers.oracle()
    full = orc.model(**PARAMS).fit(p.X, p.y, ts=p.ts, entity=p.entity, horizon=1)
    edges = [full.edges(f) for f in range(p.n_features)]
    n = 1000
    own = orc.model(**PARAMS).fit(p.X[:n], p.y[:n], ts=p.ts[:n], entity=p.entity[:n], horizon=1)
    fixed = orc.model(**PARAMS).set_edges(edges).fit(p.X[:n], p.y[:n], ts=p.ts[:n], entity=p.entity[:n],
                                                      horizon=1)
    assert all(np.array_equal(fixed.edges(f), edges[f]) for f in range(p.n_features))
    assert any(not np.array_equal(own.edges(f), edges[f]) for f in range(p.n_features))
    # thresholds index the supplied edges: every split bin is within the feature's bin range
    for t in fixed.trees():
        for node in t["ints"]:
            if node[0] == 0:
                assert 1 <= node[4] <= len(edges[node[2]]) + 1
