"""The tree comparison helper accepts exactly one characterized difference and nothing else.

The exception: missing_go_left of an axis split whose node saw no training row with the split
feature missing (the reference decides it from FP64 subtraction residue in an empty bin).  A
flipped flag there is accepted only when the training bins are supplied; a flipped flag on a
node that does route missing rows, or any other field, is still reported.
"""
import copy

import numpy as np

import checkers
from tidegbm_amd import synth


def _fit():
    p = synth.panel_large(4, n_dates=150, n_entities=120, n_features=20)
    o = checkers.oracle().model(n_estimators=2, max_depth=7, learning_rate=0.1, goss=0, pair_splits=0)
    o.fit(p.X, p.y, ts=p.ts, entity=p.entity, horizon=1)
    xb = checkers.bins_of(p.X, [o.edges(f) for f in range(p.n_features)])
    return o.trees(), xb


def _nodes(trees, xb, want_missing):
    for t, tr in enumerate(trees):
        ints = tr["ints"]
        for k in range(ints.shape[0]):
            if ints[k, 0] == 0 and ints[k, 1] == 0:
                m = checkers.rows_reaching(ints, xb, k)
                if bool(np.any(xb[m, ints[k, 2]] == 0)) == want_missing:
                    yield t, k


def test_flag_flip_on_node_without_missing_rows_is_accepted():
    trees, xb = _fit()
    t, k = next(_nodes(trees, xb, want_missing=False))
    other = copy.deepcopy(trees)
    other[t]["ints"][k, 7] ^= 1
    assert checkers.compare_trees(trees, other, xb=xb) is None
    assert checkers.compare_trees(trees, other) is not None   # without bins: strict


def test_flag_flip_on_node_with_missing_rows_is_reported():
    trees, xb = _fit()
    t, k = next(_nodes(trees, xb, want_missing=True))
    other = copy.deepcopy(trees)
    other[t]["ints"][k, 7] ^= 1
    assert "missing_go_left" in checkers.compare_trees(trees, other, xb=xb)


def test_other_fields_are_reported():
    trees, xb = _fit()
    t, k = next(_nodes(trees, xb, want_missing=False))
    other = copy.deepcopy(trees)
    other[t]["ints"][k, 4] += 1
    assert "threshold" in checkers.compare_trees(trees, other, xb=xb)
