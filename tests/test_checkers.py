"""The tree comparison helper accepts exactly two characterized differences and nothing else.

1. missing_go_left of an axis split whose node saw no training row with the split feature
   missing (the reference decides it from FP64 subtraction residue in an empty bin).  A flipped
   flag there is accepted only when the training bins are supplied; a flipped flag on a node that
   does route missing rows, or any other field, is still reported.
2. a different split feature when both splits send the same training rows of the node left
   (equal gains in exact arithmetic; the reference's FP64 sums differ by ulps).  A feature whose
   split partitions the rows differently is still reported.
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


def _equal_partition_panel():
    # feature 3 is a copy of feature 0 with its values shifted, so both give the same bins order
    p = synth.panel_small(21, n_dates=60, n_entities=20, n_features=5)
    X = p.X.copy()
    X[:, 3] = X[:, 0] * 2.0 + 1.0
    o = checkers.oracle().model(n_estimators=1, max_depth=4, learning_rate=0.1, goss=0, pair_splits=0)
    o.fit(X, p.y, ts=p.ts, entity=p.entity, horizon=1)
    xb = checkers.bins_of(X, [o.edges(f) for f in range(p.n_features)])
    return o.trees(), xb


def test_equal_partition_feature_swap_is_accepted():
    trees, xb = _equal_partition_panel()
    ints = trees[0]["ints"]
    k = next(k for k in range(ints.shape[0]) if ints[k, 0] == 0 and ints[k, 2] in (0, 3))
    other = copy.deepcopy(trees)
    f = ints[k, 2]
    other[0]["ints"][k, 2] = 3 if f == 0 else 0   # the twin feature: same bins, same threshold
    assert checkers.compare_trees(trees, other, xb=xb) is None
    assert checkers.compare_trees(trees, other) is not None   # without bins: strict


def test_feature_swap_with_another_partition_is_reported():
    trees, xb = _equal_partition_panel()
    ints = trees[0]["ints"]
    k = next(k for k in range(ints.shape[0]) if ints[k, 0] == 0)
    other = copy.deepcopy(trees)
    other[0]["ints"][k, 2] = (ints[k, 2] + 1) % 5 if (ints[k, 2] + 1) % 5 not in (0, 3) else (ints[k, 2] + 2) % 5
    assert "feature" in checkers.compare_trees(trees, other, xb=xb)


def test_mirrored_tie_split_with_equal_function_is_accepted():
    # feature 3 = -feature 0: "f3 <= t'" sends left exactly the rows "f0 <= t" sends right, so a
    # split on either cuts the same two sets; the tree with f3 and its children swapped computes
    # the same function on the training rows
    p = synth.panel_small(22, n_dates=60, n_entities=20, n_features=5)
    X = p.X.copy()
    X[:, 3] = -X[:, 0]
    o = checkers.oracle().model(n_estimators=1, max_depth=4, learning_rate=0.1, goss=0, pair_splits=0)
    o.fit(X, p.y, ts=p.ts, entity=p.entity, horizon=1)
    trees = o.trees()
    xb = checkers.bins_of(X, [o.edges(f) for f in range(p.n_features)])
    ints = trees[0]["ints"]
    done = False
    for k in range(ints.shape[0]):
        if ints[k, 0] or ints[k, 2] != 0:
            continue
        m = checkers.rows_reaching(ints, xb, k)
        right = m & (xb[:, 0] > ints[k, 4])
        if not right.any():
            continue
        t3 = int(xb[right, 3].max())
        if not np.array_equal(m & (xb[:, 3] <= t3), right):
            continue
        other = copy.deepcopy(trees)
        oi = other[0]["ints"]
        oi[k, 2], oi[k, 4], oi[k, 7] = 3, t3, 0
        oi[k, 5], oi[k, 6] = ints[k, 6], ints[k, 5]
        assert checkers.compare_trees(trees, other, xb=xb) is None
        # the children swapped without the mirrored split: another function, reported
        bad = copy.deepcopy(trees)
        bad[0]["ints"][k, 5], bad[0]["ints"][k, 6] = ints[k, 6], ints[k, 5]
        assert checkers.compare_trees(trees, bad, xb=xb) is not None
        done = True
        break
    assert done, "no split on feature 0 to mirror"
