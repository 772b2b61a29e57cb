"""Multi-leaf growth against one split per step, on the GPU.

Device-driven growth may split several leaves per step (spec_width > 1): each step splits the
leaf best-first takes next plus the next-best unsplit leaves of the frontier, and the tree is the
best-first replay over the splits made (reference grow_tree, tree.cpp:343-463, one split at a
time).  Bar: every tree identical field for field (node ints, gains, values, covers, TORB
prefixes) to spec_width = 1, and predictions bit-identical.  Cases: the plain panel, weighted rows
(h varies), GOSS samples, a depth limit that binds, a split-gain floor, and features with most rows
in one bin.  Width 2 with 31 leaves needs more steps than the first captured loop holds, so the
step-by-step continuation runs too.
The trees of spec_width = 1 are themselves checked against the oracle and the reference core by
the other GPU tests.
"""
import sys

import numpy as np
import pytest

import tidegbm_amd as t
from tidegbm_amd import synth

pytestmark = pytest.mark.gpu
KEYS = ("ints", "dbl", "values", "cover", "prefix")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fit(p, width, w=None, **kw):
    params = dict(n_estimators=25, max_depth=8, learning_rate=0.05, max_bin=255, autotune=0, pair_splits=0)
    params.update(kw)
    b = t.Booster(spec_width=width, **params)
    b.fit(p.X, p.y, timestamps=p.ts, entity=p.entity, weight=w, horizon=1)
    return b


CASES = {
    "plain": dict(goss=0),
    "skewed": dict(goss=0, skewed=True),
    "weighted": dict(goss=0, weighted=True),
    "goss": dict(goss=1),
    "depth4": dict(goss=0, max_depth=4),
    "gain_floor": dict(goss=0, min_split_gain=5.0),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_multi_leaf_steps_equal_one_split_per_step(case):
    kw = dict(CASES[case])
    weighted = kw.pop("weighted", False)
    skewed = kw.pop("skewed", False)
    p = synth.panel_large(3, n_dates=750)   # 300,000 rows x 128 features
    if skewed:
        # 8 features with 85% of rows on one value: every child build's rounds put more rows
        # into one bin than a packed LDS word holds unless the rows alternate between copies
        rng = np.random.default_rng(11)
        for f in range(0, 128, 16):
            p.X[rng.random(p.n) < 0.85, f] = 0.0
    w = np.random.default_rng(5).uniform(0.5, 2.0, p.n) if weighted else None
    _say(f"[case] multi-leaf {case}: {p.n} x {p.n_features}, {kw}")
    base = _fit(p, 1, w, **kw)
    tb = base.trees()
    pb = base.predict(p.X[::7])
    for width in (2, 4, 8):
        b = _fit(p, width, w, **kw)
        tw = b.trees()
        assert len(tw) == len(tb)
        for i, (x, y) in enumerate(zip(tb, tw)):
            for k in KEYS:
                assert np.array_equal(x[k], y[k]), f"{case} width {width} tree {i}: {k} differs"
        assert np.array_equal(b.predict(p.X[::7]), pb), f"{case} width {width}: predictions differ"
        _say(f"  width {width}: {len(tw)} trees identical to one split per step")


def test_multi_leaf_width_is_validated():
    p = synth.panel_small(0)
    with pytest.raises(Exception, match="spec_width"):
        t.Booster(spec_width=9).fit(p.X, p.y, timestamps=p.ts, entity=p.entity, horizon=1)
