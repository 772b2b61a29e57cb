"""Monotone constraints in the oracle against the reference core (bit-exact).

The reference skips a numeric split candidate whose leaf weights break the feature's constraint
(best_numeric_missing, histogram.cpp:177-180).  Pinned here per function (random histograms,
including constraints that bind) and per fit (trees that the constraints change)."""
import numpy as np
import pytest

import checkers
from monotone_cases import MONO_CASES, mono_panel

REF = checkers.reference()
pytestmark = pytest.mark.skipif(REF is None, reason="reference core (oracle/_ref) not built")


def _scan(lib_fn, G, H, C, f, mono, lam=1.0, gamma=0.0, min_leaf=3, min_hess=1e-3, l1=0.0):
    F, nb = G.shape
    gain = np.zeros(1)
    b = np.zeros(1, np.int32)
    ml = np.zeros(1, np.int32)
    rc = lib_fn(G.ctypes.data, H.ctypes.data, C.ctypes.data, F, nb, f, lam, gamma, min_leaf, min_hess, mono, l1,
                gain.ctypes.data, b.ctypes.data, ml.ctypes.data)
    assert rc == 0
    return gain[0], int(b[0]), int(ml[0])


@pytest.mark.parametrize("l1", [0.0, 0.7])
def test_scan_with_constraints_matches_reference(l1):
    rng = np.random.default_rng(5)
    orc = checkers.oracle().lib.orc_scan_feature
    ref = REF.lib.tref_best_numeric
    bound = 0
    for trial in range(60):
        F, nb = 3, 40
        C = rng.integers(0, 12, size=(F, nb)).astype(np.int64)
        C[:, rng.random(nb) < 0.2] = 0
        # gradients that trend with the bin (so one direction is favoured) plus noise
        trend = np.linspace(-1, 1, nb) * rng.choice([-1.0, 1.0])
        G = (C * (trend + 0.6 * rng.normal(size=(F, nb)))).astype(np.float64)
        H = (C * rng.uniform(0.5, 1.5, size=(F, nb))).astype(np.float64)
        G, H = np.ascontiguousarray(G), np.ascontiguousarray(H)
        for f in range(F):
            free = _scan(orc, G, H, C, f, 0, l1=l1)
            for mono in (-1, 0, 1):
                a = _scan(orc, G, H, C, f, mono, l1=l1)
                b = _scan(ref, G, H, C, f, mono, l1=l1)
                assert a == b, (trial, f, mono, a, b)
                bound += a != free
    assert bound > 20, "the cases should include constraints that change the best split"


@pytest.mark.parametrize("case", sorted(MONO_CASES))
def test_monotone_fit_matches_reference(case):
    c = MONO_CASES[case]
    p, mono = mono_panel()
    rg = p.ts if c["objective"] == 5 else None
    fits = {}
    for name, lib, m in (("orc", checkers.oracle(), mono), ("ref", REF, mono), ("free", checkers.oracle(), None)):
        model = lib.model(objective=c["objective"], learning_rate=0.1, goss=0, **c["params"])
        if m is not None:
            model.set_monotone(m)
        model.fit(p.X, p.y, ts=None if c["iid"] else p.ts, entity=None if c["iid"] else p.entity,
                  rank_groups=rg, iid=c["iid"], horizon=1)
        fits[name] = model
    ta, tb, tf = fits["orc"].trees(), fits["ref"].trees(), fits["free"].trees()
    assert len(ta) == len(tb)
    for t, (x, y) in enumerate(zip(ta, tb)):
        for k in ("ints", "dbl", "values", "cover", "prefix"):
            assert np.array_equal(x[k], y[k]), f"tree {t} field {k}"
    assert np.array_equal(fits["orc"].predict(p.X), fits["ref"].predict(p.X))
    changed = any(not np.array_equal(x["ints"], y["ints"]) for x, y in zip(ta, tf))
    assert changed, "the constraints should change the trees of this case"
