"""Monotone constraints on the GPU against the oracle and the reference core.

The device split scan skips a numeric candidate whose leaf weights break the feature's
constraint (reference best_numeric_missing, histogram.cpp:177-180), in both the device-driven and
the host-driven (PAIR) growers.  Bar: tree structure bit-exact (the characterized exceptions of
checkers.compare_trees aside), leaf and prefix values within rtol 1e-7, predictions within 1e-5."""
import sys

import numpy as np
import pytest

import checkers
from monotone_cases import MONO_CASES, mono_panel
from tidegbm_amd import Booster

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", sorted(MONO_CASES))
def test_monotone_fit_matches_oracle_and_reference(case):
    c = MONO_CASES[case]
    p, mono = mono_panel()
    print(f"[case] monotone {case} n={p.n}", flush=True)
    sys.stdout.flush()
    rg = p.ts if c["objective"] == 5 else None
    kw = dict(ts=None if c["iid"] else p.ts, entity=None if c["iid"] else p.entity, rank_groups=rg, iid=c["iid"])
    params = dict(objective=c["objective"], learning_rate=0.1, goss=0, **c["params"])
    gpu = Booster(monotone=mono, **params)
    gpu.fit(p.X, p.y, timestamps=kw["ts"], entity=kw["entity"], rank_groups=rg, iid=c["iid"], horizon=1)
    xb = checkers.bins_of(p.X, [gpu.edges(f) for f in range(p.n_features)])
    nb = gpu.model_info()["n_bins"]
    libs = [checkers.oracle()] + ([checkers.reference()] if checkers.reference() is not None else [])
    for lib in libs:
        m = lib.model(**params).set_monotone(mono)
        m.fit(p.X, p.y, horizon=1, **kw)
        diff = checkers.compare_trees(gpu.trees(), m.trees(), rtol=1e-7, atol=1e-9, xb=xb, n_bins=nb)
        assert diff is None, f"{case} vs {lib.prefix}: {diff}"
        a, b = gpu.predict(p.X), m.predict(p.X)
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < 1e-5
    free = Booster(**params)
    free.fit(p.X, p.y, timestamps=kw["ts"], entity=kw["entity"], rank_groups=rg, iid=c["iid"], horizon=1)
    assert any(not np.array_equal(x["ints"], y["ints"]) for x, y in zip(gpu.trees(), free.trees())), \
        "the constraints should change the trees"
