"""Monotone-constraint cases shared by the oracle pin tests and the GPU parity tests."""
import numpy as np

from tidegbm_amd import synth

# Booster::fit's monotone argument (booster.cpp:250-251) per case; the target bends against the
# constraint on features 0 and 5, so the constrained scan must skip candidates and the trees change
MONO_CASES = {
    "time_pair_default": dict(iid=False, objective=0, params=dict(n_estimators=10, max_depth=6)),
    "iid_device_driven": dict(iid=True, objective=0, params=dict(n_estimators=10, max_depth=5, pair_splits=0)),
    "pairwise": dict(iid=False, objective=5, params=dict(n_estimators=6, max_depth=5, pair_splits=0)),
}


def mono_panel(seed=21):
    p = synth.panel_small(seed)
    X = p.X
    y = p.y + 0.8 * np.sin(2.0 * X[:, 0]) - 0.3 * X[:, 5] ** 2 + 0.2 * X[:, 2]
    mono = np.zeros(p.n_features, np.int32)
    mono[[0, 2]] = 1
    mono[5] = -1
    return synth.Panel(X, y, p.ts, p.entity), mono
