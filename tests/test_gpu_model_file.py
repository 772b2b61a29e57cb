"""TGBM v3 interop on the GPU: a model fitted here and saved is read by the reference core with
the same predictions, and a model the reference saved predicts the same here; the reference
entry points tide_fit and tide_predict_binned (reference c_api.h:18-24) run the device path.

Bar: predictions of the same trees agree to 1e-12 relative (both sum init + lr * leaf value in
tree order without FMA); trees fitted on the GPU match the oracle as in test_gpu_fit."""
import ctypes
import sys

import numpy as np
import pytest

import checkers
import tidegbm_amd as t
from cat_cases import cat_panel
from tidegbm_amd import Booster, synth

pytestmark = pytest.mark.gpu
REF = checkers.reference()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12)))


def _held_out(p, seed, nan_frac=0.05):
    """Rows the fit never saw: fresh draws of the same panel generator with NaNs sprinkled in."""
    rng = np.random.default_rng(seed)
    X = p.X[rng.integers(0, p.n, 3000)] + rng.normal(scale=0.3, size=(3000, p.n_features))
    X[rng.random(X.shape) < nan_frac] = np.nan
    return X


@pytest.mark.skipif(REF is None, reason="reference core (oracle/_ref) not built")
@pytest.mark.parametrize("kind", ["numeric_time", "categorical"])
def test_file_written_here_loads_in_reference(tmp_path, kind):
    _say(f"[case] save here, load in reference: {kind}")
    if kind == "categorical":
        p, is_cat = cat_panel(5)
        b = Booster(n_estimators=8, max_depth=5, learning_rate=0.1, goss=0)
        b.fit(p.X, p.y, timestamps=p.ts, entity=p.entity, horizon=2, is_cat=is_cat)
    else:
        p = synth.panel_small(8)
        b = Booster(n_estimators=12, max_depth=6, learning_rate=0.08, goss=0, pair_splits=0,
                    feature_groups=np.r_[[0, 0], -np.ones(p.n_features - 2, np.int32)])
        b.fit(p.X, p.y, timestamps=p.ts, entity=p.entity, horizon=1)
    path = tmp_path / "gpu.tgbm"
    b.save(path)
    r = checkers.CpuModel.load(REF, path)
    for X in (p.X[:4000], _held_out(p, 1)):
        assert _rel(b.predict(X), r.predict(X)) <= 1e-12
    # the reference writes the same bytes back
    r.save(tmp_path / "again.tgbm")
    assert (tmp_path / "again.tgbm").read_bytes() == path.read_bytes()


@pytest.mark.skipif(REF is None, reason="reference core (oracle/_ref) not built")
def test_reference_file_predicts_the_same_here(tmp_path):
    _say("[case] reference model loaded here")
    p, is_cat = cat_panel(6)
    r = REF.model(n_estimators=10, max_depth=6, learning_rate=0.1, goss=0).set_cat(is_cat)
    r.fit(p.X, p.y, ts=p.ts, entity=p.entity, horizon=1)
    r.save(tmp_path / "ref.tgbm")
    b = Booster.load(tmp_path / "ref.tgbm")
    for X in (p.X, _held_out(p, 2)):
        assert _rel(b.predict(X), r.predict(X)) <= 1e-12


def test_reference_signature_fit_and_binned_predict():
    _say("[case] tide_fit / tide_predict_binned")
    p = synth.panel_small(9)
    L = t._lib
    h = ctypes.c_void_p()
    assert L.tide_create(ctypes.byref(h)) == 0
    try:
        for k, v in (("n_estimators", 15), ("max_depth", 5)):
            assert L.tide_set_int(h, k.encode(), v) == 0
        assert L.tide_set_double(h, b"learning_rate", 0.1) == 0
        X = np.ascontiguousarray(p.X)
        y = np.ascontiguousarray(p.y)
        ts = np.ascontiguousarray(p.ts, np.int64)
        # reference tide_fit: no entity, no rank groups (c_api.cpp:79-80); time mode, horizon 1
        rc = L.tide_fit(h, X.ctypes.data, y.ctypes.data, p.n, p.n_features, ts.ctypes.data, None, None, 0, 1)
        assert rc == 0, L.tide_last_error()
        gpu = Booster(handle=h)
        h = None
        orc = checkers.oracle().model(n_estimators=15, max_depth=5, learning_rate=0.1)
        orc.fit(p.X, p.y, ts=p.ts, horizon=1)
        xb = checkers.bins_of(p.X, [gpu.edges(f) for f in range(p.n_features)])
        diff = checkers.compare_trees(gpu.trees(), orc.trees(), rtol=1e-7, atol=1e-9, xb=xb)
        assert diff is None, diff
        # binned predict (u16 bins, reference tide_predict_binned) equals the raw-feature predict
        Xh = _held_out(p, 3)
        xbh = checkers.bins_of(Xh, [gpu.edges(f) for f in range(p.n_features)]).astype(np.uint16)
        a = gpu.predict_binned(xbh)
        assert np.array_equal(a, gpu.predict(Xh))
        assert _rel(a, orc.predict(Xh)) < 1e-5
    finally:
        if h:
            L.tide_free(h)


def test_device_forest_rejects_links_that_do_not_point_forward():
    """The device forest builder refuses a split whose child is itself or an earlier node (the
    growers append children after their parent, reference tree.cpp:447-453), so no device walk
    can loop; a well-formed tree of the same shape is accepted."""
    from pathlib import Path
    hip = ctypes.CDLL(str(Path(t.__file__).resolve().parent / "lib" / "libtgbm_hip.so"))
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    hip.tgbm_forest_create.argtypes = [P, I, I, I, P, P, P, P, D, P, P, P]
    hip.tgbm_forest_destroy.argtypes = [P]
    hip.tgbm_forest_destroy.restype = None
    hip.tgbm_last_error.restype = ctypes.c_char_p
    F = 2
    E = np.zeros((F, 256))
    E[:, 0] = 0.0
    ne = np.ones(F, np.int32)
    vals = np.array([0.0, -1.0, 1.0])
    lr = np.array([0.1])

    def create(left, right):
        # {is_leaf, split_type, feature, feature2, thr, left, right, missing_go_left, n_samples, depth}
        ints = np.array([[0, 0, 0, -1, 1, left, right, 1, 10, 0],
                         [1, 0, -1, -1, -1, -1, -1, 1, 5, 1],
                         [1, 0, -1, -1, -1, -1, -1, 1, 5, 1]], np.int32)
        counts = np.array([3], np.int32)
        fo = P()
        rc = hip.tgbm_forest_create(ctypes.byref(fo), F, 3, 1, counts.ctypes.data, ints.ctypes.data, vals.ctypes.data,
                                    lr.ctypes.data, 0.0, E.ctypes.data, ne.ctypes.data, None)
        if rc == 0:
            hip.tgbm_forest_destroy(fo)
        return rc

    assert create(1, 2) == 0, hip.tgbm_last_error()
    for left, right in [(0, 2), (1, 0)]:
        assert create(left, right) != 0
        assert b"malformed tree node" in hip.tgbm_last_error()
