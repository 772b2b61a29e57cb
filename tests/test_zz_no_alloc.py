"""tgbm_forest_predict_noalloc (the predict entry point that allocates nothing) against
tgbm_forest_predict, through the device C ABI: bit-identical scores on two calls in a row, the
second with fewer rows (the fixed scratch keeps the first call's contents), and a clear error
for more rows than the scratch holds."""
import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest

import tidegbm_amd as t
from tidegbm_amd import synth

pytestmark = pytest.mark.gpu

HIP = Path(__file__).resolve().parent.parent / "tidegbm_amd" / "lib" / "libtgbm_hip.so"
_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
SCRATCH_BYTES = 80 << 20


def _lib():
    L = ctypes.CDLL(str(HIP))
    L.tgbm_forest_create.argtypes = [_P, _I, _I, _I, _P, _P, _P, _P, _D, _P, _P, _P]
    L.tgbm_forest_destroy.argtypes = [_P]
    L.tgbm_forest_destroy.restype = None
    L.tgbm_forest_predict.argtypes = [_P, _P, _I, _P]
    L.tgbm_forest_predict_noalloc.argtypes = [_P, _P, _I, _P]
    L.tgbm_last_error.restype = ctypes.c_char_p
    return L


def _forest(L, b, F):
    """A device forest of a fitted Booster's trees (node layout of include/tgbm_device.h)."""
    trees = b.trees()
    info = b.model_info()
    counts = np.array([tr["ints"].shape[0] for tr in trees], np.int32)
    ints = np.ascontiguousarray(np.concatenate([tr["ints"] for tr in trees]).astype(np.int32))
    vals = np.ascontiguousarray(np.concatenate([tr["values"] for tr in trees]))
    lr = np.full(len(trees), 0.1)
    E = np.zeros((F, 256))
    ne = np.zeros(F, np.int32)
    for f in range(F):
        e = b.edges(f)
        E[f, :e.size] = e
        ne[f] = e.size
    fo = _P()
    rc = L.tgbm_forest_create(ctypes.byref(fo), F, info["n_bins"], len(trees), counts.ctypes.data, ints.ctypes.data,
                              vals.ctypes.data, lr.ctypes.data, info["init_score"], E.ctypes.data, ne.ctypes.data, None)
    assert rc == 0, L.tgbm_last_error()
    return fo


def _predict(L, fn, fo, X):
    X = np.ascontiguousarray(X)
    out = np.full(X.shape[0], np.nan)
    rc = fn(fo, X.ctypes.data, X.shape[0], out.ctypes.data)
    assert rc == 0, L.tgbm_last_error()
    return out


def test_noalloc_predict_matches_predict():
    print("[case] no-allocation predict", flush=True)
    sys.stdout.flush()
    p = synth.panel_small(5, n_dates=300)
    b = t.Booster(n_estimators=12, max_depth=6, learning_rate=0.1, goss=0, pair_splits=0)
    b.fit(p.X, p.y, timestamps=p.ts, entity=p.entity, horizon=1)
    L = _lib()
    F = p.n_features
    fo = _forest(L, b, F)
    try:
        rng = np.random.default_rng(1)
        X1 = p.X[rng.integers(0, p.n, 20000)] + rng.normal(scale=0.2, size=(20000, F))
        X1[rng.random(X1.shape) < 0.03] = np.nan
        X2 = p.X[rng.integers(0, p.n, 333)] - 0.1
        for X in (X1, X2):   # the second call has fewer rows than the first
            a = _predict(L, L.tgbm_forest_predict_noalloc, fo, X)
            r = _predict(L, L.tgbm_forest_predict, fo, X)
            assert np.array_equal(a, r)
            assert np.array_equal(a, b.predict(X))
        # more rows than the scratch holds: rejected on the host before any device work
        fp = (F + 15) // 16 * 16
        max_rows = (SCRATCH_BYTES - 1040 - 64 * fp) // (8 * F + fp + 8)
        out = np.zeros(4)
        rc = L.tgbm_forest_predict_noalloc(fo, X2.ctypes.data, max_rows + 1, out.ctypes.data)
        assert rc == -1
        assert b"exceed" in L.tgbm_last_error() and str(max_rows).encode() in L.tgbm_last_error()
        # and the largest accepted size runs
        Xm = np.ascontiguousarray(np.resize(X1, (max_rows, F)))
        assert np.array_equal(_predict(L, L.tgbm_forest_predict_noalloc, fo, Xm),
                              _predict(L, L.tgbm_forest_predict, fo, Xm))
    finally:
        L.tgbm_forest_destroy(fo)
