"""The C-ABI libraries load and export every entry point their headers declare; host-side error
paths behave like the reference's (no GPU needed)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import tidegbm_amd as t

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "tidegbm_amd" / "lib"


def _declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:tide|tgbm)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("tidegbm_amd.h", "libtidegbm_amd.so"), ("tgbm_device.h", "libtgbm_hip.so")])
def test_every_declared_symbol_is_exported(header, lib):
    so = ctypes.CDLL(str(LIB / lib))
    names = _declared(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_reference_signatures_present():
    # the twelve reference entry points (reference cpp/include/tidegbm/c_api.h:9-28)
    ref = ["tide_create", "tide_free", "tide_set_int", "tide_set_double", "tide_fit", "tide_predict",
           "tide_predict_raw", "tide_predict_binned", "tide_save_json", "tide_save", "tide_load",
           "tide_last_error"]
    assert set(ref) <= set(_declared("tidegbm_amd.h"))


def test_setter_errors_match_reference_messages():
    h = ctypes.c_void_p()
    L = t._lib
    assert L.tide_create(ctypes.byref(h)) == 0
    assert L.tide_set_int(h, b"no_such_key", 1) == -1
    assert L.tide_last_error() == b"unknown int key"
    assert L.tide_set_double(h, b"no_such_key", 1.0) == -1
    assert L.tide_last_error() == b"unknown double key"
    assert L.tide_set_int(h, b"n_estimators", 3) == 0
    assert L.tide_set_string(h, b"objective", b"huber") == -1
    assert b"not supported" in L.tide_last_error()
    L.tide_free(h)


def test_gpu_compiled_flag_and_missing_device_message():
    # mirrors the reference's test_cuda_compiled_matches_device_gpu (tests/test_markets.py:22-34)
    assert t.gpu_compiled is True
    rng = np.random.default_rng(0)
    X = rng.normal(size=(40, 3))
    y = X[:, 0]
    m = t.TideRegressor(n_estimators=5, iid_mode=True, device="gpu")
    if t.device_count() == 0:
        with pytest.raises(ValueError, match="GPU"):
            m.fit(X, y)
    else:
        pytest.skip("a GPU is present: the fit path is covered by the gpu tests")


def test_invalid_configs_fail_loudly():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(50, 3))
    y = X[:, 0]
    ts = np.repeat(np.arange(10), 5)
    for kw, msg in [(dict(device="cpu"), "device=cpu"), (dict(max_bin=256), "max_bin"),
                    (dict(subsample=0.5), "subsample"), (dict(halo=1), "halo"),
                    (dict(n_estimators=0), "n_estimators")]:
        b = t.Booster(**kw)
        with pytest.raises(ValueError, match=msg):
            b.fit(X, y, timestamps=ts, horizon=1)
    with pytest.raises(ValueError, match="horizon"):
        t.Booster().fit(X, y, timestamps=ts, horizon=-1)


def test_estimator_requires_timestamp_and_horizon():
    X = np.zeros((10, 2))
    y = np.zeros(10)
    with pytest.raises(ValueError, match="timestamp"):
        t.TideRegressor().fit(X, y)
    with pytest.raises(ValueError, match="horizon"):
        t.TideRegressor().fit(X, y, timestamp=np.arange(10))


def test_parse_horizon_units():
    ts = np.array([0, 5, 10, 20, 25], dtype=np.int64)  # median positive step 5
    assert t.parse_horizon(3, ts) == 15
    assert t.parse_horizon(None, ts) == 0
    assert t.parse_horizon(2.0, ts) == 10


def test_extension_setters_validate_arguments():
    h = ctypes.c_void_p()
    L = t._lib
    assert L.tide_create(ctypes.byref(h)) == 0
    try:
        # eval set: X and y are required for n_eval > 0; n_eval = 0 clears
        assert L.tide_set_eval(h, None, None, None, 5, 3) == -1
        assert b"eval set" in L.tide_last_error()
        X = np.zeros((4, 3))
        y = np.zeros(4)
        assert L.tide_set_eval(h, X.ctypes.data, y.ctypes.data, None, 4, 3) == 0
        assert L.tide_set_eval(h, None, None, None, 0, 0) == 0
        # groups: a pointer is required for n > 0
        assert L.tide_set_feature_groups(h, None, 3) == -1
        assert L.tide_set_interaction_groups(h, None, 3) == -1
        g = np.array([0, 0, -1], np.int32)
        assert L.tide_set_feature_groups(h, g.ctypes.data, 3) == 0
        assert L.tide_set_interaction_groups(h, g.ctypes.data, 3) == 0
        assert L.tide_set_int(h, b"regime_feature", 1) == 0
        assert L.tide_set_string(h, b"eval_metric", b"ic") == 0
        assert L.tide_set_double(h, b"cat_smooth", 2.0) == 0
        # an unfitted model has no trees and no categorical values
        assert L.tide_num_trees(h) == 0
        buf = np.zeros(4)
        assert L.tide_get_cat_values(h, 0, buf.ctypes.data, 4) == 0
        out = np.zeros(8, np.uint64)
        assert L.tide_tree_cat(h, 0, out.ctypes.data) == -1
    finally:
        L.tide_free(h)


def test_python_booster_rejects_bad_inputs():
    b = t.Booster()
    with pytest.raises(t.TideError, match="one entry per feature"):
        b.fit(np.zeros((10, 3)), np.zeros(10), iid=True, is_cat=[1, 0])
    with pytest.raises(t.TideError, match="same length"):
        b.fit(np.zeros((10, 3)), np.zeros(10), iid=True, eval_set=(np.zeros((4, 3)), np.zeros(5)))
    assert {"early_stopping_rounds", "eval_metric", "cat_smooth", "feature_groups"} <= t.KNOWN_KEYS
