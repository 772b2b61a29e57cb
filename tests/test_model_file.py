"""TGBM v3 files: this library reads what the reference core writes and writes it back byte for
byte; corrupt or unsupported files are rejected with a message (no GPU needed).

The reference writer and reader are cpp/src/serialize.cpp:65-255 (JSON dump :266-324).  Files
come from the reference core built in oracle/_ref; tests/tgbm_file.py is an independent Python
statement of the layout used to edit them."""
import ctypes

import numpy as np
import pytest

import checkers
import tgbm_file
import tidegbm_amd as t
from cat_cases import cat_panel
from monotone_cases import mono_panel

REF = checkers.reference()
pytestmark = pytest.mark.skipif(REF is None, reason="reference core (oracle/_ref) not built")
L = t._lib


def _ref_model(kind):
    """A fitted reference model: categorical features with feature groups, or monotone
    constraints in iid mode, or the pairwise objective."""
    if kind == "categorical_groups":
        p, is_cat = cat_panel(3)
        m = REF.model(n_estimators=6, max_depth=5, learning_rate=0.1, goss=0)
        m.set_cat(is_cat)
        grp = np.full(p.n_features, -1, np.int32)
        grp[[0, 2]] = 0
        m.set_groups(grp)
        m.fit(p.X, p.y, ts=p.ts, entity=p.entity, horizon=3)
    elif kind == "monotone_iid":
        p, mono = mono_panel()
        m = REF.model(n_estimators=5, max_depth=5, learning_rate=0.1, goss=0, pair_splits=0)
        m.set_monotone(mono)
        m.fit(p.X, p.y, iid=True)
    else:
        p, _ = mono_panel(4)
        m = REF.model(n_estimators=4, max_depth=4, learning_rate=0.1, goss=0, objective=5, pair_splits=0)
        m.fit(p.X, p.y, ts=p.ts, entity=p.entity, rank_groups=p.ts, horizon=1)
    return m


def _load(path):
    h = ctypes.c_void_p()
    rc = L.tide_load(ctypes.byref(h), str(path).encode())
    return rc, h


KINDS = ["categorical_groups", "monotone_iid", "pairwise"]


@pytest.mark.parametrize("kind", KINDS)
def test_reference_file_reads_and_writes_back_unchanged(tmp_path, kind):
    m = _ref_model(kind)
    ref_bin, ref_json = tmp_path / "ref.tgbm", tmp_path / "ref.json"
    m.save(ref_bin)
    m.save(ref_json, json=True)
    blob = ref_bin.read_bytes()
    # the Python layout reproduces the reference file exactly (it is used below to edit files)
    assert tgbm_file.write(tgbm_file.read(blob)) == blob
    rc, h = _load(ref_bin)
    assert rc == 0, L.tide_last_error()
    try:
        out_bin, out_json = tmp_path / "ours.tgbm", tmp_path / "ours.json"
        assert L.tide_save(h, str(out_bin).encode()) == 0
        assert L.tide_save_json(h, str(out_json).encode()) == 0
        assert out_bin.read_bytes() == blob
        assert out_json.read_text() == ref_json.read_text()
    finally:
        L.tide_free(h)
    # and the reference reads the file written here
    back = checkers.CpuModel.load(REF, out_bin)
    X = np.random.default_rng(0).normal(size=(50, tgbm_file.read(blob)["n_features"]))
    assert np.array_equal(back.predict(X), m.predict(X))


def _expect_reject(tmp_path, blob, msg):
    path = tmp_path / "bad.tgbm"
    path.write_bytes(blob)
    rc, h = _load(path)
    if rc == 0:
        L.tide_free(h)
    assert rc == -1
    assert msg in L.tide_last_error().decode(), L.tide_last_error()


def test_corrupt_files_are_rejected(tmp_path):
    m = _ref_model("categorical_groups")
    m.save(tmp_path / "ref.tgbm")
    blob = (tmp_path / "ref.tgbm").read_bytes()
    good = tgbm_file.read(blob)
    _expect_reject(tmp_path, b"XGBM" + blob[4:], "bad magic")
    _expect_reject(tmp_path, blob[:4] + (2).to_bytes(4, "little") + blob[8:], "version 2")
    _expect_reject(tmp_path, blob[:-7], "truncated")
    # a node whose child link points at itself or back up the tree: a walk would never end
    for child in (0, -5):
        d = tgbm_file.read(blob)
        root = d["trees"][1]["nodes"][0]
        assert not root["is_leaf"]
        root["left"] = child
        _expect_reject(tmp_path, tgbm_file.write(d), "child links")
    d = tgbm_file.read(blob)
    n1 = d["trees"][0]["nodes"][1]
    if not n1["is_leaf"]:
        n1["right"] = 0
        _expect_reject(tmp_path, tgbm_file.write(d), "child links")
    # categorical flags for fewer features than the model has
    d = tgbm_file.read(blob)
    d["is_cat"] = d["is_cat"][:-1]
    _expect_reject(tmp_path, tgbm_file.write(d), "categorical flags")
    d = tgbm_file.read(blob)
    d["group_of"] = d["group_of"][:3]
    _expect_reject(tmp_path, tgbm_file.write(d), "feature groups")
    d = tgbm_file.read(blob)
    d["features"] = d["features"][:-1]
    _expect_reject(tmp_path, tgbm_file.write(d), "binner")
    d = tgbm_file.read(blob)
    d["trees"][0]["n_targets"] = 2
    _expect_reject(tmp_path, tgbm_file.write(d), "multi-head")
    d = tgbm_file.read(blob)
    d["trees"][0]["nodes"][0]["feature"] = good["n_features"]
    _expect_reject(tmp_path, tgbm_file.write(d), "feature")
    # the unedited file still loads
    path = tmp_path / "good.tgbm"
    path.write_bytes(tgbm_file.write(good))
    rc, h = _load(path)
    assert rc == 0
    L.tide_free(h)


def test_feature_groups_survive_a_round_trip(tmp_path):
    m = _ref_model("categorical_groups")
    m.save(tmp_path / "ref.tgbm")
    d = tgbm_file.read((tmp_path / "ref.tgbm").read_bytes())
    assert d["group_of"][0] == 0 and d["group_of"][2] == 0
    b = t.Booster.load(tmp_path / "ref.tgbm")
    b.save(tmp_path / "again.tgbm")
    assert tgbm_file.read((tmp_path / "again.tgbm").read_bytes())["group_of"] == d["group_of"]
