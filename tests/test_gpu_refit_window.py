"""tgbm_task_refit_window against the calls it stands for, through the device C ABI.

A refit slot's admission (the serving loop's path for reference Booster::update on a row window,
booster.cpp:844-884) is one call, tgbm_task_refit_window(task, row0, n, y, forest).  It must leave
the task exactly as tgbm_task_bind + tgbm_task_predict_forest + tgbm_task_set_rows(y, one time
bucket) leave it.  Two tasks of one session take the same window, one each way, and then grow
trees together: the gradients after the first grad launch and every exported tree (node ints,
values, TORB prefix values) must be identical bit for bit.  The windows start off a 16-row
boundary, and a second admission of both slots uses a smaller window (the chunk list is
reused).  The forest is a random one built here (numeric splits, depth <= 5).
"""
import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest

import checkers

pytestmark = pytest.mark.gpu

HIP = Path(__file__).resolve().parent.parent / "tidegbm_amd" / "lib" / "libtgbm_hip.so"
_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


class GrowParams(ctypes.Structure):
    _fields_ = [("max_depth", _I), ("max_leaves", _I), ("min_samples_leaf", _I), ("min_child_weight", _D),
                ("l2", _D), ("l1", _D), ("max_delta_step", _D), ("min_split_gain", _D), ("pair_l2", _D),
                ("cat_smooth", _D), ("hist_mode", _I), ("host_driven", _I)]


def _lib():
    L = ctypes.CDLL(str(HIP))
    L.tgbm_session_create.argtypes = [_P, _I, _I, _I, _I, _P, _I, _I]
    L.tgbm_session_destroy.argtypes = [_P]
    L.tgbm_session_destroy.restype = None
    L.tgbm_session_bin.argtypes = [_P, _P, _P, _P]
    L.tgbm_session_set_spec.argtypes = [_P, _I]
    L.tgbm_task_bind.argtypes = [_P, _I, _I, _I]
    L.tgbm_task_predict_forest.argtypes = [_P, _I, _P]
    L.tgbm_task_refit_window.argtypes = [_P, _I, _I, _I, _P, _P]
    L.tgbm_task_set_rows.argtypes = [_P, _I, _P, _P, _P, _P, _P, _I, _P, _P]
    L.tgbm_task_grad.argtypes = [_P, _I, _I]
    L.tgbm_task_get_grad.argtypes = [_P, _I, _P, _P]
    L.tgbm_tree_begin.argtypes = [_P, _P]
    L.tgbm_tree_grow_device.argtypes = [_P, _P]
    L.tgbm_tree_end.argtypes = [_P, _D, _D, _I, _P, _P, _P, _P, _P, _P, _P]
    L.tgbm_forest_create.argtypes = [_P, _I, _I, _I, _P, _P, _P, _P, _D, _P, _P, _P]
    L.tgbm_forest_destroy.argtypes = [_P]
    L.tgbm_forest_destroy.restype = None
    L.tgbm_last_error.restype = ctypes.c_char_p
    return L


def _ok(L, rc):
    assert rc == 0, L.tgbm_last_error().decode()


def _random_forest(rng, n_trees, F):
    """Numeric trees in the export layout: children after their parent, depth <= 5."""
    sizes, ints, vals = [], [], []
    for _ in range(n_trees):
        nodes = [[1, 0, 0, 0, 0, -1, -1, 0, 0, 0]]
        k = 0
        while k < len(nodes):
            if nodes[k][9] < 5 and len(nodes) + 2 <= 63 and rng.random() < 0.75:
                d = nodes[k][9] + 1
                nodes[k][0:8] = [0, 0, int(rng.integers(F)), 0, int(rng.integers(1, 254)), len(nodes),
                                 len(nodes) + 1, int(rng.integers(2))]
                nodes.append([1, 0, 0, 0, 0, -1, -1, 0, 0, d])
                nodes.append([1, 0, 0, 0, 0, -1, -1, 0, 0, d])
            k += 1
        sizes.append(len(nodes))
        ints.extend(nodes)
        vals.extend(rng.normal(0.0, 0.05, len(nodes)).tolist())
    return (np.array(sizes, np.int32), np.ascontiguousarray(np.array(ints, np.int32)),
            np.array(vals), np.full(n_trees, 0.1))


def _edges(X, max_bin=255, bin_sample=262144):
    n, F = X.shape
    E = np.zeros((F, 256))
    ne = np.zeros(F, np.int32)
    checkers.oracle().lib.orc_fit_edges(X.ctypes.data, n, F, max_bin, bin_sample, 0, E.ctypes.data,
                                        ne.ctypes.data)
    return E, ne


def test_refit_window_equals_bind_predict_set_rows():
    L = _lib()
    rng = np.random.default_rng(11)
    n, F, cap = 60000, 24, 40000
    X = rng.normal(size=(n, F))
    y = X[:, 0] * 0.3 - X[:, 3] * 0.2 + rng.normal(0.0, 0.5, n)
    E, ne = _edges(X)
    sizes, ints, vals, lrs = _random_forest(rng, 40, F)
    s = _P()
    caps = np.array([cap, cap], np.int32)
    _ok(L, L.tgbm_session_create(ctypes.byref(s), n, F, 255, 2, caps.ctypes.data, 15, 1))
    f = _P()
    try:
        _ok(L, L.tgbm_session_bin(s, X.ctypes.data, E.ctypes.data, ne.ctypes.data))
        _ok(L, L.tgbm_session_set_spec(s, 8))
        _ok(L, L.tgbm_forest_create(ctypes.byref(f), F, 255, len(sizes), sizes.ctypes.data, ints.ctypes.data,
                                    vals.ctypes.data, lrs.ctypes.data, 0.25, E.ctypes.data, ne.ctypes.data, None))
        P = GrowParams(6, 15, 20, 1e-3, 1.0, 0.0, 0.0, 0.0, 2.0, 1.0, 0, 0)
        mx = 29
        for row0, m in ((1237, 37001), (20003, 9999)):   # windows off the 16-row boundary; then smaller
            print(f"window row0={row0} n={m}", flush=True)
            yw = np.ascontiguousarray(y[row0:row0 + m])
            _ok(L, L.tgbm_task_bind(s, 0, row0, m))
            _ok(L, L.tgbm_task_predict_forest(s, 0, f))
            _ok(L, L.tgbm_task_set_rows(s, 0, yw.ctypes.data, None, None, None, None, 1, None, None))
            ycopy = yw.copy()
            _ok(L, L.tgbm_task_refit_window(s, 1, row0, m, ycopy.ctypes.data, f))
            ycopy[:] = np.nan   # the call staged the targets: the caller may reuse its buffer at once
            for t in range(3):
                _ok(L, L.tgbm_task_grad(s, -1, 0))
                if t == 0:
                    g0, g1 = np.empty(m), np.empty(m)
                    _ok(L, L.tgbm_task_get_grad(s, 0, g0.ctypes.data, None))
                    _ok(L, L.tgbm_task_get_grad(s, 1, g1.ctypes.data, None))
                    assert np.array_equal(g0.view(np.int64), g1.view(np.int64)), "gradients differ"
                _ok(L, L.tgbm_tree_begin(s, ctypes.byref(P)))
                _ok(L, L.tgbm_tree_grow_device(s, ctypes.byref(P)))
                nn = np.zeros(2, np.int32)
                ti = np.zeros((2, mx, 10), np.int32)
                td = np.zeros((2, mx, 2))
                tv = np.zeros((2, mx))
                tc = np.zeros((2, mx), np.int32)
                tp = np.zeros((2, mx * 256))
                _ok(L, L.tgbm_tree_end(s, 0.1, 0.0, 1, ctypes.byref(P), nn.ctypes.data, ti.ctypes.data,
                                       td.ctypes.data, tv.ctypes.data, tc.ctypes.data, tp.ctypes.data))
                assert nn[0] == nn[1] and nn[0] > 1, f"tree {t}: node counts {nn}"
                k = int(nn[0])
                assert np.array_equal(ti[0, :k], ti[1, :k]), f"tree {t}: node ints differ"
                assert np.array_equal(tv[0, :k].view(np.int64), tv[1, :k].view(np.int64)), f"tree {t}: values differ"
                assert np.array_equal(tp[0].view(np.int64), tp[1].view(np.int64)), f"tree {t}: prefix values differ"
                print(f"  tree {t}: {k} nodes identical", flush=True)
                sys.stdout.flush()
    finally:
        if f:
            L.tgbm_forest_destroy(f)
        L.tgbm_session_destroy(s)
