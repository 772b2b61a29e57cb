"""Device bins and node histograms against the oracle, through the device C ABI.

Bins (reference Binner::transform, binning.cpp:104-128): the device's uint8 bins
(tgbm_session_get_bins) must equal the oracle's byte for byte, at the 1,000,000 x 128 panel size
and on an edge-case matrix (+-0, +-inf, NaN, values exactly on an edge, a feature with 254
distinct values and so 254 edges, a constant and an all-NaN feature, categorical exact matches).

Histograms (reference build_hist_col, histogram.cpp:44-76, and the parent - sibling subtraction,
:108-117): for every node checked, read back with tgbm_task_node_hist,
  * counts equal the oracle's per (feature, bin) exactly;
  * the fixed-point g sums equal sum(rint(g * 2^s_g)) over the node's rows exactly (integer
    arithmetic, checked in numpy), and likewise the h sums when h varies;
  * the histogram sums are within 1e-6 relative of the oracle's sequential FP64 sums,
    |G_dev - G_fp64| <= 1e-6 * |G_fp64| for every bin holding rows (denominator: the FP64 sum of
    that bin itself), and the same for H.  Bins whose FP64 sum is exactly 0 must come out as 0.
The worst relative error per node is printed.
"""
import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest

import checkers
from tidegbm_amd import synth

pytestmark = pytest.mark.gpu

HIP = Path(__file__).resolve().parent.parent / "tidegbm_amd" / "lib" / "libtgbm_hip.so"
_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
HIST_RTOL = 1e-6


class GrowParams(ctypes.Structure):
    _fields_ = [("max_depth", _I), ("max_leaves", _I), ("min_samples_leaf", _I), ("min_child_weight", _D),
                ("l2", _D), ("l1", _D), ("max_delta_step", _D), ("min_split_gain", _D), ("pair_l2", _D),
                ("cat_smooth", _D), ("hist_mode", _I), ("host_driven", _I)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _lib():
    L = ctypes.CDLL(str(HIP))
    L.tgbm_session_create.argtypes = [_P, _I, _I, _I, _I, _P, _I, _I]
    L.tgbm_session_destroy.argtypes = [_P]
    L.tgbm_session_destroy.restype = None
    L.tgbm_session_bin.argtypes = [_P, _P, _P, _P]
    L.tgbm_session_get_bins.argtypes = [_P, _P]
    L.tgbm_task_set_rows.argtypes = [_P, _I, _P, _P, _P, _P, _P, _I, _P, _P]
    L.tgbm_task_grad.argtypes = [_P, _I, _I]
    L.tgbm_task_get_grad.argtypes = [_P, _I, _P, _P]
    L.tgbm_tree_begin.argtypes = [_P, _P]
    L.tgbm_tree_grow_device.argtypes = [_P, _P]
    L.tgbm_tree_end.argtypes = [_P, _D, _D, _I, _P, _P, _P, _P, _P, _P, _P]
    L.tgbm_task_node_hist.argtypes = [_P, _I, _I, _P, _P, _P, _P, _P]
    L.tgbm_last_error.restype = ctypes.c_char_p
    return L


def _ok(L, rc):
    assert rc == 0, L.tgbm_last_error().decode()


def _edges(X, max_bin=255, bin_sample=262144):
    O = checkers.oracle().lib
    n, F = X.shape
    E = np.zeros((F, 256))
    ne = np.zeros(F, np.int32)
    O.orc_fit_edges(X.ctypes.data, n, F, max_bin, bin_sample, 0, E.ctypes.data, ne.ctypes.data)
    return E, ne


def _oracle_bins(X, E, ne):
    n, F = X.shape
    out = np.empty((n, F), np.uint8)
    checkers.oracle().lib.orc_bin_transform(X.ctypes.data, n, F, E.ctypes.data, ne.ctypes.data, out.ctypes.data)
    return out


class Session:
    def __init__(self, L, X, E, ne, hconst, max_leaves=31):
        self.L = L
        self.n, self.F = X.shape
        self.s = _P()
        rows = np.array([self.n], np.int32)
        self.max_nodes = 2 * max_leaves - 1
        _ok(L, L.tgbm_session_create(ctypes.byref(self.s), self.n, self.F, 255, 1, rows.ctypes.data, max_leaves,
                                     1 if hconst else 0))
        _ok(L, L.tgbm_session_bin(self.s, X.ctypes.data, E.ctypes.data, ne.ctypes.data))

    def bins(self):
        out = np.empty((self.n, self.F), np.uint8)
        _ok(self.L, self.L.tgbm_session_get_bins(self.s, out.ctypes.data))
        return out

    def node_hist(self, node):
        F = self.F
        qg = np.empty((F, 256), np.int64)
        qh = np.empty((F, 256), np.int64)
        c = np.empty((F, 256), np.int32)
        sg, sh = _I(), _I()
        _ok(self.L, self.L.tgbm_task_node_hist(self.s, 0, node, qg.ctypes.data, c.ctypes.data, qh.ctypes.data,
                                               ctypes.byref(sg), ctypes.byref(sh)))
        return qg, c, qh, sg.value, sh.value

    def close(self):
        self.L.tgbm_session_destroy(self.s)


def _exact_sums(xb_rows, q):
    """Per (feature, bin) integer sums of q over rows (xb_rows: m x F uint8), exact in int64."""
    m, F = xb_rows.shape
    out = np.zeros((F, 256), np.int64)
    for f in range(F):
        b = xb_rows[:, f]
        order = np.argsort(b, kind="stable")
        bs, qs = b[order], q[order]
        starts = np.flatnonzero(np.r_[True, bs[1:] != bs[:-1]])
        if m:
            out[f, bs[starts]] = np.add.reduceat(qs, starts)
    return out


def _check_node(name, node, xb, rows, g, h, hconst, dev):
    qg, c, qh, s_g, s_h = dev
    F = xb.shape[1]
    idx = np.ascontiguousarray(rows, np.int32)
    G = np.zeros((F, 256))
    H = np.zeros((F, 256))
    C = np.zeros((F, 256), np.int64)
    checkers.oracle().lib.orc_hist(xb.ctypes.data, F, idx.ctypes.data, idx.size, g.ctypes.data, h.ctypes.data, 256,
                                   G.ctypes.data, H.ctypes.data, C.ctypes.data)
    assert np.array_equal(c.astype(np.int64), C), f"{name} node {node}: counts differ"
    sub = xb[idx]
    q_exact = _exact_sums(sub, np.rint(np.ldexp(g[idx], s_g)).astype(np.int64))
    assert np.array_equal(qg, q_exact), f"{name} node {node}: fixed-point g sums differ from the exact integer sums"
    Gd = np.ldexp(qg.astype(np.float64), -s_g)
    if hconst:
        assert np.array_equal(qh, c.astype(np.int64)) and s_h == 0
        Hd = qh.astype(np.float64)
    else:
        qh_exact = _exact_sums(sub, np.rint(np.ldexp(h[idx], s_h)).astype(np.int64))
        assert np.array_equal(qh, qh_exact), f"{name} node {node}: fixed-point h sums differ"
        Hd = np.ldexp(qh.astype(np.float64), -s_h)
    held = C > 0
    out = {}
    for label, dv, rv in (("G", Gd, G), ("H", Hd, H)):
        zero = held & (rv == 0)
        assert np.all(dv[zero] == 0), f"{name} node {node}: {label} of a bin summing to exactly 0"
        nz = held & (rv != 0)
        rel = np.abs(dv[nz] - rv[nz]) / np.abs(rv[nz])
        worst = float(rel.max()) if rel.size else 0.0
        out[label] = worst
        assert worst <= HIST_RTOL, f"{name} node {node}: {label} relative error {worst:.3e} > {HIST_RTOL}"
    return out


def _grow_and_check(name, X, y, w, n_trees, check_trees, nodes_per_tree):
    L = _lib()
    n, F = X.shape
    E, ne = _edges(X)
    xb = _oracle_bins(X, E, ne)
    hconst = w is None
    S = Session(L, X, E, ne, hconst)
    try:
        assert np.array_equal(S.bins(), xb), f"{name}: device bins differ from the oracle's"
        pred0 = np.full(n, float(np.mean(y)))
        zeros = np.zeros(n, np.int32)
        seg = np.zeros(n, np.uint8)
        wp = None if w is None else w.ctypes.data
        _ok(L, L.tgbm_task_set_rows(S.s, 0, y.ctypes.data, wp, pred0.ctypes.data, zeros.ctypes.data,
                                    seg.ctypes.data, 1, None, None))
        P = GrowParams(8, 31, 20, 1e-3, 1.0, 0.0, 0.0, 0.0, 2.0, 1.0, 0, 0)
        mx = S.max_nodes
        nn = np.zeros(1, np.int32)
        ints = np.zeros((mx, 10), np.int32)
        dbl = np.zeros((mx, 2))
        vals = np.zeros(mx)
        cov = np.zeros(mx, np.int32)
        pref = np.zeros(mx * 256)
        for t in range(n_trees):
            _ok(L, L.tgbm_task_grad(S.s, -1, 0))
            g, h = np.empty(n), np.empty(n)
            _ok(L, L.tgbm_task_get_grad(S.s, 0, g.ctypes.data, h.ctypes.data))
            _ok(L, L.tgbm_tree_begin(S.s, ctypes.byref(P)))
            _ok(L, L.tgbm_tree_grow_device(S.s, ctypes.byref(P)))
            _ok(L, L.tgbm_tree_end(S.s, 0.1, 0.0, 1, ctypes.byref(P), nn.ctypes.data, ints.ctypes.data,
                                   dbl.ctypes.data, vals.ctypes.data, cov.ctypes.data, pref.ctypes.data))
            if t not in check_trees:
                continue
            k = int(nn[0])
            tint = ints[:k].copy()
            # nodes: the root, then split children in id order (built smaller children and
            # larger children derived as parent - sibling), then the deepest leaves
            order = sorted(range(k), key=lambda i: (tint[i, 9], i))
            pick = order[:nodes_per_tree // 2] + order[-(nodes_per_tree - nodes_per_tree // 2):]
            worst = {"G": 0.0, "H": 0.0}
            for node in sorted(set(pick)):
                rows = np.flatnonzero(checkers.rows_reaching(tint, xb, node))
                r = _check_node(name, node, xb, rows, g, h, hconst, S.node_hist(node))
                for key in worst:
                    worst[key] = max(worst[key], r[key])
            _say(f"[hist] {name} tree {t}: {k} nodes, {len(set(pick))} checked, worst rel err "
                 f"G {worst['G']:.2e} H {worst['H']:.2e}")
    finally:
        S.close()


def test_device_bins_edge_cases():
    _say("[case] bins edge cases")
    rng = np.random.default_rng(3)
    n = 6000
    X = rng.normal(size=(n, 6))
    X[:, 1] = rng.integers(0, 254, n) * 0.37 - 20.0           # 254 distinct values -> 254 edges
    X[:, 2] = rng.integers(0, 300, n) * 1.5                   # > 254 values -> quantile edges
    X[:, 3] = 2.5                                             # constant
    X[:, 4] = np.nan                                          # all missing
    X[:, 5] = rng.integers(0, 12, n).astype(np.float64)       # categorical codes
    E, ne = _edges(X)
    assert ne[1] == 254
    # exact-edge values, signed zeros, infinities and NaN in the numeric features
    for f in (0, 1, 2):
        k = rng.integers(0, ne[f], 200)
        X[rng.integers(0, n, 200), f] = E[f, k]
        X[rng.integers(0, n, 20), f] = 0.0
        X[rng.integers(0, n, 20), f] = -0.0
        X[rng.integers(0, n, 5), f] = np.inf
        X[rng.integers(0, n, 5), f] = -np.inf
        X[rng.integers(0, n, 30), f] = np.nan
    want = _oracle_bins(X, E, ne)
    # categorical feature 5: category values as edges with the categorical flag; a value bins to
    # 1 + its index among them, an unlisted value or NaN to 0 (binning.cpp:107-112)
    cats = np.array([0.0, 1.0, 2.0, 3.0, 5.0, 8.0, 11.0])
    E[5, :] = 0
    E[5, :cats.size] = cats
    ne[5] = cats.size | 0x10000
    X[rng.integers(0, n, 30), 5] = np.nan
    pos = {v: i + 1 for i, v in enumerate(cats)}
    want[:, 5] = [pos.get(v, 0) for v in X[:, 5]]
    L = _lib()
    S = Session(L, np.ascontiguousarray(X), E, ne, True)
    try:
        got = S.bins()
    finally:
        S.close()
    assert np.array_equal(got, want)
    assert got[:, 1].max() == 255 and (got[:, 4] == 0).all()


def test_histograms_full_size_panel():
    # the 1,000,000 x 128 panel (1% NaN, Student-t features), unit weights: h == 1, H = counts
    _say("[case] histograms 1M x 128")
    p = synth.panel_large(1)
    X = np.ascontiguousarray(p.X)
    _grow_and_check("1Mx128", X, np.ascontiguousarray(p.y), None, n_trees=3, check_trees=(0, 2), nodes_per_tree=10)


def test_histograms_weighted_rows():
    # weights make h vary: the h sums go through their own fixed-point plane
    _say("[case] histograms weighted 200k x 32")
    p = synth.panel_large(2, n_dates=500, n_entities=400, n_features=32)
    w = np.random.default_rng(4).uniform(0.2, 3.0, p.n)
    _grow_and_check("weighted", np.ascontiguousarray(p.X), np.ascontiguousarray(p.y), w, n_trees=2,
                    check_trees=(0, 1), nodes_per_tree=14)
