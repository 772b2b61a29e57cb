"""ctypes front ends of the two CPU checkers (test infrastructure only).

* ``oracle/liboracle.so``      — the scalar C++ restatement of the reference path (prefix orc_)
* ``oracle/_ref/libtide_ref.so`` — the reference core built from its own sources (prefix tref_)

Both expose the same shape of API, so tests can run one fit through either and compare trees
field by field with the device path.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libtide_ref.so"

_P, _I, _D, _LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong


def _ptr(a):
    return None if a is None else a.ctypes.data


class CpuModel:
    """One fitted model of a CPU checker library."""

    def __init__(self, lib: "CpuLib", **params):
        self.lib = lib
        self.h = _P(lib.fn("create")())
        for k, v in params.items():
            if lib.fn("set")(self.h, k.encode(), float(v)) != 0:
                raise ValueError(lib.error())

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.fn("free")(self.h)
            self.h = None

    def fit(self, X, y, ts=None, w=None, entity=None, rank_groups=None, iid=False, horizon=1):
        X = np.ascontiguousarray(X, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        ts = None if ts is None else np.ascontiguousarray(ts, np.int64)
        w = None if w is None else np.ascontiguousarray(w, np.float64)
        entity = None if entity is None else np.ascontiguousarray(entity, np.int64)
        rank_groups = None if rank_groups is None else np.ascontiguousarray(rank_groups, np.int64)
        rc = self.lib.fn("fit")(self.h, _ptr(X), _ptr(y), X.shape[0], X.shape[1], _ptr(ts), _ptr(w),
                                _ptr(entity), _ptr(rank_groups), 1 if iid else 0, int(horizon))
        if rc != 0:
            raise ValueError(self.lib.error())
        return self

    def set_edges(self, edges):
        """Oracle only: fix the bin edges of later fits (list of per-feature edge arrays)."""
        F = len(edges)
        E = np.zeros((F, 256))
        ne = np.zeros(F, np.int32)
        for f, e in enumerate(edges):
            E[f, :len(e)] = e
            ne[f] = len(e)
        if self.lib.lib.orc_set_edges(self.h, _ptr(E), _ptr(ne), F) != 0:
            raise ValueError(self.lib.error())
        return self

    def update(self, X, y, extra, w=None):
        X = np.ascontiguousarray(X, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        w = None if w is None else np.ascontiguousarray(w, np.float64)
        if self.lib.fn("update")(self.h, _ptr(X), _ptr(y), X.shape[0], X.shape[1], _ptr(w), int(extra)) != 0:
            raise ValueError(self.lib.error())

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float64)
        out = np.empty(X.shape[0])
        if self.lib.fn("predict")(self.h, _ptr(X), X.shape[0], X.shape[1], _ptr(out)) != 0:
            raise ValueError(self.lib.error())
        return out

    def trees(self):
        out = []
        for t in range(self.lib.fn("n_trees")(self.h)):
            nn, nb = _I(), _I()
            self.lib.fn("tree_info")(self.h, t, ctypes.byref(nn), ctypes.byref(nb))
            ints = np.empty((nn.value, 10), np.int32)
            dbl = np.empty((nn.value, 2))
            vals = np.empty(nn.value)
            cov = np.empty(nn.value, np.int32)
            pref = np.empty(nn.value * nb.value)
            self.lib.fn("tree_get")(self.h, t, _ptr(ints), _ptr(dbl), _ptr(vals), _ptr(cov), _ptr(pref))
            cat = np.zeros((nn.value, 4), np.uint64)
            self.lib.fn("tree_cat")(self.h, t, _ptr(cat))
            out.append(dict(ints=ints, dbl=dbl, values=vals, cover=cov, prefix=pref, n_buckets=nb.value, cat=cat))
        return out

    def edges(self, f):
        buf = np.empty(256)
        m = self.lib.fn("edges")(self.h, f, _ptr(buf), 256)
        return buf[:m].copy()

    def n_bins(self):
        return self.lib.fn("n_bins")(self.h)

    def init_prefix(self):
        buf = np.empty(512)
        m = self.lib.fn("init_prefix")(self.h, _ptr(buf), 512)
        return buf[:m].copy()

    def init_score(self):
        return self.lib.fn("init_score")(self.h)

    def evals_train(self):
        """Training loss after each tree of the last fit (reference evals.train)."""
        n = self.lib.fn("evals_train")(self.h, None, 0)
        out = np.empty(n)
        self.lib.fn("evals_train")(self.h, _ptr(out), n)
        return out

    def evals_valid(self):
        """Validation metric after each tree of the last fit (reference evals.valid)."""
        n = self.lib.fn("evals_valid")(self.h, None, 0)
        out = np.empty(n)
        self.lib.fn("evals_valid")(self.h, _ptr(out), n)
        return out

    def best_iteration(self):
        return self.lib.fn("best_iteration")(self.h)

    def set_cat(self, is_cat):
        """Categorical flags of later fits (Booster::fit's is_cat); None clears."""
        if is_cat is None:
            self.lib.fn("set_cat")(self.h, None, 0)
            return self
        self._cat = np.ascontiguousarray(is_cat, np.int32)
        self.lib.fn("set_cat")(self.h, _ptr(self._cat), self._cat.size)
        return self

    def set_monotone(self, monotone):
        """Booster::fit's monotone constraints of later fits (+1 / -1 / 0 per feature)."""
        m = np.ascontiguousarray(monotone, np.int32)
        if self.lib.fn("set_monotone")(self.h, _ptr(m), m.size) != 0:
            raise ValueError(self.lib.error())
        return self

    def save(self, path, json=False):
        """Reference core only: its TGBM v3 writer (serialize.cpp:65-148) or JSON dump."""
        if self.lib.fn("save")(self.h, str(path).encode(), 1 if json else 0) != 0:
            raise ValueError(self.lib.error())

    @classmethod
    def load(cls, lib, path):
        """Reference core only: a model read by its TGBM v3 reader (serialize.cpp:150-255)."""
        h = lib.lib.tref_load(str(path).encode())
        if not h:
            raise ValueError(lib.error())
        m = cls.__new__(cls)
        m.lib = lib
        m.h = _P(h)
        return m

    def set_groups(self, group_of=None, interaction=None):
        """Feature groups (Booster::fit's group_of) and interaction groups
        (Config::interaction_group) of later fits; None clears."""
        self._grp = None if group_of is None else np.ascontiguousarray(group_of, np.int32)
        self._ix = None if interaction is None else np.ascontiguousarray(interaction, np.int32)
        self.lib.fn("set_groups")(self.h, _ptr(self._grp), 0 if self._grp is None else self._grp.size)
        self.lib.fn("set_interaction")(self.h, _ptr(self._ix), 0 if self._ix is None else self._ix.size)
        return self

    def cat_values(self, f):
        buf = np.empty(256)
        m = self.lib.fn("cat_values")(self.h, f, _ptr(buf), 256)
        return buf[:m].copy()

    def set_eval(self, X, y, ts=None):
        """Eval set of later fits (Booster::fit's X_eval / y_eval / ts_eval); X=None clears."""
        if X is None:
            self.lib.fn("set_eval")(self.h, None, None, None, 0, 0)
            return self
        self._eval = (np.ascontiguousarray(X, np.float64), np.ascontiguousarray(y, np.float64),
                      None if ts is None else np.ascontiguousarray(ts, np.int64))
        Xe, ye, te = self._eval
        if self.lib.fn("set_eval")(self.h, _ptr(Xe), _ptr(ye), _ptr(te), Xe.shape[0], Xe.shape[1]) != 0:
            raise ValueError(self.lib.error())
        return self


class CpuLib:
    def __init__(self, path: Path, prefix: str):
        self.path = path
        self.prefix = prefix
        self.lib = ctypes.CDLL(str(path))
        sig = {
            "create": (_P, []), "free": (None, [_P]), "set": (_I, [_P, ctypes.c_char_p, _D]),
            "fit": (_I, [_P, _P, _P, _I, _I, _P, _P, _P, _P, _I, _LL]),
            "update": (_I, [_P, _P, _P, _I, _I, _P, _I]),
            "predict": (_I, [_P, _P, _I, _I, _P]), "n_trees": (_I, [_P]),
            "tree_info": (_I, [_P, _I, _P, _P]), "tree_get": (_I, [_P, _I, _P, _P, _P, _P, _P]),
            "edges": (_I, [_P, _I, _P, _I]), "n_bins": (_I, [_P]), "init_prefix": (_I, [_P, _P, _I]),
            "init_score": (_D, [_P]), "last_error": (ctypes.c_char_p, []),
            "evals_train": (_I, [_P, _P, _I]), "evals_valid": (_I, [_P, _P, _I]),
            "best_iteration": (_I, [_P]), "set_eval": (_I, [_P, _P, _P, _P, _I, _I]),
            "tree_cat": (_I, [_P, _I, _P]), "set_cat": (_I, [_P, _P, _I]), "cat_values": (_I, [_P, _I, _P, _I]),
            "set_groups": (_I, [_P, _P, _I]), "set_interaction": (_I, [_P, _P, _I]),
            "set_monotone": (_I, [_P, _P, _I]),
        }
        for k, (res, args) in sig.items():
            f = getattr(self.lib, f"{prefix}_{k}")
            f.restype = res
            f.argtypes = args

    def fn(self, name):
        return getattr(self.lib, f"{self.prefix}_{name}")

    def error(self):
        return self.fn("last_error")().decode(errors="replace")

    def model(self, **params):
        return CpuModel(self, **params)


def oracle() -> CpuLib:
    lib = CpuLib(ORACLE_SO, "orc")
    L = lib.lib
    L.orc_fit_edges.argtypes = [_P, _I, _I, _I, _I, ctypes.c_uint, _P, _P]
    L.orc_fit_edges.restype = _I
    L.orc_bin_transform.argtypes = [_P, _I, _I, _P, _P, _P]
    L.orc_hist.argtypes = [_P, _I, _P, _I, _P, _P, _I, _P, _P, _P]
    L.orc_scan.argtypes = [_P, _P, _P, _I, _I, _D, _D, _I, _D, _D, _P, _P, _P, _P]
    L.orc_grad.argtypes = [_I, _P, _P, _P, _P, _I, _P, _P]
    L.orc_set_edges.argtypes = [_P, _P, _P, _I]
    L.orc_scan_feature.argtypes = [_P, _P, _P, _I, _I, _I, _D, _D, _I, _D, _I, _D, _P, _P, _P]
    return lib


def reference() -> CpuLib | None:
    if not REF_SO.exists():
        return None
    lib = CpuLib(REF_SO, "tref")
    L = lib.lib
    L.tref_bins.argtypes = [_P, _P, _I, _I, _P]
    L.tref_build_hist_col.argtypes = [_P, _I, _I, _P, _I, _P, _P, _I, _P, _P, _P]
    L.tref_best_numeric.argtypes = [_P, _P, _P, _I, _I, _I, _D, _D, _I, _D, _I, _D, _P, _P, _P]
    L.tref_grad.argtypes = [_I, _P, _P, _P, _P, _I, _P, _P]
    L.tref_save.argtypes = [_P, ctypes.c_char_p, _I]
    L.tref_load.restype = _P
    L.tref_load.argtypes = [ctypes.c_char_p]
    return lib


STRUCT_COLS = ["is_leaf", "split_type", "feature", "feature2", "threshold", "left", "right",
               "missing_go_left", "n_samples", "depth"]


def bins_of(X, edges, cat_values=None):
    """uint8-range bins of raw rows (reference Binner::transform, binning.cpp:104-128): NaN -> 0,
    else 1 + #edges <= x; a feature with category values (cat_values[f] not None) bins to
    1 + the value's index, 0 when the value is not listed."""
    xb = np.zeros(X.shape, np.int64)
    for f, e in enumerate(edges):
        if cat_values is not None and cat_values[f] is not None:
            v = np.asarray(cat_values[f])
            i = np.searchsorted(v, X[:, f], side="left")
            hit = (i < v.size) & (v[np.minimum(i, max(v.size - 1, 0))] == X[:, f]) if v.size else np.zeros(X.shape[0], bool)
            xb[:, f] = np.where(hit, i + 1, 0)
            continue
        b = 1 + np.searchsorted(e, X[:, f], side="right")
        b[np.isnan(X[:, f])] = 0
        xb[:, f] = b
    return xb


def node_left(ints, k, xb, n_bins=256, cat=None):
    """Rows that node k sends left (reference apply_one, tree.cpp:15-41): Axis: bin 0 -> missing
    direction, else bin <= threshold; Pair: either bin 0 -> missing direction, else the
    pair_bin hash (histogram.hpp:81-85) <= threshold; Cat / Regime: the bin's bit in the mask."""
    f, f2, thr, ml = ints[k, 2], ints[k, 3], ints[k, 4], ints[k, 7]
    b = xb[:, f]
    if ints[k, 1] == 1:
        b2 = xb[:, f2]
        z = ((b.astype(np.uint64) * 2654435761 + b2.astype(np.uint64) * 40503) & 0xFFFFFFFF) % n_bins
        return np.where((b == 0) | (b2 == 0), ml == 1, z.astype(np.int64) <= thr)
    if ints[k, 1] in (2, 3):
        words = np.asarray(cat[k], np.uint64)[np.minimum(b, 255) >> 6]
        bit = (words >> (b & 63).astype(np.uint64)) & np.uint64(1)
        return np.where(b >= 256, ml == 1, bit == 1)
    return np.where(b == 0, ml == 1, b <= thr)


def rows_reaching(ints, xb, node, n_bins=256, cat=None):
    """Boolean mask of the rows whose path through the tree passes `node`."""
    reach = np.zeros(xb.shape[0], bool)
    stack = [(0, np.ones(xb.shape[0], bool))]
    while stack:
        k, m = stack.pop()
        if k == node:
            return m
        if ints[k, 0]:
            continue
        left = node_left(ints, k, xb, n_bins, cat)
        stack.append((ints[k, 5], m & left))
        stack.append((ints[k, 6], m & ~left))
    return reach


def leaves_of(ints, xb, n_bins=256, cat=None):
    """Leaf node of every row (reference apply_one, tree.cpp:15-55)."""
    leaf = np.zeros(xb.shape[0], np.int64)
    stack = [(0, np.ones(xb.shape[0], bool))]
    while stack:
        k, m = stack.pop()
        if not m.any():
            continue
        if ints[k, 0]:
            leaf[m] = k
            continue
        left = node_left(ints, k, xb, n_bins, cat)
        stack.append((ints[k, 5], m & left))
        stack.append((ints[k, 6], m & ~left))
    return leaf


def tie_split(x, y, xb, k, n_bins=256):
    """Node k is an axis or categorical split in both trees and the two splits cut the rows
    reaching it into the same two sets (equal, or with the sides swapped): their gains are equal
    in exact arithmetic, and only rounding in the reference's FP64 sums can order them."""
    xi, yi = x["ints"], y["ints"]
    xc, yc = x.get("cat"), y.get("cat")
    if xi[k, 0] or yi[k, 0] or xi[k, 1] == 1 or yi[k, 1] == 1:
        return False
    m = rows_reaching(xi, xb, k, n_bins, xc)
    if not np.array_equal(m, rows_reaching(yi, xb, k, n_bins, yc)):
        return False
    lx = m & node_left(xi, k, xb, n_bins, xc)
    ly = m & node_left(yi, k, xb, n_bins, yc)
    return bool(np.array_equal(lx, ly) or np.array_equal(lx, m & ~ly))


def compare_trees(a, b, rtol=1e-6, atol=1e-9, check_values=True, xb=None, n_bins=256, stats=None):
    """First structural difference between two tree lists, or None.

    `stats`, when given (a dict), counts the accepted exceptions below by kind and lists them
    as (tree, node, kind) under "where".

    Integer node fields must match exactly; value/gain/values/prefix within tolerance.

    With the training rows' bins `xb`, two exceptions are accepted, both verified on the rows:
    - missing_go_left of an axis split whose node received no training row with the split
      feature missing.  There the reference's flag is set by the rounding residue that its FP64
      parent - sibling histogram subtraction leaves in the (empty) missing bin; exact sums leave
      no residue and keep missing-right, the first maximum.  No training row is routed by such a
      flag.
    - a different feature (with its threshold and missing direction) at an axis split when both
      splits send exactly the same training rows of the node left.  Their gains are then equal
      in exact arithmetic, and exact sums pick the lower feature index; the reference's
      sequential FP64 sums group the same rows by each feature's bins and differ by a few ulps,
      which decide its choice.  Everything below such a node sees the same rows.
    When the two splits send the same sets to opposite sides, the subtrees are swapped and node
    numbers below differ; such a tree is accepted when every training row reaches leaves of equal
    value and cover in both trees (the next trees then see the same predictions and must match).
    """
    if len(a) != len(b):
        return f"tree count {len(a)} != {len(b)}"

    def note(t, k, kind):
        if stats is not None:
            stats[kind] = stats.get(kind, 0) + 1
            stats.setdefault("where", []).append((t, int(k), kind))

    for t, (x, y) in enumerate(zip(a, b)):
        if x["ints"].shape != y["ints"].shape:
            return f"tree {t}: node count {x['ints'].shape[0]} != {y['ints'].shape[0]}"
        bad = np.argwhere(x["ints"] != y["ints"])
        if bad.size and xb is not None:
            xi, yi = x["ints"], y["ints"]
            tied = set()
            for k in {int(k) for k, c in bad if c == 2}:
                if xi[k, 0] or yi[k, 0] or xi[k, 1] != 0 or yi[k, 1] != 0:
                    continue
                m = rows_reaching(xi, xb, k, n_bins, x.get("cat"))
                if not np.array_equal(m, rows_reaching(yi, xb, k, n_bins, y.get("cat"))):
                    continue
                def goes_left(ints):
                    b = xb[:, ints[k, 2]]
                    return np.where(b == 0, ints[k, 7] == 1, b <= ints[k, 4])
                if np.array_equal(m & goes_left(xi), m & goes_left(yi)):
                    tied.add(k)
            keep = []
            for k in sorted(tied):
                note(t, k, "feature_tie")
            for k, c in bad:
                if k in tied and c in (2, 4, 7):
                    continue
                if c == 7 and xi[k, 1] == 0 and xi[k, 0] == 0:
                    m = rows_reaching(xi, xb, k, n_bins, x.get("cat"))
                    if not np.any(xb[m, xi[k, 2]] == 0):
                        note(t, k, "missing_go_left")
                        continue
                keep.append((k, c))
            bad = np.array(keep).reshape(-1, 2)
        xc, yc = x.get("cat"), y.get("cat")
        if xc is not None and yc is not None:
            # categorical masks of the nodes that match so far
            for k in np.flatnonzero(np.isin(x["ints"][:, 1], (2, 3)) & (x["ints"][:, 0] == 0)):
                if k not in set(bad[:, 0].tolist()) and not np.array_equal(xc[k], yc[k]):
                    bad = np.vstack([bad.reshape(-1, 2), [[k, 1]]])
        if bad.size and xb is not None:
            # a split decided by an exact-arithmetic tie: the same function on the training rows
            k = int(bad[:, 0].min())
            if tie_split(x, y, xb, k, n_bins):
                lx, ly = leaves_of(x["ints"], xb, n_bins, xc), leaves_of(y["ints"], xb, n_bins, yc)
                if (np.allclose(x["values"][lx], y["values"][ly], rtol=rtol, atol=atol)
                        and np.array_equal(x["cover"][lx], y["cover"][ly])):
                    note(t, k, "mirrored_or_categorical_tie")
                    continue
        if bad.size:
            k, c = bad[np.argmin(bad[:, 0])]
            return f"tree {t} node {k} {STRUCT_COLS[c]}: {x['ints'][k, c]} != {y['ints'][k, c]} (or its category mask)"
        if not np.array_equal(x["cover"], y["cover"]):
            return f"tree {t}: cover differs"
        if check_values:
            for key in ("values", "prefix"):
                if not np.allclose(x[key], y[key], rtol=rtol, atol=atol):
                    d = np.max(np.abs(x[key] - y[key]))
                    return f"tree {t}: {key} differ (max abs {d:g})"
            if not np.allclose(x["dbl"], y["dbl"], rtol=rtol, atol=atol):
                d = np.max(np.abs(x["dbl"] - y["dbl"]))
                return f"tree {t}: node value/gain differ (max abs {d:g})"
    return None


def leaf_values(tree, xb, n_bins=256):
    """Per row of bins xb, the value of the leaf the tree sends it to (values[leaf])."""
    return tree["values"][leaves_of(tree["ints"], xb, n_bins, tree.get("cat"))]


def held_out_differences(a, b, xb, rtol=1e-7, atol=1e-9, stats=None, n_bins=256):
    """Compare two tree lists on rows they were not fitted on (bins xb), tree by tree.

    Returns (trees whose leaf values differ on some row although the checker accepted no exception
    in them, boolean mask of rows whose leaf value differs in some tree).  A row may differ only in
    a tree where compare_trees accepted an exception (stats["where"])."""
    excepted = {t for t, _, _ in (stats or {}).get("where", [])}
    unexplained, rows = [], np.zeros(xb.shape[0], bool)
    for t, (x, y) in enumerate(zip(a, b)):
        va, vb = leaf_values(x, xb, n_bins), leaf_values(y, xb, n_bins)
        d = ~np.isclose(va, vb, rtol=rtol, atol=atol)
        if d.any():
            rows |= d
            if t not in excepted:
                unexplained.append(t)
    return unexplained, rows
