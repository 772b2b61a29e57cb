"""MI355X-native histogram-GBDT path of TideGBM, behind the reference's C API.

The native library (lib/libtidegbm_amd.so, C++ host + gfx950 HIP kernels) is loaded with ctypes;
there is no Python or CPU fallback: if the library is missing, importing this package raises.

`TideRegressor` / `TideRanker` mirror the reference estimators' data-prep contract for this path
(reference src/tidegbm/native.py:30-220): horizon/embargo converted to timestamp units
(utils.py:47-59), rows ordered by (timestamp, entity code) (native.py:92-102), rank groups =
timestamps for ranking objectives (native.py:134-137), native errors re-raised as ValueError
(native.py:216-217), and the `device` keyword (api.py:79) whose "gpu" value selects this path.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_LIBDIR = Path(__file__).resolve().parent / "lib"
_LIBPATH = _LIBDIR / "libtidegbm_amd.so"

if not _LIBPATH.exists():
    raise ImportError(
        f"{_LIBPATH} is missing: build the native libraries first (python tidegbm_amd/build.py)")

_lib = ctypes.CDLL(str(_LIBPATH))

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_LL = ctypes.c_longlong


def _sig(name, res, *args):
    fn = getattr(_lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_sig("tide_create", _I, ctypes.POINTER(_P))
_sig("tide_free", None, _P)
_sig("tide_set_int", _I, _P, ctypes.c_char_p, _I)
_sig("tide_set_double", _I, _P, ctypes.c_char_p, _D)
_sig("tide_set_string", _I, _P, ctypes.c_char_p, ctypes.c_char_p)
_sig("tide_fit", _I, _P, _P, _P, _I, _I, _P, _P, _P, _I, _I)
_sig("tide_fit_ex", _I, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P, _I, _LL)
_sig("tide_update", _I, _P, _P, _P, _I, _I, _P, _I)
_sig("tide_predict", _I, _P, _P, _I, _I, _P)
_sig("tide_predict_raw", _I, _P, _P, _I, _I, _P)
_sig("tide_predict_binned", _I, _P, _P, _I, _P)
_sig("tide_save", _I, _P, ctypes.c_char_p)
_sig("tide_save_json", _I, _P, ctypes.c_char_p)
_sig("tide_load", _I, ctypes.POINTER(_P), ctypes.c_char_p)
_sig("tide_last_error", ctypes.c_char_p)
_sig("tide_device_compiled", _I)
_sig("tide_device_count", _I)
_sig("tide_num_trees", _I, _P)
_sig("tide_tree_info", _I, _P, _I, _P, _P)
_sig("tide_tree_get", _I, _P, _I, _P, _P, _P, _P, _P)
_sig("tide_model_info", _I, _P, _P, _P, _P, _P, _P)
_sig("tide_get_edges", _I, _P, _I, _P, _I)
_sig("tide_get_init_prefix", _I, _P, _P, _I)
_sig("tide_profile", _I, _P, _I, _P, _P, _P)
_sig("tide_fit_times", _I, _P, _P, _P)
_sig("tide_goss_stats", _I, _P, _P, _P, _P)
_sig("tide_hist_variant", _I, _P)
_sig("tide_evals_train", _I, _P, _P, _I)
_sig("tide_evals_valid", _I, _P, _P, _I)
_sig("tide_tree_cat", _I, _P, _I, _P)
_sig("tide_set_feature_groups", _I, _P, _P, _I)
_sig("tide_set_monotone", _I, _P, _P, _I)
_sig("tide_set_interaction_groups", _I, _P, _P, _I)
_sig("tide_get_cat_values", _I, _P, _I, _P, _I)
_sig("tide_set_eval", _I, _P, _P, _P, _P, _I, _I)
_sig("tide_profile_log", _I, _P, _I, _P, _P, _I)
_sig("tide_tree_times", _I, _P, _P, _I)
_sig("tide_fit_batch", _I, _P, _P, _I, _I, _P, _P, _I, _P, _P, _P, _I, _P)
_sig("tide_server_create", _I, ctypes.POINTER(_P), _P, _I, _P, _I, _I, _LL, _I, _I, _I)
_sig("tide_server_free", None, _P)
_sig("tide_server_predict", _LL, _P, _I, _P, _I)
_sig("tide_server_predict_x", _LL, _P, _I, _P, _I)
_sig("tide_server_refit", _LL, _P, _I, _I, _I, _P, _I)
_sig("tide_server_wait", _I, _P, _LL, _P, _I, _P, _P)
_sig("tide_server_model", _I, _P, _I, _I, ctypes.POINTER(_P), _P)
_sig("tide_server_stats", _I, _P, _P, _P)

#: True: the gfx950 kernels are compiled into the loaded library (reference: tidegbm.cuda_compiled).
gpu_compiled: bool = bool(_lib.tide_device_compiled())
cuda_compiled = gpu_compiled  # the reference's flag name

OBJECTIVES = {"mse": 0, "auto": 0, "l2": 0, "regression": 0, "pairwise": 5, "lambdarank": 5}
_INT_KEYS = {"n_estimators", "max_depth", "max_leaves", "max_bin", "n_time_buckets", "random_state",
             "n_jobs", "iid_mode", "halo", "verbose", "horizon", "embargo", "goss", "cat_min_count",
             "col_major_hist", "min_samples_leaf", "pair_splits", "pair_top_k",
             "max_pair_splits_per_tree", "bin_sample", "finer_torb", "profile", "hist_mode", "autotune", "spec_width",
             "force_host_driver", "early_stopping_rounds", "regime_feature", "regime_priority"}
_DOUBLE_KEYS = {"learning_rate", "l2", "l1", "subsample", "colsample", "min_child_weight", "cat_smooth",
                "min_split_gain", "pair_trigger", "pair_l2", "goss_top", "goss_other",
                "max_delta_step", "torb_min_cover", "time_decay", "tree_decay"}


#: Every configuration key Booster.set_params accepts.
KNOWN_KEYS = frozenset(_INT_KEYS | _DOUBLE_KEYS | {"device", "eval_metric", "objective", "interaction_group",
                                                   "feature_groups", "monotone"})


class TideError(ValueError):
    pass


def _err() -> str:
    return _lib.tide_last_error().decode(errors="replace")


def _ptr(a):
    return None if a is None else a.ctypes.data


def device_count() -> int:
    return int(_lib.tide_device_count())


class Booster:
    """Owner of one native handle (the reference's Booster behind its C API)."""

    def __init__(self, handle=None, **params):
        h = _P()
        if handle is None:
            if _lib.tide_create(ctypes.byref(h)) != 0:
                raise TideError(_err())
            self._h = h
        else:
            self._h = handle
        self.set_params(**params)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.tide_free(h)
            self._h = None

    def set_params(self, **params):
        for k, v in params.items():
            if v is None:
                continue
            kb = k.encode()
            if k in ("device", "eval_metric"):
                rc = _lib.tide_set_string(self._h, kb, str(v).encode())
            elif k in ("interaction_group", "feature_groups", "monotone"):
                a = np.ascontiguousarray(v, dtype=np.int32)
                fn = {"interaction_group": _lib.tide_set_interaction_groups,
                      "feature_groups": _lib.tide_set_feature_groups, "monotone": _lib.tide_set_monotone}[k]
                rc = fn(self._h, _ptr(a), a.size)
            elif k == "objective":
                if isinstance(v, (int, np.integer)):
                    rc = _lib.tide_set_int(self._h, kb, int(v))
                else:
                    rc = _lib.tide_set_string(self._h, kb, str(v).encode())
            elif k in _INT_KEYS:
                rc = _lib.tide_set_int(self._h, kb, int(v))
            elif k in _DOUBLE_KEYS:
                rc = _lib.tide_set_double(self._h, kb, float(v))
            else:
                raise TideError(f"unknown parameter {k}")
            if rc != 0:
                raise TideError(_err())
        return self

    def fit(self, X, y, *, timestamps=None, weight=None, entity=None, rank_groups=None,
            iid: bool = False, horizon: int = -1, eval_set=None, eval_timestamp=None, is_cat=None):
        """Booster::fit.  eval_set=(X_eval, y_eval) (with eval_timestamp outside iid mode) scores
        every tree on those rows and enables early_stopping_rounds, as the reference does;
        is_cat flags categorical features (one entry per feature)."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        n, F = X.shape
        if eval_set is not None:
            Xe = np.ascontiguousarray(eval_set[0], dtype=np.float64)
            ye = np.ascontiguousarray(eval_set[1], dtype=np.float64)
            te = None if eval_timestamp is None else np.ascontiguousarray(eval_timestamp, dtype=np.int64)
            if Xe.ndim != 2 or Xe.shape[0] != ye.shape[0] or (te is not None and te.shape[0] != ye.shape[0]):
                raise TideError("eval_set rows, targets and timestamps must have the same length")
            rc = _lib.tide_set_eval(self._h, _ptr(Xe), _ptr(ye), _ptr(te), Xe.shape[0], Xe.shape[1])
        else:
            rc = _lib.tide_set_eval(self._h, None, None, None, 0, 0)
        if rc != 0:
            raise TideError(_err())
        ts = None if timestamps is None else np.ascontiguousarray(timestamps, dtype=np.int64)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
        ent = None if entity is None else np.ascontiguousarray(entity, dtype=np.int64)
        rg = None if rank_groups is None else np.ascontiguousarray(rank_groups, dtype=np.int64)
        cat = None if is_cat is None else np.ascontiguousarray(np.asarray(is_cat) != 0, dtype=np.int32)
        if cat is not None and cat.shape != (F,):
            raise TideError("is_cat needs one entry per feature")
        rc = _lib.tide_fit_ex(self._h, _ptr(X), _ptr(y), n, F, _ptr(ts), _ptr(w), _ptr(cat), _ptr(ent),
                              _ptr(rg), 1 if iid else 0, int(horizon))
        if rc != 0:
            raise TideError(_err())
        return self

    def update(self, X, y, extra_trees: int, weight=None):
        X = np.ascontiguousarray(X, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
        if _lib.tide_update(self._h, _ptr(X), _ptr(y), X.shape[0], X.shape[1], _ptr(w), int(extra_trees)) != 0:
            raise TideError(_err())
        return self

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        out = np.empty(X.shape[0], dtype=np.float64)
        if _lib.tide_predict(self._h, _ptr(X), X.shape[0], X.shape[1], _ptr(out)) != 0:
            raise TideError(_err())
        return out

    def predict_binned(self, xb):
        xb = np.ascontiguousarray(xb, dtype=np.uint16)
        out = np.empty(xb.shape[0], dtype=np.float64)
        if _lib.tide_predict_binned(self._h, _ptr(xb), xb.shape[0], _ptr(out)) != 0:
            raise TideError(_err())
        return out

    def save(self, path, json: bool = False):
        fn = _lib.tide_save_json if json else _lib.tide_save
        if fn(self._h, str(path).encode()) != 0:
            raise TideError(_err())

    @classmethod
    def load(cls, path):
        h = _P()
        if _lib.tide_load(ctypes.byref(h), str(path).encode()) != 0:
            raise TideError(_err())
        return cls(handle=h)

    # ---- inspection
    def model_info(self):
        nf, nb, nt, bi = _I(), _I(), _I(), _I()
        init = _D()
        _lib.tide_model_info(self._h, ctypes.byref(nf), ctypes.byref(nb), ctypes.byref(nt), ctypes.byref(bi),
                             ctypes.byref(init))
        return dict(n_features=nf.value, n_bins=nb.value, n_time_buckets=nt.value,
                    best_iteration=bi.value, init_score=init.value)

    def trees(self):
        return [tree_arrays(self._h, _lib.tide_tree_info, _lib.tide_tree_get, t, _lib.tide_tree_cat)
                for t in range(_lib.tide_num_trees(self._h))]

    def cat_values(self, f):
        """Sorted category values of categorical feature f (empty for a numeric one)."""
        buf = np.empty(256)
        m = _lib.tide_get_cat_values(self._h, f, _ptr(buf), 256)
        return buf[:m].copy()

    def edges(self, f):
        buf = np.empty(256)
        m = _lib.tide_get_edges(self._h, f, _ptr(buf), 256)
        if m < 0:
            raise TideError(_err())
        return buf[:m].copy()

    def init_prefix(self):
        buf = np.empty(512)
        m = _lib.tide_get_init_prefix(self._h, _ptr(buf), 512)
        return buf[:m].copy()

    def profile(self, family: int):
        ms, n, b = _D(), _LL(), _D()
        _lib.tide_profile(self._h, family, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b))
        return ms.value, n.value, b.value

    def profile_log(self, family: int):
        """(ms per launch, rows per launch) of a profiled fit's kernel family, launch order."""
        m = _lib.tide_profile_log(self._h, family, None, None, 0)
        if m < 0:
            raise TideError(_err())
        ms, rows = np.empty(m), np.empty(m)
        _lib.tide_profile_log(self._h, family, _ptr(ms), _ptr(rows), m)
        return ms, rows

    def tree_times(self):
        buf = np.empty(1 << 16)
        m = _lib.tide_tree_times(self._h, _ptr(buf), buf.size)
        return buf[:m].copy()

    def fit_times(self):
        a, b = _D(), _D()
        _lib.tide_fit_times(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def evals_result(self) -> dict:
        """The reference's evals_result(): training loss and validation metric per tree."""
        res = {}
        for key, fn in (("train", _lib.tide_evals_train), ("valid", _lib.tide_evals_valid)):
            n = int(fn(self._h, None, 0))
            out = np.empty(n)
            if n:
                fn(self._h, _ptr(out), n)
            res[key] = out
        return res

    @property
    def best_iteration(self) -> int:
        return self.model_info()["best_iteration"]

    def hist_variant(self) -> int:
        """Histogram kernel variant the last fit ended with (tuned or forced; same trees)."""
        return int(_lib.tide_hist_variant(self._h))

    def goss_stats(self):
        """(trees sampled on the host because of equal |g|, seconds waiting for the generator
        draws, seconds in the sample step) of the last GOSS fit."""
        n, w, d = ctypes.c_int(), _D(), _D()
        _lib.tide_goss_stats(self._h, ctypes.byref(n), ctypes.byref(w), ctypes.byref(d))
        return n.value, w.value, d.value


def fit_batch(X, timestamps, task_rows, targets, *, entity=None, horizons=None,
              schedule: str = "batched", **params):
    """Walk-forward fits over row prefixes of one shared binned matrix.

    Rows of X must be sorted by (timestamps, entity); task t trains on rows [0, task_rows[t])
    with target targets[t] and horizon horizons[t].  schedule "batched" grows every task's trees
    together; "sequential" fits the tasks one after another on the same resident bins (the
    plain one-fit-per-call pattern).  Returns one fitted Booster per task.
    """
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, F = X.shape
    ts = None if timestamps is None else np.ascontiguousarray(timestamps, dtype=np.int64)
    ent = None if entity is None else np.ascontiguousarray(entity, dtype=np.int64)
    rows = np.ascontiguousarray(task_rows, dtype=np.int32)
    T = len(rows)
    ys = [np.ascontiguousarray(y, dtype=np.float64) for y in targets]
    if len(ys) != T:
        raise TideError("one target per task is required")
    for t in range(T):
        if ys[t].shape[0] != rows[t]:
            raise TideError(f"target {t} has {ys[t].shape[0]} values for {rows[t]} rows")
    yp = (ctypes.c_void_p * T)(*[y.ctypes.data for y in ys])
    hz = None if horizons is None else np.ascontiguousarray(horizons, dtype=np.int64)
    if hz is not None and hz.shape[0] != T:
        raise TideError("one horizon per task is required")
    sched = {"batched": 0, "sequential": 1}.get(schedule)
    if sched is None:
        raise TideError(f"unknown schedule {schedule!r}")
    proto = Booster(**params)
    out = (_P * T)()
    rc = _lib.tide_fit_batch(proto._h, _ptr(X), n, F, _ptr(ts), _ptr(ent), T, _ptr(rows), yp, _ptr(hz),
                             sched, out)
    if rc != 0:
        raise TideError(_err())
    return [Booster(handle=_P(out[t])) for t in range(T)]


class Server:
    """Serving loop over resident ensembles and a resident binned panel (see tide_server_create).

    predict / predict_x / refit return tickets at once; wait(ticket) blocks for the result:
    (scores or None, version, latency seconds).
    """

    def __init__(self, models, panel_X, *, predict_budget=1 << 30, refit_slots=16,
                 refit_max_rows=262144, batched=True):
        X = np.ascontiguousarray(panel_X, dtype=np.float64)
        self._X_rows, self._F = X.shape
        hs = (_P * len(models))(*[m._h for m in models])
        self._models = list(models)   # keep the handles alive while the server copies them
        h = _P()
        rc = _lib.tide_server_create(ctypes.byref(h), hs, len(models), _ptr(X), X.shape[0], X.shape[1],
                                     int(predict_budget), int(refit_slots), int(refit_max_rows),
                                     1 if batched else 0)
        if rc != 0:
            raise TideError(_err())
        self._h = h
        self._n = {}

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.tide_server_free(h)
            self._h = None

    def __del__(self):
        self.close()

    def _ticket(self, t, n, kind):
        if t < 0:
            raise TideError(_err())
        self._n[t] = (n, kind)
        return t

    def predict(self, model: int, rows):
        r = np.ascontiguousarray(rows, dtype=np.int32)
        return self._ticket(_lib.tide_server_predict(self._h, int(model), _ptr(r), r.shape[0]), r.shape[0], 0)

    def predict_x(self, model: int, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        if X.ndim != 2 or X.shape[1] != self._F:
            raise TideError("X must be n x n_features")
        return self._ticket(_lib.tide_server_predict_x(self._h, int(model), _ptr(X), X.shape[0]), X.shape[0], 0)

    def refit(self, model: int, row0: int, y, extra_trees: int = 50):
        y = np.ascontiguousarray(y, dtype=np.float64)
        t = _lib.tide_server_refit(self._h, int(model), int(row0), y.shape[0], _ptr(y), int(extra_trees))
        return self._ticket(t, y.shape[0], 2)

    def wait(self, ticket):
        n, kind = self._n.pop(ticket)
        out = np.empty(n) if kind != 2 else None
        ver, lat = _I(), _D()
        rc = _lib.tide_server_wait(self._h, ticket, _ptr(out), n if out is not None else 0, ctypes.byref(ver),
                                   ctypes.byref(lat))
        if rc < 0:
            raise TideError(_err())
        return out, ver.value, lat.value

    def model(self, model: int, version: int = -1):
        """(Booster snapshot of a version, parent version)."""
        h, parent = _P(), _I()
        if _lib.tide_server_model(self._h, int(model), int(version), ctypes.byref(h), ctypes.byref(parent)) != 0:
            raise TideError(_err())
        return Booster(handle=h), parent.value

    def stats(self):
        c = np.zeros(5, np.int64)
        b = np.zeros(5)
        _lib.tide_server_stats(self._h, _ptr(c), _ptr(b))
        return dict(predict_launches=int(c[0]), predict_requests=int(c[1]), predict_rows=int(c[2]),
                    refit_rounds=int(c[3]), refits=int(c[4]), predict_busy_s=float(b[0]),
                    refit_busy_s=float(b[1]), refit_admit_s=float(b[2]), refit_round_s=float(b[3]),
                    refit_publish_s=float(b[4]))


def tree_arrays(h, info_fn, get_fn, t, cat_fn=None):
    """Trees as dicts of numpy arrays: ints (nn,10), dbl (nn,2), values, cover, prefix (nb*nn),
    and with cat_fn the categorical bin masks cat (nn,4) uint64."""
    nn, nb = _I(), _I()
    info_fn(h, t, ctypes.byref(nn), ctypes.byref(nb))
    ints = np.empty((nn.value, 10), np.int32)
    dbl = np.empty((nn.value, 2))
    vals = np.empty(nn.value)
    cov = np.empty(nn.value, np.int32)
    pref = np.empty(nn.value * nb.value)
    get_fn(h, t, _ptr(ints), _ptr(dbl), _ptr(vals), _ptr(cov), _ptr(pref))
    out = dict(ints=ints, dbl=dbl, values=vals, cover=cov, prefix=pref, n_buckets=nb.value)
    if cat_fn is not None:
        cat = np.zeros((nn.value, 4), np.uint64)
        cat_fn(h, t, _ptr(cat))
        out["cat"] = cat
    return out


# ---------------------------------------------------------------- estimator mirror
def _median_step(ts: np.ndarray) -> int:  # reference utils.py:62-70
    u = np.unique(ts)
    if u.size < 2:
        return 1
    d = np.diff(np.sort(u))
    d = d[d > 0]
    return 1 if d.size == 0 else int(np.median(d))


def parse_horizon(value, ts: np.ndarray) -> int:
    """Horizon/embargo in timestamp units (reference utils.py:47-59, integer timestamps)."""
    if value is None:
        return 0
    if isinstance(value, (int, np.integer)):
        return int(value) * max(_median_step(ts), 1)
    if isinstance(value, float) and value.is_integer():
        return parse_horizon(int(value), ts)
    raise TideError("horizon must be an integer number of timestamp steps on this path")


class TideRegressor:
    """Estimator front end for the GPU path (reference TideEstimator, api.py:38-300)."""

    _objective = "mse"

    def __init__(self, n_estimators=100, learning_rate=0.1, max_depth=6, max_leaves=31,
                 min_samples_leaf=20, l2=1.0, max_bin=255, horizon=None, embargo=None,
                 iid_mode=False, pair_splits=True, goss=-1, random_state=0, device="gpu",
                 objective=None, **kw):
        self.params = dict(n_estimators=n_estimators, learning_rate=learning_rate, max_depth=max_depth,
                           max_leaves=max_leaves, min_samples_leaf=min_samples_leaf, l2=l2,
                           max_bin=max_bin, pair_splits=pair_splits, goss=goss,
                           random_state=random_state, device=device,
                           objective=objective or self._objective, **kw)
        self.horizon = horizon
        self.embargo = embargo
        self.iid_mode = iid_mode
        self.booster_: Booster | None = None

    def fit(self, X, y, *, timestamp=None, entity=None, sample_weight=None, eval_set=None,
            eval_timestamp=None, categorical_feature=None):
        """categorical_feature: indices of categorical columns (reference data.py:66-78)."""
        X = np.asarray(X, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64).reshape(-1)
        if not self.iid_mode and timestamp is None:
            raise ValueError("timestamp is required. Pass timestamp=... or set iid_mode=True "
                             "(iid_mode disables TORB and is unsafe for market data).")
        w = np.ones(len(y)) if sample_weight is None else np.asarray(sample_weight, dtype=np.float64)
        ts = ent = rg = None
        horizon = -1
        p = dict(self.params)
        if not self.iid_mode:
            ts = np.asarray(timestamp).astype(np.int64)
            if self.horizon is None:
                raise ValueError("horizon is required in time mode (label length in timestamp units).")
            horizon = parse_horizon(self.horizon, ts)
            p["embargo"] = parse_horizon(self.embargo, ts)
            ent = np.zeros(len(y), dtype=np.int64)
            if entity is not None:
                _, ent = np.unique(np.asarray(entity), return_inverse=True)
                ent = ent.astype(np.int64)
            order = np.lexsort((ent, ts))
            if not np.all(order == np.arange(order.size)):
                X, y, ts, w, ent = X[order], y[order], ts[order], w[order], ent[order]
            if OBJECTIVES.get(str(p["objective"])) == 5:
                rg = ts
        if p.get("early_stopping_rounds") is None:  # api.py:23-24
            p["early_stopping_rounds"] = 0
        te = None
        if eval_set is not None and not self.iid_mode:   # native.py:140-152
            if eval_timestamp is None:
                raise ValueError("eval_timestamp is required when iid_mode=False")
            te = np.asarray(eval_timestamp).astype(np.int64)
        is_cat = None
        if categorical_feature:
            is_cat = np.zeros(X.shape[1], np.int32)
            for c in categorical_feature:
                if not isinstance(c, (int, np.integer)) or not 0 <= int(c) < X.shape[1]:
                    raise ValueError(f"categorical feature {c!r} not in X")
                is_cat[int(c)] = 1
        b = Booster(**p)
        b.fit(X, y, timestamps=ts, weight=None if np.all(w == 1.0) else w, entity=ent, is_cat=is_cat,
              rank_groups=rg, iid=self.iid_mode, horizon=horizon,
              eval_set=None if eval_set is None else (np.asarray(eval_set[0], dtype=np.float64),
                                                      np.asarray(eval_set[1], dtype=np.float64).reshape(-1)),
              eval_timestamp=te)
        self.booster_ = b
        return self

    def predict(self, X):
        if self.booster_ is None:
            raise ValueError("not fitted")
        return self.booster_.predict(X)

    @property
    def evals_result_(self) -> dict:
        return self.booster_.evals_result()

    @property
    def best_iteration_(self) -> int:
        return self.booster_.best_iteration


class TideRanker(TideRegressor):
    _objective = "pairwise"


__all__ = ["Booster", "TideRegressor", "TideRanker", "TideError", "gpu_compiled", "cuda_compiled",
           "device_count", "parse_horizon", "tree_arrays", "OBJECTIVES", "KNOWN_KEYS", "Server", "fit_batch"]
