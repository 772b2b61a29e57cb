"""Gradients of the GPU path against the oracle's restatement of the reference objectives
(objective.cpp:105-119 MSE, :179-219 within-group pairwise), called through the device C ABI
(tgbm_task_set_rows, tgbm_task_grad, tgbm_task_get_grad).

Bar: MSE g and h bit-exact.  Pairwise: the device replays each row's pairs in the reference's
accumulation order, so g and h are bit-exact whenever every exp a row needs agrees bit for bit
between the device (ROCm OCML) and glibc.  Predictions that are equal within each group make
every exp argument 0 (exp(0) = 1 in both libraries): then g and h must be bit-exact, which pins
the pair sampling (step = m / 24 above 48 rows), skipped ties, weights, the accumulation order,
groups larger than the kernel's LDS stage and unsorted group keys.  With free predictions the
last-ulp exp differences are allowed: g and h within 1e-12 of the row's |g| + h scale, and more than
80% of rows still bit-identical.
"""
import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest

import checkers

pytestmark = pytest.mark.gpu

HIP = Path(__file__).resolve().parent.parent / "tidegbm_amd" / "lib" / "libtgbm_hip.so"
_P, _I = ctypes.c_void_p, ctypes.c_int
# group sizes: pairs of every row (<= 48), step 2 (49..71), step 41, the LDS stage edge, larger
SIZES = [1, 2, 3, 17, 48, 49, 71, 72, 500, 1000, 4096, 4097, 6000]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _lib():
    L = ctypes.CDLL(str(HIP))
    L.tgbm_session_create.argtypes = [_P, _I, _I, _I, _I, _P, _I, _I]
    L.tgbm_session_destroy.argtypes = [_P]
    L.tgbm_session_destroy.restype = None
    L.tgbm_task_set_rows.argtypes = [_P, _I, _P, _P, _P, _P, _P, _I, _P, _P]
    L.tgbm_task_grad.argtypes = [_P, _I, _I]
    L.tgbm_task_get_grad.argtypes = [_P, _I, _P, _P]
    L.tgbm_last_error.restype = ctypes.c_char_p
    return L


def _device_grad(L, y, pred, w, groups, objective):
    n = y.size
    s = ctypes.c_void_p()
    rows = np.array([n], np.int32)
    assert L.tgbm_session_create(ctypes.byref(s), n, 1, 256, 1, rows.ctypes.data, 31, 0) == 0, L.tgbm_last_error()
    try:
        gp = None if groups is None else groups.ctypes.data
        assert L.tgbm_task_set_rows(s, 0, y.ctypes.data, w.ctypes.data, pred.ctypes.data, None, None, 1, None,
                                    gp) == 0, L.tgbm_last_error()
        assert L.tgbm_task_grad(s, -1, objective) == 0, L.tgbm_last_error()
        g, h = np.empty(n), np.empty(n)
        assert L.tgbm_task_get_grad(s, 0, g.ctypes.data, h.ctypes.data) == 0, L.tgbm_last_error()
        return g, h
    finally:
        L.tgbm_session_destroy(s)


def _oracle_grad(y, pred, w, groups, objective):
    O = checkers.oracle().lib
    n = y.size
    g, h = np.empty(n), np.empty(n)
    gp = None if groups is None else groups.ctypes.data
    O.orc_grad(objective, y.ctypes.data, pred.ctypes.data, w.ctypes.data, gp, n, g.ctypes.data, h.ctypes.data)
    return g, h


def _panel(seed, shuffle_keys, equal_pred_in_group):
    rng = np.random.default_rng(seed)
    keys = np.repeat(np.arange(len(SIZES), dtype=np.int64) * 7 + 3, SIZES)
    n = keys.size
    y = rng.integers(0, 5, n).astype(np.float64)          # many ties
    cont = rng.random(n) < 0.2
    y[cont] += rng.normal(size=int(cont.sum()))
    w = rng.uniform(0.5, 2.0, n)
    if equal_pred_in_group:
        per = rng.normal(size=len(SIZES))
        pred = per[np.searchsorted(np.cumsum(SIZES), np.arange(n), side="right")]
    else:
        pred = rng.normal(scale=2.0, size=n)
        pred[rng.random(n) < 0.05] *= 40.0                  # past the +-30 clamp
    if shuffle_keys:
        perm = rng.permutation(n)
        keys, y, w, pred = keys[perm], y[perm], w[perm], pred[perm]
    return y, np.ascontiguousarray(pred), w, np.ascontiguousarray(keys)


def test_mse_gradients_bit_exact():
    L = _lib()
    rng = np.random.default_rng(1)
    n = 100_003
    y, pred, w = rng.normal(size=n), rng.normal(size=n), rng.uniform(0.1, 3.0, n)
    _say(f"[grad] mse n={n}")
    g, h = _device_grad(L, y, pred, w, None, 0)
    go, ho = _oracle_grad(y, pred, w, None, 0)
    assert np.array_equal(g, go) and np.array_equal(h, ho)


@pytest.mark.parametrize("shuffle_keys", [False, True], ids=["sorted-keys", "unsorted-keys"])
def test_pairwise_gradients_bit_exact_without_exp_rounding(shuffle_keys):
    L = _lib()
    y, pred, w, keys = _panel(2, shuffle_keys, equal_pred_in_group=True)
    _say(f"[grad] pairwise equal predictions per group, n={y.size}, shuffled keys={shuffle_keys}")
    g, h = _device_grad(L, y, pred, w, keys, 5)
    go, ho = _oracle_grad(y, pred, w, keys, 5)
    bad = np.flatnonzero((g != go) | (h != ho))
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}: g {g[bad[:3]]} vs {go[bad[:3]]}"
    assert np.count_nonzero(g) > y.size // 2   # the pairs did contribute


def test_pairwise_gradients_close_with_free_predictions():
    L = _lib()
    y, pred, w, keys = _panel(3, False, equal_pred_in_group=False)
    _say(f"[grad] pairwise free predictions, n={y.size}")
    g, h = _device_grad(L, y, pred, w, keys, 5)
    go, ho = _oracle_grad(y, pred, w, keys, 5)
    scale = np.abs(go) + ho
    assert np.all(np.abs(g - go) <= 1e-12 * scale)
    assert np.all(np.abs(h - ho) <= 1e-12 * scale)
    same = np.mean((g == go) & (h == ho))
    _say(f"[grad] rows bit-identical: {same:.4f}")
    # a row differs only where one of its ~24-48 exp values differs in the last ulp (0.93 measured)
    assert same > 0.8
