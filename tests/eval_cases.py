"""Eval-set cases shared by the oracle pin tests and the GPU parity tests."""
import numpy as np

# name: (iid, objective, eval_metric code 0 auto / 1 loss / 2 ic, early_stopping_rounds, NaN target)
EVAL_CASES = {
    "time_auto_ic_es5": (False, 0, 0, 5, False),
    "iid_auto_loss_es3": (True, 0, 0, 3, False),
    "iid_ic_no_ts": (True, 0, 2, 0, False),
    "time_loss_no_es": (False, 0, 1, 0, False),
    "pairwise_auto_ic": (False, 5, 0, 4, False),
    "nan_target_grows_all": (False, 0, 1, 3, True),
}
METRIC_NAMES = {0: "auto", 1: "loss", 2: "ic"}


def split_by_date(p, frac=0.7, gap=1):
    """Rows before the cut date train; rows from cut + gap on validate (later dates)."""
    dates = np.unique(p.ts)
    cut = dates[int(len(dates) * frac)]
    return p.ts < cut, p.ts >= cut + gap


def eval_target(p, va, nan_y):
    ye = p.y[va].copy()
    if nan_y:
        ye[5] = np.nan
    return ye
