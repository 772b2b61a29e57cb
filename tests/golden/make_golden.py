"""Write the golden fixtures of tests/golden/ from the reference core (oracle/_ref).

Inputs are regenerated from the seeded panels in tidegbm_amd.synth, so only outputs (and an input
checksum that detects generator drift) are stored.  Run from the repository root after
`python tidegbm_amd/build.py` on a machine where the reference tree was present:

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import checkers  # noqa: E402
from cat_cases import cat_panel  # noqa: E402
from tidegbm_amd import synth  # noqa: E402

OUT = Path(__file__).resolve().parent


def cases():
    """(name, panel, rank_groups, iid, horizon, params, opts) of every fixture; opts may hold
    is_cat, feature groups / interaction groups and an eval set (train mask, eval mask)."""
    small = synth.panel_small(0, n_dates=120)
    rng = np.random.default_rng(0)
    xor = synth.Panel(small.X, np.sign(small.X[:, 0] * small.X[:, 1]) + 0.1 * rng.normal(size=small.n),
                      small.ts, small.entity)
    rank = synth.panel_small(3, n_dates=60, n_entities=70)
    nan = synth.panel_large(5, n_dates=60, n_entities=50, n_features=10, nan_frac=0.05)
    cat, is_cat = cat_panel(35, n_dates=80, n_entities=30)
    grp = np.full(cat.n_features, -1, np.int32)
    grp[[0, 1, 2]] = 0
    ixg = np.full(cat.n_features, -1, np.int32)
    ixg[[19, 15]] = 0
    ixg[[18, 13]] = 1
    dates = np.unique(small.ts)
    cut = dates[int(len(dates) * 0.7)]
    cases = [
        ("mse_time", small, None, False, 1, dict(n_estimators=8, max_depth=6, learning_rate=0.08, goss=0)),
        ("mse_pair", xor, None, False, 1, dict(n_estimators=8, max_depth=6, learning_rate=0.08, goss=0)),
        ("mse_iid", small, None, True, 1, dict(n_estimators=6, max_depth=5, learning_rate=0.1, goss=0)),
        ("pairwise", rank, rank.ts, False, 1, dict(n_estimators=5, max_depth=6, learning_rate=0.08, goss=0,
                                                    objective=5)),
        ("goss", small, None, False, 1, dict(n_estimators=5, max_depth=5, learning_rate=0.1, goss=1)),
        ("nan_h10", nan, None, False, 10, dict(n_estimators=5, max_depth=6, learning_rate=0.1, goss=0,
                                               bin_sample=500, pair_splits=0)),
        ("categorical", cat, None, False, 1, dict(n_estimators=5, max_depth=5, learning_rate=0.1, goss=0),
         dict(is_cat=is_cat)),
        ("groups_regime", cat, None, False, 1, dict(n_estimators=5, max_depth=5, learning_rate=0.1, goss=0,
                                                    regime_feature=7),
         dict(is_cat=is_cat, group_of=grp, interaction=ixg)),
        ("eval_ic_es", small, None, False, 1, dict(n_estimators=8, max_depth=5, learning_rate=0.1, goss=0,
                                                   early_stopping_rounds=3),
         dict(train=small.ts < cut, eval=small.ts >= cut + 1)),
    ]
    return [c if len(c) == 7 else c + ({},) for c in cases]


def fit(lib, name, p, rg, iid, horizon, params, opts=None):
    opts = opts or {}
    m = lib.model(**params)
    if "is_cat" in opts:
        m.set_cat(opts["is_cat"])
    if "group_of" in opts or "interaction" in opts:
        m.set_groups(opts.get("group_of"), opts.get("interaction"))
    rows = opts.get("train", np.ones(p.n, bool))
    if "eval" in opts:
        va = opts["eval"]
        m.set_eval(p.X[va], p.y[va], None if iid else p.ts[va])
    m.fit(p.X[rows], p.y[rows], ts=None if iid else p.ts[rows], entity=None if iid else p.entity[rows],
          rank_groups=None if rg is None else rg[rows], iid=iid, horizon=horizon)
    return m


def pack(m, p):
    trees = m.trees()
    nn = np.array([t["ints"].shape[0] for t in trees], np.int32)
    nb = np.array([t["n_buckets"] for t in trees], np.int32)
    edges = [m.edges(f) for f in range(p.n_features)]
    return dict(
        input_checksum=np.array([np.nansum(p.X), np.sum(p.y)]),
        n_nodes=nn, n_buckets=nb,
        ints=np.concatenate([t["ints"] for t in trees]),
        dbl=np.concatenate([t["dbl"] for t in trees]),
        values=np.concatenate([t["values"] for t in trees]),
        cover=np.concatenate([t["cover"] for t in trees]),
        prefix=np.concatenate([t["prefix"] for t in trees]),
        n_edges=np.array([e.size for e in edges], np.int32),
        edges=np.concatenate(edges),
        init_prefix=m.init_prefix(),
        pred=m.predict(p.X[:400]),
        cat=np.concatenate([t["cat"] for t in trees]),
        cat_values=np.concatenate([m.cat_values(f) for f in range(p.n_features)]),
        evals_train=m.evals_train(),
        evals_valid=m.evals_valid(),
        best_iteration=np.array([m.best_iteration()], np.int32),
    )


def toolchain():
    """The library versions whose implementation-defined behaviour the fixtures capture
    (std::sort / std::shuffle tie orders in libstdc++, exp in glibc)."""
    import subprocess
    out = []
    for cmd in (["g++", "--version"], ["ldd", "--version"]):
        try:
            out.append(subprocess.run(cmd, capture_output=True, text=True).stdout.splitlines()[0])
        except (OSError, IndexError):
            out.append(f"{cmd[0]}: unavailable")
    return "\n".join(out) + "\n"


def main():
    ref = checkers.reference()
    if ref is None:
        raise SystemExit("oracle/_ref/libtide_ref.so is not built")
    (OUT / "TOOLCHAIN.txt").write_text(toolchain())
    for name, p, rg, iid, h, params, opts in cases():
        np.savez_compressed(OUT / f"{name}.npz", **pack(fit(ref, name, p, rg, iid, h, params, opts), p))
        print("wrote", name)


if __name__ == "__main__":
    main()
