"""Whole-fit parity of the GPU path against the CPU oracle and the reference core.

Bar (stated per test): tree structure (every integer node field, leaf cover) bit-exact; leaf and
TORB prefix values within rtol 1e-7 (they come from exact fixed-point sums instead of sequential
FP64 folds); predictions within rtol 1e-5; the per-tree training loss within rtol 1e-10.
"""
import dataclasses
import sys

import numpy as np
import pytest

import checkers
from tidegbm_amd import Booster, synth

pytestmark = pytest.mark.gpu

VALUE_RTOL = 1e-7
PRED_RTOL = 1e-5
LOSS_RTOL = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fit_gpu(p, rg=None, iid=False, horizon=1, w=None, **params):
    b = Booster(**params)
    b.fit(p.X, p.y, timestamps=None if iid else p.ts, weight=w, entity=None if iid else p.entity,
          rank_groups=rg, iid=iid, horizon=horizon)
    return b


def _fit_cpu(lib, p, rg=None, iid=False, horizon=1, w=None, **params):
    m = lib.model(**params)
    m.fit(p.X, p.y, ts=None if iid else p.ts, w=w, entity=None if iid else p.entity, rank_groups=rg,
          iid=iid, horizon=horizon)
    return m


def _pred_close(a, b):
    scale = np.maximum(np.abs(b), 1e-3 * np.max(np.abs(b)) + 1e-12)
    return float(np.max(np.abs(a - b) / scale))


def _check(name, p, rg=None, iid=False, n_pred=2000, w=None, horizon=1, **params):
    _say(f"[case] {name} n={p.n} F={p.n_features} params={params}")
    gpu = _fit_gpu(p, rg=rg, iid=iid, w=w, horizon=horizon, **params)
    orc = _fit_cpu(checkers.oracle(), p, rg=rg, iid=iid, w=w, horizon=horizon, **params)
    # training-row bins: lets the comparison accept the one characterized flag difference
    xb = checkers.bins_of(p.X, [gpu.edges(f) for f in range(p.n_features)])
    nb = gpu.model_info()["n_bins"]
    diff = checkers.compare_trees(gpu.trees(), orc.trees(), rtol=VALUE_RTOL, atol=1e-9, xb=xb, n_bins=nb)
    assert diff is None, f"{name}: GPU vs oracle: {diff}"
    Xp = p.X[:n_pred]
    rel = _pred_close(gpu.predict(Xp), orc.predict(Xp))
    assert rel < PRED_RTOL, f"{name}: prediction rel err {rel}"
    # evals.train: the device reduces the weighted squared errors in a fixed parallel order (the
    # reference sums sequentially) and the pairwise groups with the reference's sequential sums
    ev, eo = gpu.evals_result()["train"], orc.evals_train()
    assert len(ev) == len(eo) == params["n_estimators"], (len(ev), len(eo))
    assert np.allclose(ev, eo, rtol=LOSS_RTOL, atol=0), f"{name}: training loss {ev} vs {eo}"
    ref = checkers.reference()
    if ref is not None:
        r = _fit_cpu(ref, p, rg=rg, iid=iid, w=w, horizon=horizon, **params)
        diff = checkers.compare_trees(gpu.trees(), r.trees(), rtol=VALUE_RTOL, atol=1e-9, xb=xb, n_bins=nb)
        assert diff is None, f"{name}: GPU vs reference: {diff}"
    return gpu, orc


def test_small_panel_mse_default():
    # the reference default path: PAIR search on (below 1M rows), host-driven growth
    p = synth.panel_small(0)
    _check("small-mse-default", p, n_estimators=40, max_depth=6, learning_rate=0.08, goss=0)


@pytest.mark.parametrize("seed", [0, 11])
def test_small_panel_pair_splits(seed):
    p = synth.panel_small(seed)
    rng = np.random.default_rng(seed)
    y = np.sign(p.X[:, 0] * p.X[:, 1]) + 0.1 * rng.normal(size=p.n)
    q = synth.Panel(p.X, y, p.ts, p.entity)
    gpu, _ = _check(f"small-pair-{seed}", q, n_estimators=20, max_depth=6, learning_rate=0.08, goss=0)
    assert sum(int((t["ints"][:, 1] == 1).sum()) for t in gpu.trees()) >= 5


def test_small_panel_mse_device_driven():
    p = synth.panel_small(1)
    _check("small-mse-device", p, n_estimators=25, max_depth=6, learning_rate=0.08, goss=0, pair_splits=0)


def test_small_panel_iid_mode():
    p = synth.panel_small(2)
    _check("small-iid", p, iid=True, n_estimators=15, max_depth=5, learning_rate=0.1, goss=0, pair_splits=0)


def test_small_panel_pairwise():
    p = synth.panel_small(3, n_dates=200, n_entities=60)
    _check("small-pairwise", p, rg=p.ts, n_estimators=10, max_depth=6, learning_rate=0.08, goss=0,
           pair_splits=0, objective=5)


def test_ranking_panel_pairwise_goss_auto():
    # The ranking configuration's shape at a tenth of its dates: uneven dates (200..1000 rows),
    # 64 features, within-date pairwise objective, reference defaults (GOSS auto-on at >= 250k
    # rows, PAIR splits allowed below 1M rows), checked against the oracle and the reference core.
    p = synth.panel_ranking(3, n_dates=500)
    assert p.n >= 250000
    gpu, _ = _check("ranking-pairwise-goss-auto", p, rg=p.ts, n_pred=20000, n_estimators=5, max_depth=6,
                    learning_rate=0.08, objective=5)
    _say(f"  GOSS trees sampled on the host: {gpu.goss_stats()[0]} of 5")


def test_small_panel_goss():
    # distinct |g| values: the sample is taken on the device
    p = synth.panel_small(4)
    gpu, _ = _check("small-goss", p, n_estimators=10, max_depth=5, learning_rate=0.1, goss=1, pair_splits=0)
    host_trees, _, _ = gpu.goss_stats()
    _say(f"  GOSS trees sampled on the host: {host_trees} of 10")
    assert host_trees < 10, "no tree took the device GOSS sample"


def test_small_panel_goss_equal_gradients():
    # targets on a 0.5 grid: rows of one leaf share |g|, the order of equal |g| decides the draw,
    # and those trees are sampled on the host with std::sort as the reference samples them
    p = synth.panel_small(6)
    p = dataclasses.replace(p, y=np.round(p.y * 2.0) / 2.0)
    gpu, _ = _check("small-goss-ties", p, n_estimators=8, max_depth=5, learning_rate=0.1, goss=1,
                    pair_splits=0)
    host_trees, _, _ = gpu.goss_stats()
    _say(f"  GOSS trees sampled on the host: {host_trees} of 8")
    assert host_trees > 0, "equal |g| values did not reach the host GOSS sample"


def test_larger_panel_depth8_missing_values():
    p = synth.panel_large(5, n_dates=250, n_entities=200, n_features=48)
    _check("large-subset-mse", p, n_estimators=8, max_depth=8, learning_rate=0.05, goss=0, pair_splits=0)


def test_full_size_panel_matches_reference_core():
    # the headline configuration at full size (1,000,000 x 128, depth 8, 31 leaves, lr 0.05, GOSS
    # off, PAIR auto-off at >= 1M rows): the device trees against the reference core itself
    ref = checkers.reference()
    if ref is None:
        pytest.skip("reference core (oracle/_ref) not built")
    p = synth.panel_large(1)
    params = dict(n_estimators=5, max_depth=8, learning_rate=0.05, goss=0)
    _say(f"[case] full-size n={p.n} F={p.n_features} params={params}")
    gpu = _fit_gpu(p, **params)
    r = _fit_cpu(ref, p, n_jobs=16, **params)
    diff = checkers.compare_trees(gpu.trees(), r.trees(), rtol=VALUE_RTOL, atol=1e-9)
    if diff is not None:   # let the checker verify a characterized difference on the rows
        xb = checkers.bins_of(p.X, [gpu.edges(f) for f in range(p.n_features)])
        diff = checkers.compare_trees(gpu.trees(), r.trees(), rtol=VALUE_RTOL, atol=1e-9, xb=xb,
                                      n_bins=gpu.model_info()["n_bins"])
    assert diff is None, diff
    for f in range(0, p.n_features, 17):
        assert np.array_equal(gpu.edges(f), r.edges(f))
    Xp = p.X[:20000]
    rel = _pred_close(gpu.predict(Xp), r.predict(Xp))
    assert rel < PRED_RTOL, rel
    assert np.allclose(gpu.evals_result()["train"], r.evals_train(), rtol=LOSS_RTOL, atol=0)


def test_wide_panel_300_features():
    # more features than a 256-thread selection block and than 16-feature tiles divide evenly
    p = synth.panel_small(13, n_dates=300, n_entities=40, n_features=300)
    _check("wide-300", p, n_estimators=5, max_depth=6, learning_rate=0.1, goss=0, pair_splits=0)


@pytest.mark.parametrize("n_dates,n_entities", [(20, 5), (7, 3)])
def test_tiny_panels(n_dates, n_entities):
    # 100 and 21 rows: nodes below min_samples_leaf, tails inside the first tile of every kernel
    p = synth.panel_small(14, n_dates=n_dates, n_entities=n_entities, n_features=5)
    _check(f"tiny-{p.n}", p, n_estimators=6, max_depth=4, learning_rate=0.1, goss=0, pair_splits=0,
           min_samples_leaf=2)


# Growth and leaf parameters over their ranges, one small panel each (seeded), against the oracle
# and the reference core: depth / leaf budget, minimum leaf rows and weight, L1 / L2, minimum gain,
# delta clipping, learning rate, coarse bins, time buckets, embargo, horizon, iid mode, weights.
SWEEP = [
    ("leaves-63-depth-9", dict(max_depth=9, max_leaves=63, min_samples_leaf=5), {}),
    ("depth-3", dict(max_depth=3, min_samples_leaf=50), {}),
    ("l1-l2", dict(l1=0.5, l2=5.0), {}),
    ("min-gain-and-weight", dict(min_split_gain=0.05, min_child_weight=8.0), {}),
    ("delta-clip", dict(max_delta_step=0.05, learning_rate=0.3), {}),
    ("coarse-bins", dict(max_bin=15), {}),
    ("bins-63-buckets-8", dict(max_bin=63, n_time_buckets=8), {}),
    ("embargo-horizon-5", dict(embargo=2), dict(horizon=5)),
    ("no-finer-torb", dict(finer_torb=0, torb_min_cover=100.0), {}),
    ("iid-weighted", dict(), dict(iid=True, weighted=True)),
    ("time-weighted", dict(min_child_weight=2.0), dict(weighted=True)),
    ("pairwise-l1", dict(objective=5, l1=0.1, max_depth=5), dict(groups=True)),
    # host-driven PAIR search with its knobs moved: every leaf counts as weak
    ("pair-knobs", dict(pair_splits=1, pair_trigger=1e6, pair_top_k=4, max_pair_splits_per_tree=5, pair_l2=0.5), {}),
    ("goss-knobs", dict(goss=1, goss_top=0.3, goss_other=0.2), {}),
    ("bin-sample-seed", dict(bin_sample=300, random_state=7), {}),
]


@pytest.mark.parametrize("name,params,opts", SWEEP, ids=[c[0] for c in SWEEP])
def test_parameter_sweep_matches_oracle(name, params, opts):
    seed = 100 + [c[0] for c in SWEEP].index(name)
    p = synth.panel_small(seed, n_dates=200, n_entities=40, n_features=12)
    w = np.random.default_rng(seed).uniform(0.2, 3.0, p.n) if opts.get("weighted") else None
    base = dict(n_estimators=8, max_depth=6, learning_rate=0.1, goss=0, pair_splits=0)
    base.update(params)
    _check(name, p, rg=p.ts if opts.get("groups") else None, iid=opts.get("iid", False), w=w,
           horizon=opts.get("horizon", 1), **base)


def test_refit_update_matches_oracle():
    p = synth.panel_small(6)
    _say("[case] update")
    params = dict(n_estimators=6, max_depth=5, learning_rate=0.1, goss=0, pair_splits=0)
    gpu = _fit_gpu(p, **params)
    orc = _fit_cpu(checkers.oracle(), p, **params)
    q = synth.panel_small(7, n_dates=100)
    gpu.update(q.X, q.y, 5)
    orc.update(q.X, q.y, 5)
    diff = checkers.compare_trees(gpu.trees(), orc.trees(), rtol=VALUE_RTOL, atol=1e-9)
    assert diff is None, diff
    assert _pred_close(gpu.predict(q.X), orc.predict(q.X)) < PRED_RTOL


def test_fit_is_deterministic():
    p = synth.panel_small(8)
    _say("[case] determinism")
    params = dict(n_estimators=10, max_depth=6, learning_rate=0.08, goss=0)
    a = _fit_gpu(p, **params)
    b = _fit_gpu(p, **params)
    for x, y in zip(a.trees(), b.trees()):
        for k in ("ints", "dbl", "values", "prefix", "cover"):
            assert np.array_equal(x[k], y[k]), k
    assert np.array_equal(a.predict(p.X), b.predict(p.X))


def test_graph_replays_match_eager_launches(monkeypatch):
    # The device-driven split loop is one captured hipGraph replayed once per tree, with eager
    # gradient / root / tree-end work and host synchronizations between replays; every tree of
    # the replayed fit must equal the eagerly launched fit bit for bit.
    p = synth.panel_large(9, n_dates=120, n_entities=150, n_features=24)
    params = dict(n_estimators=12, max_depth=7, learning_rate=0.1, goss=0, pair_splits=0)
    _say("[case] graph vs eager")
    monkeypatch.setenv("TGBM_NO_GRAPH", "1")
    eager = _fit_gpu(p, **params)
    monkeypatch.setenv("TGBM_NO_GRAPH", "0")
    graph = _fit_gpu(p, **params)
    for t, (x, y) in enumerate(zip(eager.trees(), graph.trees())):
        for k in ("ints", "dbl", "values", "prefix", "cover"):
            assert np.array_equal(x[k], y[k]), f"tree {t} {k}"
    assert np.array_equal(eager.predict(p.X), graph.predict(p.X))


def test_missing_values_panel_matches_oracle():
    # A panel with 1% NaN where a split's node receives no row with its feature missing: the
    # comparison accepts only the characterized missing_go_left difference (see checkers).
    p = synth.panel_large(4, n_dates=150, n_entities=120, n_features=20)
    params = dict(n_estimators=10, max_depth=7, learning_rate=0.1, goss=0, pair_splits=0)
    _say("[case] NaN panel, seed 4")
    gpu = _fit_gpu(p, **params)
    orc = _fit_cpu(checkers.oracle(), p, **params)
    xb = checkers.bins_of(p.X, [gpu.edges(f) for f in range(p.n_features)])
    nb = gpu.model_info()["n_bins"]
    diff = checkers.compare_trees(gpu.trees(), orc.trees(), rtol=VALUE_RTOL, atol=1e-9, xb=xb, n_bins=nb)
    assert diff is None, diff
    assert _pred_close(gpu.predict(p.X), orc.predict(p.X)) < PRED_RTOL
