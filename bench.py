#!/usr/bin/env python3
"""Benchmark of the MI355X histogram-GBDT path (one GPU).

Workload (default, "plain"): the 1,000,000-row x 128-feature panel (2,500 dates x 400 entities,
AR(1) features with phi ~ U(0.5, 0.99), 13 Student-t(3) features, 1% NaN, linear + depth-3 tree
target demeaned within date); MSE, depth 8, lr 0.05, max_bin 255, GOSS off, max_leaves 31.
A step is one boosting iteration (one tree).  `--warmup W` trees are grown untimed, then the next
`--steps K` trees are timed from the per-tree wall clock of the native loop (every tree ends with
a device synchronization).  Inputs are resident in HBM before the timed trees start.

Roofline: the dominant kernel by device time is the child histogram build (k_hist_spec, one launch
per multi-leaf step, building the smaller child of each of the step's splits).  A child of m rows
costs m*F bin bytes + m*8 packed-gradient bytes + m*4 row-index bytes read and F*n_bins*24
histogram bytes written (the algorithmic bytes, with the reference's 24-byte HBin), summed over
the splits the finished trees keep; HIP events around every launch on the session stream give
the average duration.  The root build (k_hist_root, m = n, no row index) is reported beside it.
`traffic` comes from the FETCH_SIZE / WRITE_SIZE counter runs summarized in
profiles/hist_traffic.json when present.

The root build is timed twice: as the fit runs (its 128 MB of bins may then sit in the 256 MB
Infinity Cache in front of HBM) and with a 512 MiB buffer read before every root build (profile
mode 2), so that its inputs come from HBM; roofline_root reports the latter.  The sweep only
reads: a sweep that rewrote the buffer would leave dirty lines whose write-back the root build
would pay for (a plain 128 MiB read: 21.2 us warm, 24.9 us after the read sweep, 45.0 us after a
rewriting sweep).

CPU baseline: the reference core (oracle/_ref, built from the reference's own C++ sources) on
the host's cores, same panel, timed as (fit of 1+C trees) - (fit of 1 tree) over C trees, after
a warm-up fit, median of 3 runs with all the threads OMP_NUM_THREADS allows, and one run on one
thread.  The ranking and the small panel carry their own (single-run) reference figures.

Usage: python bench.py [--gpus 1] [--steps 190] [--warmup 10] [--cpu-trees 6]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

METRIC = "trees/sec + hist HBM GB/s at 1M×128; 64-fold batch speedup; predict p50/p99"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _ref_fit_seconds(ref, p, params, n_trees, threads, fit_kw):
    m = ref.model(**{**params, "n_estimators": n_trees, "n_jobs": threads})
    t0 = time.perf_counter()
    m.fit(p.X, p.y, **fit_kw)
    return time.perf_counter() - t0


def ref_trees_per_s(ref, p, params, cpu_trees, threads, reps, fit_kw, tag):
    """Reference core trees/s: (fit(1 + C) - fit(1)) / C, the median over `reps` runs."""
    per = []
    for _ in range(reps):
        t1 = _ref_fit_seconds(ref, p, params, 1, threads, fit_kw)
        tn = _ref_fit_seconds(ref, p, params, 1 + cpu_trees, threads, fit_kw)
        per.append((tn - t1) / cpu_trees)
        log(f"[cpu] {tag}: fit(1) {t1:.2f}s fit({1 + cpu_trees}) {tn:.2f}s on {threads} threads")
    per_tree = float(sorted(per)[len(per) // 2])
    return 1.0 / per_tree, per


def cpu_baseline(p, params, cpu_trees):
    """Reference core on the host's cores (and on one thread); returns a cpu_baseline dict."""
    try:
        import checkers
    except Exception:
        return None
    ref = checkers.reference()
    if ref is None:
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    fit_kw = dict(ts=p.ts, entity=p.entity, horizon=1)
    _ref_fit_seconds(ref, p, params, 1, threads, fit_kw)   # warm-up (page cache, thread pool)
    v, runs = ref_trees_per_s(ref, p, params, cpu_trees, threads, 3, fit_kw, "reference 1M x 128")
    v1, _ = ref_trees_per_s(ref, p, params, 1, 1, 1, fit_kw, "reference 1M x 128, one thread")
    return {"value": round(v, 4), "unit": "trees/s", "cores": threads, "kind": "reference",
            "sample": f"reference core, panel 1,000,000 x 128, {cpu_trees} trees after the first "
                      f"(fit({1 + cpu_trees}) - fit(1)), {threads} OpenMP threads, median of 3 after a warm-up fit",
            "runs_s_per_tree": [round(x, 4) for x in runs], "one_thread": round(v1, 4),
            "nproc": os.cpu_count(), "cpu_model": _cpu_model()}


def _traffic():
    """Per-launch HBM bytes of the histogram kernels from the committed counter summary."""
    f = ROOT / "profiles" / "hist_traffic.json"
    try:
        d = json.loads(f.read_text())
        return {k: v["bytes_per_launch"] for k, v in d["kernels"].items()}
    except Exception:
        return {}


def run_plain(args, p):
    import tidegbm_amd as t

    W, K = args.warmup, args.steps
    params = dict(max_depth=8, learning_rate=0.05, max_bin=255, goss=0)
    b = t.Booster(n_estimators=W + K, **params)
    log(f"[bench] fit {W}+{K} trees")
    t0 = time.perf_counter()
    b.fit(p.X, p.y, timestamps=p.ts, entity=p.entity, horizon=1)
    wall = time.perf_counter() - t0
    tt = b.tree_times()
    t_start = tt[W - 1] if W > 0 else 0.0
    timed = float(tt[W + K - 1] - t_start)
    trees_per_s = K / timed
    setup_s, boost_s = b.fit_times()
    # Kernel durations: a second fit of the same panel with HIP events around every histogram
    # launch on the session stream (eager launches; the timed fit above replays graphs).
    n_prof = max(4, min(W + K, args.profile_trees))
    bp = t.Booster(n_estimators=n_prof, profile=1, **params)
    log(f"[bench] profiled fit {n_prof} trees")
    bp.fit(p.X, p.y, timestamps=p.ts, entity=p.entity, horizon=1)
    for k_a, k_b in zip(b.trees()[:n_prof], bp.trees()):
        if not (k_a["ints"] == k_b["ints"]).all():
            raise SystemExit("profiled fit diverged from the timed fit")
    ms0, n0, by0 = bp.profile(0)   # root histogram builds
    ms5, n5, by5 = bp.profile(5)   # child histogram builds
    ms1, n1, _ = bp.profile(1)     # partitions (with the split selection)
    # the root builds again with the Infinity Cache flushed before each (profile mode 2)
    bf = t.Booster(n_estimators=n_prof, profile=2, **params)
    log(f"[bench] profiled fit {n_prof} trees, caches flushed before each root build")
    bf.fit(p.X, p.y, timestamps=p.ts, entity=p.entity, horizon=1)
    msf, nf, byf = bf.profile(0)
    root_bytes = byf / max(nf, 1)
    # The first root build of a fit also records the per-bin row counts of the panel, which every
    # later root build reuses: the averages cover the later builds, the first is reported apart.
    logf, _ = bf.profile_log(0)
    log0, _ = bp.profile_log(0)
    root_ms = float(logf[1:].mean())
    warm_root_ms = float(log0[1:].mean())
    root_gbs = root_bytes / (root_ms * 1e-3) / 1e9
    child_ms = ms5 / max(n5, 1)
    child_bytes = by5 / max(n5, 1)
    achieved = child_bytes / (child_ms * 1e-3) / 1e9
    tree_gbs = (by0 + by5) / ((ms0 + ms5) * 1e-3) / 1e9
    traffic = _traffic()
    out = {
        "metric": METRIC,
        "value": round(trees_per_s, 3),
        "unit": "trees/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * timed / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (int64 fixed-point histogram sums)",
        "data": "synthetic",
        "config": {
            "workload": "panel_large: 2,500 dates x 400 entities = 1,000,000 rows x 128 float32 features, "
                        "mse, depth 8, max_leaves 31, lr 0.05, max_bin 255, GOSS off, horizon 1",
            "rows": p.n, "features": p.n_features, "trees_timed": K,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic.get("k_hist_spec"),
            "kernel": "child histogram build (k_hist_spec), the smaller children of a multi-leaf step, 128 features",
            "bytes_per_launch": round(child_bytes),
            "avg_launch_ms": round(child_ms, 5),
            "launches": n5,
        },
        "roofline_root": {
            "bound": "hbm", "achieved": round(root_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(root_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic.get("k_hist_root_flushed", traffic.get("k_hist_root")),
            "kernel": "root histogram build (k_hist_root), 1,000,000 x 128, Infinity Cache flushed before each; "
                      "builds after the first of the fit (the first records the panel's per-bin counts)",
            "bytes_per_launch": root_bytes, "avg_launch_ms": round(root_ms, 5), "launches": nf - 1,
            "first_launch_ms": round(float(logf[0]), 5),
            "avg_launch_ms_unflushed": round(warm_root_ms, 5),
            "frac_unflushed": round(root_bytes / (warm_root_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "hist_tree_gbs": round(tree_gbs, 1),
        "device_ms_per_tree": {"root_hist": round(ms0 / max(n0, 1), 4),
                               "child_hist": round(ms5 / max(n0, 1), 4),
                               "partition_select": round(ms1 / max(n0, 1), 4)},
        "fit_setup_s": round(setup_s, 3),
        "fit_wall_s": round(wall, 3),
        "trees_per_s_end_to_end": round((W + K) / (setup_s + boost_s), 3),
        "hist_variant": b.hist_variant(),
    }
    if not args.no_cpu:
        cb = cpu_baseline(p, params, args.cpu_trees)
        if cb is not None:
            out["cpu_baseline"] = cb
    return out


def copy_probe(mib):
    """A plain device copy of known size: the reference point that scales rocprofv3's FETCH_SIZE
    and WRITE_SIZE counters to bytes (each dispatch reads and writes exactly mib MiB).  The HIP
    runtime's device-to-device hipMemcpy runs as its own copy kernel (__amd_rocclr_copyBuffer)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), nbytes) != 0:
            raise SystemExit("copy probe: hipMalloc failed")
    if hip.hipMemset(a, 1, nbytes) != 0:
        raise SystemExit("copy probe: hipMemset failed")
    for _ in range(5):
        if hip.hipMemcpy(b, a, nbytes, 3) != 0:   # hipMemcpyDeviceToDevice
            raise SystemExit("copy probe: hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_mib": mib, "dispatches": 5}), flush=True)


def run_walkforward(args, p):
    """64 walk-forward fits (16 expanding folds x horizons 1, 3, 5, 10) over the shared binned
    1M x 128 matrix: batched (every kernel launch covers all fits) against the same fits run one
    after another over the same resident bins.  Both schedules must give identical trees."""
    import numpy as np
    import tidegbm_amd as t
    from tidegbm_amd import synth

    tasks = synth.walk_forward_tasks(p, n_folds=16, horizons=(1, 3, 5, 10), first_dates=250)
    params = dict(n_estimators=args.wf_trees, max_depth=6, learning_rate=0.05, max_bin=255, goss=0,
                  pair_splits=0)
    kw = dict(entity=p.entity, horizons=[k.horizon for k in tasks], **params)
    rows = [k.n_train for k in tasks]
    ys = [k.y for k in tasks]
    log(f"[walk-forward] {len(tasks)} fits, rows {min(rows)}..{max(rows)}, {args.wf_trees} trees each: batched")
    mb = t.fit_batch(p.X, p.ts, rows, ys, schedule="batched", **kw)
    sb, bb = mb[0].fit_times()
    if args.wf_batched_only:   # one schedule per process (kernel-trace profiles of the batched launches)
        return {"batched_s": round(bb, 3), "setup_s": round(sb, 3), "fits": len(tasks)}
    log(f"[walk-forward] batched boost {bb:.3f}s (setup {sb:.3f}s); sequential")
    ms = t.fit_batch(p.X, p.ts, rows, ys, schedule="sequential", **kw)
    ss, bs = ms[0].fit_times()
    log(f"[walk-forward] sequential boost {bs:.3f}s")
    same = True
    for a, b in zip(mb, ms):
        for x, y in zip(a.trees(), b.trees()):
            same &= all(np.array_equal(x[k], y[k]) for k in ("ints", "dbl", "values", "prefix"))
    if not same:
        raise SystemExit("batched and sequential walk-forward fits differ")
    n_trees = len(tasks) * args.wf_trees
    return {
        "workload": "64 walk-forward fits of the 1M x 128 panel: 16 expanding folds (train 250..2500 "
                    "dates minus the purge gap) x horizons 1,3,5,10; mse, depth 6, max_leaves 31, lr 0.05, "
                    f"{args.wf_trees} trees each, goss=0, pair_splits=0, shared bins",
        "fits": len(tasks), "rows_min": min(rows), "rows_max": max(rows), "rows_total": int(sum(rows)),
        "batched_s": round(bb, 3), "sequential_s": round(bs, 3),
        "speedup": round(bs / bb, 3),
        "batched_trees_per_s": round(n_trees / bb, 1), "sequential_trees_per_s": round(n_trees / bs, 1),
        "setup_s": round(sb, 3), "identical_trees": bool(same),
        "hist_variant_batched": mb[0].hist_variant(),
    }


def _serve_stream(p, seconds, rate, n_models, seed):
    """Seeded request stream: Poisson arrivals; 90% predicts of log-uniform 1..65,536 panel rows
    (a contiguous block at a random start), 10% refits of 20k..200k rows x 50 trees."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        m = int(rng.integers(n_models))
        if rng.random() < 0.9:
            n = int(min(65536, max(1, np.floor(np.exp(rng.uniform(0.0, np.log(65536.0)))))))
            out.append((t, "predict", m, int(rng.integers(0, p.n - n + 1)), n))
        else:
            n = int(rng.integers(20000, 200001))
            out.append((t, "refit", m, int(rng.integers(0, p.n - n + 1)), n))
    return out


def _serve_run(t, p, models, stream, batched, slots, burst=0):
    """Serves the stream on time and reports latencies.  With burst > 0 it then submits that many
    refits of the stream's size distribution at once and reports the refits completed per second
    while they drain (the refit capacity, above any offered rate)."""
    import gc
    import numpy as np
    srv = t.Server(models, p.X, predict_budget=2 << 30, refit_slots=slots, refit_max_rows=200000,
                   batched=batched)
    tickets = []
    # the stream is submitted on time: no Python garbage-collection pause (after earlier workloads
    # the heap is large) may hold requests back and release them in a burst
    gc.collect()
    gc.disable()
    t0 = time.perf_counter()
    for (at, kind, m, r0, n) in stream:
        while True:
            d = t0 + at - time.perf_counter()
            if d <= 0:
                break
            time.sleep(min(d, 0.0005))
        if kind == "predict":
            tickets.append((kind, n, srv.predict(m, np.arange(r0, r0 + n, dtype=np.int32))))
        else:
            tickets.append((kind, n, srv.refit(m, r0, p.y[r0:r0 + n], extra_trees=50)))
    t_sub = time.perf_counter() - t0
    lat = {"predict": [], "refit": []}
    small = []
    rows = 0
    for kind, n, tk in tickets:
        _, _, l = srv.wait(tk)
        lat[kind].append(l)
        if kind == "predict":
            rows += n
            if n <= 4096:
                small.append(l)
    wall = time.perf_counter() - t0
    st = srv.stats()
    cap = None
    if burst:
        rng = np.random.default_rng(12345)
        reqs = []
        for _ in range(burst):
            n = int(rng.integers(20000, 200001))
            reqs.append((int(rng.integers(len(models))), int(rng.integers(0, p.n - n + 1)), n))
        tb = time.perf_counter()
        tks = [srv.refit(m, r0, p.y[r0:r0 + n], extra_trees=50) for (m, r0, n) in reqs]
        for tk in tks:
            srv.wait(tk)
        cap = round(burst / (time.perf_counter() - tb), 1)
    gc.enable()
    srv.close()
    pct = lambda a, q: round(1e3 * float(np.percentile(a, q)), 3) if len(a) else None
    return {
        "requests": len(stream), "predicts": len(lat["predict"]), "refits": len(lat["refit"]),
        "submit_s": round(t_sub, 3), "drain_s": round(wall, 3),
        "predict_p50_ms": pct(lat["predict"], 50), "predict_p99_ms": pct(lat["predict"], 99),
        "predict_le4096_p50_ms": pct(small, 50), "predict_le4096_p99_ms": pct(small, 99),
        "refit_p50_ms": pct(lat["refit"], 50), "refit_p99_ms": pct(lat["refit"], 99),
        "predict_rows_per_s": round(rows / wall, 1), "predict_launches": st["predict_launches"],
        "refit_rounds": st["refit_rounds"], "refit_admit_s": round(st["refit_admit_s"], 3),
        "refit_round_s": round(st["refit_round_s"], 3), "refit_publish_s": round(st["refit_publish_s"], 3),
        "refit_burst": burst or None, "refit_capacity_per_s": cap,
    }


def run_serve(args, p):
    """Serving loop: 8 resident ensembles (500 trees, depth 8, max_leaves 31, 128 features) over the
    resident 1M x 128 binned panel; a seeded Poisson stream at 2,000 requests/s, served batched
    (coalesced predict launches, refits grown together in slots) and unbatched (one request at a
    time, arrival order)."""
    import tidegbm_amd as t
    from tidegbm_amd import synth

    tasks = synth.walk_forward_tasks(p, n_folds=2, horizons=(1, 3, 5, 10), first_dates=2200)
    params = dict(n_estimators=args.serve_trees, max_depth=8, learning_rate=0.05, max_bin=255, goss=0,
                  pair_splits=0)
    log(f"[serve] training {len(tasks)} ensembles x {args.serve_trees} trees (batched fit)")
    t_fit = time.perf_counter()
    models = t.fit_batch(p.X, p.ts, [k.n_train for k in tasks], [k.y for k in tasks], entity=p.entity,
                         horizons=[k.horizon for k in tasks], **params)
    t_fit = time.perf_counter() - t_fit
    stream = _serve_stream(p, args.serve_seconds, 2000.0, len(models), seed=7)
    log(f"[serve] {len(stream)} requests over {args.serve_seconds}s: batched")
    b = _serve_run(t, p, models, stream, True, args.serve_slots, burst=4 * args.serve_slots)
    log(f"[serve] batched: {b}")
    u = None
    if not args.serve_batched_only:
        log("[serve] unbatched")
        u = _serve_run(t, p, models, stream, False, 1)
        log(f"[serve] unbatched: {u}")
    return {
        "workload": f"8 resident ensembles ({args.serve_trees} trees, depth 8, max_leaves 31, 128 features) over "
                    "the resident 1M x 128 binned panel; Poisson arrivals at 2,000 req/s for "
                    f"{args.serve_seconds}s; 90% predicts of log-uniform 1..65,536 panel rows, 10% refits of "
                    "20k..200k rows x 50 trees; predict staging budget 2 GiB, "
                    f"{args.serve_slots} refit slots of up to 200,000 rows; refit capacity: a burst of "
                    f"{4 * args.serve_slots} such refits submitted at once after the stream (batched)",
        "ensembles_fit_s": round(t_fit, 3), "batched": b, "unbatched": u,
    }


def run_ranking(args):
    """Pairwise ranking fit at full size: 5,000 dates with U(200, 1000) entities (~3M rows) x 64
    AR(1) features, within-date pairwise objective, 100 trees, depth 6, lr 0.08, reference
    defaults (GOSS auto-on at >= 250k rows; PAIR splits auto-off at >= 1M rows)."""
    import tidegbm_amd as t
    from tidegbm_amd import synth

    log("[ranking] generating the ranking panel")
    p = synth.panel_ranking(2)
    b = t.Booster(n_estimators=args.rank_trees, max_depth=6, learning_rate=0.08, objective=5, profile=0)
    t0 = time.perf_counter()
    b.fit(p.X, p.y, timestamps=p.ts, entity=p.entity, rank_groups=p.ts, horizon=1)
    wall = time.perf_counter() - t0
    setup_s, boost_s = b.fit_times()
    tt = b.tree_times()
    host_trees, wait_s, sample_s = b.goss_stats()
    cpu = None
    if not args.no_cpu:
        try:
            import checkers
            ref = checkers.reference()
        except Exception:
            ref = None
        if ref is not None:
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            rp = dict(max_depth=6, learning_rate=0.08, objective=5)
            v, _ = ref_trees_per_s(ref, p, rp, 2, threads, 1,
                                   dict(ts=p.ts, entity=p.entity, rank_groups=p.ts, horizon=1), "reference ranking")
            cpu = {"value": round(v, 4), "unit": "trees/s", "cores": threads, "kind": "reference",
                   "sample": "reference core, ranking panel, fit(3) - fit(1), one run (memory-mapped row-major "
                             "histogram path at this size)"}
    return {
        "workload": f"ranking panel: 5,000 dates x U(200,1000) entities = {p.n:,} rows x 64 features, "
                    f"pairwise within dates, {args.rank_trees} trees, depth 6, lr 0.08, GOSS auto (on), "
                    "PAIR auto (off at >= 1M rows)",
        "rows": p.n, "trees": args.rank_trees, "fit_wall_s": round(wall, 3), "setup_s": round(setup_s, 3),
        "boost_s": round(boost_s, 3), "trees_per_s": round(args.rank_trees / boost_s, 2),
        "ms_per_tree_steady": round(1e3 * float(tt[-1] - tt[min(4, len(tt) - 1)]) / max(1, len(tt) - 5), 3),
        "goss_host_trees": host_trees, "goss_wait_ms_per_tree": round(1e3 * wait_s / args.rank_trees, 3),
        "goss_sample_ms_per_tree": round(1e3 * sample_s / args.rank_trees, 3),
        "cpu_baseline": cpu,
    }


def run_small(args):
    """The reference's CPU-runnable case: 500 dates x 30 entities = 15,000 rows x 20 AR(1) features
    (phi 0.9), mse, 80 trees, depth 6, lr 0.08, max_bin 255, GOSS off, horizon 1, PAIR splits on
    (the reference default below 1M rows: growth is driven from the host, with device histograms
    and scans)."""
    import tidegbm_amd as t
    from tidegbm_amd import synth

    p = synth.panel_small(0)
    params = dict(max_depth=6, learning_rate=0.08, max_bin=255, goss=0)
    b = t.Booster(n_estimators=80, **params)
    t0 = time.perf_counter()
    b.fit(p.X, p.y, timestamps=p.ts, entity=p.entity, horizon=1)
    wall = time.perf_counter() - t0
    setup_s, boost_s = b.fit_times()
    n_pair = sum(int((tr["ints"][:, 1] == 1).sum()) for tr in b.trees())
    cpu = None
    if not args.no_cpu:
        try:
            import checkers
            ref = checkers.reference()
        except Exception:
            ref = None
        if ref is not None:
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            v, _ = ref_trees_per_s(ref, p, params, 79, threads, 3, dict(ts=p.ts, entity=p.entity, horizon=1),
                                   "reference small panel")
            cpu = {"value": round(v, 2), "unit": "trees/s", "cores": threads, "kind": "reference",
                   "sample": "reference core, 15,000 x 20 panel, fit(80) - fit(1), median of 3"}
    return {"workload": "small panel: 500 dates x 30 entities = 15,000 rows x 20 features, mse, 80 trees, "
                        "depth 6, lr 0.08, GOSS off, PAIR splits on (host-driven growth)",
            "trees_per_s": round(80 / boost_s, 1), "fit_wall_s": round(wall, 4), "setup_s": round(setup_s, 4),
            "pair_splits": n_pair, "cpu_baseline": cpu}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=190)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-trees", type=int, default=6)
    ap.add_argument("--profile-trees", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["all", "plain", "walkforward", "serve", "ranking", "small"],
                    default="all")
    ap.add_argument("--rank-trees", type=int, default=100, help="trees of the ranking workload")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="length of the serving request stream")
    ap.add_argument("--serve-trees", type=int, default=500, help="trees per served ensemble")
    ap.add_argument("--serve-slots", type=int, default=64, help="concurrent refits when batched")
    ap.add_argument("--wf-trees", type=int, default=100, help="trees per walk-forward fit")
    ap.add_argument("--serve-batched-only", action="store_true",
                    help="serving workload: run the batched server only (no unbatched baseline)")
    ap.add_argument("--wf-batched-only", action="store_true",
                    help="walk-forward workload: run the batched schedule only (no sequential fits, no speedup)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): this library runs a fit on one "
              "GPU; sharding rows across GPUs (one histogram all-reduce per built node) is not built.",
              file=sys.stderr)
        raise SystemExit(2)
    if args.warmup < 1:
        args.warmup = 1
    import tidegbm_amd as t
    from tidegbm_amd import synth
    if t.device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    if args.workload == "ranking":
        print(json.dumps({"ranking": run_ranking(args)}), flush=True)
        return
    if args.workload == "small":
        print(json.dumps({"small": run_small(args)}), flush=True)
        return
    log("[bench] generating the 1M x 128 panel")
    p = synth.panel_large(1)
    if args.workload == "walkforward":
        print(json.dumps({"walk_forward": run_walkforward(args, p)}), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps({"serve": run_serve(args, p)}), flush=True)
        return
    out = run_plain(args, p)
    if args.workload == "all":
        out["walk_forward"] = run_walkforward(args, p)
        out["serve"] = run_serve(args, p)
        del p
        out["ranking"] = run_ranking(args)
        out["small"] = run_small(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
