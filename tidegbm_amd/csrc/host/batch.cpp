// Batched walk-forward fits: many independent fits over row prefixes of one shared binned
// matrix, grown together so every fit kernel launch covers all of them.
//
// The reference refits each walk-forward fold x horizon separately (cv.py:45-53 defines the
// expanding folds with a purge gap of `horizon` dates; every fold is a full Booster::fit,
// booster.cpp:231-742).  Here the folds share the bins of the largest training window: rows are
// in (timestamp, entity) order, so each fold trains on a row prefix [0, rows_t) of that matrix,
// with its own target, its own TORB layout (the horizon sets the embargo) and its own trees.
// The data preparation of each fold is exactly Booster::fit's (time_layout); only the bin
// edges are shared.
//
// Schedules (same kernels, same per-fold results):
//   batched    : all folds active; one launch of each kernel per tree round covers every fold
//   sequential : one fold active at a time, all of its trees, then the next fold (the plain
//                one-fit-per-call pattern, over the same resident bins)
#include "booster.h"
#include "tuner.h"

#include <algorithm>
#include <cmath>
#include <exception>
#include <memory>

#include <omp.h>

namespace tide {

namespace {

struct SessionOwner {
    tgbm_session* s = nullptr;
    ~SessionOwner() {
        if (s) tgbm_session_destroy(s);
    }
};

}  // namespace

BatchStats fit_batch(const Config& proto, const double* X, int n, int F, const long long* ts,
                     const long long* entity, int n_tasks, const int* task_rows, const double* const* ys,
                     const long long* horizons, int schedule, std::vector<std::unique_ptr<Booster>>& out) {
    BatchStats st;
    const double t0 = now_s();
    if (n <= 0 || F <= 0) throw Error("empty input");
    if (n_tasks < 1) throw Error("n_tasks must be >= 1");
    if (schedule != 0 && schedule != 1) throw Error("schedule must be 0 (batched) or 1 (sequential)");
    for (int t = 0; t < n_tasks; ++t) {
        Config c = proto;
        if (horizons) c.horizon = horizons[t];
        c.validate();
    }
    if (proto.pair_splits || proto.force_host_driver)
        throw Error("batched fits grow every tree on the device: set pair_splits=0");
    if (!proto.iid_mode && !ts) throw Error("timestamp is required unless iid_mode");
    int n_fit = 0;
    for (int t = 0; t < n_tasks; ++t) {
        if (task_rows[t] < 1 || task_rows[t] > n) throw Error("task rows out of range");
        if (!ys[t]) throw Error("every task needs a target");
        n_fit = std::max(n_fit, task_rows[t]);
        const bool goss = (proto.goss > 0) || (proto.goss < 0 && task_rows[t] >= 250000);
        if (goss && proto.goss_top < 1.0) throw Error("batched fits need goss=0 (GOSS samples on the host per tree)");
    }
    if (ts) {
        for (int i = 1; i < n_fit; ++i) {
            const bool ok = ts[i] > ts[i - 1] || (ts[i] == ts[i - 1] && (!entity || entity[i] >= entity[i - 1]));
            if (!ok) throw Error("batched fits need rows sorted by (timestamp, entity)");
        }
    }
    if (tgbm_device_count() < 1) throw Error("GPU device not available: no HIP device was found for device=gpu");
    // ---- shared edges over the largest training window, bins on the device
    std::vector<std::vector<double>> edges;
    fit_edges(X, n_fit, F, proto.max_bin, proto.bin_sample, (uint32_t)proto.random_state, edges);
    int n_bins = proto.max_bin;
    for (auto& e : edges) n_bins = std::max(n_bins, (int)e.size() + 2);
    const bool hconst = proto.objective == 0;
    SessionOwner so;
    dev_check(tgbm_session_create(&so.s, n_fit, F, n_bins, n_tasks, task_rows, proto.max_leaves, hconst ? 1 : 0));
    if (!proto.monotone.empty()) {
        if ((int)proto.monotone.size() != F) throw Error("monotone constraints need one entry per feature");
        dev_check(tgbm_session_set_monotone(so.s, proto.monotone.data()));
    }
    // multi-leaf steps for every fold, batched or one after another
    dev_check(tgbm_session_set_spec(so.s, proto.spec_width > 1 ? proto.spec_width : 0));
    bin_on_device(so.s, X, edges);
    // ---- per-fold preparation, exactly as a plain fit of that fold; the folds' time layouts are
    // computed in parallel waves of one fold per host thread (the first failing fold's error is
    // reported, as in fold order)
    out.clear();
    std::vector<double> ones(n_fit, 1.0);
    const int wave = std::max(1, omp_get_max_threads());
    std::vector<TimeLayout> layouts;
    std::vector<std::exception_ptr> errs;
    for (int t = 0; t < n_tasks; ++t) {
        const int w0 = t - t % wave;
        if (t == w0) {
            const int nw = std::min(wave, n_tasks - w0);
            layouts.assign(nw, TimeLayout());
            errs.assign(nw, nullptr);
#pragma omp parallel for schedule(dynamic) if (nw > 1)
            for (int q = 0; q < nw; ++q) {
                try {
                    Config c = proto;
                    if (horizons) c.horizon = horizons[w0 + q];
                    layouts[q] = time_layout(c, ts, ys[w0 + q], ones.data(), task_rows[w0 + q]);
                } catch (...) {
                    errs[q] = std::current_exception();
                }
            }
            for (auto& e : errs)
                if (e) std::rethrow_exception(e);
        }
        auto b = std::make_unique<Booster>();
        b->cfg = proto;
        if (horizons) b->cfg.horizon = horizons[t];
        TimeLayout tl = std::move(layouts[t - w0]);
        const long long* groups = proto.objective == 5 ? ts : nullptr;  // rank groups = dates (native.py:134-137)
        if (proto.objective == 5 && !groups) throw Error("the pairwise objective needs timestamps as groups");
        dev_check(tgbm_task_set_rows(so.s, t, ys[t], nullptr, tl.pred.data(), tl.bucket.data(), nullptr, tl.n_buckets,
                                     b->cfg.iid_mode ? nullptr : tl.cuts.data(), groups));
        b->n_features = F;
        b->n_bins = n_bins;
        b->edges = edges;
        b->n_time_buckets = tl.n_buckets;
        b->init_prefix = tl.init_prefix;
        b->init_score = tl.init_score;
        out.push_back(std::move(b));
    }
    if (proto.profile) dev_check(tgbm_timer_enable(so.s, proto.profile));
    st.setup_s = now_s() - t0;
    // ---- boosting
    const double t1 = now_s();
    const tgbm_grow_params gp0 = grow_params(proto);
    const double min_cover =
        proto.torb_min_cover > 0.0 ? proto.torb_min_cover : (double)(2 * std::max(proto.min_samples_leaf, 1));
    const int MN = 2 * proto.max_leaves - 1;
    const int iid = proto.iid_mode ? 1 : 0;
    auto run = [&](int first, int count) {
        dev_check(tgbm_session_set_active(so.s, first, count));
        std::vector<int> nn(count), ints((size_t)count * MN * TGBM_NODE_INTS), cov((size_t)count * MN);
        std::vector<double> dbl((size_t)count * MN * 2), vals((size_t)count * MN), pref((size_t)count * MN * 256);
        // histogram kernel variant: forced, or tuned over the first trees (identical trees either
        // way); shape: the active tasks and their largest launch size
        int max_rows = 0;
        for (int t = first; t < first + count; ++t) max_rows = std::max(max_rows, task_rows[t]);
        const std::string shape = "batch:" + std::to_string(count) + ":" + std::to_string(launch_rows_of(max_rows)) +
                                  ":" + std::to_string(F) + ":" + std::to_string(proto.objective) + ":" +
                                  std::to_string(proto.max_leaves);
        HistTuner tuner(shape, proto.n_estimators, proto.autotune != 0 && proto.hist_mode == 0);
        tgbm_grow_params gp = gp0;
        for (int m = 0; m < proto.n_estimators; ++m) {
            if (proto.hist_mode == 0) gp.hist_mode = tuner.mode(m);
            dev_check(tgbm_task_grad(so.s, -1, proto.objective));
            const double t_tree = now_s();
            dev_check(tgbm_tree_begin(so.s, &gp));
            dev_check(tgbm_tree_grow_device(so.s, &gp));
            dev_check(tgbm_tree_end(so.s, proto.learning_rate, min_cover, iid, &gp, nn.data(), ints.data(), dbl.data(),
                                    vals.data(), cov.data(), pref.data()));
            tuner.record(m, now_s() - t_tree);   // tree_end synchronized
            for (int i = 0; i < count; ++i) {
                Booster& b = *out[first + i];
                b.hist_variant = gp.hist_mode;
                double loss = 0.0;
                dev_check(tgbm_task_train_loss(so.s, first + i, &loss));
                b.evals_train.push_back(loss);
                const size_t o = (size_t)i * MN;
                b.trees.push_back(unpack_tree(nn[i], &ints[o * TGBM_NODE_INTS], &dbl[o * 2], &vals[o], &cov[o],
                                              &pref[o * 256], b.n_time_buckets));
            }
        }
    };
    if (schedule == 0) {
        run(0, n_tasks);
    } else {
        for (int t = 0; t < n_tasks; ++t) run(t, 1);
    }
    st.boost_s = now_s() - t1;
    if (proto.profile)
        for (int k = 0; k < 8; ++k) dev_check(tgbm_timer_read(so.s, k, &st.prof_ms[k], &st.prof_n[k], &st.prof_bytes[k]));
    for (auto& b : out) {
        b->best_iteration = proto.n_estimators;
        b->fitted = true;
        b->t_setup = st.setup_s;
        b->t_boost = st.boost_s;
        for (int k = 0; k < 8; ++k) {
            b->prof_ms[k] = st.prof_ms[k];
            b->prof_n[k] = st.prof_n[k];
            b->prof_bytes[k] = st.prof_bytes[k];
        }
    }
    return st;
}

}  // namespace tide
