// The drop-in C API (include/tidegbm_amd.h).  Conventions follow the reference
// cpp/src/c_api.cpp:12-158: 0 / -1 returns, a thread-local error message, opaque handles,
// "unknown int key" / "unknown double key" on bad setter keys.
#include <cstdio>
#include <fstream>
#include <iterator>
#include <memory>
#include <string>

#include "../../../include/tidegbm_amd.h"
#include "booster.h"
#include "server.h"

using tide::Booster;

static thread_local std::string g_err;

template <class Fn>
static int guard(Fn&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

extern "C" {

const char* tide_last_error(void) { return g_err.c_str(); }

int tide_create(TideHandle* out) {
    return guard([&] { *out = new Booster(); });
}

void tide_free(TideHandle h) { delete static_cast<Booster*>(h); }

// Keys of the reference setter (c_api.cpp:35-49) plus the extensions listed in the header.
int tide_set_int(TideHandle h, const char* key, int v) {
    tide::Config& c = static_cast<Booster*>(h)->cfg;
    const std::string k(key);
    if (k == "n_estimators") c.n_estimators = v;
    else if (k == "max_depth") c.max_depth = v;
    else if (k == "max_leaves") c.max_leaves = v;
    else if (k == "max_bin") c.max_bin = v;
    else if (k == "n_time_buckets") c.n_time_buckets = v;
    else if (k == "random_state") c.random_state = v;
    else if (k == "n_jobs") c.n_jobs = v;
    else if (k == "iid_mode") c.iid_mode = v != 0;
    else if (k == "halo") c.halo = v != 0;
    else if (k == "verbose") c.verbose = v;
    else if (k == "horizon") c.horizon = v;
    else if (k == "embargo") c.embargo = v;
    else if (k == "goss") c.goss = v;
    else if (k == "cat_min_count") c.cat_min_count = v;
    else if (k == "col_major_hist") c.col_major_hist = v != 0;
    // extensions
    else if (k == "min_samples_leaf") c.min_samples_leaf = v;
    else if (k == "pair_splits") c.pair_splits = v != 0;
    else if (k == "pair_top_k") c.pair_top_k = v;
    else if (k == "max_pair_splits_per_tree") c.max_pair_splits_per_tree = v;
    else if (k == "bin_sample") c.bin_sample = v;
    else if (k == "finer_torb") c.finer_torb = v != 0;
    else if (k == "objective") c.objective = v;
    else if (k == "profile") c.profile = v;
    else if (k == "hist_mode") c.hist_mode = v;
    else if (k == "autotune") c.autotune = v;
    else if (k == "spec_width") {
        if (v < 1 || v > 8) {
            g_err = "spec_width must be in [1, 8]";
            return -1;
        }
        c.spec_width = v;
    }
    else if (k == "force_host_driver") c.force_host_driver = v;
    else if (k == "early_stopping_rounds") c.early_stopping_rounds = v;
    else if (k == "regime_feature") c.regime_feature = v;
    else if (k == "regime_priority") c.regime_priority = v != 0;
    else {
        g_err = "unknown int key";
        return -1;
    }
    return 0;
}

int tide_set_double(TideHandle h, const char* key, double v) {
    tide::Config& c = static_cast<Booster*>(h)->cfg;
    const std::string k(key);
    if (k == "learning_rate") c.learning_rate = v;
    else if (k == "l2") c.l2 = v;
    else if (k == "l1") c.l1 = v;
    else if (k == "subsample") c.subsample = v;
    else if (k == "colsample") c.colsample = v;
    // extensions
    else if (k == "min_child_weight") c.min_child_weight = v;
    else if (k == "min_split_gain") c.min_split_gain = v;
    else if (k == "pair_trigger") c.pair_trigger = v;
    else if (k == "pair_l2") c.pair_l2 = v;
    else if (k == "goss_top") c.goss_top = v;
    else if (k == "goss_other") c.goss_other = v;
    else if (k == "max_delta_step") c.max_delta_step = v;
    else if (k == "torb_min_cover") c.torb_min_cover = v;
    else if (k == "time_decay") c.time_decay = v;
    else if (k == "cat_smooth") c.cat_smooth = v;
    else if (k == "tree_decay") c.tree_decay = v;
    else {
        g_err = "unknown double key";
        return -1;
    }
    return 0;
}

int tide_set_string(TideHandle h, const char* key, const char* value) {
    tide::Config& c = static_cast<Booster*>(h)->cfg;
    const std::string k(key), v(value ? value : "");
    if (k == "device") {
        c.device = v;
        return 0;
    }
    if (k == "eval_metric") {  // config.hpp:76 (any other value scores the loss, booster.cpp:492-495)
        c.eval_metric = v;
        return 0;
    }
    if (k == "objective") {  // names of bindings.cpp:108-119
        if (v == "mse" || v == "auto" || v == "l2" || v == "regression") c.objective = 0;
        else if (v == "pairwise" || v == "lambdarank") c.objective = 5;
        else if (v == "huber" || v == "quantile" || v == "pinball" || v == "logloss" || v == "binary" ||
                 v == "classification" || v == "rank" || v == "listnet" || v == "multiclass" || v == "custom") {
            g_err = "objective " + v + " is not supported on the GPU path";
            return -1;
        } else {
            g_err = "unknown objective: " + v;
            return -1;
        }
        return 0;
    }
    g_err = "unknown string key";
    return -1;
}

int tide_fit(TideHandle h, const double* X, const double* y, int n, int n_features, const long long* timestamps,
             const double* weight, const int* is_cat, int iid, int horizon) {
    return tide_fit_ex(h, X, y, n, n_features, timestamps, weight, is_cat, nullptr, nullptr, iid, horizon);
}

int tide_fit_ex(TideHandle h, const double* X, const double* y, int n, int n_features, const long long* timestamps,
                const double* weight, const int* is_cat, const long long* entity, const long long* rank_groups,
                int iid, long long horizon) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] {
        b->cfg.iid_mode = iid != 0;
        if (!b->cfg.iid_mode) b->cfg.horizon = horizon;
        b->fit(X, y, n, n_features, timestamps, weight, is_cat, entity, rank_groups);
    });
}

int tide_update(TideHandle h, const double* X, const double* y, int n, int n_features, const double* weight,
                int extra_trees) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] { b->update(X, y, n, n_features, weight, extra_trees); });
}

int tide_predict(TideHandle h, const double* X, int n, int n_features, double* out) {
    // MSE and pairwise finalize as the identity (objective.cpp:331-341)
    return tide_predict_raw(h, X, n, n_features, out);
}

int tide_predict_raw(TideHandle h, const double* X, int n, int n_features, double* out) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] { b->predict_raw(X, n, n_features, out); });
}

int tide_predict_binned(TideHandle h, const unsigned short* xb, int n, double* out) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] { b->predict_binned_u16(xb, n, out); });
}

int tide_save_json(TideHandle h, const char* path) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] {
        std::ofstream f(path);
        if (!f) throw tide::Error("cannot open file");
        f << b->save_json();
    });
}

int tide_save(TideHandle h, const char* path) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] {
        const std::string blob = b->save_binary();
        std::ofstream f(path, std::ios::binary);
        if (!f) throw tide::Error("cannot open file");
        f.write(blob.data(), (std::streamsize)blob.size());
    });
}

int tide_load(TideHandle* out, const char* path) {
    return guard([&] {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw tide::Error("cannot open file");
        std::string blob((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        *out = Booster::load_binary(blob);
    });
}

int tide_device_compiled(void) { return 1; }
int tide_device_count(void) { return tgbm_device_count(); }

int tide_num_trees(TideHandle h) { return (int)static_cast<Booster*>(h)->trees.size(); }

int tide_tree_info(TideHandle h, int t, int* n_nodes, int* n_buckets) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] {
        const tide::Tree& tr = b->trees.at(t);
        *n_nodes = (int)tr.nodes.size();
        *n_buckets = tr.n_buckets;
    });
}

int tide_tree_get(TideHandle h, int t, int* ints, double* dbl, double* values, int* cover, double* prefix) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] {
        const tide::Tree& tr = b->trees.at(t);
        for (size_t k = 0; k < tr.nodes.size(); ++k) {
            const tide::Node& d = tr.nodes[k];
            const int v[10] = {d.is_leaf, d.split_type, d.feature, d.feature2, d.threshold, d.left, d.right,
                               d.missing_go_left, d.n_samples, d.depth};
            for (int j = 0; j < 10; ++j) ints[10 * k + j] = v[j];
            dbl[2 * k] = d.value;
            dbl[2 * k + 1] = d.gain;
            values[k] = tr.values[k];
            cover[k] = tr.cover[k];
        }
        for (size_t i = 0; i < tr.prefix_values.size(); ++i) prefix[i] = tr.prefix_values[i];
    });
}

int tide_model_info(TideHandle h, int* n_features, int* n_bins, int* n_time_buckets, int* best_iteration,
                    double* init_score) {
    auto* b = static_cast<Booster*>(h);
    *n_features = b->n_features;
    *n_bins = b->n_bins;
    *n_time_buckets = b->n_time_buckets;
    *best_iteration = b->best_iteration;
    *init_score = b->init_score;
    return 0;
}

int tide_get_edges(TideHandle h, int f, double* out, int cap) {
    auto* b = static_cast<Booster*>(h);
    if (f < 0 || f >= (int)b->edges.size()) {
        g_err = "feature out of range";
        return -1;
    }
    const auto& e = b->edges[f];
    for (int i = 0; i < (int)e.size() && i < cap; ++i) out[i] = e[i];
    return (int)e.size();
}

int tide_get_init_prefix(TideHandle h, double* out, int cap) {
    auto* b = static_cast<Booster*>(h);
    const auto& v = b->init_prefix;
    for (int i = 0; i < (int)v.size() && i < cap; ++i) out[i] = v[i];
    return (int)v.size();
}

int tide_profile(TideHandle h, int family, double* ms, long long* launches, double* bytes) {
    auto* b = static_cast<Booster*>(h);
    if (family < 0 || family >= 8) {
        g_err = "family out of range";
        return -1;
    }
    *ms = b->prof_ms[family];
    *launches = b->prof_n[family];
    *bytes = b->prof_bytes[family];
    return 0;
}

int tide_tree_times(TideHandle h, double* out, int cap) {
    auto* b = static_cast<Booster*>(h);
    const int m = (int)b->tree_end_times.size();
    for (int i = 0; i < m && i < cap; ++i) out[i] = b->tree_end_times[i];
    return m;
}

int tide_profile_log(TideHandle h, int family, double* ms, double* rows, int cap) {
    auto* b = static_cast<Booster*>(h);
    if (family < 0 || family >= 8) {
        g_err = "profile family out of range";
        return -1;
    }
    const int m = (int)b->prof_log_ms[family].size();
    for (int i = 0; i < m && i < cap; ++i) {
        ms[i] = b->prof_log_ms[family][i];
        rows[i] = b->prof_log_rows[family][i];
    }
    return m;
}

int tide_fit_times(TideHandle h, double* setup_s, double* boost_s) {
    auto* b = static_cast<Booster*>(h);
    *setup_s = b->t_setup;
    *boost_s = b->t_boost;
    return 0;
}

int tide_hist_variant(TideHandle h) { return static_cast<Booster*>(h)->hist_variant; }

int tide_set_feature_groups(TideHandle h, const int* group_of, int n_features) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] {
        if (n_features < 0 || (n_features > 0 && !group_of)) throw tide::Error("bad feature groups");
        b->group_of.assign(group_of, group_of + n_features);
    });
}

int tide_set_monotone(TideHandle h, const int* monotone, int n_features) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] {
        if (n_features < 0 || (n_features > 0 && !monotone)) throw tide::Error("bad monotone constraints");
        b->cfg.monotone.clear();
        for (int f = 0; f < n_features; ++f) b->cfg.monotone.push_back(monotone[f] > 0 ? 1 : (monotone[f] < 0 ? -1 : 0));
    });
}

int tide_set_interaction_groups(TideHandle h, const int* groups, int n_features) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] {
        if (n_features < 0 || (n_features > 0 && !groups)) throw tide::Error("bad interaction groups");
        b->cfg.interaction_group.assign(groups, groups + n_features);
    });
}

int tide_tree_cat(TideHandle h, int tree, unsigned long long* out) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] {
        const tide::Tree& t = b->trees.at(tree);
        for (size_t k = 0; k < t.nodes.size(); ++k) std::copy(t.nodes[k].cat_left, t.nodes[k].cat_left + 4, out + 4 * k);
    });
}

int tide_get_cat_values(TideHandle h, int feature, double* out, int cap) {
    auto* b = static_cast<Booster*>(h);
    if (feature < 0 || feature >= (int)b->cat_values.size()) return 0;
    const auto& v = b->cat_values[feature];
    for (int i = 0; i < (int)v.size() && i < cap; ++i) out[i] = v[i];
    return (int)v.size();
}

int tide_set_eval(TideHandle h, const double* X_eval, const double* y_eval, const long long* ts_eval, int n_eval,
                  int n_features) {
    auto* b = static_cast<Booster*>(h);
    return guard([&] { b->set_eval(X_eval, y_eval, ts_eval, n_eval, n_features); });
}

int tide_evals_valid(TideHandle h, double* out, int cap) {
    const auto& e = static_cast<Booster*>(h)->evals_valid;
    for (int i = 0; i < (int)e.size() && i < cap; ++i) out[i] = e[i];
    return (int)e.size();
}

int tide_evals_train(TideHandle h, double* out, int cap) {
    const auto& e = static_cast<Booster*>(h)->evals_train;
    for (int i = 0; i < (int)e.size() && i < cap; ++i) out[i] = e[i];
    return (int)e.size();
}

int tide_goss_stats(TideHandle h, int* host_trees, double* wait_s, double* sample_s) {
    auto* b = static_cast<Booster*>(h);
    *host_trees = b->goss_host_trees;
    *wait_s = b->goss_wait_s;
    *sample_s = b->goss_sample_s;
    return 0;
}

int tide_fit_batch(TideHandle proto, const double* X, int n, int n_features, const long long* timestamps,
                   const long long* entity, int n_tasks, const int* task_rows, const double* const* y,
                   const long long* horizons, int schedule, TideHandle* out) {
    auto* p = static_cast<Booster*>(proto);
    return guard([&] {
        if (!p || !X || !task_rows || !y || !out) throw tide::Error("null argument");
        if (p->n_eval > 0) throw tide::Error("eval sets are not supported by batched fits");
        if (!p->group_of.empty() || !p->cfg.interaction_group.empty() || p->cfg.regime_feature >= 0)
            throw tide::Error("feature groups, interaction groups and a regime feature need the host-driven grower; "
                              "batched fits grow on the device");
        std::vector<std::unique_ptr<Booster>> res;
        tide::fit_batch(p->cfg, X, n, n_features, timestamps, entity, n_tasks, task_rows, y, horizons, schedule, res);
        for (int t = 0; t < n_tasks; ++t) out[t] = res[t].release();
    });
}

int tide_server_create(TideServer* out, const TideHandle* models, int n_models, const double* panel_X, int n_rows,
                       int n_features, long long predict_budget_bytes, int refit_slots, int refit_max_rows,
                       int batched) {
    return guard([&] {
        *out = nullptr;
        if (!models || n_models < 1) throw tide::Error("a server needs at least one ensemble");
        if (predict_budget_bytes <= 0) throw tide::Error("predict_budget_bytes must be positive");
        std::vector<const Booster*> ms(n_models);
        for (int i = 0; i < n_models; ++i) ms[i] = static_cast<const Booster*>(models[i]);
        *out = new tide::Server(ms, panel_X, n_rows, n_features, (size_t)predict_budget_bytes, refit_slots,
                                refit_max_rows, batched != 0);
    });
}

void tide_server_free(TideServer s) { delete static_cast<tide::Server*>(s); }

long long tide_server_predict(TideServer s, int model, const int* rows, int n) {
    long long id = -1;
    if (guard([&] { id = static_cast<tide::Server*>(s)->submit_predict(model, rows, n); }) != 0) return -1;
    return id;
}

long long tide_server_predict_x(TideServer s, int model, const double* X, int n) {
    long long id = -1;
    if (guard([&] { id = static_cast<tide::Server*>(s)->submit_predict_x(model, X, n); }) != 0) return -1;
    return id;
}

long long tide_server_refit(TideServer s, int model, int row0, int n, const double* y, int extra_trees) {
    long long id = -1;
    if (guard([&] { id = static_cast<tide::Server*>(s)->submit_refit(model, row0, n, y, extra_trees); }) != 0)
        return -1;
    return id;
}

int tide_server_wait(TideServer s, long long ticket, double* out, int cap, int* version, double* latency_s) {
    int n = -1;
    if (guard([&] { n = static_cast<tide::Server*>(s)->wait(ticket, out, cap, version, latency_s); }) != 0) return -1;
    return n;
}

int tide_server_model(TideServer s, int model, int version, TideHandle* out, int* parent) {
    return guard([&] {
        auto m = static_cast<tide::Server*>(s)->model_version(model, version, parent);
        Booster* b = new Booster();
        *out = b;
        b->cfg = m->cfg;
        b->n_features = m->n_features;
        b->n_bins = m->n_bins;
        b->n_time_buckets = m->n_time_buckets;
        b->best_iteration = m->best_iteration;
        b->init_score = m->init_score;
        b->init_prefix = m->init_prefix;
        b->edges = m->edges;
        b->trees = m->trees;
        b->fitted = m->fitted;
    });
}

int tide_server_stats(TideServer s, long long* counts5, double* busy2) {
    return guard([&] {
        const tide::ServerStats st = static_cast<tide::Server*>(s)->stats();
        counts5[0] = st.predict_batches;
        counts5[1] = st.predict_requests;
        counts5[2] = st.predict_rows;
        counts5[3] = st.refit_rounds;
        counts5[4] = st.refit_requests;
        busy2[0] = st.predict_busy_s;
        busy2[1] = st.refit_busy_s;
        busy2[2] = st.refit_admit_s;
        busy2[3] = st.refit_round_s;
        busy2[4] = st.refit_publish_s;
    });
}

}  // extern "C"
