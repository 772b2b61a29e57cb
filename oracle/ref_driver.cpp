// Test infrastructure only (a checker, never shipped or measured as the product).
//
// Thin extern "C" driver over the REFERENCE TideGBM C++ core, compiled from the
// reference's own sources where they lie (see oracle/Makefile.ref).  The reference's
// C API (cpp/include/tidegbm/c_api.h:9-28) cannot pass rank_groups/entity
// (cpp/src/c_api.cpp:79-80) nor set objective/pair_splits/min_samples_leaf
// (cpp/src/c_api.cpp:32-70), so this driver calls tidegbm::Booster::fit
// (cpp/include/tidegbm/booster.hpp:45-48) directly and exposes the reference's public
// per-function entry points (build_hist_col, best_numeric_missing, Objective::grad_hess,
// Binner) so the restated oracle and the device path can be pinned against the
// reference itself.
#include "tidegbm/booster.hpp"
#include "tidegbm/histogram.hpp"
#include "tidegbm/objective.hpp"
#include "tidegbm/tree.hpp"

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace tidegbm;

static thread_local std::string g_ref_err;

// Eval sets handed to Booster::fit (X_eval, y_eval, ts_eval, n_eval), kept per handle.
struct EvalSet {
    std::vector<double> X, y;
    std::vector<long long> ts;
    int n = 0;
    std::vector<int> is_cat;   // Booster::fit's is_cat of later fits (empty: none)
    std::vector<int> group_of; // Booster::fit's group_of of later fits (empty: none)
    std::vector<int8_t> monotone; // Booster::fit's monotone of later fits (empty: none)
};
static std::map<void*, EvalSet> g_eval;
static std::mutex g_eval_mu;

template <class F>
static int guarded(F&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception& e) {
        g_ref_err = e.what();
        return -1;
    }
}

extern "C" {

const char* tref_last_error(void) { return g_ref_err.c_str(); }

void* tref_create(void) { return new Booster(); }
void tref_free(void* h) {
    {
        std::lock_guard<std::mutex> lk(g_eval_mu);
        g_eval.erase(h);
    }
    delete static_cast<Booster*>(h);
}

// Feature groups of later fits (Booster::fit's group_of); nf = 0 clears.
int tref_set_groups(void* h, const int* group_of, int nf) {
    std::lock_guard<std::mutex> lk(g_eval_mu);
    g_eval[h].group_of.assign(group_of, group_of + (nf > 0 ? nf : 0));
    return 0;
}

// Monotone constraints of later fits (Booster::fit's monotone, +1 / -1 / 0); nf = 0 clears.
int tref_set_monotone(void* h, const int* monotone, int nf) {
    std::lock_guard<std::mutex> lk(g_eval_mu);
    g_eval[h].monotone.assign(monotone, monotone + (nf > 0 ? nf : 0));
    return 0;
}

// Config::interaction_group (config.hpp:59); n = 0 clears.
int tref_set_interaction(void* h, const int* groups, int n) {
    static_cast<Booster*>(h)->cfg.interaction_group.assign(groups, groups + (n > 0 ? n : 0));
    return 0;
}

// Categorical flags of later fits (Booster::fit's is_cat); nf = 0 clears.
int tref_set_cat(void* h, const int* is_cat, int nf) {
    std::lock_guard<std::mutex> lk(g_eval_mu);
    g_eval[h].is_cat.assign(is_cat, is_cat + (nf > 0 ? nf : 0));
    return 0;
}

// Eval set of later fits: X (n x nf, row-major), y, ts (may be NULL); n = 0 clears.
int tref_set_eval(void* h, const double* X, const double* y, const long long* ts, int n, int nf) {
    std::lock_guard<std::mutex> lk(g_eval_mu);
    EvalSet& e = g_eval[h];
    e.n = n > 0 ? n : 0;  // keeps e.is_cat
    e.X.assign(X, X + (size_t)e.n * (nf > 0 ? nf : 0));
    e.y.assign(y, y + e.n);
    if (ts) e.ts.assign(ts, ts + e.n);
    else e.ts.clear();
    return 0;
}

// Numeric config setter; keys are the reference Config field names (config.hpp:12-76).
int tref_set(void* h, const char* key, double v) {
    Config& c = static_cast<Booster*>(h)->cfg;
    const std::string k(key);
    const int iv = static_cast<int>(v);
    if (k == "n_estimators") c.n_estimators = iv;
    else if (k == "learning_rate") c.learning_rate = v;
    else if (k == "max_depth") c.max_depth = iv;
    else if (k == "max_leaves") c.max_leaves = iv;
    else if (k == "min_samples_leaf") c.min_samples_leaf = iv;
    else if (k == "min_child_weight") c.min_child_weight = v;
    else if (k == "l2") c.l2 = v;
    else if (k == "l1") c.l1 = v;
    else if (k == "min_split_gain") c.min_split_gain = v;
    else if (k == "max_bin") c.max_bin = iv;
    else if (k == "n_time_buckets") c.n_time_buckets = iv;
    else if (k == "embargo") c.embargo = static_cast<i64>(v);
    else if (k == "pair_splits") c.pair_splits = iv != 0;
    else if (k == "pair_trigger") c.pair_trigger = v;
    else if (k == "pair_top_k") c.pair_top_k = iv;
    else if (k == "max_pair_splits_per_tree") c.max_pair_splits_per_tree = iv;
    else if (k == "pair_l2") c.pair_l2 = v;
    else if (k == "objective") c.objective = static_cast<ObjectiveKind>(iv);
    else if (k == "n_jobs") c.n_jobs = iv;
    else if (k == "random_state") c.random_state = iv;
    else if (k == "goss") c.goss = iv;
    else if (k == "goss_top") c.goss_top = v;
    else if (k == "goss_other") c.goss_other = v;
    else if (k == "bin_sample") c.bin_sample = iv;
    else if (k == "finer_torb") c.finer_torb = iv != 0;
    else if (k == "torb_min_cover") c.torb_min_cover = v;
    else if (k == "max_delta_step") c.max_delta_step = v;
    else if (k == "col_major_hist") c.col_major_hist = iv != 0;
    else if (k == "ooc_bytes") c.ooc_bytes = static_cast<size_t>(v);
    else if (k == "early_stopping_rounds") c.early_stopping_rounds = iv;
    else if (k == "cat_min_count") c.cat_min_count = iv;
    else if (k == "regime_feature") c.regime_feature = iv;
    else if (k == "regime_priority") c.regime_priority = iv != 0;
    else if (k == "cat_smooth") c.cat_smooth = v;
    else if (k == "eval_metric") c.eval_metric = iv == 1 ? "loss" : (iv == 2 ? "ic" : "auto");
    else {
        g_ref_err = "unknown key " + k;
        return -1;
    }
    return 0;
}

int tref_fit(void* h, const double* X, const double* y, int n, int nf, const long long* ts,
             const double* w, const long long* entity, const long long* rank_groups, int iid,
             long long horizon) {
    auto* b = static_cast<Booster*>(h);
    EvalSet ev;
    {
        std::lock_guard<std::mutex> lk(g_eval_mu);
        auto it = g_eval.find(h);
        if (it != g_eval.end()) ev = it->second;
    }
    return guarded([&] {
        b->cfg.iid_mode = iid != 0;
        if (!b->cfg.iid_mode) b->cfg.horizon = horizon;
        b->fit(X, y, n, nf, reinterpret_cast<const i64*>(ts), w,
               (int)ev.is_cat.size() == nf ? ev.is_cat.data() : nullptr,
               (int)ev.group_of.size() == nf ? reinterpret_cast<const i32*>(ev.group_of.data()) : nullptr,
               (int)ev.monotone.size() == nf ? ev.monotone.data() : nullptr,
               reinterpret_cast<const i64*>(rank_groups), ev.n > 0 ? ev.X.data() : nullptr,
               ev.n > 0 ? ev.y.data() : nullptr,
               ev.n > 0 && !ev.ts.empty() ? reinterpret_cast<const i64*>(ev.ts.data()) : nullptr, ev.n,
               reinterpret_cast<const i64*>(entity), nullptr);
    });
}

int tref_update(void* h, const double* X, const double* y, int n, int nf, const double* w,
                int extra_trees) {
    auto* b = static_cast<Booster*>(h);
    return guarded([&] { b->update(X, y, n, nf, w, nullptr, extra_trees); });
}

int tref_predict(void* h, const double* X, int n, int nf, double* out) {
    auto* b = static_cast<Booster*>(h);
    return guarded([&] { b->predict(X, n, nf, out); });
}

int tref_bins(void* h, const double* X, int n, int nf, unsigned short* out) {
    auto* b = static_cast<Booster*>(h);
    return guarded([&] {
        std::vector<u16> xb;
        b->bin_transform(X, n, nf, xb);
        std::memcpy(out, xb.data(), xb.size() * sizeof(u16));
    });
}

int tref_n_bins(void* h) { return static_cast<Booster*>(h)->n_bins; }
int tref_n_time_buckets(void* h) { return static_cast<Booster*>(h)->n_time_buckets; }
double tref_init_score(void* h) { return static_cast<Booster*>(h)->init_score; }
int tref_best_iteration(void* h) { return static_cast<Booster*>(h)->best_iteration; }
// Categorical split masks of tree t: 4 words per node.
int tref_tree_cat(void* h, int t, unsigned long long* out) {
    const Tree& tr = static_cast<Booster*>(h)->trees.at(t);
    for (size_t k = 0; k < tr.nodes.size(); ++k)
        for (int w = 0; w < 4; ++w) out[4 * k + w] = tr.nodes[k].cat_left[w];
    return (int)tr.nodes.size();
}

// Category vocabulary of feature f of the last fit.
int tref_cat_values(void* h, int f, double* out, int cap) {
    const auto& v = static_cast<Booster*>(h)->binner.cat_values.at(f);
    for (int i = 0; i < (int)v.size() && i < cap; ++i) out[i] = v[i];
    return (int)v.size();
}

int tref_evals_valid(void* h, double* out, int cap) {
    const auto& e = static_cast<Booster*>(h)->evals.valid;
    for (int i = 0; i < (int)e.size() && i < cap; ++i) out[i] = e[i];
    return (int)e.size();
}
int tref_evals_train(void* h, double* out, int cap) {
    const auto& e = static_cast<Booster*>(h)->evals.train;
    for (int i = 0; i < (int)e.size() && i < cap; ++i) out[i] = e[i];
    return (int)e.size();
}

int tref_init_prefix(void* h, double* out, int cap) {
    const auto& v = static_cast<Booster*>(h)->init_prefix;
    const int m = static_cast<int>(v.size());
    for (int i = 0; i < m && i < cap; ++i) out[i] = v[i];
    return m;
}

int tref_edges(void* h, int f, double* out, int cap) {
    const auto& e = static_cast<Booster*>(h)->binner.edges[f];
    const int m = static_cast<int>(e.size());
    for (int i = 0; i < m && i < cap; ++i) out[i] = e[i];
    return m;
}

int tref_n_trees(void* h) { return static_cast<int>(static_cast<Booster*>(h)->trees.size()); }

int tref_tree_info(void* h, int t, int* n_nodes, int* n_buckets) {
    const Tree& tr = static_cast<Booster*>(h)->trees.at(t);
    *n_nodes = static_cast<int>(tr.nodes.size());
    *n_buckets = tr.n_buckets;
    return 0;
}

// ints: per node 10 x int32 {is_leaf, split_type, feature, feature2, threshold_bin, left,
// right, missing_go_left, n_samples, depth}; dbl: per node {value, gain}.
int tref_tree_get(void* h, int t, int* ints, double* dbl, double* values, int* cover,
                  double* prefix) {
    const Tree& tr = static_cast<Booster*>(h)->trees.at(t);
    const int nn = static_cast<int>(tr.nodes.size());
    for (int k = 0; k < nn; ++k) {
        const Node& nd = tr.nodes[k];
        int* o = ints + 10 * k;
        o[0] = nd.is_leaf ? 1 : 0;
        o[1] = static_cast<int>(nd.split_type);
        o[2] = nd.feature;
        o[3] = nd.feature2;
        o[4] = nd.threshold_bin;
        o[5] = nd.left;
        o[6] = nd.right;
        o[7] = nd.missing_go_left;
        o[8] = nd.n_samples;
        o[9] = nd.depth;
        dbl[2 * k] = nd.value;
        dbl[2 * k + 1] = nd.gain;
        values[k] = tr.values[k];
        cover[k] = tr.cover[k];
    }
    for (size_t i = 0; i < tr.prefix_values.size(); ++i) prefix[i] = tr.prefix_values[i];
    return 0;
}

// Reference histogram build (histogram.cpp:44-76) over a column-major u16 bin matrix.
int tref_build_hist_col(const unsigned short* xcol, int n_rows, int nf, const int* idx, int n_idx,
                        const double* g, const double* h, int n_bins, double* G, double* H,
                        long long* C) {
    return guarded([&] {
        Hist hist;
        hist.resize(nf, n_bins);
        build_hist_col(xcol, n_rows, nf, idx, n_idx, g, h, hist, 1);
        for (size_t i = 0; i < hist.B.size(); ++i) {
            G[i] = hist.B[i].g;
            H[i] = hist.B[i].h;
            C[i] = hist.B[i].c;
        }
    });
}

// Reference per-feature split scan (histogram.cpp:144-193) over a planar histogram.
int tref_best_numeric(const double* G, const double* H, const long long* C, int nf, int n_bins,
                      int f, double lam, double gamma, int min_leaf, double min_hess, int monotone,
                      double l1, double* gain, int* bin, int* miss_left) {
    return guarded([&] {
        Hist hist;
        hist.resize(nf, n_bins);
        for (size_t i = 0; i < hist.B.size(); ++i) {
            hist.B[i].g = G[i];
            hist.B[i].h = H[i];
            hist.B[i].c = C[i];
        }
        SplitCand s = best_numeric_missing(hist, f, lam, gamma, min_leaf, min_hess, monotone, l1);
        *gain = s.gain;
        *bin = s.threshold_bin;
        *miss_left = s.missing_go_left;
    });
}

// Reference gradients (objective.cpp:102-225); kind is ObjectiveKind (0 MSE, 5 pairwise).
int tref_grad(int kind, const double* y, const double* pred, const double* w,
              const long long* groups, int n, double* g, double* h) {
    return guarded([&] {
        Objective o;
        o.kind = static_cast<ObjectiveKind>(kind);
        o.grad_hess(y, pred, w, reinterpret_cast<const i64*>(groups), n, g, h);
    });
}

int tref_save(void* h, const char* path, int json) {
    auto* b = static_cast<Booster*>(h);
    return guarded([&] {
        const std::string blob = json ? b->save_json() : b->save_binary();
        FILE* f = std::fopen(path, "wb");
        if (!f) throw TideError("cannot open file");
        std::fwrite(blob.data(), 1, blob.size(), f);
        std::fclose(f);
    });
}

void* tref_load(const char* path) {
    try {
        FILE* f = std::fopen(path, "rb");
        if (!f) throw TideError("cannot open file");
        std::string blob;
        char buf[65536];
        size_t r;
        while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) blob.append(buf, r);
        std::fclose(f);
        return new Booster(Booster::load_binary(blob));
    } catch (const std::exception& e) {
        g_ref_err = e.what();
        return nullptr;
    }
}

}  // extern "C"
