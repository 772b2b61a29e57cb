// Host side of the MI355X histogram-GBDT path: the C++ booster that prepares data exactly like
// the reference Booster::fit (reference cpp/src/booster.cpp:231-742) and drives the gfx950
// kernels through the thin device ABI (include/tgbm_device.h).
#pragma once

#include <cstdint>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/tgbm_device.h"

namespace tide {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

// Configuration: names and defaults follow the reference Config (config.hpp:12-113) and the
// kwargs defaults of bindings.cpp:43-126.  Fields the device path does not implement are kept
// so that setting them to a non-default value fails loudly at fit time.
struct Config {
    int n_estimators = 100;
    double learning_rate = 0.1;
    int max_depth = 6;
    int max_leaves = 31;
    int min_samples_leaf = 20;
    double min_child_weight = 1e-3;
    double l2 = 1.0;
    double l1 = 0.0;
    double max_delta_step = 0.0;
    double min_split_gain = 0.0;
    int max_bin = 255;
    double subsample = 1.0;
    double colsample = 1.0;
    int n_time_buckets = 0;
    long long horizon = -1;
    long long embargo = 0;
    bool iid_mode = false;
    bool halo = false;
    bool pair_splits = true;
    double pair_trigger = 0.25;
    int pair_top_k = 6;
    int max_pair_splits_per_tree = 3;
    double pair_l2 = 2.0;
    double time_decay = 1.0;
    double tree_decay = 1.0;
    int objective = 0;  // ObjectiveKind: 0 MSE, 5 RankPairwise (types.hpp:26-35)
    int n_jobs = 0;
    int random_state = 0;
    int verbose = 0;
    int goss = -1;
    double goss_top = 0.2;
    double goss_other = 0.1;
    int bin_sample = 262144;
    bool finer_torb = true;
    double torb_min_cover = 0.0;
    int cat_min_count = 5;
    double cat_smooth = 1.0;   // config.hpp:70: Laplace term of the categorical G/H ranking
    int regime_feature = -1;   // config.hpp:55-56: split on this feature when it is half as good
    int regime_priority = 1;
    std::vector<int> interaction_group;   // config.hpp:59: per feature, -1 = unconstrained
    // Booster::fit's monotone argument (booster.cpp:250-251): per feature +1 / -1 / 0; a split
    // whose leaf weights break the order is skipped (histogram.cpp:179-180).  Empty: none.
    std::vector<int> monotone;
    // objective parameters of the model file that this path does not use (config.hpp); kept so a
    // loaded file saves back unchanged
    double quantile = 0.5, huber_delta = 1.0, rank_temperature = 1.0;
    bool col_major_hist = true;
    std::string device = "gpu";
    int profile = 0;
    int hist_mode = 0;       // histogram kernel variant; 0: defaults, tuned per fit when autotune
    int autotune = 1;        // choose the variant online over the first trees (tuner.h)
    int spec_width = 8;      // leaves split per step of device-driven growth (1: one at a time)
    int force_host_driver = 0;
    int early_stopping_rounds = 0;   // config.hpp early_stopping_rounds (0: never stop early)
    std::string eval_metric = "auto";  // config.hpp:76: auto | loss | ic

    void validate() const;  // config.hpp:78-97 plus the device path's limits
};

struct Node {  // reference Node (tree.hpp:11-25)
    int is_leaf = 1, split_type = 0, feature = -1, feature2 = -1, threshold = -1, left = -1,
        right = -1, missing_go_left = 1, n_samples = 0, depth = 0;
    double value = 0.0, gain = 0.0;
    unsigned long long cat_left[4] = {0, 0, 0, 0};   // categorical split: bins that go left
};

struct Tree {  // reference Tree (tree.hpp:32-41)
    std::vector<Node> nodes;
    std::vector<double> values;
    std::vector<int> cover;
    std::vector<double> prefix_values;
    int n_buckets = 1;
};

// TORB layout and initial scores of one time-sorted fit.
struct TimeLayout {
    int n_buckets = 1;
    std::vector<int> bucket, cuts;       // per row / per bucket
    std::vector<double> init_prefix;     // per bucket
    double init_score = 0.0;
    std::vector<double> pred;            // initial prediction per row
};
TimeLayout time_layout(const Config& cfg, const long long* ts, const double* y, const double* w, int n);
void fit_edges(const double* X, int n, int F, int max_bin, int bin_sample, uint32_t seed,
               std::vector<std::vector<double>>& edges);
// Sorted category values of categorical feature f (reference Binner::fit, binning.cpp:48-82).
std::vector<double> category_values(const double* X, int n, int F, int f, int max_bin, int cat_min_count);
void bin_on_device(tgbm_session* s, const double* X, const std::vector<std::vector<double>>& edges,
                   const std::vector<int>& is_cat = {}, const std::vector<std::vector<double>>& cats = {});
Tree unpack_tree(int nn, const int* ints, const double* dbl, const double* vals, const int* cov,
                 const double* pref, int n_buckets);
tgbm_grow_params grow_params(const Config& c);
long long launch_rows_of(long long n);
class Booster;
// Device-resident forest of a fitted model's first best_iteration trees (lr per tree = cfg lr).
tgbm_forest* make_forest(const Booster& b);
// One tree in the export node layout (TGBM_NODE_INTS ints per node) with its leaf values.
void tree_node_ints(const Tree& t, std::vector<int>& ints, std::vector<double>& vals);
// Edges as the device takes them: edges[f*256 + k], k < n_edges[f]; a categorical feature
// passes its category values with TGBM_CAT_EDGES set in n_edges[f].
void flat_edges(const std::vector<std::vector<double>>& edges, std::vector<double>& E, std::vector<int>& ne,
                const std::vector<int>& is_cat = {}, const std::vector<std::vector<double>>& cats = {});
// Bin masks of a tree's categorical nodes, 4 words per node (zeros elsewhere), or empty when the
// tree has none.
std::vector<unsigned long long> tree_cat_masks(const Tree& t);
void dev_check(int rc);
double now_s();

class Booster {
public:
    Config cfg;
    int n_features = 0;
    int n_bins = 255;
    int n_time_buckets = 1;
    int best_iteration = 0;
    double init_score = 0.0;
    std::vector<double> init_prefix;
    std::vector<std::vector<double>> edges;
    std::vector<int> is_cat;                       // per feature (empty: none categorical)
    std::vector<int> group_of;                     // feature groups of later fits (Booster::fit's group_of)
    std::vector<std::string> feature_names;        // names in the model file (f0 .. f{F-1} at fit)
    std::vector<std::vector<double>> cat_values;   // categorical features' sorted category values
    std::vector<Tree> trees;
    bool fitted = false;
    bool has_cat() const {
        for (int v : is_cat)
            if (v) return true;
        return false;
    }
    // last fit diagnostics
    double t_setup = 0.0, t_boost = 0.0;
    std::vector<double> tree_end_times;  // seconds since the boosting loop started, per tree
    std::vector<double> evals_train;     // training loss after each tree (reference evals.train)
    std::vector<double> evals_valid;     // validation metric after each tree (reference evals.valid)
    // Eval set of later fits (Booster::fit's X_eval, y_eval, ts_eval, n_eval; booster.cpp:231-234)
    std::vector<double> eval_X, eval_y;
    std::vector<long long> eval_ts;
    int n_eval = 0, eval_features = 0;
    int goss_host_trees = 0;             // GOSS trees of the last fit sampled on the host (equal |g|)
    int hist_variant = 0;                // histogram kernel variant the last fit ended with
    double goss_wait_s = 0.0, goss_sample_s = 0.0;  // GOSS: waiting for draws, the sample step
    double prof_ms[8] = {0}, prof_bytes[8] = {0};
    long long prof_n[8] = {0};
    std::vector<double> prof_log_ms[8], prof_log_rows[8];

    ~Booster();
    Booster() = default;
    Booster(const Booster&) = delete;
    Booster& operator=(const Booster&) = delete;

    void fit(const double* X, const double* y, int n, int nf, const long long* ts, const double* w,
             const int* is_cat, const long long* entity, const long long* rank_groups);
    void update(const double* X, const double* y, int n, int nf, const double* w, int extra_trees);
    void predict_raw(const double* X, int n, int nf, double* out);
    void predict_binned_u16(const unsigned short* xb, int n, double* out);
    // Keep an eval set for later fits (X: n x nf row-major; ts may be null); n = 0 clears it.
    void set_eval(const double* X, const double* y, const long long* ts, int n, int nf);

    // Copy of the fitted model (configuration, edges, trees, initial scores); no device state.
    std::shared_ptr<Booster> clone_model() const;

    std::string save_binary() const;
    std::string save_json() const;
    static Booster* load_binary(const std::string& blob);

    struct EvalRun;   // device eval set and early-stopping state of one fit

private:
    EvalRun* eval_run_ = nullptr;   // set while a fit with an eval set grows its trees
    bool eval_step(int m);          // after tree m: score, record, decide; true = stop early
    tgbm_forest* forest_ = nullptr;
    int forest_trees_ = -1;
    void ensure_forest();
    void drop_forest();
    void grow_trees(tgbm_session* s, int n, int n_trees, bool time_mode, const std::vector<int>& cuts,
                    std::mt19937& rng, bool goss, const double* gsample_unused, bool for_update);
};

struct BatchStats {
    double setup_s = 0.0, boost_s = 0.0;
    double prof_ms[8] = {0}, prof_bytes[8] = {0};
    long long prof_n[8] = {0};
};

// Walk-forward batch: n_tasks fits over row prefixes [0, task_rows[t]) of one matrix whose rows
// are sorted by (timestamp, entity), with per-task targets ys[t] and horizons; schedule 0 grows
// all tasks' trees together, 1 one task after another.  Fills out[t] with fitted boosters.
BatchStats fit_batch(const Config& proto, const double* X, int n, int F, const long long* ts,
                     const long long* entity, int n_tasks, const int* task_rows, const double* const* ys,
                     const long long* horizons, int schedule, std::vector<std::unique_ptr<Booster>>& out);

}  // namespace tide
