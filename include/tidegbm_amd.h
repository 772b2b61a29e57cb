/* Drop-in C API of the MI355X histogram-GBDT path (libtidegbm_amd.so).
 *
 * The first twelve entry points keep the exact signatures and conventions of the reference C API
 * (reference cpp/include/tidegbm/c_api.h:9-28, implemented in cpp/src/c_api.cpp:18-158):
 * return 0 on success and -1 on failure with the message in a thread-local string returned by
 * tide_last_error(); handles are opaque and not re-entrant.  Behind them the fit, gradients,
 * histograms, split search, row partition, TORB leaf statistics and predict run as gfx950 HIP
 * kernels.  A handle's device defaults to "gpu"; there is no CPU fallback: with no GPU present
 * tide_fit/tide_predict fail with a message that names the GPU (the reference's behaviour for
 * device="gpu" without GPU support, config.hpp:89-94).
 *
 * The extensions below are additive; the reference C API has no way to set the objective,
 * pair_splits, min_samples_leaf or device, nor to pass entity / rank_groups
 * (c_api.cpp:32-86).
 */
#ifndef TIDEGBM_AMD_H
#define TIDEGBM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef void* TideHandle;

/* ---- reference C API (c_api.h:9-28) ---- */
int tide_create(TideHandle* out);                                         /* c_api.h:11 */
void tide_free(TideHandle h);                                             /* c_api.h:12 */
int tide_set_int(TideHandle h, const char* key, int value);               /* c_api.h:15 */
int tide_set_double(TideHandle h, const char* key, double value);         /* c_api.h:16 */
int tide_fit(TideHandle h, const double* X, const double* y, int n, int n_features,
             const long long* timestamps, const double* weight, const int* is_cat,
             int iid, int horizon);                                       /* c_api.h:18-20 */
int tide_predict(TideHandle h, const double* X, int n, int n_features, double* out);     /* :22 */
int tide_predict_raw(TideHandle h, const double* X, int n, int n_features, double* out); /* :23 */
int tide_predict_binned(TideHandle h, const unsigned short* xb, int n, double* out);     /* :24 */
int tide_save_json(TideHandle h, const char* path);                      /* c_api.h:25 */
int tide_save(TideHandle h, const char* path);                           /* c_api.h:27 */
int tide_load(TideHandle* out, const char* path);                        /* c_api.h:28 */
const char* tide_last_error(void);                                        /* c_api.h:30 */

/* ---- extensions ---- */
/* 1 when the gfx950 kernels are compiled in (the reference's cuda_compiled flag,
 * bindings.cpp:324-328); the device may still be absent at run time. */
int tide_device_compiled(void);
/* Number of visible HIP devices (0 when none). */
int tide_device_count(void);
/* String keys: "device" ("gpu" | "cpu"; "cpu" is rejected at fit: this library has no CPU
 * path) and "objective" ("mse" | "pairwise" and the reference's aliases, bindings.cpp:108-119). */
int tide_set_string(TideHandle h, const char* key, const char* value);
/* tide_fit plus the entity and rank_groups arrays Booster::fit accepts (booster.hpp:45-48);
 * horizon in timestamp units as a 64-bit value.  Any pointer except X/y may be NULL. */
int tide_fit_ex(TideHandle h, const double* X, const double* y, int n, int n_features,
                const long long* timestamps, const double* weight, const int* is_cat,
                const long long* entity, const long long* rank_groups, int iid,
                long long horizon);
/* Booster::update (booster.cpp:844-884): grow extra_trees more trees on new rows. */
int tide_update(TideHandle h, const double* X, const double* y, int n, int n_features,
                const double* weight, int extra_trees);
/* Model inspection, node layout as in include/tgbm_device.h (10 ints + {value, gain}). */
int tide_num_trees(TideHandle h);
int tide_tree_info(TideHandle h, int tree, int* n_nodes, int* n_buckets);
int tide_tree_get(TideHandle h, int tree, int* node_ints, double* node_dbl, double* values,
                  int* cover, double* prefix_values);
int tide_model_info(TideHandle h, int* n_features, int* n_bins, int* n_time_buckets,
                    int* best_iteration, double* init_score);
int tide_get_edges(TideHandle h, int feature, double* out, int cap);
int tide_get_init_prefix(TideHandle h, double* out, int cap);
/* Per-kernel-family device time of the last fit (ms, launches, algorithmic bytes); family 0 =
 * root histograms, 5 = child histograms, 1 = partitions, 3 = tree end, 4 = gradients.
 * Collected only when the int key "profile" was set to 1 before the fit (2: as 1, and the
 * Infinity Cache is flushed before every root histogram build, outside its timed region). */
int tide_profile(TideHandle h, int family, double* ms, long long* launches, double* bytes);
/* Per-launch log of a profiled fit's kernel family: milliseconds and rows per launch, in launch
 * order.  Returns the number of launches (at most cap copied). */
int tide_profile_log(TideHandle h, int family, double* ms, double* rows, int cap);
/* Fit timing breakdown of the last fit: seconds for setup (sort, bins, uploads) and for the
 * boosting loop. */
int tide_fit_times(TideHandle h, double* setup_s, double* boost_s);
/* GOSS diagnostics of the last fit: trees whose sample was drawn on the host (the device sample
 * needs distinct |g| values past the kept top rows; a tree with equal ones is ordered by
 * std::sort on the host, as the reference orders it, booster.cpp:561-562), seconds the loop
 * waited for the generator draws, and seconds spent in the sample step (device or host). */
int tide_goss_stats(TideHandle h, int* host_trees, double* wait_s, double* sample_s);
/* Histogram kernel variant the last fit ended with: the tuner's choice (tide_set_int "autotune",
 * on by default for fits of 54 trees or more; the variants give identical trees), or the forced
 * "hist_mode". */
int tide_hist_variant(TideHandle h);
/* Training loss after each tree of the last fit, the reference's evals_result()["train"]
 * (booster.cpp:704-706, bindings.cpp:318-321).  Returns the number of trees (at most cap copied). */
int tide_evals_train(TideHandle h, double* out, int cap);
/* Feature groups of the following fits (Booster::fit's group_of, booster.hpp:45-48; -1 = none):
 * a node considers only the best-gain feature of each group (tree.cpp:135-158).  n_features = 0
 * clears them. */
int tide_set_feature_groups(TideHandle h, const int* group_of, int n_features);
/* Monotone constraints of the following fits (Booster::fit's monotone, booster.cpp:250-251; the
 * Python layer's monotone_constraints): per feature +1 (leaf weights must not decrease across a
 * split: left >= right), -1 (left <= right) or 0, checked by the split scan as in
 * best_numeric_missing (histogram.cpp:177-180).  n_features = 0 clears them. */
int tide_set_monotone(TideHandle h, const int* monotone, int n_features);
/* Interaction groups (Config::interaction_group, config.hpp:59; -1 = unconstrained): below a
 * split on a feature of group g, nodes split only on unconstrained features and those of g
 * (tree.cpp:304-317, 436-445).  n_features = 0 clears them.  With the int keys
 * "regime_feature" / "regime_priority" (tree.cpp:168-176) these run in the host-driven grower. */
int tide_set_interaction_groups(TideHandle h, const int* groups, int n_features);
/* Bin masks of tree `tree`'s categorical nodes (split_type 2): 4 words per node, bit b of word
 * b >> 6 set when bin b goes left (reference Node::cat_left, tree.hpp:24-30); zeros elsewhere. */
int tide_tree_cat(TideHandle h, int tree, unsigned long long* out);
/* Sorted category values of a categorical feature (reference Binner::cat_values, binning.cpp:48-82);
 * returns their number (0 for a numeric feature), at most cap copied. */
int tide_get_cat_values(TideHandle h, int feature, double* out, int cap);
/* Eval set of the following fits: the X_eval, y_eval, ts_eval, n_eval arguments of the reference's
 * Booster::fit (booster.hpp:45-48, booster.cpp:231-234), which its C API cannot pass.  X_eval is
 * n_eval x n_features row-major; ts_eval may be NULL in iid mode; n_eval = 0 clears the set.
 * Each tree is then scored on these rows on the GPU ("eval_metric" auto | loss | ic, tide_set_string)
 * and "early_stopping_rounds" (tide_set_int) applies as in booster.cpp:707-732. */
int tide_set_eval(TideHandle h, const double* X_eval, const double* y_eval, const long long* ts_eval,
                  int n_eval, int n_features);
/* Validation metric after each tree of the last fit, the reference's evals_result()["valid"].
 * Returns the number of values (at most cap copied). */
int tide_evals_valid(TideHandle h, double* out, int cap);
/* Wall time (s since the boosting loop began) at the end of each tree of the last fit; each
 * tree ends with a device synchronization.  Returns the number of trees. */
int tide_tree_times(TideHandle h, double* out, int cap);

/* Batched walk-forward fits (no counterpart in the reference C API; the reference refits every
 * fold x horizon of cv.py:45-53 with its own Booster::fit).  n_tasks fits share one binned
 * matrix: rows of X (n x n_features, row-major) must be sorted by (timestamps, entity); task t
 * trains on rows [0, task_rows[t]) with target y[t] (task_rows[t] values) and horizon
 * horizons[t] (NULL: the proto's horizon), every other setting taken from proto.  Bin edges
 * are fitted once over the largest task's rows.  schedule 0 grows all tasks' trees together
 * (one launch per kernel covers every task); schedule 1 fits the tasks one after another on the
 * same resident bins.  Both give identical trees.  out[t] receives a new fitted handle (free
 * each with tide_free).  Requires pair_splits=0 and goss off (goss=0, or every task under the
 * 250k-row GOSS threshold with goss=-1). */
int tide_fit_batch(TideHandle proto, const double* X, int n, int n_features,
                   const long long* timestamps, const long long* entity, int n_tasks,
                   const int* task_rows, const double* const* y, const long long* horizons,
                   int schedule, TideHandle* out);

/* ---- serving loop (no counterpart in the reference, whose request-level calls are
 * tide_predict and Booster::update, booster.cpp:801-884) ----
 * A server keeps ensembles and a binned panel resident on the GPU and serves asynchronous
 * predict and refit requests.  Batched mode coalesces every queued predict that fits the
 * predict_budget staging bytes into one launch and grows up to refit_slots refits together;
 * unbatched mode serves one request at a time in arrival order (baseline).  The ensembles must
 * be fitted MSE models with pair_splits=0 that share bin edges and growth settings (e.g. the
 * outputs of one tide_fit_batch); panel_X (n_rows x n_features, row-major) is binned once with
 * those edges.  Submissions return a ticket (>0) or -1 with tide_last_error() set. */
typedef void* TideServer;
int tide_server_create(TideServer* out, const TideHandle* models, int n_models,
                       const double* panel_X, int n_rows, int n_features,
                       long long predict_budget_bytes, int refit_slots, int refit_max_rows,
                       int batched);
void tide_server_free(TideServer s);
/* Score panel rows (ids into panel_X) with ensemble `model`. */
long long tide_server_predict(TideServer s, int model, const int* rows, int n);
/* Score raw feature rows (n x n_features); binned on the GPU with the shared edges. */
long long tide_server_predict_x(TideServer s, int model, const double* X, int n);
/* Booster::update on panel rows [row0, row0 + n) with targets y: extra_trees more trees,
 * published as the model's next version when done. */
long long tide_server_refit(TideServer s, int model, int row0, int n, const double* y,
                           int extra_trees);
/* Wait for a ticket: predict scores into out (cap values), the version used (predict) or
 * published (refit), submit-to-done seconds.  Returns the request's row count, -1 on error. */
int tide_server_wait(TideServer s, long long ticket, double* out, int cap, int* version,
                     double* latency_s);
/* New handle holding a version of a model (-1 = current) and that version's parent. */
int tide_server_model(TideServer s, int model, int version, TideHandle* out, int* parent);
/* Counters: predict launches, predict requests, predict rows, refit rounds, refits done; and
 * seconds: predict thread busy, refit thread busy, and the refit thread's admitting (base
 * predictions, slot binding, uploads), tree rounds and version publishing. */
int tide_server_stats(TideServer s, long long* counts5, double* seconds5);

#ifdef __cplusplus
}
#endif
#endif
