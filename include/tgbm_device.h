/* Thin C ABI between the C++ host booster (libtidegbm_amd.so) and the gfx950 HIP kernels
 * (libtgbm_hip.so).  Plain pointers, sizes and POD structs only; every call returns 0 on
 * success and -1 on failure with the message in tgbm_last_error().
 *
 * It replaces the reference's only device seam, build_hist_gpu
 * (reference cpp/include/tidegbm/histogram.hpp:56-58, cpp/src/histogram_cuda.cu:81-167), which
 * re-uploaded g/h/row indices and downloaded a histogram on every node.  Here a session keeps
 * the binned matrix, gradients, row partitions, histograms and tree state resident in HBM; the
 * host only exchanges split decisions and finished trees.
 *
 * A session holds one shared binned matrix (n_rows x n_features, uint8 bins) and n_tasks
 * independent fits ("tasks") over row prefixes [0, task_rows[t]) of it.  A plain fit is one task.
 */
#ifndef TGBM_DEVICE_H
#define TGBM_DEVICE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tgbm_session tgbm_session;

/* Per-node export layout shared by every tree accessor in this project:
 * ints[10] = {is_leaf, split_type, feature, feature2, threshold_bin, left, right,
 *             missing_go_left, n_samples, depth}; dbl[2] = {value, gain}.
 * Mirrors reference Node (cpp/include/tidegbm/tree.hpp:11-25). */
#define TGBM_NODE_INTS 10

typedef struct tgbm_grow_params {
    int max_depth;          /* reference Config::max_depth (config.hpp:15) */
    int max_leaves;         /* Config::max_leaves, <= 128 on device */
    int min_samples_leaf;   /* Config::min_samples_leaf */
    double min_child_weight;
    double l2, l1, max_delta_step, min_split_gain;
    double pair_l2;         /* Config::pair_l2, used by PAIR scans */
    double cat_smooth;      /* Config::cat_smooth (config.hpp:70), categorical ratio ranking */
    int hist_mode;          /* histogram kernel variant, 0 = defaults (all variants give the same
                               sums): bits 0-7 child tile (>= 16: 16 << (v-16)/8 features and
                               256 << (v-16)%8 minimum rows per block; 2: the row kernel also for
                               the root), bits 12-13 child partial cap (32, 64, 16, 8), bits
                               20-29 root kernel variant; chosen per fit by the host tuner */
    int host_driven;        /* 1: the host decides splits (tgbm_apply_split); 0: the device does */
} tgbm_grow_params;

/* One split decision made by the host (PAIR-capable driver). */
typedef struct tgbm_decision {
    int leaf;               /* node id to split, -1 = stop growing */
    int type;               /* 0 axis, 1 pair, 2 categorical, 3 regime (2 and 3: the bins of the
                               leaf's best split of categorical `feature`, as its scan found
                               them, go left) */
    int feature, feature2, threshold, missing_go_left;
    double gain;
} tgbm_decision;

/* Leaf candidate as cached by the grower (reference SplitCand, histogram.hpp:65-74). */
typedef struct tgbm_cand {
    double gain;
    int feature, feature2, threshold, type, missing_go_left, valid;
} tgbm_cand;

int tgbm_device_count(void);
const char* tgbm_last_error(void);

/* Session lifetime.  n_bins is the histogram width used by the PAIR hash (reference
 * Booster::n_bins, binning.cpp:130-134).  h_const=1 declares h == 1.0 for every row of every
 * task (unit-weight MSE without GOSS), which lets hessian sums be taken from counts. */
int tgbm_session_create(tgbm_session** out, int n_rows, int n_features, int n_bins, int n_tasks,
                        const int* task_rows, int max_leaves, int h_const);
void tgbm_session_destroy(tgbm_session* s);

/* Upload raw features (host row-major f64 n_rows x n_features) and bin them on the device
 * with per-feature right edges edges[f*256 + k], k < n_edges[f] (reference
 * Binner::transform, binning.cpp:104-128): NaN -> 0, else 1 + #edges <= x.  A categorical
 * feature passes its sorted category values as edges with TGBM_CAT_EDGES set in n_edges[f]:
 * a value bins to 1 + its index, an unlisted value or NaN to 0.  Every call that takes edges
 * (forests, eval sets) reads them this way. */
#define TGBM_CAT_EDGES 0x10000
int tgbm_session_bin(tgbm_session* s, const double* X, const double* edges, const int* n_edges);
/* Mark categorical features (is_cat[f] != 0; reference Booster::fit's is_cat, booster.cpp:246):
 * their split search becomes best_categorical (histogram.cpp:195-244) and tgbm_decision type 2
 * routes by the bin mask the leaf's scan found.  Once per session, before the first tree;
 * growth must then be host-driven. */
int tgbm_session_set_cat(tgbm_session* s, const int* is_cat);
/* Monotone constraints of the numeric split scan (reference Booster::fit's monotone,
 * booster.cpp:250-251; checked in best_numeric_missing, histogram.cpp:177-180): for feature f with
 * monotone[f] > 0 a candidate whose left leaf weight is below the right one is skipped, for
 * monotone[f] < 0 one whose left weight is above.  Once per session, before the first tree. */
int tgbm_session_set_monotone(tgbm_session* s, const int* monotone);
/* Leaves split per step of tgbm_tree_grow_device (0 or 1: one split per step, the reference's
 * loop, tree.cpp:343-463).  With width w > 1 a step splits the leaf best-first takes next and up
 * to w - 1 next-best leaves of the frontier; the tree is the best-first replay over the splits
 * made, so nodes, ids and values equal one-split-per-step growth (a split the replay does not
 * take is dropped).  Takes effect in one-task sessions with max_leaves <= 64; w <= 8. */
int tgbm_session_set_spec(tgbm_session* s, int width);
/* Bin masks (4 words per node, bins that go left) of the task's last exported tree, for the
 * nodes whose split_type is 2. */
int tgbm_task_cat_masks(tgbm_session* s, int task, int n_nodes, unsigned long long* out);
/* Restrict every later gradient, tree_begin/grow/end launch to tasks [first, first + count)
 * (all tasks after create).  Batched fits keep every task active; the one-fit-at-a-time
 * schedule activates one task at a time over the same resident bins.  The results of a task
 * do not depend on which other tasks share its launches. */
int tgbm_session_set_active(tgbm_session* s, int first, int count);
/* Rebind task `task` to rows [row0, row0 + n_rows) of the shared matrix (n_rows at most the
 * task's row count at creation); resets its sample, TORB layout (one bucket) and groups, so
 * tgbm_task_set_rows must follow.  Used to recycle task slots for refits of row windows. */
int tgbm_task_bind(tgbm_session* s, int task, int row0, int n_rows);
/* Exchange the state of tasks a and b (row windows, per-row arrays, trees in progress), so
 * the tasks a caller wants to grow can be kept in one contiguous active range. */
int tgbm_task_swap(tgbm_session* s, int a, int b);
/* Histogram of node `node` of the task's current (or last finished) tree, for checking: per
 * (feature, bin), F x 256 entries, the exact fixed-point sums the split scan used.  The sum of
 * g over the node's rows in the bin is qg * 2^-s_g; qh * 2^-s_h the sum of h (when h == 1 for
 * every row, qh holds the counts and s_h = 0); count the rows.  qh may be NULL.  Valid for the
 * nodes the tree has (built or derived as parent - sibling) until the next tgbm_tree_begin. */
int tgbm_task_node_hist(tgbm_session* s, int task, int node, long long* qg, int* count, long long* qh,
                        int* s_g, int* s_h);
/* Row-major uint8 bins back to the host (n_rows x n_features), for checking. */
int tgbm_session_get_bins(tgbm_session* s, unsigned char* out);

/* Per-task row data (host arrays of task_rows[t] entries).  w may be NULL (all 1); pred0 may be
 * NULL (keep the task's device predictions, e.g. from tgbm_task_predict_forest); bucket may be
 * NULL (keep the device buckets).  bucket and sweep_seg drive TORB: sweep_seg[i] = number of
 * bucket cuts <= i (0..n_buckets). */
int tgbm_task_set_rows(tgbm_session* s, int task, const double* y, const double* w,
                       const double* pred0, const int* bucket, const unsigned char* sweep_seg,
                       int n_buckets, const int* cuts, const long long* groups);

/* Gradients for the current predictions: objective 0 = MSE (objective.cpp:105-119),
 * 5 = within-group pairwise (objective.cpp:179-219). */
int tgbm_task_grad(tgbm_session* s, int task, int objective);
/* Copy g (and h) of a task back to the host (GOSS selection runs on the host). */
int tgbm_task_get_grad(tgbm_session* s, int task, double* g, double* h);
/* Restrict the next tree to sorted sample rows; rows with scaled[i]!=0 use g*scale, h*scale
 * (GOSS, booster.cpp:554-585).  sample == NULL restores "all rows". */
int tgbm_task_set_sample(tgbm_session* s, int task, const int* sample, int n_sample,
                         const unsigned char* scaled, double scale);

/* GOSS on the device (booster.cpp:554-585) for the next tree of `task`, from the gradients of
 * the last tgbm_task_grad: keep the ceil(top*S) rows of largest |g|, draw the n_other =
 * ceil(other*S) rows rest[perm[k]] from the remaining rows in |g|-descending order, scale their
 * g and h by (1-top)/other.  perm is the first n_other entries of the permutation std::shuffle
 * applies to the rest (it depends on the rest's size and the generator only).  When two rows
 * share a |g| value, *tie is set and nothing changes (order such rows with std::sort on the host
 * and call tgbm_task_set_sample). */
int tgbm_task_goss(tgbm_session* s, int task, double goss_top, double goss_other, const int* perm,
                   int n_perm, int* tie);
/* Start a tree for every task: quantize g/h, build the root histogram, scan it. */
int tgbm_tree_begin(tgbm_session* s, const tgbm_grow_params* p);
/* Device-driven best-first growth of every task's tree to completion (no host round trips). */
int tgbm_tree_grow_device(tgbm_session* s, const tgbm_grow_params* p);

/* Host-driven growth (one task), used when PAIR splits are enabled. */
int tgbm_leaf_scan_results(tgbm_session* s, int task, int node, double* gain_f, int* thr_f,
                           int* ml_f, double* energy, int* n_rows);
int tgbm_pair_search(tgbm_session* s, int task, int node, const int* f1, const int* f2,
                     int n_pairs, const tgbm_grow_params* p, tgbm_cand* out);
int tgbm_apply_split(tgbm_session* s, int task, const tgbm_decision* d,
                     const tgbm_grow_params* p, int* left_id);
/* tgbm_apply_split followed by tgbm_leaf_scan_results of both new children, with one wait: the
 * split of the reference's best-first loop (tree.cpp:445-456) and the candidates of its two
 * children (tree.cpp:457-461).  gain_f / thr_f / ml_f hold 2F entries (left child's F, then the
 * right's), energy and n_rows 2.  left_id = -1 when the decision named no splittable leaf. */
int tgbm_apply_split_fetch(tgbm_session* s, int task, const tgbm_decision* d,
                           const tgbm_grow_params* p, int* left_id, double* gain_f, int* thr_f,
                           int* ml_f, double* energy, int* n_rows);

/* Finish every active task's tree: leaf ids for all rows, leaf refit and TORB prefix values
 * (booster.cpp:589-648), pred += lr * prefix (booster.cpp:655-664).  Writes the tree of the
 * i-th active task into the caller's buffers at offset i*max_nodes (prefix: i*max_nodes*256). */
int tgbm_tree_end(tgbm_session* s, double learning_rate, double min_cover, int iid,
                  const tgbm_grow_params* p, int* n_nodes, int* node_ints, double* node_dbl,
                  double* values, int* cover, double* prefix);
/* tgbm_tree_end in two halves, so the host can enqueue the next tree before it reads this one:
 * tgbm_tree_end_async enqueues the tree end and the copy of the finished trees (at most two may
 * wait); tgbm_tree_collect waits for the oldest and unpacks it like tgbm_tree_end.  Between the
 * two only gradient, tree-begin, device-growth and further async tree-end calls may be made
 * (host-driven growth and the task readers need the pending trees collected first). */
int tgbm_tree_end_async(tgbm_session* s, double learning_rate, double min_cover, int iid,
                        const tgbm_grow_params* p);
int tgbm_tree_collect(tgbm_session* s, const tgbm_grow_params* p, int* n_nodes, int* node_ints,
                      double* node_dbl, double* values, int* cover, double* prefix);

int tgbm_task_get_pred(tgbm_session* s, int task, double* pred);
/* Training loss of `task` after its last finished tree (reference Objective::loss,
 * objective.cpp:262-329, recorded at booster.cpp:704-706): weighted mean squared error, or for
 * the pairwise objective minus the mean within-group Pearson correlation.  Computed on the device
 * from the updated predictions and exported with the tree. */
int tgbm_task_train_loss(tgbm_session* s, int task, double* loss);

/* Timers (HIP events on the session stream) summed per kernel family since the last reset:
 * which = 0 root histogram, 1 partition (with the split selection), 2 split scan, 3 tree end,
 * 4 gradients, 5 child histograms.  on = 2 also reads and rewrites a 512 MiB buffer before every
 * root histogram build (outside its timed region), so the root build's inputs come from HBM and
 * not from the 256 MB Infinity Cache. */
int tgbm_timer_enable(tgbm_session* s, int on);
int tgbm_timer_read(tgbm_session* s, int which, double* ms, long long* launches,
                    double* bytes);
/* Per-launch log of a family (launch order): milliseconds and the rows the launch covered
 * (root: the tree's sample; child histograms: the smaller child, 0 once the tree stopped
 * growing; -1 where unknown).  Returns the number of launches logged (at most cap copied). */
int tgbm_timer_log(tgbm_session* s, int which, double* ms, double* rows, int cap);

/* ---------------- predict / serving: resident ensembles ---------------- */
typedef struct tgbm_forest tgbm_forest;

/* Flatten trees (node arrays as in the export layout) into a device-resident forest.
 * cat_masks: 4 words per node (read for categorical nodes, split_type 2 or 3) or NULL.  The
 * batched predict kernel (tgbm_predict_batch, tgbm_task_predict_forest) refuses forests with
 * categorical nodes; tgbm_forest_predict routes them. */
int tgbm_forest_create(tgbm_forest** out, int n_features, int n_bins, int n_trees,
                       const int* tree_nodes, const int* node_ints, const double* values,
                       const double* lr, double init_score, const double* edges,
                       const int* n_edges, const unsigned long long* cat_masks);
void tgbm_forest_destroy(tgbm_forest* f);
/* New forest = the first keep_trees trees of `base` (copied on the device) followed by n_trees
 * new trees (same node layout as tgbm_forest_create); features, bins, edges and init score are
 * base's.  Lets a refit publish a version without re-uploading its parent's trees. */
int tgbm_forest_append(tgbm_forest** out, const tgbm_forest* base, int keep_trees, int n_trees,
                       const int* tree_nodes, const int* node_ints, const double* values,
                       const double* lr, const unsigned long long* cat_masks);
/* Predict raw scores for host rows X (row-major f64) or for device-resident uint8 bins. */
int tgbm_forest_predict(tgbm_forest* f, const double* X, int n, double* out);
int tgbm_forest_predict_binned_u16(tgbm_forest* f, const unsigned short* xb, int n, double* out);
/* tgbm_forest_predict that allocates nothing (no counterpart in the reference): for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same kernels and bit-identical scores; the staging of rows, bins and
 * scores lives in a fixed TGBM_NOALLOC_SCRATCH_BYTES array on each GPU, held until the process
 * ends, so n may be at most (TGBM_NOALLOC_SCRATCH_BYTES - 1040 - 64 fp) / (8 F + fp + 8) rows,
 * fp = F rounded up to 16 (72,307 rows at F = 128); larger n is rejected with an error.  Calls
 * on one GPU share the array and would race if they overlapped: the library runs them one at a
 * time.  The forest must be resident already (tgbm_forest_create). */
#define TGBM_NOALLOC_SCRATCH_BYTES (80u << 20)
int tgbm_forest_predict_noalloc(tgbm_forest* f, const double* X, int n, double* out);

/* ---- batched predict over many resident ensembles (serving) ----
 * A predictor owns a stream and pinned/device staging memory.  tgbm_predict_batch scores all
 * items in one kernel launch: each item names an ensemble and either rows of a session's
 * resident binned panel (rows != NULL, ids into the panel) or raw feature rows X (row-major
 * f64, binned with the ensemble's edges).  Scores are bit-identical to tgbm_forest_predict
 * (the reference's predict_raw_binned order, booster.cpp:744-773) whatever the batch holds. */
typedef struct tgbm_predictor tgbm_predictor;
typedef struct tgbm_predict_item {
    const tgbm_forest* forest;
    const int* rows;        /* panel row ids, or NULL */
    const double* X;        /* n x n_features raw rows when rows == NULL */
    int n;
    double* out;            /* n scores (host) */
} tgbm_predict_item;
int tgbm_predictor_create(tgbm_predictor** out, size_t staging_bytes);
void tgbm_predictor_destroy(tgbm_predictor* p);
/* Staging bytes one item needs (admission accounting); a batch needs their sum plus 256 per item. */
size_t tgbm_predict_bytes(const tgbm_forest* f, int from_x, int n);
int tgbm_predict_batch(tgbm_predictor* p, const tgbm_session* panel, const tgbm_predict_item* items,
                       int n_items);
/* Overwrite task `task`'s device predictions with forest f's raw scores of the task's panel row
 * window (the same sums as tgbm_predict_batch), on the session stream: a refit's starting
 * predictions without a host round trip.  Call after tgbm_task_bind. */
int tgbm_task_predict_forest(tgbm_session* s, int task, const tgbm_forest* f);
/* One call for a refit slot (reference Booster::update on a row window, booster.cpp:844-884):
 * the effect of tgbm_task_bind(task, row0, n_rows), tgbm_task_predict_forest(task, f) and
 * tgbm_task_set_rows(task, y, NULL, NULL, NULL, NULL, 1, NULL, NULL), all queued on the session
 * stream with no host wait (y is staged through pinned host memory of the session, so the
 * caller may reuse it on return).  No reference counterpart: the serving loop's admission path. */
int tgbm_task_refit_window(tgbm_session* s, int task, int row0, int n_rows, const double* y,
                           const tgbm_forest* f);

/* ---- validation set scored tree by tree during a fit (reference Booster::fit with X_eval,
 * booster.cpp:459-478 and :707-732) ----
 * The rows are binned once with the fit's edges (edges[f*256 + k], k < n_edges[f]).  A device
 * prediction per row accumulates init + lr * values[leaf] tree by tree, the running sum of
 * predict_multi_binned (booster.cpp:775-799).  For the per-date correlation metric, order lists
 * the rows of every date of at least 3 rows, date after date, and date g is
 * order[group_start[g] .. group_start[g+1]) (grouped_ic, booster.cpp:155-174); n_groups may be 0. */
typedef struct tgbm_evalset tgbm_evalset;
int tgbm_eval_create(tgbm_evalset** out, int n_features, int n_bins, const double* edges,
                     const int* n_edges, const double* X, const double* y, int n,
                     const int* order, const int* group_start, int n_groups);
void tgbm_eval_destroy(tgbm_evalset* e);
/* Every row's prediction = init_score. */
int tgbm_eval_reset(tgbm_evalset* e, double init_score);
/* Add one tree (export node layout, TGBM_NODE_INTS ints per node; values per node) times lr. */
int tgbm_eval_add_tree(tgbm_evalset* e, int n_nodes, const int* node_ints, const double* values,
                       double lr, const unsigned long long* cat_masks);
/* metric 0: mean squared error, unit weights, rows summed in order (objective.cpp:323-328);
 * 1: mean over the listed dates of the Pearson correlation of y and prediction (0 without dates). */
int tgbm_eval_metric(tgbm_evalset* e, int metric, double* out);
/* The current predictions (first cap rows). */
int tgbm_eval_predictions(tgbm_evalset* e, double* out, int cap);

#ifdef __cplusplus
}
#endif
#endif
