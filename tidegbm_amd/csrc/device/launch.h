// Host-side launch wrappers of the fit and predict kernels (defined next to the kernels).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include "dev_common.h"

struct tgbm_session;

namespace tgbm {

// Resident binned panel of a session (row-major bins, fp bytes per row, n rows, F features).
void session_panel_info(const tgbm_session* s, const unsigned char** bins_row, int* fp, int* n, int* F);
// A task's device predictions, row window and the session stream, plus a device scratch of at
// least scratch_bytes owned by the session.
void session_task_window(tgbm_session* s, int task, double** pred, int* row0, int* n_rows, hipStream_t* stream,
                         void** scratch, size_t scratch_bytes);
// Rebind a task to a refit window with its targets (one time bucket, no weights, no groups),
// queued on the session stream without a wait; returns the task's prediction array and the
// stream for the window's base predictions.
void session_refit_window(tgbm_session* s, int task, int row0, int n_rows, const double* y, double** pred,
                          hipStream_t* stream);

void init_kernel_attributes();
void launch_bin(hipStream_t st, const double* X, int n, int F, const double* edges,
                const int* n_edges, unsigned char* bins_row, int fp, unsigned char* bins_col,
                int n_pad);
void launch_grad(hipStream_t st, Task* tasks, int n_tasks, int objective, int max_rows,
                 int max_groups);
void launch_tree_start(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, int max_rows,
                       int max_nodes);
// e0/e1 (optional): events recorded at the kernel's start and end.
void launch_hist(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, bool root,
                 bool contiguous_root, bool hconst, int max_count, int variant,
                 hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// mode 0: the root; 1: the smaller child of T.cur_* and its sibling; 2: every slot of a
// multi-leaf step (partials of slot j from partial blk0_j; hconst: h == 1 for every task).
void launch_scan(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, const GrowP& P, int mode,
                 bool hconst = true);
// Child-build tiles of a histogram variant: lanes per row, minimum rows per block, partials cap.
void child_tiles(int variant, bool hconst, int& lpr, int& min_rows, int& pcap);
// Multi-leaf growth: the step's selection (histogram tiles of at least min_per rows, at most
// rb_target row blocks and step_parts partials per task step), its child builds (a fixed grid of
// `blocks` blocks, fbs feature blocks per row block), and the tree's final numbering after the
// last step.
void launch_spec_select(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, const GrowP& P, int width,
                        int min_per, int rb_target, int step_parts);
void spec_child_shape(bool hconst, int n_cu, int fp, int& blocks, int& fbs);
void launch_hist_spec(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, bool hconst, int blocks,
                      hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_spec_finalize(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, const GrowP& P);
// out (device-visible) = {every active task's growth done, the most steps a task took}.
void launch_spec_status(hipStream_t st, const Task* tasks, int n_tasks, int* out);
// One block per task selects the next split into T.cur_* (for launches over many tasks).
void launch_select(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, const GrowP& P);
// mode 0: the partition launch selects the split itself; mode 1: the decision is in T.cur_*;
// mode 2: the slots of a multi-leaf step (max_count: the task's sample rows).
void launch_partition(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, int max_count, const GrowP& P,
                      int mode);
void launch_energy(hipStream_t st, Task* tasks, int task, int node, int count);
void launch_pair(hipStream_t st, const Shared& S, Task* tasks, int task, int node, int count,
                 const int* pf, int n_pairs, long long* ph_g, int* ph_c, long long* ph_h,
                 bool hconst, const GrowP& P, FCand* out);
void launch_tree_end(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, bool walk,
                     int max_rows, const TorbChunk* const* chunks, const int* n_chunks,
                     int max_chunks, const GrowP& P, double min_cover, int iid, double lr, int pairwise);
void launch_export(hipStream_t st, Task* tasks, int n_tasks, const ExportLayout& L, char* out, int pairwise);
// Group correlations of the pairwise training loss from the updated predictions (k_export
// finishes the loss; the mean squared error comes from the prediction update).
void launch_loss(hipStream_t st, Task* tasks, int n_tasks, int max_groups);
// GOSS selection (see goss.hip): temp storage bytes for n rows, and the launch sequence.
size_t goss_temp_bytes(int n);
hipError_t launch_goss(hipStream_t st, const double* g, int n, int n_top, int n_other, const int* perm,
                       unsigned long long* k0, unsigned long long* k1, int* v0, int* v1, unsigned char* keep,
                       unsigned char* scaled, int* sample_out, int* n_sample_out, int* tie, void* tmp,
                       size_t tmp_bytes);

}  // namespace tgbm
