// Thin C ABI (include/tgbm_device.h) over the gfx950 fit kernels: device memory ownership,
// kernel sequencing for host-driven and device-driven best-first growth, and tree export.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <deque>
#include <map>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/tgbm_device.h"
#include "dev_common.h"
#include "launch.h"

using namespace tgbm;

thread_local std::string g_dev_err;  // shared by every entry point of this library

#define HIP_OK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) +   \
                                     " at " #expr);                                          \
    } while (0)

template <class Fn>
static int dev_guard(Fn&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception& e) {
        g_dev_err = e.what();
        return -1;
    }
}

template <class T>
static T* dalloc(size_t n) {
    if (n == 0) n = 1;
    void* p = nullptr;
    HIP_OK(hipMalloc(&p, n * sizeof(T)));
    return static_cast<T*>(p);
}

struct TaskHost {
    int n_rows = 0;
    int row0 = 0;
    int cap_rows = 0;
    int n_buckets = 1;
    std::vector<TorbChunk> chunks;
    TorbChunk* chunks_d = nullptr;   // device copy of chunks (capacity chunks_cap)
    int chunks_cap = 0;
    double* w_d = nullptr;           // weights (allocated on first use, cap_rows)
    // device GOSS buffers (allocated on first use, cap_rows)
    unsigned long long* gk0 = nullptr;
    unsigned long long* gk1 = nullptr;
    int* gv0 = nullptr;
    int* gv1 = nullptr;
    unsigned char* gkeep = nullptr;
    int* gperm = nullptr;
    int* gsmall = nullptr;           // [0] selected count, [1] tie flag
    void* gtmp = nullptr;
    size_t gtmp_bytes = 0;
    int* gstart_d = nullptr;         // pairwise groups (allocated on first use)
    int* grows_d = nullptr;
    int n_groups = 0;
    double* loss_d = nullptr;        // loss partial sums (allocated on first use, loss_cap)
    int loss_cap = 0;
    double last_loss = 0.0;          // training loss after the last exported tree
    bool has_sample = false;
    int n_sample = 0;
    bool goss = false;
    int max_group = 0;
};

struct EventPair {
    hipEvent_t a, b;
    int family;
    double bytes;
};

struct tgbm_session {
    int n = 0, F = 0, fp = 0, n_pad = 0, nb = 255, T = 1, max_leaves = 31, max_nodes = 61;
    int a0 = 0, an = 1;     // active tasks [a0, a0 + an): every fit launch covers exactly these
    bool hconst = true;
    int objective = 0;      // objective of the last gradient launch (the tree end's loss)
    hipStream_t stream = nullptr;
    Shared S{};
    unsigned char* bins_row = nullptr;
    unsigned char* bins_col = nullptr;
    Task* tasks_d = nullptr;
    std::vector<Task> tasks_h;      // pointer mirror (scalars are device-owned)
    std::vector<TaskHost> th;
    std::vector<void*> allocs;
    // TORB chunk lists (device): array of per-task pointers
    TorbChunk** chunk_ptrs_d = nullptr;
    int* n_chunks_d = nullptr;
    int max_chunks = 1;
    // PAIR scratch
    int* pair_f_d = nullptr;
    long long* ph_g = nullptr;
    int* ph_c = nullptr;
    long long* ph_h = nullptr;
    // host-driven growth: a mapped pinned buffer the device writes a split's results into (one
    // wait per split), and the row count of every node grown so far (-1: unknown)
    char* rb_h = nullptr;
    char* rb_d = nullptr;
    std::vector<int> node_rows_h;
    void ensure_readback() {
        if (rb_h) return;
        // room for two nodes' candidates, or a PAIR search's (at most 64 pairs)
        HIP_OK(hipHostMalloc((void**)&rb_h, 40 + sizeof(FCand) * std::max(2 * (size_t)F, (size_t)64),
                             hipHostMallocMapped));
        HIP_OK(hipHostGetDevicePointer((void**)&rb_d, rb_h, 0));
    }
    // captured device-driven split loops (one graph launch per tree), keyed by what the capture
    // baked in: largest sample, leaf budget, histogram variant and the active task range
    struct CachedGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        long long last_use = 0;
    };
    std::map<std::vector<long long>, CachedGraph> graphs;
    long long graph_clock = 0;
    bool use_graphs = true;
    // pinned staging for refit windows' targets: a ring of kStage buffers of the largest task
    // capacity, allocated together on first use; an event marks the last copy out of each
    static constexpr int kStage = 4;
    double* stage[kStage] = {};
    hipEvent_t stage_ev[kStage] = {};
    int stage_next = 0;
    bool spec_graph = false;    // multi-leaf loops from a captured graph instead of direct launches
    // multi-leaf growth (one-task sessions with max_leaves <= 64): leaves per step (0: off), the
    // node capacity and partial slots it was allocated with, the steps of the launched loop (the
    // most splitting steps a recent tree needed), whether the current tree grows this way, and a pinned
    // copy of the tree's {done, steps}
    int n_cu = 256;         // compute units of the device
    int spec_width = 0;
    int spec_cap = 0;
    int step_parts = 0;     // partials one task's multi-leaf step may write (all slots)
    int max_task_rows = 1;
    int spec_k = 8;
    bool spec_tree = false;
    int* flag_h = nullptr;      // pinned, mapped: per task {status, steps} from the final numbering,
                                // then {every task done, most steps} from the status kernel
    int* flag_d = nullptr;      // its device address
    bool spec_launched = false;     // the current tree's captured loop was launched without a wait
    bool spec_wait = false;         // a multi-leaf tree's queued end has not been waited for
    int spec_k_used = 0;            // the steps of that tree's captured loop
    int spec_hist[32] = {0};        // steps the last trees needed (ring)
    int spec_hist_n = 0;
    struct EndArgs {                // what a tree end was queued with
        GrowP P;
        tgbm_grow_params p;
        double lr, min_cover;
        int iid, max_groups, slot;
        bool walk, pairwise, spec;
        ExportLayout L;
    };
    EndArgs spec_end{};             // the unresolved multi-leaf tree's end
    // small device scratch for descriptor tables of launches made on behalf of other modules
    char* small_scratch = nullptr;
    size_t small_scratch_bytes = 0;
    // finished-tree staging (device + pinned host), sized for every task at the largest bucket count
    // two mapped host slots (a tree is read while the next runs), written by k_export directly
    // through their device addresses export_hd
    char* export_h[2] = {nullptr, nullptr};
    char* export_hd[2] = {nullptr, nullptr};
    hipEvent_t export_ev[2] = {nullptr, nullptr};
    size_t export_bytes = 0;
    struct PendingExport {
        int slot, a0, an;
        ExportLayout L;
        bool spec;   // grown by multi-leaf steps (their rows were logged per step)
    };
    std::deque<PendingExport> pending;       // enqueued tree ends not yet collected (at most 2)
    int next_slot = 0;
    // timers
    bool timing = false;
    // timing mode 2: a 512 MiB buffer read and rewritten before every root histogram build, so
    // the build's inputs come from HBM and not from the 256 MB Infinity Cache
    bool flush = false;
    unsigned long long* flush_buf = nullptr;
    static constexpr size_t kFlushBytes = size_t(512) << 20;
    std::vector<EventPair> events;
    double t_ms[8] = {0};
    long long t_n[8] = {0};
    double t_bytes[8] = {0};
    std::vector<double> log_ms[8], log_rows[8];   // per launch, in launch order

    template <class T>
    T* alloc(size_t n) {
        T* p = dalloc<T>(n);
        allocs.push_back(p);
        return p;
    }
    void release(void* p) {
        for (auto it = allocs.begin(); it != allocs.end(); ++it)
            if (*it == p) {
                allocs.erase(it);
                HIP_OK(hipFree(p));
                return;
            }
    }
    ~tgbm_session() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (flush_buf) (void)hipFree(flush_buf);
        if (flag_h) (void)hipHostFree(flag_h);
        if (rb_h) (void)hipHostFree(rb_h);
        for (auto& kv : graphs) {
            (void)hipGraphExecDestroy(kv.second.ge);
            (void)hipGraphDestroy(kv.second.g);
        }
        for (auto& e : events) {
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
        for (int k = 0; k < kStage; ++k) {
            if (stage[k]) (void)hipHostFree(stage[k]);
            if (stage_ev[k]) (void)hipEventDestroy(stage_ev[k]);
        }
        for (void* p : allocs) (void)hipFree(p);
        for (int k = 0; k < 2; ++k) {
            if (export_h[k]) (void)hipHostFree(export_h[k]);
            if (export_ev[k]) (void)hipEventDestroy(export_ev[k]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }
    void set_field_int(int t, size_t off, int v) {
        HIP_OK(hipMemcpyAsync(reinterpret_cast<char*>(tasks_d + t) + off, &v, sizeof v, hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }
    void set_field_double(int t, size_t off, double v) {
        HIP_OK(hipMemcpyAsync(reinterpret_cast<char*>(tasks_d + t) + off, &v, sizeof v, hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }
    void set_field_ptr(int t, size_t off, const void* v) {
        HIP_OK(hipMemcpyAsync(reinterpret_cast<char*>(tasks_d + t) + off, &v, sizeof v, hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }
    // TORB (segment, node) sums of task t to zero (on the session stream).
    void clear_torb(int t) {
        const Task& T = tasks_h[t];
        const size_t tw = (size_t)(MAX_BUCKETS + 1) * MAX_NODES;
        HIP_OK(hipMemsetAsync(T.tg, 0, sizeof(long long) * tw, stream));
        if (T.th) HIP_OK(hipMemsetAsync(T.th, 0, sizeof(long long) * tw, stream));
        HIP_OK(hipMemsetAsync(T.tc, 0, sizeof(int) * tw, stream));
    }
    Task read_task(int t) {
        Task out;
        HIP_OK(hipMemcpyAsync(&out, tasks_d + t, sizeof(Task), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return out;
    }
    GrowP growp(const tgbm_grow_params* p) const {
        GrowP g;
        g.max_depth = p->max_depth;
        g.max_leaves = p->max_leaves;
        g.min_leaf = p->min_samples_leaf;
        g.min_hess = p->min_child_weight;
        g.l2 = p->l2;
        g.l1 = p->l1;
        g.mds = p->max_delta_step;
        g.gamma = p->min_split_gain;
        g.pair_l2 = p->pair_l2;
        g.cat_smooth = p->cat_smooth;
        return g;
    }
    Task* td() const { return tasks_d + a0; }   // first active task (device)
    int max_rows() const {
        int m = 0;
        for (int t = a0; t < a0 + an; ++t) m = std::max(m, th[t].n_rows);
        return m;
    }
    // Grid sizing row count of the split loop: the largest active sample rounded up to a power of
    // two, so captured loops are shared by every task mix of similar size (surplus blocks exit).
    int launch_rows() const {
        const int m = max_sample();
        int r = 1;
        while (r < m && r < (1 << 30)) r <<= 1;
        return r;
    }
    int max_sample() const {
        int m = 0;
        for (int t = a0; t < a0 + an; ++t) m = std::max(m, th[t].has_sample ? th[t].n_sample : th[t].n_rows);
        return m;
    }
    bool all_contiguous() const {
        for (int t = a0; t < a0 + an; ++t)
            if (th[t].has_sample) return false;
        return true;
    }
    // Upload task t's TORB chunk list and its entry in the per-task pointer table.
    void upload_chunks(int t) {
        TaskHost& H = th[t];
        const int m = std::max<int>(1, (int)H.chunks.size());
        if (m > H.chunks_cap) {
            H.chunks_d = alloc<TorbChunk>(m);
            H.chunks_cap = m;
        }
        HIP_OK(hipMemcpy(H.chunks_d, H.chunks.data(), sizeof(TorbChunk) * H.chunks.size(), hipMemcpyHostToDevice));
        const int cnt = (int)H.chunks.size();
        HIP_OK(hipMemcpy(chunk_ptrs_d + t, &H.chunks_d, sizeof(TorbChunk*), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(n_chunks_d + t, &cnt, sizeof(int), hipMemcpyHostToDevice));
        max_chunks = 1;
        for (auto& x : th) max_chunks = std::max(max_chunks, (int)x.chunks.size());
    }
    // -------- timers
    void tic(int family, double bytes, EventPair& ep) {
        ep.family = family;
        ep.bytes = bytes;
        HIP_OK(hipEventCreate(&ep.a));
        HIP_OK(hipEventCreate(&ep.b));
        HIP_OK(hipEventRecord(ep.a, stream));
    }
    void toc(EventPair& ep) {
        HIP_OK(hipEventRecord(ep.b, stream));
        events.push_back(ep);
    }
    // events that an extended launch records at the kernel's own start and end
    void kernel_events(int family, double bytes, EventPair& ep) {
        ep.family = family;
        ep.bytes = bytes;
        HIP_OK(hipEventCreate(&ep.a));
        HIP_OK(hipEventCreate(&ep.b));
    }
    void drain_events() {
        if (events.empty()) return;
        HIP_OK(hipStreamSynchronize(stream));
        for (auto& e : events) {
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, e.a, e.b));
            t_ms[e.family] += ms;
            t_n[e.family] += 1;
            t_bytes[e.family] += e.bytes;
            log_ms[e.family].push_back(ms);
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
        events.clear();
    }
};

static void spec_resolve(tgbm_session* s);

namespace tgbm {
void session_panel_info(const tgbm_session* s, const unsigned char** bins_row, int* fp, int* n, int* F) {
    *bins_row = s->bins_row;
    *fp = s->fp;
    *n = s->n;
    *F = s->F;
}

void session_task_window(tgbm_session* s, int task, double** pred, int* row0, int* n_rows, hipStream_t* stream,
                         void** scratch, size_t scratch_bytes) {
    if (task < 0 || task >= s->T) throw std::runtime_error("task out of range");
    *pred = s->tasks_h[task].pred;
    *row0 = s->th[task].row0;
    *n_rows = s->th[task].n_rows;
    *stream = s->stream;
    if (scratch_bytes > s->small_scratch_bytes) {
        s->small_scratch = s->alloc<char>(scratch_bytes);
        s->small_scratch_bytes = scratch_bytes;
    }
    *scratch = s->small_scratch;
}

// A GOSS sample's fields of the task descriptor, set in stream order from the selection's result
// (gsmall = {rows selected, equal-|g| flag}).  Equal |g| values or an unexpected row count leave
// them unchanged: the host then orders the rows itself (ties) or reports the error.
__global__ void k_goss_commit(Task* T, const int* gsmall, int expect, double scale) {
    if (threadIdx.x != 0 || gsmall[1] != 0 || gsmall[0] != expect) return;
    T->n_sample = gsmall[0];
    T->contiguous = 0;
    T->goss = 1;
    T->goss_scale = scale;
}

// Device side of a refit window (tgbm_task_bind + tgbm_task_set_rows with one time bucket, no
// weights and no groups): the descriptor's window fields, zero bucket ids and TORB sums, the
// sweep-chunk list of 2,048-row chunks in segment 1 and its table entry.
__global__ __launch_bounds__(256) void k_task_window(Task* tasks, int task, int row0, int n_rows, TorbChunk* chunks,
                                                     int n_chunks, TorbChunk** chunk_ptrs, int* chunk_counts) {
    Task& T = tasks[task];
    const int gt = blockIdx.x * blockDim.x + threadIdx.x, gs = gridDim.x * blockDim.x;
    const size_t tw = (size_t)(MAX_BUCKETS + 1) * MAX_NODES;
    int* const bucket = (int*)T.bucket;
    for (int i = gt; i < n_rows; i += gs) bucket[i] = 0;
    for (size_t i = gt; i < tw; i += gs) {
        T.tg[i] = 0;
        if (T.th) T.th[i] = 0;
        T.tc[i] = 0;
    }
    for (int k = gt; k < n_chunks; k += gs)
        chunks[k] = TorbChunk{k * 2048, min(n_rows, (k + 1) * 2048), 1, 0};
    if (gt == 0) {
        T.n_rows = n_rows;
        T.row0 = row0;
        T.n_sample = n_rows;
        T.contiguous = 1;
        T.goss = 0;
        T.n_buckets = 1;
        T.has_groups = 0;
        T.w = nullptr;
        ((int*)T.cuts)[0] = 0;
        chunk_ptrs[task] = chunks;
        chunk_counts[task] = n_chunks;
    }
}

// Host side of a refit window: the task's bookkeeping, the device fields (k_task_window) and the
// targets (through the task's pinned staging), all queued on the session stream without a wait.
// Returns the task's prediction array and the stream, for the caller to fill the window's base
// predictions behind them.
void session_refit_window(tgbm_session* s, int task, int row0, int n_rows, const double* y, double** pred,
                          hipStream_t* stream) {
    spec_resolve(s);
    if (task < 0 || task >= s->T) throw std::runtime_error("task out of range");
    TaskHost& H = s->th[task];
    if (n_rows < 1 || n_rows > H.cap_rows) throw std::runtime_error("task rows exceed the task's capacity");
    if (row0 < 0 || (long long)row0 + n_rows > s->n) throw std::runtime_error("task row window outside the shared matrix");
    if (!s->pending.empty()) throw std::runtime_error("collect the pending tree ends first");
    const int n_chunks = (n_rows + 2047) / 2048;
    if (n_chunks > H.chunks_cap) {
        HIP_OK(hipStreamSynchronize(s->stream));   // the old list may still be read
        H.chunks_d = s->alloc<TorbChunk>(n_chunks);
        H.chunks_cap = n_chunks;
    }
    if (!s->stage[0]) {
        int cap = 1;
        for (auto& x : s->th) cap = std::max(cap, x.cap_rows);
        for (int k = 0; k < tgbm_session::kStage; ++k) {
            HIP_OK(hipHostMalloc((void**)&s->stage[k], sizeof(double) * cap, hipHostMallocDefault));
            HIP_OK(hipEventCreateWithFlags(&s->stage_ev[k], hipEventDisableTiming));
        }
    }
    const int sk = s->stage_next;
    s->stage_next = (sk + 1) % tgbm_session::kStage;
    double* const ypin = s->stage[sk];
    HIP_OK(hipEventSynchronize(s->stage_ev[sk]));   // that buffer's last copy has run
    H.n_rows = n_rows;
    H.row0 = row0;
    H.has_sample = false;
    H.n_sample = n_rows;
    H.goss = false;
    H.n_buckets = 1;
    H.chunks.clear();
    for (int k = 0; k < n_chunks; ++k) H.chunks.push_back(TorbChunk{k * 2048, std::min(n_rows, (k + 1) * 2048), 1, 0});
    s->max_chunks = 1;
    for (auto& x : s->th) s->max_chunks = std::max(s->max_chunks, (int)x.chunks.size());
    Task& T = s->tasks_h[task];
    T.n_rows = n_rows;
    T.row0 = row0;
    T.w = nullptr;
    const int blocks = std::min(64, (std::max(n_rows, (MAX_BUCKETS + 1) * MAX_NODES) + 255) / 256);
    hipLaunchKernelGGL(k_task_window, dim3(blocks), dim3(256), 0, s->stream, s->tasks_d, task, row0, n_rows, H.chunks_d,
                       n_chunks, s->chunk_ptrs_d, s->n_chunks_d);
    HIP_OK(hipGetLastError());
    std::memcpy(ypin, y, sizeof(double) * n_rows);
    HIP_OK(hipMemcpyAsync((void*)T.y, ypin, sizeof(double) * n_rows, hipMemcpyHostToDevice, s->stream));
    HIP_OK(hipEventRecord(s->stage_ev[sk], s->stream));
    *pred = T.pred;
    *stream = s->stream;
}
}  // namespace tgbm

static void spec_resolve(tgbm_session* s);

// One node's readback (layout of k_split_readback, the node in the L slot): its scan results,
// energy, row count and the task's energy scale.
__global__ __launch_bounds__(256) void k_node_readback(const Task* tasks, int task, int node, int F, char* out) {
    const Task& T = tasks[task];
    int* hdr = (int*)out;
    long long* en = (long long*)(out + 24);
    FCand* fc = (FCand*)(out + 40);
    if (threadIdx.x == 0) {
        hdr[0] = 0;
        hdr[1] = node;
        hdr[2] = T.s_e;
        hdr[3] = T.nodes[node].v[N_NS];
        hdr[4] = 0;
        en[0] = T.energy[node];
        en[1] = 0;
    }
    for (int j = threadIdx.x; j < F; j += blockDim.x) fc[j] = T.fcand[(size_t)node * F + j];
}

extern "C" {

const char* tgbm_last_error(void) { return g_dev_err.c_str(); }

int tgbm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        g_dev_err = "no HIP GPU device available";
        return 0;
    }
    return n;
}

int tgbm_session_create(tgbm_session** out, int n_rows, int n_features, int n_bins, int n_tasks,
                        const int* task_rows, int max_leaves, int h_const) {
    *out = nullptr;
    return dev_guard([&] {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
            throw std::runtime_error("GPU device not available: no HIP device found");
        if (n_rows <= 0 || n_features <= 0) throw std::runtime_error("empty input");
        if (n_features > MAX_FEATURES) throw std::runtime_error("too many features for the device path");
        if (n_bins < 2 || n_bins > NBIN) throw std::runtime_error("n_bins must be in [2, 256] on the device path");
        if (max_leaves < 2 || 2 * max_leaves - 1 > MAX_NODES)
            throw std::runtime_error("max_leaves must be in [2, 128] on the device path");
        if (n_tasks < 1) throw std::runtime_error("n_tasks must be >= 1");
        auto* s = new tgbm_session();
        std::unique_ptr<tgbm_session> guard(s);
        s->n = n_rows;
        s->F = n_features;
        s->fp = (n_features + 15) / 16 * 16;
        s->n_pad = (n_rows + 63) / 64 * 64 + 64;
        s->nb = n_bins;
        s->T = n_tasks;
        s->a0 = 0;
        s->an = n_tasks;
        s->max_leaves = max_leaves;
        s->max_nodes = 2 * max_leaves - 1;
        s->hconst = h_const != 0;
        if (const char* e = std::getenv("TGBM_NO_GRAPH")) s->use_graphs = !(e[0] == '1');
        if (const char* e = std::getenv("TGBM_SPEC_GRAPH")) s->spec_graph = e[0] == '1';
        init_kernel_attributes();
        {
            int dev = 0, cu = 0;
            HIP_OK(hipGetDevice(&dev));
            if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cu > 0) s->n_cu = cu;
        }
        HIP_OK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->bins_row = s->alloc<unsigned char>((size_t)n_rows * s->fp);
        s->bins_col = s->alloc<unsigned char>((size_t)n_features * s->n_pad);
        HIP_OK(hipMemsetAsync(s->bins_row, 0, (size_t)n_rows * s->fp, s->stream));
        HIP_OK(hipMemsetAsync(s->bins_col, 0, (size_t)n_features * s->n_pad, s->stream));
        s->S.n_rows = n_rows;
        s->S.n_features = n_features;
        s->S.fp = s->fp;
        s->S.n_pad = s->n_pad;
        s->S.n_bins = n_bins;
        s->S.bins_row = s->bins_row;
        s->S.bins_col = s->bins_col;
        int max_task_rows = 1;
        for (int t = 0; t < n_tasks; ++t) max_task_rows = std::max(max_task_rows, task_rows ? task_rows[t] : n_rows);
        // Partials per launch: launches of the split loop are sized for the largest sample rounded
        // up to a power of two (< 2x the rows), and no block may fold more than MAX_BLOCK_ROWS rows.
        s->S.part_cap = std::max(64, (int)((2LL * max_task_rows + MAX_BLOCK_ROWS - 1) / MAX_BLOCK_ROWS) + 1);
        s->max_task_rows = max_task_rows;
        const int node_cap = s->max_nodes;
        s->tasks_h.assign(n_tasks, Task{});
        s->th.assign(n_tasks, TaskHost{});
        const size_t per_node = (size_t)n_features * NBIN;
        for (int t = 0; t < n_tasks; ++t) {
            const int m = task_rows ? task_rows[t] : n_rows;
            if (m <= 0 || m > n_rows) throw std::runtime_error("task rows out of range");
            Task& T = s->tasks_h[t];
            std::memset(&T, 0, sizeof T);
            s->th[t].n_rows = m;
            s->th[t].cap_rows = m;
            T.n_rows = m;
            T.row0 = 0;
            T.cap_rows = m;
            T.n_buckets = 1;
            T.hconst = s->hconst;
            T.n_sample = m;
            T.contiguous = 1;
            T.goss_scale = 1.0;
            T.y = s->alloc<double>(m);
            T.w = nullptr;
            T.pred = s->alloc<double>(m);
            T.g = s->alloc<double>(m);
            T.h = s->alloc<double>(m);
            T.pk = s->alloc<unsigned long long>(m);
            T.qh = s->hconst ? nullptr : s->alloc<unsigned long long>(m);
            T.bucket = s->alloc<int>(m);
            T.nid = s->alloc<int>(m);
            T.idx = s->alloc<int>(m);
            T.idx_tmp = s->alloc<int>(m);
            T.chunk_cnt = s->alloc<int>(m / 2048 + 2);
            T.red = s->alloc<unsigned long long>(2 * MAX_RED_BLOCKS);
            T.nodes = s->alloc<DNode>(MAX_NODES);
            T.cands = s->alloc<DCand>(MAX_NODES);
            T.fcand = s->alloc<FCand>((size_t)MAX_NODES * n_features);
            T.hg = s->alloc<long long>(per_node * node_cap);
            T.hc = s->alloc<int>(per_node * node_cap);
            T.hh = s->hconst ? nullptr : s->alloc<long long>(per_node * node_cap);
            T.slots = nullptr;   // set to the task's inline table once the descriptors are allocated
            T.tnodes = s->alloc<DNode>(MAX_NODES);
            T.tmap = s->alloc<int>(MAX_NODES);
            T.rs_node = s->alloc<int>(64);
            T.rs_sim = s->alloc<int>(64);
            T.spec_cap = node_cap;
            T.host_flag = nullptr;
            T.energy = s->alloc<long long>(MAX_NODES);
            T.part = s->alloc<unsigned long long>((size_t)s->S.part_cap * per_node);
            T.part_c = s->alloc<unsigned int>((size_t)s->S.part_cap * per_node);
            T.part_h = s->hconst ? nullptr : s->alloc<unsigned long long>((size_t)s->S.part_cap * per_node);
            T.root_cnt = s->alloc<int>(per_node);
            T.cnt_row0 = -1;
            T.cnt_n = -1;
            T.cuts = s->alloc<int>(MAX_BUCKETS);
            const size_t tw = (size_t)(MAX_BUCKETS + 1) * MAX_NODES;
            T.tg = s->alloc<long long>(tw);
            T.th = s->hconst ? nullptr : s->alloc<long long>(tw);
            T.tc = s->alloc<int>(tw);
            s->clear_torb(t);   // k_torb_final clears what each tree adds; they start at zero
            T.values = s->alloc<double>(MAX_NODES);
            T.prefix = s->alloc<double>((size_t)MAX_BUCKETS * MAX_NODES);
            T.cover = s->alloc<int>(MAX_NODES);
            HIP_OK(hipMemsetAsync((void*)T.bucket, 0, sizeof(int) * m, s->stream));
            HIP_OK(hipMemsetAsync((void*)T.cuts, 0, sizeof(int) * MAX_BUCKETS, s->stream));
            s->th[t].chunks.push_back(TorbChunk{0, m, 1, 0});  // iid default: one segment past cut 0
        }
        s->tasks_d = s->alloc<Task>(n_tasks);
        for (int t = 0; t < n_tasks; ++t) s->tasks_h[t].slots = s->tasks_d[t].slot_tab;
        HIP_OK(hipMemcpyAsync(s->tasks_d, s->tasks_h.data(), sizeof(Task) * n_tasks, hipMemcpyHostToDevice, s->stream));
        s->chunk_ptrs_d = s->alloc<TorbChunk*>(n_tasks);
        s->n_chunks_d = s->alloc<int>(n_tasks);
        for (int t = 0; t < n_tasks; ++t) s->upload_chunks(t);
        s->pair_f_d = s->alloc<int>(2 * 64);
        s->ph_g = s->alloc<long long>(64 * NBIN);
        s->ph_c = s->alloc<int>(64 * NBIN);
        s->ph_h = s->alloc<long long>(64 * NBIN);
        HIP_OK(hipStreamSynchronize(s->stream));
        *out = guard.release();
    });
}

void tgbm_session_destroy(tgbm_session* s) { delete s; }

static_assert(CAT_EDGE_FLAG == TGBM_CAT_EDGES, "categorical edge flag");

int tgbm_session_set_cat(tgbm_session* s, const int* is_cat) {
    return dev_guard([&] {
        spec_resolve(s);
        if (s->S.is_cat) throw std::runtime_error("categorical flags are set once per session");
        std::vector<unsigned char> flags(s->F, 0);
        bool any = false;
        for (int f = 0; f < s->F; ++f) {
            flags[f] = is_cat && is_cat[f] ? 1 : 0;
            any |= flags[f] != 0;
        }
        if (!any) return;
        unsigned char* d = s->alloc<unsigned char>(s->F);
        HIP_OK(hipMemcpy(d, flags.data(), s->F, hipMemcpyHostToDevice));
        for (int t = 0; t < s->T; ++t) {
            Task& T = s->tasks_h[t];
            T.fmask = s->alloc<unsigned long long>((size_t)MAX_NODES * s->F * 4);
            T.nmask = s->alloc<unsigned long long>((size_t)MAX_NODES * 4);
            HIP_OK(hipMemsetAsync(T.nmask, 0, sizeof(unsigned long long) * MAX_NODES * 4, s->stream));
            s->set_field_ptr(t, offsetof(Task, fmask), T.fmask);
            s->set_field_ptr(t, offsetof(Task, nmask), T.nmask);
        }
        s->S.is_cat = d;
    });
}

int tgbm_session_set_spec(tgbm_session* s, int width) {
    return dev_guard([&] {
        spec_resolve(s);
        if (width < 0 || width > SPEC_W) throw std::runtime_error("multi-leaf width must be in [0, 8]");
        if (width > 1 && s->spec_cap == 0 && s->max_leaves <= 64 && s->T <= 1024) {
            // Room for the grown nodes (at most twice the final tree's) and for the partials of
            // one step of a task, SPEC_W * sp of them over all its slots (one task: 512; several:
            // 128, enough for one active task's step to fill the GPU).  The smaller children of a
            // step hold at most half the task's rows, so sp above their MAX_BLOCK_ROWS tiles plus
            // one per slot bounds every step.
            HIP_OK(hipStreamSynchronize(s->stream));
            const int cap = std::min(MAX_NODES, 2 * s->max_nodes);
            const int sp = std::max(s->T == 1 ? 64 : 16,
                                    (int)(((long long)s->max_task_rows / 2 + MAX_BLOCK_ROWS - 1) / MAX_BLOCK_ROWS) + 1);
            const int pc = std::max(s->S.part_cap, SPEC_W * sp);
            const size_t per_node = (size_t)s->F * NBIN;
            int* fh = nullptr;
            HIP_OK(hipHostMalloc((void**)&fh, sizeof(int) * (2 * s->T + 2), hipHostMallocMapped | hipHostMallocCoherent));
            s->flag_h = fh;
            HIP_OK(hipHostGetDevicePointer((void**)&s->flag_d, s->flag_h, 0));
            for (int t = 0; t < s->T; ++t) {
                Task& T = s->tasks_h[t];
                auto swap_buf = [&](auto*& field, size_t n, size_t off) {
                    using E = std::remove_pointer_t<std::remove_reference_t<decltype(field)>>;
                    if (!field) return;
                    s->release((void*)field);
                    field = s->alloc<std::remove_const_t<E>>(n);
                    s->set_field_ptr(t, off, field);
                };
                swap_buf(T.hg, per_node * cap, offsetof(Task, hg));
                swap_buf(T.hc, per_node * cap, offsetof(Task, hc));
                swap_buf(T.hh, per_node * cap, offsetof(Task, hh));
                swap_buf(T.part, per_node * pc, offsetof(Task, part));
                swap_buf(T.part_c, per_node * pc, offsetof(Task, part_c));
                swap_buf(T.part_h, per_node * pc, offsetof(Task, part_h));
                T.spec_cap = cap;
                s->set_field_int(t, offsetof(Task, spec_cap), cap);
                T.host_flag = s->flag_d + 2 * t;
                s->set_field_ptr(t, offsetof(Task, host_flag), T.host_flag);
            }
            s->spec_cap = cap;
            s->step_parts = SPEC_W * sp;
            s->S.part_cap = pc;
            s->S.spec_ctr = s->alloc<int>(8);
            HIP_OK(hipMemsetAsync(s->S.spec_ctr, 0, 8 * sizeof(int), s->stream));
            HIP_OK(hipStreamSynchronize(s->stream));
        }
        s->spec_width = s->spec_cap > 0 ? width : 0;
    });
}

int tgbm_session_set_monotone(tgbm_session* s, const int* monotone) {
    return dev_guard([&] {
        spec_resolve(s);
        if (s->S.monotone) throw std::runtime_error("monotone constraints are set once per session");
        if (!s->graphs.empty()) throw std::runtime_error("set monotone constraints before the first tree");
        std::vector<signed char> m(s->F, 0);
        bool any = false;
        for (int f = 0; f < s->F; ++f) {
            const int v = monotone ? monotone[f] : 0;
            m[f] = (signed char)(v > 0 ? 1 : (v < 0 ? -1 : 0));
            any |= m[f] != 0;
        }
        if (!any) return;
        signed char* d = s->alloc<signed char>(s->F);
        HIP_OK(hipMemcpy(d, m.data(), s->F, hipMemcpyHostToDevice));
        s->S.monotone = d;
    });
}

int tgbm_task_cat_masks(tgbm_session* s, int task, int n_nodes, unsigned long long* out) {
    return dev_guard([&] {
        spec_resolve(s);
        if (task < 0 || task >= s->T) throw std::runtime_error("task out of range");
        const Task& T = s->tasks_h[task];
        if (!T.nmask) throw std::runtime_error("the session has no categorical feature");
        if (!s->pending.empty()) throw std::runtime_error("collect the pending tree ends first");
        if (n_nodes < 0 || n_nodes > MAX_NODES) throw std::runtime_error("node count out of range");
        HIP_OK(hipMemcpyAsync(out, T.nmask, sizeof(unsigned long long) * 4 * n_nodes, hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
    });
}

int tgbm_session_bin(tgbm_session* s, const double* X, const double* edges, const int* n_edges) {
    return dev_guard([&] {
        spec_resolve(s);
        for (int f = 0; f < s->F; ++f) {
            const int ne = n_edges[f] & ~CAT_EDGE_FLAG;
            if (ne < 0 || ne > NBIN - 2 || (n_edges[f] & ~(CAT_EDGE_FLAG | (CAT_EDGE_FLAG - 1))))
                throw std::runtime_error("a feature has more than 254 bin edges; the device path stores uint8 bins (max_bin <= 255)");
        }
        // stream X in row blocks to bound device memory
        const int rows_blk = std::max(1, (int)std::min<size_t>((size_t)s->n, (512ull << 20) / (sizeof(double) * s->F)));
        double* Xd = dalloc<double>((size_t)rows_blk * s->F);
        double* Ed = dalloc<double>((size_t)s->F * NBIN);
        int* Nd = dalloc<int>(s->F);
        try {
            HIP_OK(hipMemcpy(Ed, edges, sizeof(double) * s->F * NBIN, hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(Nd, n_edges, sizeof(int) * s->F, hipMemcpyHostToDevice));
            for (int r0 = 0; r0 < s->n; r0 += rows_blk) {
                const int m = std::min(rows_blk, s->n - r0);
                HIP_OK(hipMemcpyAsync(Xd, X + (size_t)r0 * s->F, sizeof(double) * m * s->F, hipMemcpyHostToDevice, s->stream));
                launch_bin(s->stream, Xd, m, s->F, Ed, Nd, s->bins_row + (size_t)r0 * s->fp, s->fp,
                           s->bins_col + r0, s->n_pad);
                HIP_OK(hipGetLastError());
                HIP_OK(hipStreamSynchronize(s->stream));
            }
            // new bins: the cached root counts no longer hold
            for (int k = 0; k < s->T; ++k) s->set_field_int(k, offsetof(Task, cnt_n), -1);
        } catch (...) {
            (void)hipFree(Xd);
            (void)hipFree(Ed);
            (void)hipFree(Nd);
            throw;
        }
        (void)hipFree(Xd);
        (void)hipFree(Ed);
        (void)hipFree(Nd);
    });
}

int tgbm_session_set_active(tgbm_session* s, int first, int count) {
    return dev_guard([&] {
        spec_resolve(s);
        if (first < 0 || count < 1 || first + count > s->T) throw std::runtime_error("active task range out of range");
        HIP_OK(hipStreamSynchronize(s->stream));
        s->a0 = first;
        s->an = count;
    });
}

int tgbm_task_bind(tgbm_session* s, int task, int row0, int n_rows) {
    return dev_guard([&] {
        spec_resolve(s);
        if (task < 0 || task >= s->T) throw std::runtime_error("task out of range");
        TaskHost& H = s->th[task];
        if (n_rows < 1 || n_rows > H.cap_rows) throw std::runtime_error("task rows exceed the task's capacity");
        if (row0 < 0 || (long long)row0 + n_rows > s->n) throw std::runtime_error("task row window outside the shared matrix");
        HIP_OK(hipStreamSynchronize(s->stream));
        H.n_rows = n_rows;
        H.row0 = row0;
        H.has_sample = false;
        H.n_sample = n_rows;
        H.goss = false;
        H.n_buckets = 1;
        H.chunks.assign(1, TorbChunk{0, n_rows, 1, 0});
        Task& T = s->tasks_h[task];
        T.n_rows = n_rows;
        T.row0 = row0;
        struct {
            int n_rows, row0;
        } win{n_rows, row0};
        static_assert(offsetof(Task, row0) - offsetof(Task, n_rows) == 4, "layout");
        HIP_OK(hipMemcpyAsync(reinterpret_cast<char*>(s->tasks_d + task) + offsetof(Task, n_rows), &win, sizeof win,
                              hipMemcpyHostToDevice, s->stream));
        s->set_field_int(task, offsetof(Task, n_sample), n_rows);
        s->set_field_int(task, offsetof(Task, contiguous), 1);
        s->set_field_int(task, offsetof(Task, goss), 0);
        s->set_field_int(task, offsetof(Task, n_buckets), 1);
        HIP_OK(hipMemsetAsync((void*)T.bucket, 0, sizeof(int) * n_rows, s->stream));
        s->clear_torb(task);
        s->upload_chunks(task);
    });
}

// Exchange two task descriptors and their chunk-table entries on the device (in stream order,
// behind the work that last wrote them); the slot table pointer and the mapped completion flag
// belong to the index and stay.
__global__ __launch_bounds__(64) void k_task_swap(Task* tasks, int a, int b, TorbChunk** chunk_ptrs, int* n_chunks) {
    static_assert(sizeof(Task) % 4 == 0, "descriptor words");
    Slot* const sa = tasks[a].slots;
    Slot* const sb = tasks[b].slots;
    int* const fa = tasks[a].host_flag;
    int* const fb = tasks[b].host_flag;
    __syncthreads();
    int* const wa = (int*)(tasks + a);
    int* const wb = (int*)(tasks + b);
    for (int j = threadIdx.x; j < (int)(sizeof(Task) / 4); j += blockDim.x) {
        const int x = wa[j], y = wb[j];
        wa[j] = y;
        wb[j] = x;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    tasks[a].slots = sa;
    tasks[b].slots = sb;
    tasks[a].host_flag = fa;
    tasks[b].host_flag = fb;
    TorbChunk* const c = chunk_ptrs[a];
    chunk_ptrs[a] = chunk_ptrs[b];
    chunk_ptrs[b] = c;
    const int n = n_chunks[a];
    n_chunks[a] = n_chunks[b];
    n_chunks[b] = n;
}

int tgbm_task_swap(tgbm_session* s, int a, int b) {
    return dev_guard([&] {
        spec_resolve(s);
        if (a < 0 || a >= s->T || b < 0 || b >= s->T) throw std::runtime_error("task out of range");
        if (a == b) return;
        if (!s->pending.empty()) throw std::runtime_error("collect the pending tree ends first");
        hipLaunchKernelGGL(k_task_swap, dim3(1), dim3(64), 0, s->stream, s->tasks_d, a, b, s->chunk_ptrs_d,
                           s->n_chunks_d);
        HIP_OK(hipGetLastError());
        std::swap(s->tasks_h[a], s->tasks_h[b]);
        std::swap(s->tasks_h[a].slots, s->tasks_h[b].slots);
        std::swap(s->tasks_h[a].host_flag, s->tasks_h[b].host_flag);
        std::swap(s->th[a], s->th[b]);
    });
}

int tgbm_task_node_hist(tgbm_session* s, int task, int node, long long* qg, int* count, long long* qh, int* s_g,
                        int* s_h) {
    return dev_guard([&] {
        spec_resolve(s);
        if (task < 0 || task >= s->T) throw std::runtime_error("task out of range");
        if (node < 0 || node >= s->max_nodes) throw std::runtime_error("node out of range");
        if (!s->pending.empty()) throw std::runtime_error("collect the pending tree ends first");
        const Task& T = s->tasks_h[task];
        const size_t words = (size_t)s->F * NBIN, off = (size_t)node * words;
        Task dev;
        HIP_OK(hipMemcpyAsync(&dev, s->tasks_d + task, sizeof(Task), hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipMemcpyAsync(qg, T.hg + off, words * sizeof(long long), hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipMemcpyAsync(count, T.hc + off, words * sizeof(int), hipMemcpyDeviceToHost, s->stream));
        if (qh && T.hh)
            HIP_OK(hipMemcpyAsync(qh, T.hh + off, words * sizeof(long long), hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        if (qh && !T.hh)
            for (size_t k = 0; k < words; ++k) qh[k] = count[k];
        *s_g = dev.s_g;
        *s_h = T.hh ? dev.s_h : 0;
    });
}

int tgbm_session_get_bins(tgbm_session* s, unsigned char* out) {
    return dev_guard([&] {
        spec_resolve(s);
        std::vector<unsigned char> tmp((size_t)s->n * s->fp);
        HIP_OK(hipMemcpy(tmp.data(), s->bins_row, tmp.size(), hipMemcpyDeviceToHost));
        for (int i = 0; i < s->n; ++i) std::memcpy(out + (size_t)i * s->F, &tmp[(size_t)i * s->fp], s->F);
    });
}

int tgbm_task_set_rows(tgbm_session* s, int task, const double* y, const double* w, const double* pred0,
                       const int* bucket, const unsigned char* sweep_seg, int n_buckets,
                       const int* cuts, const long long* groups) {
    (void)sweep_seg;
    return dev_guard([&] {
        spec_resolve(s);
        if (task < 0 || task >= s->T) throw std::runtime_error("task out of range");
        if (n_buckets < 1 || n_buckets > MAX_BUCKETS) throw std::runtime_error("n_buckets out of range for the device path");
        Task& T = s->tasks_h[task];
        TaskHost& H = s->th[task];
        s->clear_torb(task);   // a fit starts from zero sums whatever an earlier one left
        const int m = H.n_rows;
        HIP_OK(hipStreamSynchronize(s->stream));
        HIP_OK(hipMemcpy((void*)T.y, y, sizeof(double) * m, hipMemcpyHostToDevice));
        if (pred0) HIP_OK(hipMemcpy(T.pred, pred0, sizeof(double) * m, hipMemcpyHostToDevice));
        if (w) {
            if (s->hconst) {
                for (int i = 0; i < m; ++i)
                    if (w[i] != 1.0) throw std::runtime_error("session declared h == 1 but weights are not all 1");
            }
            if (!H.w_d) H.w_d = s->alloc<double>(H.cap_rows);
            HIP_OK(hipMemcpy(H.w_d, w, sizeof(double) * m, hipMemcpyHostToDevice));
            T.w = H.w_d;
            s->set_field_ptr(task, offsetof(Task, w), H.w_d);
        } else if (T.w) {
            T.w = nullptr;
            s->set_field_ptr(task, offsetof(Task, w), nullptr);
        }
        if (bucket) HIP_OK(hipMemcpy((void*)T.bucket, bucket, sizeof(int) * m, hipMemcpyHostToDevice));
        H.n_buckets = n_buckets;
        s->set_field_int(task, offsetof(Task, n_buckets), n_buckets);
        std::vector<int> cut(n_buckets, 0);
        if (cuts) std::copy(cuts, cuts + n_buckets, cut.begin());
        HIP_OK(hipMemcpy((void*)T.cuts, cut.data(), sizeof(int) * n_buckets, hipMemcpyHostToDevice));
        // sweep-segment chunk list: segment s covers [cut_{s-1}, cut_s), s = n_buckets past the last
        H.chunks.clear();
        const int CH = 2048;
        int prev = 0;
        for (int sg = 0; sg <= n_buckets; ++sg) {
            const int end = sg < n_buckets ? std::min(std::max(cut[sg], prev), m) : m;
            for (int r = prev; r < end; r += CH) H.chunks.push_back(TorbChunk{r, std::min(end, r + CH), sg, 0});
            prev = std::max(prev, end);
        }
        // groups: runs of equal keys in stable key order (objective.cpp:18-39)
        if (groups) {
            std::vector<int> ord(m);
            for (int i = 0; i < m; ++i) ord[i] = i;
            bool sorted = true;
            for (int i = 1; i < m; ++i)
                if (groups[i] < groups[i - 1]) sorted = false;
            if (!sorted) std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return groups[a] < groups[b]; });
            std::vector<int> gs;
            int i = 0, mx = 0;
            while (i < m) {
                int j = i + 1;
                while (j < m && groups[ord[j]] == groups[ord[i]]) ++j;
                gs.push_back(i);
                mx = std::max(mx, j - i);
                i = j;
            }
            gs.push_back(m);
            if (!H.gstart_d) H.gstart_d = s->alloc<int>((size_t)H.cap_rows + 1);
            int* gsd = H.gstart_d;
            HIP_OK(hipMemcpy(gsd, gs.data(), sizeof(int) * gs.size(), hipMemcpyHostToDevice));
            s->set_field_ptr(task, offsetof(Task, group_start), gsd);
            s->set_field_int(task, offsetof(Task, n_groups), (int)gs.size() - 1);
            H.n_groups = (int)gs.size() - 1;
            s->set_field_int(task, offsetof(Task, has_groups), 1);
            if (!sorted) {
                if (!H.grows_d) H.grows_d = s->alloc<int>(H.cap_rows);
                HIP_OK(hipMemcpy(H.grows_d, ord.data(), sizeof(int) * m, hipMemcpyHostToDevice));
                s->set_field_ptr(task, offsetof(Task, group_rows), H.grows_d);
            } else {
                s->set_field_ptr(task, offsetof(Task, group_rows), nullptr);
            }
            H.max_group = mx;
        } else {
            s->set_field_int(task, offsetof(Task, has_groups), 0);
        }
        s->upload_chunks(task);
    });
}

int tgbm_task_grad(tgbm_session* s, int task, int objective) {
    return dev_guard([&] {
        spec_resolve(s);
        if (objective != 0 && objective != 5) throw std::runtime_error("objective not supported on the device path");
        if (task >= s->T) throw std::runtime_error("task out of range");
        const int t0 = task < 0 ? s->a0 : task, nt = task < 0 ? s->an : 1;
        int mg = 1;
        for (int t = t0; t < t0 + nt; ++t) mg = std::max(mg, (int)std::max<size_t>(1, s->th[t].n_rows));
        EventPair ep;
        if (s->timing) s->tic(4, 0.0, ep);
        launch_grad(s->stream, s->tasks_d + t0, nt, objective, mg, mg);
        s->objective = objective;
        if (s->timing) s->toc(ep);
        HIP_OK(hipGetLastError());
    });
}

int tgbm_task_get_grad(tgbm_session* s, int task, double* g, double* h) {
    return dev_guard([&] {
        spec_resolve(s);
        const Task& T = s->tasks_h[task];
        const int m = s->th[task].n_rows;
        HIP_OK(hipStreamSynchronize(s->stream));
        if (g) HIP_OK(hipMemcpy(g, T.g, sizeof(double) * m, hipMemcpyDeviceToHost));
        // h == 1 sessions (MSE without weights or GOSS) never store h on the device
        if (h && s->hconst) std::fill(h, h + m, 1.0);
        else if (h) HIP_OK(hipMemcpy(h, T.h, sizeof(double) * m, hipMemcpyDeviceToHost));
    });
}

int tgbm_task_get_pred(tgbm_session* s, int task, double* pred) {
    return dev_guard([&] {
        spec_resolve(s);
        HIP_OK(hipStreamSynchronize(s->stream));
        HIP_OK(hipMemcpy(pred, s->tasks_h[task].pred, sizeof(double) * s->th[task].n_rows, hipMemcpyDeviceToHost));
    });
}

int tgbm_task_set_sample(tgbm_session* s, int task, const int* sample, int n_sample,
                         const unsigned char* scaled, double scale) {
    return dev_guard([&] {
        spec_resolve(s);
        Task& T = s->tasks_h[task];
        TaskHost& H = s->th[task];
        HIP_OK(hipStreamSynchronize(s->stream));
        if (!sample) {
            H.has_sample = false;
            H.n_sample = H.n_rows;
            H.goss = false;
            s->set_field_int(task, offsetof(Task, n_sample), H.n_rows);
            s->set_field_int(task, offsetof(Task, contiguous), 1);
            s->set_field_int(task, offsetof(Task, goss), 0);
            return;
        }
        if (n_sample < 0 || n_sample > H.n_rows) throw std::runtime_error("sample size out of range");
        for (int k = 0; k < n_sample; ++k)
            if (sample[k] < 0 || sample[k] >= H.n_rows || (k && sample[k] <= sample[k - 1]))
                throw std::runtime_error("sample rows must be sorted, unique and in range");
        HIP_OK(hipMemcpy(T.idx, sample, sizeof(int) * n_sample, hipMemcpyHostToDevice));
        H.has_sample = true;
        H.n_sample = n_sample;
        s->set_field_int(task, offsetof(Task, n_sample), n_sample);
        s->set_field_int(task, offsetof(Task, contiguous), 0);
        if (scaled) {
            if (s->hconst) throw std::runtime_error("scaled sample rows need a session without h == 1");
            unsigned char* sd = const_cast<unsigned char*>(T.scaled);
            if (!sd) {
                sd = s->alloc<unsigned char>(H.n_rows);
                T.scaled = sd;
                s->set_field_ptr(task, offsetof(Task, scaled), sd);
            }
            HIP_OK(hipMemcpy(sd, scaled, H.n_rows, hipMemcpyHostToDevice));
            H.goss = true;
            s->set_field_int(task, offsetof(Task, goss), 1);
            s->set_field_double(task, offsetof(Task, goss_scale), scale);
        } else {
            H.goss = false;
            s->set_field_int(task, offsetof(Task, goss), 0);
        }
    });
}

int tgbm_task_goss(tgbm_session* s, int task, double goss_top, double goss_other, const int* perm, int n_perm,
                   int* tie) {
    return dev_guard([&] {
        spec_resolve(s);
        if (task < 0 || task >= s->T) throw std::runtime_error("task out of range");
        if (s->hconst) throw std::runtime_error("GOSS scales h: the session must not declare h == 1");
        Task& T = s->tasks_h[task];
        TaskHost& H = s->th[task];
        const int S = H.n_rows;
        const int n_top = std::min(S, std::max(1, (int)std::ceil(goss_top * S)));
        const int R = S - n_top;
        const int n_other = std::min((int)std::ceil(goss_other * S), R);
        if (n_perm != n_other) throw std::runtime_error("GOSS draw size does not match the rows");
        for (int k = 0; k < n_perm; ++k)
            if (perm[k] < 0 || perm[k] >= R) throw std::runtime_error("GOSS draw index out of range");
        if (!H.gk0) {
            const size_t c = (size_t)H.cap_rows;
            H.gk0 = s->alloc<unsigned long long>(c);
            H.gk1 = s->alloc<unsigned long long>(c);
            H.gv0 = s->alloc<int>(c);
            H.gv1 = s->alloc<int>(c);
            H.gkeep = s->alloc<unsigned char>(c);
            H.gperm = s->alloc<int>(c);
            H.gsmall = s->alloc<int>(2);
            H.gtmp_bytes = goss_temp_bytes(H.cap_rows);
            H.gtmp = s->alloc<char>(H.gtmp_bytes);
        }
        unsigned char* sd = const_cast<unsigned char*>(T.scaled);
        if (!sd) {
            sd = s->alloc<unsigned char>(H.cap_rows);
            T.scaled = sd;
            s->set_field_ptr(task, offsetof(Task, scaled), sd);
        }
        if (n_other > 0)
            HIP_OK(hipMemcpyAsync(H.gperm, perm, sizeof(int) * n_other, hipMemcpyHostToDevice, s->stream));
        HIP_OK(hipMemsetAsync(H.gsmall, 0, sizeof(int) * 2, s->stream));
        HIP_OK(launch_goss(s->stream, T.g, S, n_top, n_other, H.gperm, H.gk0, H.gk1, H.gv0, H.gv1, H.gkeep, sd,
                           T.idx, H.gsmall, H.gsmall + 1, H.gtmp, H.gtmp_bytes));
        // the descriptor takes the sample in stream order (no synchronous field writes after the wait)
        const double scale = (1.0 - goss_top) / std::max(goss_other, 1e-12);
        hipLaunchKernelGGL(k_goss_commit, dim3(1), dim3(64), 0, s->stream, s->tasks_d + task, H.gsmall,
                           n_top + n_other, scale);
        HIP_OK(hipGetLastError());
        int res[2] = {0, 0};
        HIP_OK(hipMemcpyAsync(res, H.gsmall, sizeof res, hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        *tie = res[1] != 0;
        if (*tie) return;   // equal |g| values: the caller orders them with std::sort on the host
        if (res[0] != n_top + n_other) throw std::runtime_error("internal: GOSS selected an unexpected row count");
        H.has_sample = true;
        H.n_sample = res[0];
        H.goss = true;
    });
}

// Reads every word of a zero-filled buffer twice the size of the Infinity Cache, so earlier data
// leaves the caches in front of HBM.  Nothing is written (the buffer stays zero, so the guarded
// store never runs): the caches keep no dirty lines whose write-back would load the next kernel's
// reads.
__global__ __launch_bounds__(256) void k_flush_cache(unsigned long long* buf, size_t words) {
    unsigned long long acc = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        acc |= buf[i];
    if (acc != 0) buf[0] = 0;
}

// Root of every task's tree: quantize, zero the node histograms, build and scan the root.  The
// first split is selected by the first split step (its partition launch in device-driven mode).
static void tree_begin_impl(tgbm_session* s, const tgbm_grow_params* p, bool energy, bool device_mode) {
    spec_resolve(s);
    const GrowP P = s->growp(p);
    {   // the root's rows (one active task); children are learned from their splits' readbacks
        const TaskHost& H = s->th[s->a0];
        s->node_rows_h.assign(1, s->an == 1 ? (H.has_sample ? H.n_sample : H.n_rows) : -1);
    }
    s->spec_tree = false;
    s->spec_launched = false;
    launch_tree_start(s->stream, s->S, s->td(), s->an, s->max_rows(), s->max_nodes);
    if (energy)
        for (int t = s->a0; t < s->a0 + s->an; ++t) HIP_OK(hipMemsetAsync(s->tasks_h[t].energy, 0, sizeof(long long) * MAX_NODES, s->stream));
    const int ms = s->max_sample();
    EventPair ep;
    if (s->timing) {
        double bytes = 0;
        for (int k = s->a0; k < s->a0 + s->an; ++k) {
            const TaskHost& t = s->th[k];
            const double m = t.has_sample ? t.n_sample : t.n_rows;
            bytes += m * s->F + m * (s->hconst ? 8.0 : 16.0) + (t.has_sample ? 4.0 * m : 0.0) + (double)s->F * s->nb * 24.0;
        }
        if (s->flush)
            hipLaunchKernelGGL(k_flush_cache, dim3(4096), dim3(256), 0, s->stream, s->flush_buf,
                               tgbm_session::kFlushBytes / 8);
        s->kernel_events(0, bytes, ep);
    }
    launch_hist(s->stream, s->S, s->td(), s->an, true, s->all_contiguous(), s->hconst, ms,
                p->hist_mode, s->timing ? ep.a : nullptr, s->timing ? ep.b : nullptr);
    if (s->timing) s->events.push_back(ep);
    launch_scan(s->stream, s->S, s->td(), s->an, P, 0);
    (void)device_mode;   // both modes begin a tree the same way
    if (energy)
        for (int t = s->a0; t < s->a0 + s->an; ++t) launch_energy(s->stream, s->tasks_d, t, 0, s->th[t].has_sample ? s->th[t].n_sample : s->th[t].n_rows);
    HIP_OK(hipGetLastError());
}

// One split step of every task: decide the split, partition the split leaf, build the smaller
// child, scan both children.  Device mode: every partition block runs the best-first selection
// over the last scan's results (with several active tasks one selection launch decides once per
// task and the partition reads the decision); host mode: the host wrote its own decision first.
static void split_step(tgbm_session* s, const GrowP& P, int mode, const tgbm_grow_params* p) {
    const int ms = s->launch_rows();
    EventPair e1, e2;
    if (s->timing) s->tic(1, 0.0, e1);
    if (mode == 0 && s->an > 1) {
        // many tasks: one selection block per task instead of one per partition block
        launch_select(s->stream, s->S, s->td(), s->an, P);
        launch_partition(s->stream, s->S, s->td(), s->an, ms, P, 1);
    } else {
        launch_partition(s->stream, s->S, s->td(), s->an, ms, P, mode);
    }
    if (s->timing) s->toc(e1);
    if (s->timing) s->kernel_events(5, 0.0, e2);  // child histogram builds (bytes filled at tree end)
    launch_hist(s->stream, s->S, s->td(), s->an, false, false, s->hconst, ms / 2 + 1, p->hist_mode,
                s->timing ? e2.a : nullptr, s->timing ? e2.b : nullptr);
    if (s->timing) s->events.push_back(e2);
    EventPair e3;
    if (s->timing) s->tic(2, 0.0, e3);
    launch_scan(s->stream, s->S, s->td(), s->an, P, 1);
    if (s->timing) s->toc(e3);
    HIP_OK(hipGetLastError());
}

// Whether every active task's growth is done, and the most steps a task took, after the stream
// drains (the status kernel writes them into the mapped words after the tasks' flags).
static bool spec_done(tgbm_session* s, int* steps) {
    int* out = s->flag_d + 2 * s->T;
    launch_spec_status(s->stream, s->td(), s->an, out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(s->stream));
    volatile int* h = s->flag_h + 2 * s->T;
    if (steps) *steps = h[1];
    return h[0] != 0;
}

// One multi-leaf step: the selection, the slots' partitions, their smaller children's builds and
// the scans of both children of every slot.
static void spec_step(tgbm_session* s, const GrowP& P, const tgbm_grow_params* p) {
    const int ms = s->launch_rows();
    // child builds: a fixed grid of resident blocks takes the step's (row block, feature block)
    // items; tiles of at least 1,024 rows, sized so the items about fill those blocks (1M x 128
    // fit: 1,290 trees/s at 1,024, 1,280 at 2,048, 1,270 at 512, 1,213 at 4,096)
    int blocks, fbs;
    spec_child_shape(s->hconst, s->n_cu, s->fp, blocks, fbs);
    EventPair e0, e1, e2, e3;
    if (s->timing) s->tic(6, 0.0, e0);   // the step's selection (timer 6)
    launch_spec_select(s->stream, s->S, s->td(), s->an, P, s->spec_width, 1024, std::max(1, blocks / (fbs * s->an)),
                       s->step_parts);
    if (s->timing) s->toc(e0);
    if (s->timing) s->tic(1, 0.0, e1);
    launch_partition(s->stream, s->S, s->td(), s->an, ms, P, 2);
    if (s->timing) s->toc(e1);
    if (s->timing) s->kernel_events(5, 0.0, e2);
    launch_hist_spec(s->stream, s->S, s->td(), s->an, s->hconst, blocks,
                     s->timing ? e2.a : nullptr, s->timing ? e2.b : nullptr);
    if (s->timing) s->events.push_back(e2);
    if (s->timing) s->tic(2, 0.0, e3);
    launch_scan(s->stream, s->S, s->td(), s->an, P, 2, s->hconst);
    if (s->timing) s->toc(e3);
    HIP_OK(hipGetLastError());
}

// Rows of the step just run (timing mode): the smaller children built and the leaves partitioned.
static void spec_log_rows(tgbm_session* s) {
    double small = 0, rows = 0, splits = 0;
    for (int k = s->a0; k < s->a0 + s->an; ++k) {
        Task T;
        HIP_OK(hipMemcpy(&T, s->tasks_d + k, sizeof(Task), hipMemcpyDeviceToHost));
        if (T.done) continue;
        Slot sl[SPEC_W];
        HIP_OK(hipMemcpy(sl, T.slots, sizeof(sl), hipMemcpyDeviceToHost));
        for (int j = 0; j < T.n_slots; ++j) {
            small += sl[j].small_count;
            rows += sl[j].count;
        }
        splits += T.n_slots;
    }
    s->log_rows[5].push_back(small);
    s->log_rows[1].push_back(rows);
    s->log_rows[6].push_back(splits);   // splits the selections made
}

// Multi-leaf growth of the current tree.  The loop of spec_k steps (the most splitting steps the
// last 32 trees needed) is only launched here: the tree end is queued behind it and then waits for it
// (tree_end_enqueue), so the device never waits for the host in between.  The steps go out as
// direct launches: a captured graph of them was slower (1M x 128 fit: 1,282 against 1,225
// trees/s), its launch call holding the host thread for most of the graph's run, with the device
// idle between some of its kernels; TGBM_SPEC_GRAPH=1 selects the graph.  A tree that needs more
// steps continues with single steps (every step makes at least one split the tree keeps, so at
// most max_leaves - 1 steps end any tree).  Timed and graph-less growth runs step by step,
// reading the tree's state after each.
static void grow_spec(tgbm_session* s, const GrowP& P, const tgbm_grow_params* p) {
    const int limit = std::max(1, p->max_leaves - 1);
    int steps = 0;
    bool done = false;
    if (!s->use_graphs || s->timing) {
        for (int it = 0; it < limit && !done; ++it) {
            spec_step(s, P, p);
            done = spec_done(s, &steps);
            if (s->timing) spec_log_rows(s);
        }
    } else {
        const int K = std::min(limit, s->spec_k);
        if (!s->spec_graph) {
            for (int k = s->a0; k < s->a0 + s->an; ++k) s->flag_h[2 * k] = 0;
            for (int k = 0; k < K; ++k) spec_step(s, P, p);
            s->spec_launched = true;
            s->spec_k_used = K;
            return;
        }
        const std::vector<long long> key{s->launch_rows(), p->max_leaves, p->hist_mode, s->a0, s->an, -K};
        auto it = s->graphs.find(key);
        if (it == s->graphs.end()) {
            HIP_OK(hipStreamSynchronize(s->stream));
            if (s->graphs.size() >= 256) {
                auto lru = s->graphs.begin();
                for (auto j = s->graphs.begin(); j != s->graphs.end(); ++j)
                    if (j->second.last_use < lru->second.last_use) lru = j;
                HIP_OK(hipGraphExecDestroy(lru->second.ge));
                HIP_OK(hipGraphDestroy(lru->second.g));
                s->graphs.erase(lru);
            }
            tgbm_session::CachedGraph c;
            HIP_OK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < K; ++k) spec_step(s, P, p);
            HIP_OK(hipStreamEndCapture(s->stream, &c.g));
            const hipError_t e = hipGraphInstantiate(&c.ge, c.g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(c.g);
                HIP_OK(e);
            }
            it = s->graphs.emplace(key, c).first;
        }
        it->second.last_use = ++s->graph_clock;
        for (int k = s->a0; k < s->a0 + s->an; ++k) s->flag_h[2 * k] = 0;
        HIP_OK(hipGraphLaunch(it->second.ge, s->stream));
        // the tree end is queued behind the loop and waited for later (spec_resolve)
        s->spec_launched = true;
        s->spec_k_used = K;
        return;
    }
    if (!done) throw std::runtime_error("internal: multi-leaf growth did not finish the tree");
}

int tgbm_tree_begin(tgbm_session* s, const tgbm_grow_params* p) {
    return dev_guard([&] { tree_begin_impl(s, p, p->host_driven != 0, p->host_driven == 0); });
}

int tgbm_tree_grow_device(tgbm_session* s, const tgbm_grow_params* p) {
    return dev_guard([&] {
        spec_resolve(s);
        const GrowP P = s->growp(p);
        s->spec_tree = s->spec_width > 0 && p->max_leaves <= 64;
        if (s->spec_tree) {
            // the multi-leaf scan is numeric only; categorical features grow on the host
            if (s->S.is_cat) throw std::runtime_error("multi-leaf growth needs numeric features only");
            grow_spec(s, P, p);
            return;
        }
        if (!s->use_graphs || s->timing) {
            for (int it = 0; it < p->max_leaves - 1; ++it) split_step(s, P, 0, p);
            return;
        }
        // The loop's kernels read every decision from device memory, so one capture serves
        // every tree with the same sample size and leaf budget.
        const std::vector<long long> key{s->launch_rows(), p->max_leaves, p->hist_mode, s->a0, s->an};
        auto it = s->graphs.find(key);
        if (it == s->graphs.end()) {
            // Capture on a drained stream, and destroy an executable graph only once no launch of
            // it can still be running: the cache is bounded by evicting the least recently used
            // graph after a stream synchronize.
            HIP_OK(hipStreamSynchronize(s->stream));
            if (s->graphs.size() >= 256) {
                auto lru = s->graphs.begin();
                for (auto j = s->graphs.begin(); j != s->graphs.end(); ++j)
                    if (j->second.last_use < lru->second.last_use) lru = j;
                HIP_OK(hipGraphExecDestroy(lru->second.ge));
                HIP_OK(hipGraphDestroy(lru->second.g));
                s->graphs.erase(lru);
            }
            tgbm_session::CachedGraph c;
            HIP_OK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < p->max_leaves - 1; ++k) split_step(s, P, 0, p);
            HIP_OK(hipStreamEndCapture(s->stream, &c.g));
            const hipError_t e = hipGraphInstantiate(&c.ge, c.g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(c.g);
                HIP_OK(e);
            }
            it = s->graphs.emplace(key, c).first;
        }
        it->second.last_use = ++s->graph_clock;
        HIP_OK(hipGraphLaunch(it->second.ge, s->stream));
    });
}

int tgbm_leaf_scan_results(tgbm_session* s, int task, int node, double* gain_f, int* thr_f, int* ml_f,
                           double* energy, int* n_rows) {
    return dev_guard([&] {
        spec_resolve(s);
        if (task < 0 || task >= s->T) throw std::runtime_error("task out of range");
        if (node < 0 || node >= MAX_NODES) throw std::runtime_error("node out of range");
        const int F = s->F;
        s->ensure_readback();
        // the node's scan results, energy, row count and energy scale, written by the device into
        // mapped host memory (one wait)
        hipLaunchKernelGGL(k_node_readback, dim3(1), dim3(256), 0, s->stream, (const Task*)s->tasks_d, task, node, F,
                           s->rb_d);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(s->stream));
        const int* hdr = (const int*)s->rb_h;
        const long long* en = (const long long*)(s->rb_h + 24);
        const FCand* fc = (const FCand*)(s->rb_h + 40);
        for (int f = 0; f < F; ++f) {
            gain_f[f] = fc[f].thr >= 0 ? fc[f].gain : -1e300;
            thr_f[f] = fc[f].thr;
            ml_f[f] = fc[f].ml;
        }
        if (energy) *energy = std::ldexp((double)en[0], -hdr[2]);
        if (n_rows) *n_rows = hdr[3];
        if (node < (int)s->node_rows_h.size()) s->node_rows_h[node] = hdr[3];
    });
}

int tgbm_pair_search(tgbm_session* s, int task, int node, const int* f1, const int* f2, int n_pairs,
                     const tgbm_grow_params* p, tgbm_cand* out) {
    return dev_guard([&] {
        spec_resolve(s);
        if (n_pairs <= 0) return;
        if (n_pairs > 32) throw std::runtime_error("too many PAIR candidates");
        std::vector<int> pf(2 * n_pairs);
        for (int k = 0; k < n_pairs; ++k) {
            if (f1[k] < 0 || f1[k] >= s->F || f2[k] < 0 || f2[k] >= s->F)
                throw std::runtime_error("PAIR feature out of range");
            pf[2 * k] = f1[k];
            pf[2 * k + 1] = f2[k];
        }
        HIP_OK(hipMemcpyAsync(s->pair_f_d, pf.data(), sizeof(int) * pf.size(), hipMemcpyHostToDevice, s->stream));
        HIP_OK(hipMemsetAsync(s->ph_g, 0, sizeof(long long) * n_pairs * NBIN, s->stream));
        HIP_OK(hipMemsetAsync(s->ph_c, 0, sizeof(int) * n_pairs * NBIN, s->stream));
        HIP_OK(hipMemsetAsync(s->ph_h, 0, sizeof(long long) * n_pairs * NBIN, s->stream));
        // the node's row count sizes the launch: known from its split's readback, else read back
        int count = node < (int)s->node_rows_h.size() ? s->node_rows_h[node] : -1;
        if (count < 0) {
            DNode nd;
            HIP_OK(hipStreamSynchronize(s->stream));
            HIP_OK(hipMemcpy(&nd, s->tasks_h[task].nodes + node, sizeof nd, hipMemcpyDeviceToHost));
            count = nd.seg_count;
        }
        // the pair scans write their results into the mapped readback buffer: one wait, no copy
        s->ensure_readback();
        launch_pair(s->stream, s->S, s->tasks_d, task, node, count, s->pair_f_d, n_pairs, s->ph_g,
                    s->ph_c, s->ph_h, s->hconst, s->growp(p), (FCand*)(s->rb_d + 40));
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(s->stream));
        const FCand* r = (const FCand*)(s->rb_h + 40);
        for (int k = 0; k < n_pairs; ++k) {
            out[k].gain = r[k].thr >= 0 ? r[k].gain : -1e300;
            out[k].feature = f1[k];
            out[k].feature2 = f2[k];
            out[k].threshold = r[k].thr;
            out[k].type = 1;
            out[k].missing_go_left = r[k].ml;
            out[k].valid = r[k].thr >= 0;
        }
    });
}

}  // extern "C"

// Host-driven split results: the energy (sum of (g 2^s_e)^2) of the new children L, L + 1
// (L = n_nodes - 2), then their scan results into the mapped readback buffer.
__global__ __launch_bounds__(256) void k_child_energy(Task* tasks, int task) {
    Task& T = tasks[task];
    if (T.done) return;
    const int node = T.n_nodes - 2 + (int)blockIdx.y;
    const DNode& n = T.nodes[node];
    const int* rows = node_rows(T, n) + n.seg_start;
    long long acc = 0;
    const double sc = T.goss_scale;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n.seg_count; k += gridDim.x * blockDim.x) {
        const int i = rows[k];
        double g = T.g[i];
        if (T.goss && T.scaled && T.scaled[i]) g *= sc;
        acc += __double2ll_rn(ldexp(g * g, T.s_e));
    }
    acc = wave_sum_ll(acc);
    __shared__ long long sacc[4];
    if ((threadIdx.x & 63) == 0) sacc[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const long long sum = sacc[0] + sacc[1] + sacc[2] + sacc[3];
        if (sum) atomicAdd((unsigned long long*)&T.energy[node], (unsigned long long)sum);
    }
}

// Readback layout: int {done, L, s_e, rows_L, rows_R, pad}, long long energy[2], FCand[2][F].
__global__ __launch_bounds__(256) void k_split_readback(const Task* tasks, int task, int F, char* out) {
    const Task& T = tasks[task];
    int* hdr = (int*)out;
    long long* en = (long long*)(out + 24);
    FCand* fc = (FCand*)(out + 40);
    const int done = T.done;
    const int L = T.n_nodes - 2;
    if (threadIdx.x == 0) {
        hdr[0] = done;
        hdr[1] = L;
        hdr[2] = T.s_e;
        hdr[3] = done ? 0 : T.nodes[L].v[N_NS];
        hdr[4] = done ? 0 : T.nodes[L + 1].v[N_NS];
        en[0] = done ? 0 : T.energy[L];
        en[1] = done ? 0 : T.energy[L + 1];
    }
    if (done) return;
    for (int j = threadIdx.x; j < 2 * F; j += blockDim.x) fc[j] = T.fcand[(size_t)L * F + j];
}

extern "C" {

int tgbm_apply_split_fetch(tgbm_session* s, int task, const tgbm_decision* d, const tgbm_grow_params* p,
                           int* left_id, double* gain_f, int* thr_f, int* ml_f, double* energy, int* n_rows) {
    return dev_guard([&] {
        spec_resolve(s);
        if (s->an != 1 || task != s->a0)
            throw std::runtime_error("host-driven growth needs exactly one active task, the one named");
        const int F = s->F;
        s->ensure_readback();
        struct {
            int leaf, type, feat, feat2, thr, mgl;
            double gain;
        } cur{d->leaf, d->type, d->feature, d->feature2, d->threshold, d->missing_go_left, d->gain};
        static_assert(offsetof(Task, cur_gain) - offsetof(Task, cur_leaf) == 24, "layout");
        HIP_OK(hipMemcpyAsync(reinterpret_cast<char*>(s->tasks_d + task) + offsetof(Task, cur_leaf), &cur,
                              sizeof cur, hipMemcpyHostToDevice, s->stream));
        const GrowP P = s->growp(p);
        split_step(s, P, 1, p);
        // the split leaf's rows bound the children's (all the task's rows when unknown)
        int parent_rows = s->th[task].has_sample ? s->th[task].n_sample : s->th[task].n_rows;
        if (d->leaf >= 0 && d->leaf < (int)s->node_rows_h.size() && s->node_rows_h[d->leaf] >= 0)
            parent_rows = s->node_rows_h[d->leaf];
        const int eb = std::max(1, std::min(256, (parent_rows + 1023) / 1024));
        hipLaunchKernelGGL(k_child_energy, dim3(eb, 2), dim3(256), 0, s->stream, s->tasks_d, task);
        hipLaunchKernelGGL(k_split_readback, dim3(1), dim3(256), 0, s->stream, (const Task*)s->tasks_d, task, F,
                           s->rb_d);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(s->stream));
        const int* hdr = (const int*)s->rb_h;
        if (hdr[0]) {
            *left_id = -1;
            return;
        }
        const int L = hdr[1];
        const long long* en = (const long long*)(s->rb_h + 24);
        const FCand* fc = (const FCand*)(s->rb_h + 40);
        for (int j = 0; j < 2 * F; ++j) {
            gain_f[j] = fc[j].thr >= 0 ? fc[j].gain : -1e300;
            thr_f[j] = fc[j].thr;
            ml_f[j] = fc[j].ml;
        }
        for (int c = 0; c < 2; ++c) {
            energy[c] = std::ldexp((double)en[c], -hdr[2]);
            n_rows[c] = hdr[3 + c];
        }
        if ((int)s->node_rows_h.size() < L + 2) s->node_rows_h.resize(L + 2, -1);
        s->node_rows_h[L] = hdr[3];
        s->node_rows_h[L + 1] = hdr[4];
        *left_id = L;
    });
}

int tgbm_apply_split(tgbm_session* s, int task, const tgbm_decision* d, const tgbm_grow_params* p,
                     int* left_id) {
    return dev_guard([&] {
        spec_resolve(s);
        if (s->an != 1 || task != s->a0)
            throw std::runtime_error("host-driven growth needs exactly one active task, the one named");
        struct {
            int leaf, type, feat, feat2, thr, mgl;
            double gain;
        } cur{d->leaf, d->type, d->feature, d->feature2, d->threshold, d->missing_go_left, d->gain};
        static_assert(offsetof(Task, cur_gain) - offsetof(Task, cur_leaf) == 24, "layout");
        HIP_OK(hipMemcpyAsync(reinterpret_cast<char*>(s->tasks_d + task) + offsetof(Task, cur_leaf), &cur,
                              sizeof cur, hipMemcpyHostToDevice, s->stream));
        const GrowP P = s->growp(p);
        split_step(s, P, 1, p);
        Task T = s->read_task(task);
        if (T.done) {
            *left_id = -1;
            return;
        }
        const int L = T.n_nodes - 2;
        DNode nd[2];
        HIP_OK(hipMemcpy(nd, s->tasks_h[task].nodes + L, sizeof nd, hipMemcpyDeviceToHost));
        launch_energy(s->stream, s->tasks_d, task, L, nd[0].seg_count);
        launch_energy(s->stream, s->tasks_d, task, L + 1, nd[1].seg_count);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(s->stream));
        *left_id = L;
    });
}

// Enqueue the end of every active task's tree: leaf table, TORB sums, prediction update, loss,
// export into the next host slot (k_export writes the mapped slot; an event marks its end).
// The tree end: final numbering (multi-leaf trees), leaf ids, TORB, predictions, loss, export
// into the tree's host slot.
static void enqueue_end(tgbm_session* s, const tgbm_session::EndArgs& E) {
    EventPair ep;
    if (s->timing) s->tic(3, 0.0, ep);
    if (E.spec) launch_spec_finalize(s->stream, s->S, s->td(), s->an, E.P);
    launch_tree_end(s->stream, s->S, s->td(), s->an, E.walk, s->max_rows(),
                    (const TorbChunk* const*)(s->chunk_ptrs_d + s->a0), s->n_chunks_d + s->a0, s->max_chunks, E.P,
                    E.min_cover, E.iid, E.lr, E.pairwise ? 1 : 0);
    if (E.pairwise && E.max_groups > 0) launch_loss(s->stream, s->td(), s->an, E.max_groups);
    if (s->timing) s->toc(ep);
    launch_export(s->stream, s->td(), s->an, E.L, s->export_hd[E.slot], E.pairwise ? 1 : 0);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(s->export_ev[E.slot], s->stream));
}

// A multi-leaf tree whose end was queued behind its captured loop: the end's first kernel (the
// final numbering) reports whether the loop finished the tree, and if not the other kernels do
// nothing.  Wait for that report (the host meanwhile did its own work), and for an unfinished
// tree grow on step by step, queue the tree end again and let it drain (so no late report meets
// the next tree's).  Every entry point that queues work or reads task state calls this first.
static void spec_resolve(tgbm_session* s) {
    if (!s->spec_wait) return;
    s->spec_wait = false;
    const tgbm_session::EndArgs E = s->spec_end;
    volatile int* flag = s->flag_h;
    const int t0 = s->a0, t1 = s->a0 + s->an;
    auto reported = [&] {
        for (int k = t0; k < t1; ++k)
            if (!flag[2 * k]) return false;
        return true;
    };
    auto all_finished = [&](int& steps) {
        bool ok = true;
        steps = 0;
        for (int k = t0; k < t1; ++k) {
            ok = ok && flag[2 * k] == 1;
            steps = std::max(steps, (int)flag[2 * k + 1]);
        }
        return ok;
    };
    // Poll the flags alone; ask the runtime (which takes its locks) only every 200 us, so the
    // wait does not slow the runtime's own work on the stream.
    auto last = std::chrono::steady_clock::now();
    while (!reported()) {
        const auto now = std::chrono::steady_clock::now();
        if (now - last < std::chrono::microseconds(200)) continue;
        last = now;
        const hipError_t q = hipEventQuery(s->export_ev[E.slot]);
        if (q == hipSuccess && !reported()) throw std::runtime_error("internal: the tree end reported nothing");
        if (q != hipSuccess && q != hipErrorNotReady) HIP_OK(q);
    }
    int steps = 0;
    bool done = all_finished(steps);
    const int limit = std::max(1, E.p.max_leaves - 1);
    if (!done) {
        // tasks whose trees have ended skip the steps and the second end (done, end_done)
        for (int k = s->spec_k_used; k < limit && !done; ++k) {
            spec_step(s, E.P, &E.p);
            done = spec_done(s, nullptr);
        }
        if (!done) throw std::runtime_error("internal: multi-leaf growth did not finish the tree");
        enqueue_end(s, E);
        HIP_OK(hipStreamSynchronize(s->stream));
        if (!all_finished(steps)) throw std::runtime_error("internal: multi-leaf growth did not finish the tree");
    }
    // the next loop's steps: the most splitting steps any of the last 32 trees needed, so a tree
    // rarely needs the step-by-step path.  No step is added for a selection that finds the tree
    // finished: the final numbering replays to the end itself (the pending candidates included)
    // and reports an unfinished tree.
    s->spec_hist[s->spec_hist_n++ % 32] = std::max(1, steps);
    int kmax = 1;
    for (int k = 0; k < std::min(s->spec_hist_n, 32); ++k) kmax = std::max(kmax, s->spec_hist[k]);
    s->spec_k = std::min(limit, kmax);
}

static void tree_end_enqueue(tgbm_session* s, double learning_rate, double min_cover, int iid,
                             const tgbm_grow_params* p) {
    spec_resolve(s);
    if (s->pending.size() >= 2) throw std::runtime_error("at most two tree ends may wait for collection");
    tgbm_session::EndArgs E;
    E.P = s->growp(p);
    E.p = *p;
    E.lr = learning_rate;
    E.min_cover = min_cover;
    E.iid = iid;
    E.walk = !s->all_contiguous();
    E.spec = s->spec_tree;
    // training-loss partials (the prediction update and the pairwise group pass write them)
    E.pairwise = s->objective == 5;
    E.max_groups = 0;
    for (int t = s->a0; t < s->a0 + s->an; ++t) {
        TaskHost& H = s->th[t];
        const int need = 2 * loss_chunks(H.cap_rows) + H.n_groups;
        if (need > H.loss_cap) {
            H.loss_d = s->alloc<double>((size_t)need);
            H.loss_cap = need;
            s->tasks_h[t].loss_part = H.loss_d;
            s->set_field_ptr(t, offsetof(Task, loss_part), H.loss_d);
        }
        E.max_groups = std::max(E.max_groups, H.n_groups);
    }
    int bmax = 1;
    for (int t = 0; t < s->T; ++t) bmax = std::max(bmax, s->th[t].n_buckets);
    E.L = export_layout(s->max_nodes, bmax);
    const size_t need = E.L.stride * s->T;
    if (need > s->export_bytes) {
        // the layout grows only when a task's bucket count does (a new fit): no tree is pending
        if (!s->pending.empty()) throw std::runtime_error("internal: export slots resized with a tree pending");
        HIP_OK(hipStreamSynchronize(s->stream));
        for (int k = 0; k < 2; ++k) {
            if (s->export_h[k]) HIP_OK(hipHostFree(s->export_h[k]));
            s->export_h[k] = nullptr;
            HIP_OK(hipHostMalloc((void**)&s->export_h[k], need, hipHostMallocMapped));
            HIP_OK(hipHostGetDevicePointer((void**)&s->export_hd[k], s->export_h[k], 0));
            if (!s->export_ev[k]) HIP_OK(hipEventCreateWithFlags(&s->export_ev[k], hipEventDisableTiming));
        }
        s->export_bytes = need;
    }
    E.slot = s->next_slot;
    enqueue_end(s, E);
    if (s->spec_launched) {
        // the captured loop ran without a wait: resolve the tree when the host next needs it
        s->spec_launched = false;
        s->spec_end = E;
        s->spec_wait = true;
    }
    s->pending.push_back(tgbm_session::PendingExport{E.slot, s->a0, s->an, E.L, E.spec});
    s->spec_tree = false;
    s->next_slot ^= 1;
}

// Wait for the oldest enqueued tree end and unpack its trees (one per task of its active range).
static void tree_collect(tgbm_session* s, const tgbm_grow_params* p, int* n_nodes, int* node_ints, double* node_dbl,
                         double* values, int* cover, double* prefix) {
    if (s->pending.empty()) throw std::runtime_error("no tree end to collect");
    const tgbm_session::PendingExport pe = s->pending.front();
    // the unresolved multi-leaf tree's export is final only once the tree is resolved
    if (s->spec_wait && s->spec_end.slot == pe.slot) spec_resolve(s);
    s->pending.pop_front();
    HIP_OK(hipEventSynchronize(s->export_ev[pe.slot]));
    const ExportLayout& L = pe.L;
    const int MN = s->max_nodes;
    for (int i = 0; i < pe.an; ++i) {
        const int t = pe.a0 + i;
        const char* base = s->export_h[pe.slot] + L.stride * i;
        const int nn = *(const int*)base;
        if (nn < 1 || nn > MN) throw std::runtime_error("internal: exported tree size out of range");
        const int err = *((const int*)base + 1);
        if (err == 1)
            throw std::runtime_error("internal: a partition placed a different number of rows left than its split's histogram counts");
        if (err != 0) throw std::runtime_error("internal: multi-leaf growth left the tree unfinished or out of node space");
        n_nodes[i] = nn;
        s->th[t].last_loss = *((const double*)base + 1);
        const int B = s->th[t].n_buckets;
        const DNode* nodes = (const DNode*)(base + L.off_nodes);
        for (int k = 0; k < nn; ++k) {
            std::memcpy(node_ints + ((size_t)i * MN + k) * TGBM_NODE_INTS, nodes[k].v, sizeof(int) * 10);
            node_dbl[((size_t)i * MN + k) * 2] = nodes[k].value;
            node_dbl[((size_t)i * MN + k) * 2 + 1] = nodes[k].gain;
        }
        std::memcpy(values + (size_t)i * MN, base + L.off_values, sizeof(double) * nn);
        std::memcpy(cover + (size_t)i * MN, base + L.off_cover, sizeof(int) * nn);
        std::memcpy(prefix + (size_t)i * MN * MAX_BUCKETS, base + L.off_prefix, sizeof(double) * nn * B);
        if (s->timing) {  // rows of the root and of every child histogram build, in launch order
            const int steps = pe.spec ? 0 : p->max_leaves - 1;
            s->log_rows[0].push_back(nodes[0].seg_count);
            for (int k = 0; k < steps; ++k) {
                const bool grew = 2 * k + 2 < nn;
                s->log_rows[5].push_back(grew ? std::min(nodes[2 * k + 1].seg_count, nodes[2 * k + 2].seg_count) : 0);
                s->log_rows[1].push_back(grew ? nodes[2 * k + 1].seg_count + nodes[2 * k + 2].seg_count : 0);
            }
            for (int k = 0; k < nn; ++k) {
                if (nodes[k].v[N_LEAF]) continue;
                const int Lc = nodes[k].v[N_LEFT], Rc = nodes[k].v[N_RIGHT];
                const double m = std::min(nodes[Lc].seg_count, nodes[Rc].seg_count);
                s->t_bytes[5] += m * s->F + m * (s->hconst ? 8.0 : 16.0) + 4.0 * m + (double)s->F * s->nb * 24.0;
            }
        }
    }
}

int tgbm_tree_end(tgbm_session* s, double learning_rate, double min_cover, int iid,
                  const tgbm_grow_params* p, int* n_nodes, int* node_ints, double* node_dbl,
                  double* values, int* cover, double* prefix) {
    return dev_guard([&] {
        spec_resolve(s);
        if (!s->pending.empty()) throw std::runtime_error("collect the pending tree ends first");
        tree_end_enqueue(s, learning_rate, min_cover, iid, p);
        tree_collect(s, p, n_nodes, node_ints, node_dbl, values, cover, prefix);
    });
}

int tgbm_tree_end_async(tgbm_session* s, double learning_rate, double min_cover, int iid, const tgbm_grow_params* p) {
    return dev_guard([&] { tree_end_enqueue(s, learning_rate, min_cover, iid, p); });
}

int tgbm_tree_collect(tgbm_session* s, const tgbm_grow_params* p, int* n_nodes, int* node_ints, double* node_dbl,
                      double* values, int* cover, double* prefix) {
    return dev_guard([&] { tree_collect(s, p, n_nodes, node_ints, node_dbl, values, cover, prefix); });
}

int tgbm_task_train_loss(tgbm_session* s, int task, double* loss) {
    return dev_guard([&] {
        if (task < 0 || task >= s->T) throw std::runtime_error("task out of range");
        *loss = s->th[task].last_loss;
    });
}

int tgbm_timer_enable(tgbm_session* s, int on) {
    return dev_guard([&] {
        spec_resolve(s);
        s->drain_events();
        s->timing = on != 0;
        s->flush = on == 2;
        if (s->flush && !s->flush_buf) {
            HIP_OK(hipMalloc(&s->flush_buf, tgbm_session::kFlushBytes));
            HIP_OK(hipMemsetAsync(s->flush_buf, 0, tgbm_session::kFlushBytes, s->stream));
        }
        for (int k = 0; k < 8; ++k) {
            s->t_ms[k] = 0;
            s->t_n[k] = 0;
            s->t_bytes[k] = 0;
            s->log_ms[k].clear();
            s->log_rows[k].clear();
        }
    });
}

int tgbm_timer_log(tgbm_session* s, int which, double* ms, double* rows, int cap) {
    int n = -1;
    if (dev_guard([&] {
            spec_resolve(s);
            if (which < 0 || which >= 8) throw std::runtime_error("timer index");
            s->drain_events();
            n = (int)s->log_ms[which].size();
            for (int i = 0; i < n && i < cap; ++i) {
                ms[i] = s->log_ms[which][i];
                rows[i] = i < (int)s->log_rows[which].size() ? s->log_rows[which][i] : -1.0;
            }
        }) != 0)
        return -1;
    return n;
}

int tgbm_timer_read(tgbm_session* s, int which, double* ms, long long* launches, double* bytes) {
    return dev_guard([&] {
        spec_resolve(s);
        if (which < 0 || which >= 8) throw std::runtime_error("timer index");
        s->drain_events();
        *ms = s->t_ms[which];
        *launches = s->t_n[which];
        *bytes = s->t_bytes[which];
    });
}

}  // extern "C"
