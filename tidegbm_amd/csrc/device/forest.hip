// Ensemble predict on gfx950 (reference Booster::predict_raw_binned, booster.cpp:744-773):
// raw score = init_score + sum over trees, in tree order, of lr_t * values_t[leaf(row)].
// Built with -ffp-contract=off so the running sum rounds exactly like the reference loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/tgbm_device.h"
#include "dev_common.h"
#include "forest.h"
#include "launch.h"

using namespace tgbm;

extern thread_local std::string g_dev_err;  // defined in session.hip

#define FHIP(expr)                                                                           \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) +   \
                                     " at " #expr);                                          \
    } while (0)

void tgbm_forest::reserve(int n) {
    if (!stream) FHIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));   // first single-call predict
    if (n <= cap) return;
    for (void* p : {(void*)Xd, (void*)bins, (void*)bins_col_dummy, (void*)outd})
        if (p) (void)hipFree(p);
    Xd = nullptr;
    bins = nullptr;
    outd = nullptr;
    cap = 0;
    const int c = std::max(n, 1024);
    FHIP(hipMalloc(&Xd, sizeof(double) * (size_t)c * F));
    FHIP(hipMalloc(&bins, (size_t)c * fp + 16));
    FHIP(hipMalloc(&outd, sizeof(double) * c));
    cap = c;
}

// Leaf reached by one binned row from node `k` (tree.cpp apply_one): bin 0 is missing and goes
// the node's missing direction; a pair node compares the hashed bin pair with its threshold.
__device__ __forceinline__ int walk_tree(const FNode* __restrict__ nodes, const unsigned char* __restrict__ flags,
                                         const unsigned long long* __restrict__ cmask,
                                         const unsigned char* __restrict__ row, int k, int nb) {
    unsigned char fl = flags[k];
    while (!(fl & 1)) {
        const FNode nd = nodes[k];
        const int b1 = row[nd.fx & 0xffff];
        bool left;
        if (((fl >> 1) & 3) >= 2) {   // categorical: the bin's bit in the node's mask
            left = (cmask[(size_t)nd.thr * 4 + (b1 >> 6)] >> (b1 & 63)) & 1ull;
        } else if (((fl >> 1) & 3) == 1) {
            const int b2 = row[nd.fx >> 16];
            if (b1 == 0 || b2 == 0) left = (fl >> 3) & 1;
            else left = (int)(((unsigned)b1 * 2654435761u + (unsigned)b2 * 40503u) % (unsigned)nb) <= nd.thr;
        } else {
            left = b1 == 0 ? ((fl >> 3) & 1) : (b1 <= nd.thr);
        }
        k = left ? nd.left : nd.right;
        fl = flags[k];
    }
    return k;
}

// One thread per row; the row's bins are re-read through L1/L2 at every visited node.
__global__ __launch_bounds__(256) void k_forest(const FNode* __restrict__ nodes,
                                                const unsigned char* __restrict__ flags,
                                                const double* __restrict__ values,
                                                const int* __restrict__ roots,
                                                const double* __restrict__ lr, int n_trees,
                                                double init, const unsigned char* __restrict__ bins,
                                                int fp, int n, int nb, const unsigned long long* __restrict__ cmask,
                                                double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned char* row = bins + (size_t)i * fp;
    double s = init;
    for (int t = 0; t < n_trees; ++t) s += lr[t] * values[walk_tree(nodes, flags, cmask, row, roots[t], nb)];
    out[i] = s;
}

__global__ void k_u16_to_u8(const unsigned short* __restrict__ in, int n, int F, unsigned char* __restrict__ out, int fp) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (size_t)n * F) return;
    const size_t i = j / F, f = j % F;
    const unsigned short v = in[j];
    out[i * fp + f] = (unsigned char)(v > 255 ? 255 : v);
}

template <class Fn>
static int fguard(Fn&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception& e) {
        g_dev_err = e.what();
        return -1;
    }
}

// Host-side node tables of n_trees trees (export layout) whose first node gets absolute id `base`.
struct HostNodes {
    std::vector<FNode> nodes;
    std::vector<SNode> snodes;
    std::vector<unsigned char> flags;
    std::vector<int> tree_nodes;   // per tree
    std::vector<unsigned long long> cmask;   // 4 words per categorical node, in node order
    std::vector<int> tree_cats;    // categorical nodes per tree
    int total = 0;
    int max_depth = 0;             // deepest node of any tree (the root at depth 0)
    bool numeric_only = true;      // no pair or categorical node
};

// cat_masks: 4 words per node (all nodes of all trees), or null when no node is categorical;
// a categorical node's mask gets index cat_base + (its rank among the categorical nodes).
static HostNodes build_nodes(int n_trees, const int* tree_nodes, const int* node_ints, int n_features, int base,
                             const unsigned long long* cat_masks = nullptr, int cat_base = 0) {
    HostNodes h;
    h.tree_cats.assign(n_trees, 0);
    for (int t = 0; t < n_trees; ++t) h.total += tree_nodes[t];
    h.nodes.resize(std::max(h.total, 1));
    h.snodes.resize(std::max(h.total, 1));
    h.flags.resize(std::max(h.total, 1));
    h.tree_nodes.assign(tree_nodes, tree_nodes + n_trees);
    int off = 0;
    std::vector<int> depth;
    for (int t = 0; t < n_trees; ++t) {
        if (tree_nodes[t] < 1 || tree_nodes[t] > 255) throw std::runtime_error("trees must have 1..255 nodes");
        depth.assign(tree_nodes[t], 0);
        for (int k = 0; k < tree_nodes[t]; ++k) {
            const int* v = node_ints + (size_t)(off + k) * TGBM_NODE_INTS;
            const bool leaf = v[0] != 0;
            FNode nd{0, 0, 0, 0};
            if (!leaf) {
                // children come after their parent (every grower appends them, tree.cpp:447-453), so
                // every walk ends; a link back or to itself would loop the device walks
                if (v[2] < 0 || v[2] >= n_features || v[5] <= k || v[6] <= k || v[5] >= tree_nodes[t] || v[6] >= tree_nodes[t])
                    throw std::runtime_error("malformed tree node");
                if (v[1] < 0 || v[1] > 3) throw std::runtime_error("unknown split type");
                const int f2 = v[1] == 1 ? v[3] : 0;
                if (f2 < 0 || f2 >= n_features) throw std::runtime_error("malformed pair node");
                nd.fx = v[2] | (f2 << 16);
                nd.thr = v[4];
                if (v[1] >= 2) {
                    if (!cat_masks) throw std::runtime_error("categorical node without its bin mask");
                    nd.thr = cat_base + (int)(h.cmask.size() / 4);
                    const unsigned long long* m = cat_masks + (size_t)(off + k) * 4;
                    h.cmask.insert(h.cmask.end(), m, m + 4);
                    h.tree_cats[t] += 1;
                }
                nd.left = base + off + v[5];
                nd.right = base + off + v[6];
                // children come after their parent, so a node's depth (its longest path from the
                // root) is final when it is reached
                depth[v[5]] = std::max(depth[v[5]], depth[k] + 1);
                depth[v[6]] = std::max(depth[v[6]], depth[k] + 1);
                h.max_depth = std::max(h.max_depth, depth[k] + 1);
                if (v[1] != 0) h.numeric_only = false;
            }
            h.nodes[off + k] = nd;
            h.flags[off + k] = (unsigned char)((leaf ? 1 : 0) | ((v[1] & 3) << 1) | ((v[7] ? 1 : 0) << 3));
            SNode sn{0, 0, 0, 0, 0, h.flags[off + k]};
            if (!leaf) {
                sn.f = (unsigned short)v[2];
                sn.left = (unsigned char)v[5];
                sn.right = (unsigned char)v[6];
                if (v[1] < 2) {
                    if (v[4] < 0 || v[4] > 255) throw std::runtime_error("split threshold outside the uint8 bin range");
                    sn.f2 = (unsigned short)(v[1] == 1 ? v[3] : 0);
                    sn.thr = (unsigned char)v[4];
                } else {
                    // categorical: f2 holds the index of the node's bin mask in the forest's table
                    if (nd.thr > 0xffff) throw std::runtime_error("more than 65536 categorical nodes in a forest");
                    sn.f2 = (unsigned short)nd.thr;
                }
            }
            h.snodes[off + k] = sn;
        }
        off += tree_nodes[t];
    }
    return h;
}

// Allocate a forest of n_trees / total nodes whose first `keep_trees` trees (keep_nodes nodes)
// are copied from `base` on the device, followed by the trees of `h`.
static tgbm_forest* assemble(const tgbm_forest* base, int keep_trees, int keep_nodes, const HostNodes& h,
                             const double* values, const double* lr, int n_features, int n_bins, double init_score,
                             const double* edges, const int* n_edges) {
    auto* f = new tgbm_forest();
    std::unique_ptr<tgbm_forest> guard(f);
    const int n_new = (int)h.tree_nodes.size();
    f->F = n_features;
    f->fp = (n_features + 15) / 16 * 16;
    f->nb = n_bins;
    f->n_trees = keep_trees + n_new;
    f->n_nodes = keep_nodes + h.total;
    f->init = init_score;
    f->toff_h.assign(base ? base->toff_h.begin() : f->toff_h.begin(), base ? base->toff_h.begin() + keep_trees + 1
                                                                            : f->toff_h.begin());
    if (f->toff_h.empty()) f->toff_h.push_back(0);
    for (int t = 0; t < n_new; ++t) f->toff_h.push_back(f->toff_h.back() + h.tree_nodes[t]);
    if (base) f->cat_off_h.assign(base->cat_off_h.begin(), base->cat_off_h.begin() + keep_trees + 1);
    if (f->cat_off_h.empty()) f->cat_off_h.push_back(0);
    const int keep_cat = f->cat_off_h.back();
    for (int t = 0; t < n_new; ++t) f->cat_off_h.push_back(f->cat_off_h.back() + h.tree_cats[t]);
    f->n_cat = f->cat_off_h.back();
    // kept trees of the base are covered by its whole-forest figures (conservative)
    f->max_depth = std::max(base && keep_trees > 0 ? base->max_depth : 0, h.max_depth);
    f->numeric_only = (!base || keep_trees == 0 || base->numeric_only) && h.numeric_only;
    f->max_chunk_nodes = 1;
    for (int t0 = 0; t0 < f->n_trees; t0 += SERVE_TREES_PER_CHUNK) {
        const int t1 = std::min(f->n_trees, t0 + SERVE_TREES_PER_CHUNK);
        f->max_chunk_nodes = std::max(f->max_chunk_nodes, f->toff_h[t1] - f->toff_h[t0]);
    }
    // Every array in one device allocation, filled by one upload of a host image of the same
    // layout (the new trees' parts and the tables) followed by device copies of the base's kept
    // parts: a forest costs one allocation and one free, which matters to a server publishing
    // versions while other threads launch work.
    const size_t NN = std::max(f->n_nodes, 1), NT = std::max(f->n_trees, 1), NF = (size_t)n_features;
    size_t bytes = 0;
    auto carve = [&](size_t b) {
        const size_t o = bytes;
        bytes = (bytes + b + 255) / 256 * 256;
        return o;
    };
    const size_t o_nodes = carve(sizeof(FNode) * NN), o_snodes = carve(sizeof(SNode) * NN),
                 o_values = carve(sizeof(double) * NN), o_lr = carve(sizeof(double) * NT),
                 o_edges = carve(sizeof(double) * NF * NBIN), o_cmask = carve(sizeof(unsigned long long) * 4 * f->n_cat),
                 o_toff = carve(sizeof(int) * (f->n_trees + 1)), o_roots = carve(sizeof(int) * NT),
                 o_nedges = carve(sizeof(int) * NF), o_flags = carve(NN);
    FHIP(hipMalloc(&f->block, bytes));
    char* const d = (char*)f->block;
    f->nodes = (FNode*)(d + o_nodes);
    f->snodes = (SNode*)(d + o_snodes);
    f->values = (double*)(d + o_values);
    f->lr = (double*)(d + o_lr);
    f->edges = (double*)(d + o_edges);
    f->cmask = f->n_cat ? (unsigned long long*)(d + o_cmask) : nullptr;
    f->tree_off = (int*)(d + o_toff);
    f->roots = (int*)(d + o_roots);
    f->n_edges = (int*)(d + o_nedges);
    f->flags = (unsigned char*)(d + o_flags);
    std::vector<char> img(bytes, 0);
    auto put = [&](size_t o, const void* src, size_t b) {
        if (b) std::memcpy(img.data() + o, src, b);
    };
    if (h.total) {
        put(o_nodes + sizeof(FNode) * keep_nodes, h.nodes.data(), sizeof(FNode) * h.total);
        put(o_snodes + sizeof(SNode) * keep_nodes, h.snodes.data(), sizeof(SNode) * h.total);
        put(o_values + sizeof(double) * keep_nodes, values, sizeof(double) * h.total);
        put(o_flags + keep_nodes, h.flags.data(), h.total);
    }
    if (n_new) put(o_lr + sizeof(double) * keep_trees, lr, sizeof(double) * n_new);
    if (!h.cmask.empty()) put(o_cmask + sizeof(unsigned long long) * 4 * keep_cat, h.cmask.data(), sizeof(unsigned long long) * h.cmask.size());
    put(o_toff, f->toff_h.data(), sizeof(int) * f->toff_h.size());
    put(o_roots, f->toff_h.data(), sizeof(int) * std::max(f->n_trees, 1));
    if (!base && edges) {
        put(o_edges, edges, sizeof(double) * NF * NBIN);
        put(o_nedges, n_edges, sizeof(int) * NF);
    }
    FHIP(hipMemcpyAsync(d, img.data(), bytes, hipMemcpyHostToDevice, nullptr));
    if (base) {   // the base's kept trees, edges and masks, copied after the image on the same stream
        auto dd = [&](void* dst, const void* src, size_t b) {
            if (b) FHIP(hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToDevice, nullptr));
        };
        if (keep_trees > 0) {
            dd(f->nodes, base->nodes, sizeof(FNode) * keep_nodes);
            dd(f->flags, base->flags, keep_nodes);
            dd(f->snodes, base->snodes, sizeof(SNode) * keep_nodes);
            dd(f->values, base->values, sizeof(double) * keep_nodes);
            dd(f->lr, base->lr, sizeof(double) * keep_trees);
        }
        if (keep_cat) dd(f->cmask, base->cmask, sizeof(unsigned long long) * 4 * keep_cat);
        dd(f->edges, base->edges, sizeof(double) * NF * NBIN);
        dd(f->n_edges, base->n_edges, sizeof(int) * NF);
    }
    FHIP(hipStreamSynchronize(nullptr));
    return guard.release();
}

extern "C" {

int tgbm_forest_create(tgbm_forest** out, int n_features, int n_bins, int n_trees, const int* tree_nodes,
                       const int* node_ints, const double* values, const double* lr, double init_score,
                       const double* edges, const int* n_edges, const unsigned long long* cat_masks) {
    *out = nullptr;
    return fguard([&] {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
            throw std::runtime_error("GPU device not available: no HIP device found");
        if (n_features <= 0 || n_trees < 0) throw std::runtime_error("bad forest shape");
        const HostNodes h = build_nodes(n_trees, tree_nodes, node_ints, n_features, 0, cat_masks, 0);
        *out = assemble(nullptr, 0, 0, h, values, lr, n_features, n_bins, init_score, edges, n_edges);
    });
}

int tgbm_forest_append(tgbm_forest** out, const tgbm_forest* base, int keep_trees, int n_trees, const int* tree_nodes,
                       const int* node_ints, const double* values, const double* lr,
                       const unsigned long long* cat_masks) {
    *out = nullptr;
    return fguard([&] {
        if (!base) throw std::runtime_error("forest append needs a base forest");
        if (keep_trees < 0 || keep_trees > base->n_trees || n_trees < 0) throw std::runtime_error("bad forest append shape");
        const int keep_nodes = base->toff_h[keep_trees];
        const HostNodes h = build_nodes(n_trees, tree_nodes, node_ints, base->F, keep_nodes, cat_masks,
                                        base->cat_off_h[keep_trees]);
        *out = assemble(base, keep_trees, keep_nodes, h, values, lr, base->F, base->nb, base->init, nullptr, nullptr);
    });
}

void tgbm_forest_destroy(tgbm_forest* f) { delete f; }

static void forest_run(tgbm_forest* f, int n, double* out) {
    if (n > 0) {
        hipLaunchKernelGGL(k_forest, dim3((n + 255) / 256), dim3(256), 0, f->stream, f->nodes, f->flags, f->values,
                           f->roots, f->lr, f->n_trees, f->init, f->bins, f->fp, n, f->nb, f->cmask, f->outd);
        FHIP(hipGetLastError());
        FHIP(hipMemcpyAsync(out, f->outd, sizeof(double) * n, hipMemcpyDeviceToHost, f->stream));
    }
    FHIP(hipStreamSynchronize(f->stream));
}

int tgbm_forest_predict(tgbm_forest* f, const double* X, int n, double* out) {
    return fguard([&] {
        if (n < 0) throw std::runtime_error("n < 0");
        if (n == 0) return;
        f->reserve(n);
        FHIP(hipMemcpyAsync(f->Xd, X, sizeof(double) * (size_t)n * f->F, hipMemcpyHostToDevice, f->stream));
        FHIP(hipMemsetAsync(f->bins, 0, (size_t)n * f->fp, f->stream));
        launch_bin(f->stream, f->Xd, n, f->F, f->edges, f->n_edges, f->bins, f->fp, nullptr, n + 64);
        FHIP(hipGetLastError());
        forest_run(f, n, out);
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Predict without allocating: the staging of tgbm_forest_predict (raw rows, their bins, the
// scores) lives in `scratch`, one file-scope array per GPU that the process holds until it ends.
// In host code `scratch` names a host-side stand-in; the array's device address comes from
// hipGetSymbolAddress, once per device, and only that address is handed to copies and kernels.
namespace {

__device__ __align__(256) unsigned char scratch[TGBM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];   // one call at a time per device (the calls share the array)
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's scratch array, looked up on the device's first call.
unsigned char* scratch_on(int dev) {
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char*)ptr;
    }
    return g_scratch_ptr[dev];
}

size_t up256(size_t x) { return (x + 255) / 256 * 256; }

// Rows the scratch holds for a forest: raw rows (8 F bytes each), bins (fp bytes each, plus 64
// rows and 16 bytes of slack) and scores (8 bytes each), each region rounded up to 256 bytes.
int noalloc_max_rows(const tgbm_forest* f) {
    const size_t per = (size_t)8 * f->F + f->fp + 8;
    const size_t n = (TGBM_NOALLOC_SCRATCH_BYTES - 4 * 256 - 16 - (size_t)64 * f->fp) / per;
    return (int)std::min<size_t>(n, 1 << 30);
}

}  // namespace

extern "C" {

int tgbm_forest_predict_noalloc(tgbm_forest* f, const double* X, int n, double* out) {
    return fguard([&] {
        if (!f) throw std::runtime_error("no forest");
        if (n < 0) throw std::runtime_error("n < 0");
        if (n == 0) return;
        const int max_rows = noalloc_max_rows(f);
        if (n > max_rows)
            throw std::runtime_error("tgbm_forest_predict_noalloc: " + std::to_string(n) + " rows exceed the " +
                                     std::to_string(max_rows) + " rows of " + std::to_string(f->F) +
                                     " features its fixed scratch holds (" +
                                     std::to_string(TGBM_NOALLOC_SCRATCH_BYTES >> 20) + " MiB)");
        int dev = 0;
        FHIP(hipGetDevice(&dev));
        if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device id out of range");
        std::lock_guard<std::mutex> lk(g_scratch_mu[dev]);
        unsigned char* base = scratch_on(dev);
        double* Xd = (double*)base;
        unsigned char* bins = base + up256(sizeof(double) * (size_t)n * f->F);
        double* outd = (double*)(bins + up256((size_t)(n + 64) * f->fp + 16));
        // the forest's own stream when a plain predict created it, else the null stream
        hipStream_t st = f->stream;
        FHIP(hipMemcpyAsync(Xd, X, sizeof(double) * (size_t)n * f->F, hipMemcpyHostToDevice, st));
        FHIP(hipMemsetAsync(bins, 0, (size_t)n * f->fp, st));
        launch_bin(st, Xd, n, f->F, f->edges, f->n_edges, bins, f->fp, nullptr, n + 64);
        FHIP(hipGetLastError());
        hipLaunchKernelGGL(k_forest, dim3((n + 255) / 256), dim3(256), 0, st, f->nodes, f->flags, f->values, f->roots,
                           f->lr, f->n_trees, f->init, bins, f->fp, n, f->nb, f->cmask, outd);
        FHIP(hipGetLastError());
        FHIP(hipMemcpyAsync(out, outd, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        FHIP(hipStreamSynchronize(st));
    });
}

int tgbm_forest_predict_binned_u16(tgbm_forest* f, const unsigned short* xb, int n, double* out) {
    return fguard([&] {
        if (n < 0) throw std::runtime_error("n < 0");
        if (n == 0) return;
        f->reserve(n);
        unsigned short* d = nullptr;
        FHIP(hipMalloc(&d, sizeof(unsigned short) * (size_t)n * f->F));
        try {
            FHIP(hipMemcpyAsync(d, xb, sizeof(unsigned short) * (size_t)n * f->F, hipMemcpyHostToDevice, f->stream));
            FHIP(hipMemsetAsync(f->bins, 0, (size_t)n * f->fp, f->stream));
            const size_t tot = (size_t)n * f->F;
            hipLaunchKernelGGL(k_u16_to_u8, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, f->stream, d, n, f->F,
                               f->bins, f->fp);
            FHIP(hipGetLastError());
            forest_run(f, n, out);
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        (void)hipFree(d);
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Validation set scored tree by tree during a fit (reference Booster::fit with an eval set,
// booster.cpp:459-478 and :707-732; scores as predict_multi_binned, booster.cpp:775-799).

struct tgbm_evalset {
    int F = 0, fp = 0, nb = 255, n = 0, n_groups = 0;
    unsigned char* bins = nullptr;    // n x fp row-major uint8 bins
    double* y = nullptr;
    double* pred = nullptr;           // init + sum of lr * leaf value over the trees added so far
    int* order = nullptr;             // rows of the listed dates, date by date
    int* gstart = nullptr;            // [n_groups + 1] offsets into order
    double* gvals = nullptr;          // per-date correlation
    double* out = nullptr;            // metric
    FNode* nodes = nullptr;           // the tree being added (at most 255 nodes)
    unsigned char* flags = nullptr;
    double* values = nullptr;
    unsigned long long* cmask = nullptr;   // its categorical nodes' bin masks
    hipStream_t stream = nullptr;
    ~tgbm_evalset() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : {(void*)bins, (void*)y, (void*)pred, (void*)order, (void*)gstart, (void*)gvals, (void*)out,
                        (void*)nodes, (void*)flags, (void*)values, (void*)cmask})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

__global__ __launch_bounds__(256) void k_eval_fill(double* __restrict__ pred, int n, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pred[i] = v;
}

// pred += lr * value of the row's leaf: the running sum of predict_multi_binned, one tree at a time.
__global__ __launch_bounds__(256) void k_eval_add(const FNode* __restrict__ nodes, const unsigned char* __restrict__ flags,
                                                  const unsigned long long* __restrict__ cmask,
                                                  const double* __restrict__ values, double lr,
                                                  const unsigned char* __restrict__ bins, int fp, int n, int nb,
                                                  double* __restrict__ pred) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pred[i] = pred[i] + lr * values[walk_tree(nodes, flags, cmask, bins + (size_t)i * fp, 0, nb)];
}

// Mean squared error with unit weights (Objective::loss, objective.cpp:323-328).  The reference
// adds the rows' terms in row order; one wave keeps that order: each step forms 64 terms at once
// and adds them lane by lane.
__global__ __launch_bounds__(64) void k_eval_mse(const double* __restrict__ y, const double* __restrict__ pred, int n,
                                                 double* __restrict__ out) {
    const int lane = threadIdx.x;
    double s = 0.0;
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane;
        double t = 0.0;
        if (i < n) {
            const double d = pred[i] - y[i];
            t = 1.0 * d * d;
        }
        const int m = min(64, n - base);
        for (int k = 0; k < m; ++k) s += read_lane_f64(t, k);
    }
    if (lane == 0) out[0] = n > 0 ? s / (double)n : 0.0;
}

// One thread per date: pearson_ix (booster.cpp:132-153) over the date's rows in the listed order.
__global__ __launch_bounds__(64) void k_eval_ic(const double* __restrict__ y, const double* __restrict__ p,
                                                const int* __restrict__ order, const int* __restrict__ gstart, int ng,
                                                double* __restrict__ vals) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int* ix = order + gstart[g];
    const int m = gstart[g + 1] - gstart[g];
    double v = 0.0;
    if (m >= 3) {
        double my = 0.0, mp = 0.0;
        for (int k = 0; k < m; ++k) {
            my += y[ix[k]];
            mp += p[ix[k]];
        }
        my /= m;
        mp /= m;
        double vy = 0.0, vp = 0.0, c = 0.0;
        for (int k = 0; k < m; ++k) {
            const double dy = y[ix[k]] - my;
            const double dp = p[ix[k]] - mp;
            vy += dy * dy;
            vp += dp * dp;
            c += dy * dp;
        }
        v = (vy < 1e-18 || vp < 1e-18) ? 0.0 : c / sqrt(vy * vp);
    }
    vals[g] = v;
}

// grouped_ic's mean over dates (booster.cpp:163-174), summed in date order.
__global__ void k_eval_ic_mean(const double* __restrict__ vals, int ng, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int g = 0; g < ng; ++g) s += vals[g];
    out[0] = ng > 0 ? s / ng : 0.0;
}

extern "C" {

int tgbm_eval_create(tgbm_evalset** out, int n_features, int n_bins, const double* edges, const int* n_edges,
                     const double* X, const double* y, int n, const int* order, const int* group_start, int n_groups) {
    *out = nullptr;
    return fguard([&] {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
            throw std::runtime_error("GPU device not available: no HIP device found");
        if (n_features <= 0 || n <= 0 || n_groups < 0) throw std::runtime_error("bad eval set shape");
        if (n_groups > 0) {
            if (!order || !group_start) throw std::runtime_error("eval set dates need order and group_start");
            if (group_start[0] < 0) throw std::runtime_error("eval date offsets out of range");
            for (int g = 0; g < n_groups; ++g)
                if (group_start[g + 1] < group_start[g]) throw std::runtime_error("eval date offsets must not decrease");
            if (group_start[n_groups] > n) throw std::runtime_error("eval date offsets out of range");
            for (int k = 0; k < group_start[n_groups]; ++k)
                if (order[k] < 0 || order[k] >= n) throw std::runtime_error("eval row id out of range");
        }
        auto* e = new tgbm_evalset();
        std::unique_ptr<tgbm_evalset> guard(e);
        e->F = n_features;
        e->fp = (n_features + 15) / 16 * 16;
        e->nb = n_bins;
        e->n = n;
        e->n_groups = n_groups;
        FHIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
        FHIP(hipMalloc(&e->bins, (size_t)n * e->fp + 16));
        FHIP(hipMalloc(&e->y, sizeof(double) * n));
        FHIP(hipMalloc(&e->pred, sizeof(double) * n));
        FHIP(hipMalloc(&e->gvals, sizeof(double) * std::max(n_groups, 1)));
        FHIP(hipMalloc(&e->out, sizeof(double)));
        FHIP(hipMalloc(&e->nodes, sizeof(FNode) * 256));
        FHIP(hipMalloc(&e->flags, 256));
        FHIP(hipMalloc(&e->values, sizeof(double) * 256));
        FHIP(hipMalloc(&e->cmask, sizeof(unsigned long long) * 4 * 256));
        if (n_groups > 0) {
            const int m = group_start[n_groups];
            FHIP(hipMalloc(&e->order, sizeof(int) * std::max(m, 1)));
            FHIP(hipMalloc(&e->gstart, sizeof(int) * (n_groups + 1)));
            if (m) FHIP(hipMemcpy(e->order, order, sizeof(int) * m, hipMemcpyHostToDevice));
            FHIP(hipMemcpy(e->gstart, group_start, sizeof(int) * (n_groups + 1), hipMemcpyHostToDevice));
        }
        FHIP(hipMemcpy(e->y, y, sizeof(double) * n, hipMemcpyHostToDevice));
        // bin the rows once with the fit's edges (same kernel as the training matrix)
        double *Xd = nullptr, *Ed = nullptr;
        int* NEd = nullptr;
        try {
            FHIP(hipMalloc(&Xd, sizeof(double) * (size_t)n * n_features));
            FHIP(hipMalloc(&Ed, sizeof(double) * (size_t)n_features * NBIN));
            FHIP(hipMalloc(&NEd, sizeof(int) * n_features));
            FHIP(hipMemcpy(Xd, X, sizeof(double) * (size_t)n * n_features, hipMemcpyHostToDevice));
            FHIP(hipMemcpy(Ed, edges, sizeof(double) * (size_t)n_features * NBIN, hipMemcpyHostToDevice));
            FHIP(hipMemcpy(NEd, n_edges, sizeof(int) * n_features, hipMemcpyHostToDevice));
            FHIP(hipMemsetAsync(e->bins, 0, (size_t)n * e->fp, e->stream));
            launch_bin(e->stream, Xd, n, n_features, Ed, NEd, e->bins, e->fp, nullptr, n + 64);
            FHIP(hipGetLastError());
            FHIP(hipStreamSynchronize(e->stream));
        } catch (...) {
            for (void* p : {(void*)Xd, (void*)Ed, (void*)NEd})
                if (p) (void)hipFree(p);
            throw;
        }
        for (void* p : {(void*)Xd, (void*)Ed, (void*)NEd}) (void)hipFree(p);
        *out = guard.release();
    });
}

void tgbm_eval_destroy(tgbm_evalset* e) { delete e; }

int tgbm_eval_reset(tgbm_evalset* e, double init_score) {
    return fguard([&] {
        hipLaunchKernelGGL(k_eval_fill, dim3((e->n + 255) / 256), dim3(256), 0, e->stream, e->pred, e->n, init_score);
        FHIP(hipGetLastError());
        FHIP(hipStreamSynchronize(e->stream));
    });
}

int tgbm_eval_add_tree(tgbm_evalset* e, int n_nodes, const int* node_ints, const double* values, double lr,
                       const unsigned long long* cat_masks) {
    return fguard([&] {
        const HostNodes h = build_nodes(1, &n_nodes, node_ints, e->F, 0, cat_masks, 0);
        if (!h.cmask.empty())
            FHIP(hipMemcpyAsync(e->cmask, h.cmask.data(), sizeof(unsigned long long) * h.cmask.size(),
                                hipMemcpyHostToDevice, e->stream));
        FHIP(hipMemcpyAsync(e->nodes, h.nodes.data(), sizeof(FNode) * n_nodes, hipMemcpyHostToDevice, e->stream));
        FHIP(hipMemcpyAsync(e->flags, h.flags.data(), n_nodes, hipMemcpyHostToDevice, e->stream));
        FHIP(hipMemcpyAsync(e->values, values, sizeof(double) * n_nodes, hipMemcpyHostToDevice, e->stream));
        hipLaunchKernelGGL(k_eval_add, dim3((e->n + 255) / 256), dim3(256), 0, e->stream, e->nodes, e->flags, e->cmask,
                           e->values, lr, e->bins, e->fp, e->n, e->nb, e->pred);
        FHIP(hipGetLastError());
        FHIP(hipStreamSynchronize(e->stream));   // the host node tables above end with this call
    });
}

int tgbm_eval_metric(tgbm_evalset* e, int metric, double* out) {
    return fguard([&] {
        if (metric == 0) {
            hipLaunchKernelGGL(k_eval_mse, dim3(1), dim3(64), 0, e->stream, e->y, e->pred, e->n, e->out);
        } else if (metric == 1) {
            if (e->n_groups > 0)
                hipLaunchKernelGGL(k_eval_ic, dim3((e->n_groups + 63) / 64), dim3(64), 0, e->stream, e->y, e->pred,
                                   e->order, e->gstart, e->n_groups, e->gvals);
            hipLaunchKernelGGL(k_eval_ic_mean, dim3(1), dim3(64), 0, e->stream, e->gvals, e->n_groups, e->out);
        } else {
            throw std::runtime_error("eval metric must be 0 (mean squared error) or 1 (per-date correlation)");
        }
        FHIP(hipGetLastError());
        FHIP(hipMemcpyAsync(out, e->out, sizeof(double), hipMemcpyDeviceToHost, e->stream));
        FHIP(hipStreamSynchronize(e->stream));
    });
}

int tgbm_eval_predictions(tgbm_evalset* e, double* out, int cap) {
    return fguard([&] {
        const int m = std::min(std::max(cap, 0), e->n);
        if (m) FHIP(hipMemcpyAsync(out, e->pred, sizeof(double) * m, hipMemcpyDeviceToHost, e->stream));
        FHIP(hipStreamSynchronize(e->stream));
    });
}

}  // extern "C"
