// Batched ensemble predict for the serving loop: one launch scores every request of a batch,
// each request naming its own ensemble and its rows (rows of the session's resident binned
// panel, or raw feature rows binned here with the ensemble's edges).
//
// Per request the result is the reference's predict_raw_binned (booster.cpp:744-773):
// init_score + sum over trees, in tree order, of lr_t * values_t[leaf(row)], with the same
// rounding (sequential FP64 adds, no FMA contraction), so a row's score does not depend on which
// other requests shared its launch.
//
// Kernel layout: a block scores SERVE_ROWS rows of one request.  It stages the rows' bins in
// LDS, then walks the ensemble SERVE_TREES_PER_CHUNK trees at a time: the chunk's packed nodes
// are copied to LDS, SERVE_LANES threads per row traverse different trees of the chunk
// (LDS-only dependent loads) and record leaf ids, and one thread per row then adds the chunk's
// leaf values in tree order.  A row's bins sit at a stride of an odd number of 4-byte words, so
// the 64 rows of a wave read 64 different LDS banks at every node.  Ensembles of numeric splits
// only walk SERVE_ILP trees per thread at once, each for the ensemble's depth (a leaf keeps its
// id), so the dependent LDS loads of those trees overlap.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/tgbm_device.h"
#include "dev_common.h"
#include "forest.h"
#include "launch.h"

using namespace tgbm;

extern thread_local std::string g_dev_err;  // defined in session.hip

#define SHIP(expr)                                                                           \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) +   \
                                     " at " #expr);                                          \
    } while (0)

// Named (not anonymous) so that profilers can show the kernels' names.
namespace tgbm_serve {

constexpr int SERVE_ROWS = 64;
constexpr int SERVE_LANES = 4;   // threads per row (256-thread blocks)
constexpr int SERVE_ILP = 4;     // trees one thread walks together (numeric ensembles)
constexpr int SERVE_TC = SERVE_TREES_PER_CHUNK;
constexpr size_t SERVE_MAX_LDS = 152 * 1024;   // dynamic LDS opt-in (below 160 KiB: the kernel has static LDS too)

struct SReq {
    const SNode* nodes;
    const unsigned long long* cmask;   // bin masks of categorical nodes (4 words each), or null
    const int* tree_off;
    const double* values;
    const double* lr;
    const unsigned char* bins;   // row-major, fp bytes per row
    const int* rows;             // row ids into bins (null: rows 0..n-1)
    double* out;
    double init;
    int n_trees, nb, fp, n, blk0;
    int depth;    // > 0: every split numeric, no node deeper than this; 0: general walk
};

__host__ __device__ constexpr size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// LDS bytes per staged row: fp (a multiple of 16) plus one word, an odd number of words.
__host__ __device__ constexpr int serve_row_bytes(int fp) { return fp + 4; }

// LDS bytes of k_serve for rows of fp bytes and chunks of at most chunk_nodes nodes.
size_t serve_lds(int fp, int chunk_nodes) {
    size_t b = (size_t)SERVE_ROWS * serve_row_bytes(fp);   // bins
    b += (size_t)SERVE_TC * SERVE_ROWS;                  // leaf ids
    b = align_up(b + sizeof(int) * (SERVE_TC + 1), 8);   // chunk tree offsets
    return b + sizeof(SNode) * (size_t)chunk_nodes;
}

// Block `blk` of request R: rows blk * SERVE_ROWS .. of the request.
__device__ __forceinline__ void serve_block(const SReq& R, int blk, int max_fp) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int r0 = blk * SERVE_ROWS;
    const int nr = min(SERVE_ROWS, R.n - r0);
    const int max_rb = serve_row_bytes(max_fp);
    unsigned char* lbins = smem;
    unsigned char* lleaf = smem + (size_t)SERVE_ROWS * max_rb;
    int* loff = (int*)(lleaf + SERVE_TC * SERVE_ROWS);
    SNode* lnodes = (SNode*)(smem + align_up((size_t)SERVE_ROWS * max_rb + SERVE_TC * SERVE_ROWS +
                                                 sizeof(int) * (SERVE_TC + 1), 8));
    const int q = R.fp / 16;                     // 16-byte words per row in HBM
    const int rw = serve_row_bytes(R.fp) / 4;    // 4-byte words per row in LDS (odd)
    for (int j = threadIdx.x; j < nr * q; j += blockDim.x) {
        const int r = j / q, c = j % q;
        const int row = R.rows ? R.rows[r0 + r] : r0 + r;
        const uint4 v = ((const uint4*)(R.bins + (size_t)row * R.fp))[c];
        unsigned int* d = (unsigned int*)lbins + r * rw + 4 * c;
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    }
    const int r = threadIdx.x % SERVE_ROWS, g = threadIdx.x / SERVE_ROWS;
    const unsigned char* row = lbins + r * (4 * rw);
    double s = R.init;
    for (int t0 = 0; t0 < R.n_trees; t0 += SERVE_TC) {
        const int nt = min(SERVE_TC, R.n_trees - t0);
        const int base = R.tree_off[t0];
        const int nn = R.tree_off[t0 + nt] - base;
        __syncthreads();   // previous chunk fully consumed
        for (int j = threadIdx.x; j <= nt; j += blockDim.x) loff[j] = R.tree_off[t0 + j] - base;
        for (int j = threadIdx.x; j < nn; j += blockDim.x)
            ((uint2*)lnodes)[j] = ((const uint2*)(R.nodes + base))[j];
        __syncthreads();
        if (r < nr && R.depth > 0) {
            // numeric splits only: SERVE_ILP trees at once, R.depth levels each
            for (int t = g; t < nt; t += SERVE_LANES * SERVE_ILP) {
                const uint2* tn[SERVE_ILP];
                int k[SERVE_ILP];
#pragma unroll
                for (int u = 0; u < SERVE_ILP; ++u) {
                    const int tt = t + u * SERVE_LANES;
                    tn[u] = (const uint2*)(lnodes + loff[tt < nt ? tt : t]);
                    k[u] = 0;
                }
                for (int d = 0; d < R.depth; ++d) {
                    uint2 nd[SERVE_ILP];
#pragma unroll
                    for (int u = 0; u < SERVE_ILP; ++u) nd[u] = tn[u][k[u]];
                    bool more = false;
#pragma unroll
                    for (int u = 0; u < SERVE_ILP; ++u) {
                        // x = f | f2 << 16; y = thr | left << 8 | right << 16 | flags << 24
                        const unsigned int fl = nd[u].y >> 24;
                        const int b = row[nd[u].x & 0xffff];
                        const bool left = b == 0 ? ((fl >> 3) & 1) : (b <= (int)(nd[u].y & 0xff));
                        const int nk = (int)((left ? nd[u].y >> 8 : nd[u].y >> 16) & 0xff);
                        k[u] = (fl & 1) ? k[u] : nk;
                        more |= !(fl & 1);
                    }
                    if (!__any(more)) break;   // every walk of the wave is at a leaf
                }
#pragma unroll
                for (int u = 0; u < SERVE_ILP; ++u) {
                    const int tt = t + u * SERVE_LANES;
                    if (tt < nt) lleaf[tt * SERVE_ROWS + r] = (unsigned char)k[u];
                }
            }
        } else if (r < nr) {
            for (int t = g; t < nt; t += SERVE_LANES) {
                const SNode* tn = lnodes + loff[t];
                int k = 0;
                SNode nd = tn[0];
                while (!(nd.flags & 1)) {
                    const int b1 = row[nd.f];
                    bool left;
                    if (((nd.flags >> 1) & 3) >= 2) {   // categorical: the bin's bit in the node's mask
                        left = (R.cmask[(size_t)nd.f2 * 4 + (b1 >> 6)] >> (b1 & 63)) & 1ull;
                    } else if (((nd.flags >> 1) & 3) == 1) {
                        const int b2 = row[nd.f2];
                        if (b1 == 0 || b2 == 0) left = (nd.flags >> 3) & 1;
                        else left = (int)(((unsigned)b1 * 2654435761u + (unsigned)b2 * 40503u) % (unsigned)R.nb) <= (int)nd.thr;
                    } else {
                        left = b1 == 0 ? ((nd.flags >> 3) & 1) : (b1 <= (int)nd.thr);
                    }
                    k = left ? nd.left : nd.right;
                    nd = tn[k];
                }
                lleaf[t * SERVE_ROWS + r] = (unsigned char)k;
            }
        }
        __syncthreads();
        if (g == 0 && r < nr) {
            for (int t = 0; t < nt; ++t) s += R.lr[t0 + t] * R.values[base + loff[t] + lleaf[t * SERVE_ROWS + r]];
        }
    }
    if (g == 0 && r < nr) R.out[r0 + r] = s;
}

// A batch of requests: block b serves the last request whose first block is <= b.
__global__ __launch_bounds__(256, 8) void k_serve(const SReq* __restrict__ reqs, int n_req, int max_fp) {
    __shared__ int s_req;
    if (threadIdx.x == 0) {
        int lo = 0, hi = n_req - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (reqs[mid].blk0 <= (int)blockIdx.x) lo = mid;
            else hi = mid - 1;
        }
        s_req = lo;
    }
    __syncthreads();
    const SReq R = reqs[s_req];
    serve_block(R, (int)blockIdx.x - R.blk0, max_fp);
}

// One request passed by value (no descriptor upload, so the caller need not wait for the launch).
__global__ __launch_bounds__(256, 8) void k_serve_one(SReq R, int max_fp) { serve_block(R, (int)blockIdx.x, max_fp); }

// A batch of at most SERVE_ARG_REQS requests whose table travels in the kernel arguments and whose
// row ids and scores live in mapped host memory: the batch is one launch, with no copy before or
// after it (each copy would wait for GPU resources of its own behind concurrent work).
constexpr int SERVE_ARG_REQS = 24;
struct SBatch {
    SReq r[SERVE_ARG_REQS];
    int n_req;
};
__global__ __launch_bounds__(256, 8) void k_serve_batch(const SBatch B, int max_fp) {
    // the block's request (the last whose first block is <= blockIdx.x), by constant indices
    int lo = 0;
#pragma unroll
    for (int k = 1; k < SERVE_ARG_REQS; ++k)
        if (k < B.n_req && B.r[k].blk0 <= (int)blockIdx.x) lo = k;
    SReq R = B.r[0];
#pragma unroll
    for (int k = 1; k < SERVE_ARG_REQS; ++k)
        if (k == lo) R = B.r[k];
    serve_block(R, (int)blockIdx.x - R.blk0, max_fp);
}

template <class Fn>
int sguard(Fn&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception& e) {
        g_dev_err = e.what();
        return -1;
    }
}

size_t item_bytes(const tgbm_forest* f, int from_x, int n) {
    size_t b = align_up(sizeof(double) * (size_t)n, 256);                 // scores
    if (from_x) b += align_up(sizeof(double) * (size_t)n * f->F, 256) + align_up((size_t)n * f->fp, 256);
    else b += align_up(sizeof(int) * (size_t)n, 256);                     // row ids
    return b;
}

}  // namespace tgbm_serve

using namespace tgbm_serve;

struct tgbm_predictor {
    hipStream_t stream = nullptr;
    char* dev = nullptr;
    char* host = nullptr;       // pinned, mapped
    char* host_dev = nullptr;   // its device address
    size_t cap = 0;
    int lds_ok = 0;
    ~tgbm_predictor() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (dev) (void)hipFree(dev);
        if (host) (void)hipHostFree(host);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

size_t tgbm_predict_bytes(const tgbm_forest* f, int from_x, int n) {
    return item_bytes(f, from_x, n) + align_up(sizeof(SReq), 256);
}

int tgbm_predictor_create(tgbm_predictor** out, size_t staging_bytes) {
    *out = nullptr;
    return sguard([&] {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
            throw std::runtime_error("GPU device not available: no HIP device found");
        auto* p = new tgbm_predictor();
        std::unique_ptr<tgbm_predictor> guard(p);
        p->cap = std::max<size_t>(staging_bytes, 1 << 20);
        SHIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
        SHIP(hipMalloc(&p->dev, p->cap));
        SHIP(hipHostMalloc((void**)&p->host, p->cap, hipHostMallocMapped));
        SHIP(hipHostGetDevicePointer((void**)&p->host_dev, p->host, 0));
        SHIP(hipFuncSetAttribute((const void*)k_serve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)SERVE_MAX_LDS));
        SHIP(hipFuncSetAttribute((const void*)k_serve_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)SERVE_MAX_LDS));
        *out = guard.release();
    });
}

void tgbm_predictor_destroy(tgbm_predictor* p) { delete p; }

int tgbm_predict_batch(tgbm_predictor* p, const tgbm_session* panel, const tgbm_predict_item* items, int n_items) {
    return sguard([&] {
        if (n_items <= 0) return;
        const unsigned char* pbins = nullptr;
        int pfp = 0, pn = 0, pF = 0;
        if (panel) session_panel_info(panel, &pbins, &pfp, &pn, &pF);
        // ---- layout: request table, every item's scores, then every item's inputs (row ids, or
        // raw rows and their bins), so the scores come back in one contiguous copy
        const size_t table = align_up(sizeof(SReq) * (size_t)n_items, 256);
        size_t total = table;
        std::vector<size_t> off(n_items), in_off(n_items);
        for (int i = 0; i < n_items; ++i) {
            off[i] = total;
            total += align_up(sizeof(double) * (size_t)std::max(items[i].n, 0), 256);
        }
        const size_t scores_end = total;
        int max_fp = 16, max_chunk = 1;
        for (int i = 0; i < n_items; ++i) {
            const tgbm_predict_item& it = items[i];
            const tgbm_forest* f = it.forest;
            if (!f) throw std::runtime_error("predict item without an ensemble");
            if (it.n < 0) throw std::runtime_error("predict item with n < 0");
            const int from_x = it.rows ? 0 : 1;
            if (!from_x) {
                if (!panel) throw std::runtime_error("row ids need a resident panel");
                if (f->F != pF) throw std::runtime_error("ensemble and panel feature counts differ");
                for (int k = 0; k < it.n; ++k)
                    if (it.rows[k] < 0 || it.rows[k] >= pn) throw std::runtime_error("panel row id out of range");
            } else if (!it.X && it.n > 0) {
                throw std::runtime_error("predict item needs row ids or feature rows");
            }
            in_off[i] = total;
            total += item_bytes(f, from_x, it.n) - align_up(sizeof(double) * (size_t)it.n, 256);
            max_fp = std::max(max_fp, f->fp);
            max_chunk = std::max(max_chunk, f->max_chunk_nodes);
        }
        if (total > p->cap) throw std::runtime_error("predict batch exceeds the predictor's staging memory");
        const size_t lds = serve_lds(max_fp, max_chunk);
        if (lds > SERVE_MAX_LDS) throw std::runtime_error("ensemble chunk too large for LDS");
        bool all_rows = n_items <= SERVE_ARG_REQS;
        for (int i = 0; i < n_items && all_rows; ++i) all_rows = items[i].rows != nullptr || items[i].n == 0;
        if (all_rows) {
            // row ids in, scores out, both in the mapped staging; the table in the arguments
            SBatch B;
            std::memset(&B, 0, sizeof B);
            B.n_req = n_items;
            int blocks = 0;
            for (int i = 0; i < n_items; ++i) {
                const tgbm_predict_item& it = items[i];
                const tgbm_forest* f = it.forest;
                SReq& R = B.r[i];
                R.nodes = f->snodes;
                R.cmask = f->cmask;
                R.tree_off = f->tree_off;
                R.values = f->values;
                R.lr = f->lr;
                R.init = f->init;
                R.n_trees = f->n_trees;
                R.nb = f->nb;
                R.fp = f->fp;
                R.n = it.n;
                R.blk0 = blocks;
                R.depth = f->numeric_only ? std::max(f->max_depth, 1) : 0;
                blocks += (it.n + SERVE_ROWS - 1) / SERVE_ROWS;
                R.out = (double*)(p->host_dev + off[i]);
                R.bins = pbins;
                R.rows = (const int*)(p->host_dev + in_off[i]);
                if (it.n) std::memcpy(p->host + in_off[i], it.rows, sizeof(int) * (size_t)it.n);
            }
            if (blocks > 0)
                hipLaunchKernelGGL(k_serve_batch, dim3(blocks), dim3(256), lds, p->stream, B, max_fp);
            SHIP(hipGetLastError());
            SHIP(hipStreamSynchronize(p->stream));
            for (int i = 0; i < n_items; ++i)
                if (items[i].n > 0) std::memcpy(items[i].out, p->host + off[i], sizeof(double) * (size_t)items[i].n);
            return;
        }
        // ---- pack inputs on the host (pinned), one upload
        std::vector<SReq> reqs(n_items);
        int blocks = 0;
        size_t upload_end = table;
        for (int i = 0; i < n_items; ++i) {
            const tgbm_predict_item& it = items[i];
            const tgbm_forest* f = it.forest;
            SReq& R = reqs[i];
            R.nodes = f->snodes;
            R.cmask = f->cmask;
            R.tree_off = f->tree_off;
            R.values = f->values;
            R.lr = f->lr;
            R.init = f->init;
            R.n_trees = f->n_trees;
            R.nb = f->nb;
            R.fp = f->fp;
            R.n = it.n;
            R.blk0 = blocks;
            R.depth = f->numeric_only ? std::max(f->max_depth, 1) : 0;
            blocks += (it.n + SERVE_ROWS - 1) / SERVE_ROWS;
            R.out = (double*)(p->dev + off[i]);
            size_t o = in_off[i];
            if (it.rows) {
                std::memcpy(p->host + o, it.rows, sizeof(int) * (size_t)it.n);
                R.rows = (const int*)(p->dev + o);
                R.bins = pbins;
                o += align_up(sizeof(int) * (size_t)it.n, 256);
            } else {
                std::memcpy(p->host + o, it.X, sizeof(double) * (size_t)it.n * f->F);
                o += align_up(sizeof(double) * (size_t)it.n * f->F, 256);
                R.rows = nullptr;
                R.bins = (const unsigned char*)(p->dev + o);
                o += align_up((size_t)it.n * f->fp, 256);
            }
            upload_end = std::max(upload_end, o);
        }
        std::memcpy(p->host, reqs.data(), sizeof(SReq) * n_items);
        SHIP(hipMemcpyAsync(p->dev, p->host, upload_end, hipMemcpyHostToDevice, p->stream));
        for (int i = 0; i < n_items; ++i) {
            const tgbm_predict_item& it = items[i];
            if (it.rows || it.n == 0) continue;
            const tgbm_forest* f = it.forest;
            const double* Xd = (const double*)(p->dev + in_off[i]);
            launch_bin(p->stream, Xd, it.n, f->F, f->edges, f->n_edges, (unsigned char*)reqs[i].bins, f->fp, nullptr,
                       it.n + 64);
        }
        if (blocks > 0)
            hipLaunchKernelGGL(k_serve, dim3(blocks), dim3(256), lds, p->stream, (const SReq*)p->dev, n_items, max_fp);
        SHIP(hipGetLastError());
        // ---- scores back: one copy over the span of all score regions
        SHIP(hipMemcpyAsync(p->host + table, p->dev + table, scores_end - table, hipMemcpyDeviceToHost, p->stream));
        SHIP(hipStreamSynchronize(p->stream));
        for (int i = 0; i < n_items; ++i)
            if (items[i].n > 0) std::memcpy(items[i].out, p->host + off[i], sizeof(double) * (size_t)items[i].n);
    });
}

int tgbm_task_refit_window(tgbm_session* s, int task, int row0, int n_rows, const double* y, const tgbm_forest* f) {
    return sguard([&] {
        if (!f) throw std::runtime_error("no ensemble");
        if (!y) throw std::runtime_error("a refit window needs its targets");
        const unsigned char* pbins = nullptr;
        int pfp = 0, pn = 0, pF = 0;
        session_panel_info(s, &pbins, &pfp, &pn, &pF);
        if (f->F != pF) throw std::runtime_error("ensemble and panel feature counts differ");
        const size_t lds = serve_lds(f->fp, f->max_chunk_nodes);
        if (lds > SERVE_MAX_LDS) throw std::runtime_error("ensemble chunk too large for LDS");
        static std::once_flag attr;
        std::call_once(attr, [] {
            SHIP(hipFuncSetAttribute((const void*)k_serve_one, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)SERVE_MAX_LDS));
        });
        double* pred = nullptr;
        hipStream_t st = nullptr;
        session_refit_window(s, task, row0, n_rows, y, &pred, &st);
        SReq R;
        R.nodes = f->snodes;
        R.cmask = f->cmask;
        R.tree_off = f->tree_off;
        R.values = f->values;
        R.lr = f->lr;
        R.bins = pbins + (size_t)row0 * pfp;   // the window's rows are contiguous
        R.rows = nullptr;
        R.out = pred;
        R.init = f->init;
        R.n_trees = f->n_trees;
        R.nb = f->nb;
        R.fp = f->fp;
        R.n = n_rows;
        R.blk0 = 0;
        R.depth = f->numeric_only ? std::max(f->max_depth, 1) : 0;
        hipLaunchKernelGGL(k_serve_one, dim3((n_rows + SERVE_ROWS - 1) / SERVE_ROWS), dim3(256), lds, st, R, f->fp);
        SHIP(hipGetLastError());
    });
}

int tgbm_task_predict_forest(tgbm_session* s, int task, const tgbm_forest* f) {
    return sguard([&] {
        if (!f) throw std::runtime_error("no ensemble");
        const unsigned char* pbins = nullptr;
        int pfp = 0, pn = 0, pF = 0;
        session_panel_info(s, &pbins, &pfp, &pn, &pF);
        if (f->F != pF) throw std::runtime_error("ensemble and panel feature counts differ");
        double* pred = nullptr;
        int row0 = 0, n = 0;
        hipStream_t st = nullptr;
        void* scratch = nullptr;
        session_task_window(s, task, &pred, &row0, &n, &st, &scratch, sizeof(SReq));
        const size_t lds = serve_lds(f->fp, f->max_chunk_nodes);
        if (lds > SERVE_MAX_LDS) throw std::runtime_error("ensemble chunk too large for LDS");
        static bool attr = false;
        if (!attr) {
            SHIP(hipFuncSetAttribute((const void*)k_serve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)SERVE_MAX_LDS));
            attr = true;
        }
        SReq R;
        R.nodes = f->snodes;
        R.cmask = f->cmask;
        R.tree_off = f->tree_off;
        R.values = f->values;
        R.lr = f->lr;
        R.bins = pbins + (size_t)row0 * pfp;   // the task's rows are a contiguous window
        R.rows = nullptr;
        R.out = pred;
        R.init = f->init;
        R.n_trees = f->n_trees;
        R.nb = f->nb;
        R.fp = f->fp;
        R.n = n;
        R.blk0 = 0;
        R.depth = f->numeric_only ? std::max(f->max_depth, 1) : 0;
        SHIP(hipMemcpyAsync(scratch, &R, sizeof R, hipMemcpyHostToDevice, st));
        const int blocks = (n + SERVE_ROWS - 1) / SERVE_ROWS;
        if (blocks > 0)
            hipLaunchKernelGGL(k_serve, dim3(blocks), dim3(256), lds, st, (const SReq*)scratch, 1, f->fp);
        SHIP(hipGetLastError());
        SHIP(hipStreamSynchronize(st));   // the descriptor lives on this call's stack
    });
}

}  // extern "C"
