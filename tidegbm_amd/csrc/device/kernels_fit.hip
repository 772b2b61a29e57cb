// gfx950 kernels of the histogram-GBDT fit path.  Every kernel takes the task table and works
// on task blockIdx.z (or .y); kernels of the best-first loop read the current split from the
// task state, so the whole loop can run without host round trips.
//
// Built with -ffp-contract=off: every floating-point expression that mirrors a reference
// formula must round like the reference's x86-64 build (no fused multiply-add).
#include "dev_common.h"
#include "launch.h"

namespace tgbm {

// =================================================================== binning
// Reference Binner::transform (binning.cpp:104-128): NaN -> 0, numeric bin = 1 + number of
// right-edges <= x (std::upper_bound).  Fuses the column-major transpose (histogram.cpp:78-87)
// and narrows to uint8.  Block = FB features x rows; the features' edges sit in LDS.
template <int FB>
__global__ __launch_bounds__(256) void k_bin(const double* __restrict__ X, int n, int F,
                                             const double* __restrict__ edges,
                                             const int* __restrict__ n_edges,
                                             unsigned char* __restrict__ bins_row, int fp,
                                             unsigned char* __restrict__ bins_col, int n_pad,
                                             int rows_per_block) {
    __shared__ double e[FB][NBIN];
    __shared__ int ne[FB];
    const int f0 = blockIdx.y * FB;
    for (int j = threadIdx.x; j < FB * NBIN; j += blockDim.x) {
        const int fl = j / NBIN, k = j % NBIN, f = f0 + fl;
        e[fl][k] = (f < F && k < n_edges[f]) ? edges[(size_t)f * NBIN + k] : 0.0;
    }
    if (threadIdx.x < FB) ne[threadIdx.x] = (f0 + (int)threadIdx.x < F) ? n_edges[f0 + threadIdx.x] : 0;
    __syncthreads();
    const int fl = threadIdx.x % FB;
    const int f = f0 + fl;
    if (f >= F) return;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(n, r0 + rows_per_block);
    const int cnt = ne[fl];
    for (int i = r0 + (int)threadIdx.x / FB; i < r1; i += blockDim.x / FB) {
        const double x = X[(size_t)i * F + f];
        int b = 0;
        if (x == x) {
            int lo = 0, hi = cnt;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (e[fl][mid] <= x) lo = mid + 1;
                else hi = mid;
            }
            b = 1 + lo;
        }
        bins_row[(size_t)i * fp + f] = (unsigned char)b;
        bins_col[(size_t)f * n_pad + i] = (unsigned char)b;
    }
}

void launch_bin(hipStream_t st, const double* X, int n, int F, const double* edges,
                const int* n_edges, unsigned char* bins_row, int fp, unsigned char* bins_col,
                int n_pad) {
    constexpr int FB = 16;
    const int rpb = 2048;
    dim3 grid((n + rpb - 1) / rpb, (F + FB - 1) / FB);
    hipLaunchKernelGGL(k_bin<FB>, grid, dim3(256), 0, st, X, n, F, edges, n_edges, bins_row, fp,
                       bins_col, n_pad, rpb);
}

// =================================================================== gradients
__device__ inline void block_max_u64(unsigned long long v, unsigned long long* dst) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(v, o, 64);
        v = t > v ? t : v;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicMax(dst, v);
}

// MSE (objective.cpp:105-119, cost terms off): g = w (pred - y), h = w.
__global__ __launch_bounds__(256) void k_grad_mse(Task* tasks) {
    Task& T = tasks[blockIdx.y];
    unsigned long long mg = 0, mh = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < T.n_rows; i += gridDim.x * blockDim.x) {
        const double wi = T.w ? T.w[i] : 1.0;
        const double g = wi * (T.pred[i] - T.y[i]);
        T.g[i] = g;
        T.h[i] = wi;
        const unsigned long long bg = (unsigned long long)__double_as_longlong(fabs(g));
        const unsigned long long bh = (unsigned long long)__double_as_longlong(fabs(wi));
        mg = bg > mg ? bg : mg;
        mh = bh > mh ? bh : mh;
    }
    block_max_u64(mg, &T.max_g_bits);
    block_max_u64(mh, &T.max_h_bits);
}

// Within-group pairwise gradient (objective.cpp:179-219).  One block per group; each row
// replays, in the reference's order, the pairs that touch it: first as the later element b of
// pairs (a, b) for ascending a, then as the earlier element a for ascending b.  Row values of
// the group are staged in LDS.
constexpr int PAIR_MAX_GROUP = 4096;
__global__ __launch_bounds__(256) void k_grad_pairwise(Task* tasks) {
    Task& T = tasks[blockIdx.y];
    __shared__ double sy[PAIR_MAX_GROUP], sp[PAIR_MAX_GROUP];
    unsigned long long mg = 0, mh = 0;
    for (int grp = blockIdx.x; grp < T.n_groups; grp += gridDim.x) {
        const int g0 = T.group_start[grp], m = T.group_start[grp + 1] - g0;
        const int* gr = T.group_rows ? T.group_rows + g0 : nullptr;  // null: rows g0.. in order
        __syncthreads();
        for (int k = threadIdx.x; k < m && k < PAIR_MAX_GROUP; k += blockDim.x) {
            const int r = gr ? gr[k] : g0 + k;
            sy[k] = T.y[r];
            sp[k] = T.pred[r];
        }
        __syncthreads();
        (void)sy;
        const int step = m > 48 ? max(1, m / 24) : 1;
        if (m > PAIR_MAX_GROUP) {  // oversized group: read the row values from global memory
            for (int p = threadIdx.x; p < m; p += blockDim.x) {
                const int rp = gr ? gr[p] : g0 + p;
                const double wp = T.w ? T.w[rp] : 1.0;
                double g = 0.0, h = 1e-6;
                const int step = max(1, m / 24);
                auto val_y = [&](int q) { return T.y[gr ? gr[q] : g0 + q]; };
                auto val_p = [&](int q) { return T.pred[gr ? gr[q] : g0 + q]; };
                auto pair = [&](int a, int b) {
                    const double ya = val_y(a), yb = val_y(b);
                    if (ya == yb) return;
                    const bool b_hi = yb > ya;
                    const int hi = b_hi ? b : a, lo = b_hi ? a : b;
                    const double s = val_p(hi) - val_p(lo);
                    const double sc = s < -30.0 ? -30.0 : (s > 30.0 ? 30.0 : s);
                    const double e = exp(sc);
                    const double dyw = fmax(fabs(val_y(hi) - val_y(lo)), 1e-6);
                    const double lam = -1.0 / (1.0 + e) * dyw;
                    const double d2 = e / ((1 + e) * (1 + e)) * dyw;
                    if (p == hi) g += wp * lam;
                    else g -= wp * lam;
                    h += wp * d2;
                };
                for (int a = p % step; a + step <= p; a += step) pair(a, p);
                for (int b = p + step; b < m; b += step) pair(p, b);
                T.g[rp] = g;
                T.h[rp] = h;
                const unsigned long long bg = (unsigned long long)__double_as_longlong(fabs(g));
                const unsigned long long bh = (unsigned long long)__double_as_longlong(fabs(h));
                mg = bg > mg ? bg : mg;
                mh = bh > mh ? bh : mh;
            }
            continue;
        }
        for (int p = threadIdx.x; p < m; p += blockDim.x) {
            const int rp = gr ? gr[p] : g0 + p;
            const double wp = T.w ? T.w[rp] : 1.0;
            double g = 0.0, h = 1e-6;
            if (m >= 2) {
                auto pair = [&](int a, int b) {  // a < b, positions within the group
                    const double ya = sy[a], yb = sy[b];
                    if (ya == yb) return;
                    const bool b_hi = yb > ya;
                    const int hi = b_hi ? b : a, lo = b_hi ? a : b;
                    const double s = sp[hi] - sp[lo];
                    const double sc = s < -30.0 ? -30.0 : (s > 30.0 ? 30.0 : s);
                    const double e = exp(sc);
                    const double dyw = fmax(fabs(sy[hi] - sy[lo]), 1e-6);
                    const double lam = -1.0 / (1.0 + e) * dyw;
                    const double d2 = e / ((1 + e) * (1 + e)) * dyw;
                    if (p == hi) g += wp * lam;
                    else g -= wp * lam;
                    h += wp * d2;
                };
                for (int a = p % step; a + step <= p; a += step) pair(a, p);
                for (int b = p + step; b < m; b += step) pair(p, b);
            }
            T.g[rp] = g;
            T.h[rp] = h;
            const unsigned long long bg = (unsigned long long)__double_as_longlong(fabs(g));
            const unsigned long long bh = (unsigned long long)__double_as_longlong(fabs(h));
            mg = bg > mg ? bg : mg;
            mh = bh > mh ? bh : mh;
        }
    }
    block_max_u64(mg, &T.max_g_bits);
    block_max_u64(mh, &T.max_h_bits);
}

// Rows outside any group keep g = 0, h = 1e-6 (objective.cpp:180-181).
__global__ void k_grad_fill(Task* tasks) {
    Task& T = tasks[blockIdx.y];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < T.n_rows; i += gridDim.x * blockDim.x) {
        T.g[i] = 0.0;
        T.h[i] = 1e-6;
    }
}

__global__ void k_reset_max(Task* tasks) {
    if (threadIdx.x == 0) {
        tasks[blockIdx.x].max_g_bits = 0;
        tasks[blockIdx.x].max_h_bits = 0;
    }
}

void launch_grad(hipStream_t st, Task* tasks, int n_tasks, int objective, int max_rows,
                 int max_groups) {
    hipLaunchKernelGGL(k_reset_max, dim3(n_tasks), dim3(64), 0, st, tasks);
    if (objective == 5) {
        hipLaunchKernelGGL(k_grad_fill, dim3(std::max(1, std::min(2048, (max_rows + 255) / 256)), n_tasks),
                           dim3(256), 0, st, tasks);
        hipLaunchKernelGGL(k_grad_pairwise, dim3(std::max(1, std::min(max_groups, 8192)), n_tasks),
                           dim3(256), 0, st, tasks);
    } else {
        hipLaunchKernelGGL(k_grad_mse, dim3(std::max(1, std::min(2048, (max_rows + 255) / 256)), n_tasks),
                           dim3(256), 0, st, tasks);
    }
}

// =================================================================== tree start
// Quantization exponents from the gradient maxima, root node over the sample.
__global__ void k_tree_init(Task* tasks, int n_features) {
    Task& T = tasks[blockIdx.x];
    if (threadIdx.x != 0) return;
    int eg = mag_exponent(T.max_g_bits);
    if (T.goss && T.goss_scale > 1.0) {
        int es;
        frexp(T.goss_scale, &es);
        eg += es;
    }
    const int eh = mag_exponent(T.max_h_bits) + ((T.goss && T.goss_scale > 1.0) ? 4 : 0);
    T.s_g = QG_BITS - eg;
    T.s_h = QH_BITS - eh;
    T.s_r = QR_BITS - mag_exponent(T.max_g_bits);
    T.s_rh = QR_BITS - mag_exponent(T.max_h_bits);
    T.s_e = QE_BITS - 2 * eg;
    T.n_nodes = 1;
    T.done = 0;
    T.pending0 = 0;
    T.pending1 = -1;
    T.cur_leaf = -1;
    DNode& r = T.nodes[0];
    for (int k = 0; k < 10; ++k) r.v[k] = -1;
    r.v[N_LEAF] = 1;
    r.v[N_TYPE] = 0;
    r.v[N_MGL] = 1;
    r.v[N_NS] = T.n_sample;
    r.v[N_DEPTH] = 0;
    r.seg_start = 0;
    r.seg_count = T.n_sample;
    r.value = 0.0;
    r.gain = 0.0;
    (void)n_features;
}

// Packed per-row histogram operands: pk = 2^48 + rint(g_use 2^s_g), qh = rint(h_use 2^s_h).
__global__ __launch_bounds__(256) void k_pack(Task* tasks) {
    Task& T = tasks[blockIdx.y];
    const int sg = T.s_g, sh = T.s_h;
    const bool gs = T.goss && T.scaled;
    const double sc = T.goss_scale;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < T.n_rows; i += gridDim.x * blockDim.x) {
        double g = T.g[i], h = T.h[i];
        if (gs && T.scaled[i]) {
            g *= sc;
            h *= sc;
        }
        const long long q = __double2ll_rn(ldexp(g, sg));
        T.pk[i] = CNT_ONE + (unsigned long long)q;
        if (!T.hconst) T.qh[i] = (unsigned long long)__double2ll_rn(ldexp(h, sh));
    }
}

// Identity partition for "all rows" samples.
__global__ void k_iota(Task* tasks) {
    Task& T = tasks[blockIdx.y];
    if (!T.contiguous) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < T.n_sample; i += gridDim.x * blockDim.x)
        T.idx[i] = i;
}

// Zero the histogram slots of `count` nodes starting at node id first (first < 0: the two
// children of the current split, i.e. ids n_nodes-2, n_nodes-1 after the split was applied).
__global__ __launch_bounds__(256) void k_zero_slots(Task* tasks, int F, int first, int count) {
    Task& T = tasks[blockIdx.y];
    int n0 = first;
    if (first < 0) {
        if (T.done) return;
        n0 = T.n_nodes - 2;
    }
    const size_t per = (size_t)F * NBIN;
    const size_t total = per * count;
    const size_t base = (size_t)n0 * per;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += (size_t)gridDim.x * blockDim.x) {
        T.hg[base + j] = 0;
        T.hc[base + j] = 0;
        if (!T.hconst) T.hh[base + j] = 0;
    }
}

// =================================================================== histogram build
// Which node a histogram launch fills: the root (node 0, rows [0, n_sample) of the partition)
// or the smaller child of the current split (reference tree.cpp:418-428).
__device__ inline bool hist_target(const Task& T, bool root, int& node, int& start, int& count) {
    if (root) {
        node = 0;
        start = 0;
        count = T.n_sample;
        return true;
    }
    if (T.done) return false;
    const int L = T.nodes[T.cur_leaf].v[N_LEFT], R = T.nodes[T.cur_leaf].v[N_RIGHT];
    const int cl = T.nodes[L].seg_count, cr = T.nodes[R].seg_count;
    node = (cl <= cr) ? L : R;
    start = T.nodes[node].seg_start;
    count = T.nodes[node].seg_count;
    return true;
}

// Fold a block's LDS histogram into the node's global histogram (exact integer atomics).
template <bool HV>
__device__ inline void hist_flush(const Task& T, const unsigned long long* lA,
                                  const unsigned long long* lB, int FB, int f0, int F, int node) {
    const size_t base = (size_t)node * F * NBIN;
    for (int j = threadIdx.x; j < FB * NBIN; j += blockDim.x) {
        const int f = f0 + j / NBIN;
        if (f >= F) break;
        const unsigned long long v = lA[j];
        if (v == 0) continue;
        const size_t o = base + (size_t)f * NBIN + (j % NBIN);
        atomicAdd((unsigned long long*)&T.hg[o], (unsigned long long)unpack_q(v));
        atomicAdd(&T.hc[o], (int)unpack_c(v));
        if (HV) atomicAdd((unsigned long long*)&T.hh[o], lB[j]);
    }
}

// Row-major kernel: each lane loads 16 consecutive feature bins of one row (128-bit load);
// LPR lanes cover the block's FB = 16*LPR features of a row.  Works on the partition index
// (gathered rows) or on contiguous rows.  One LDS word per (feature, bin) packs count and the
// fixed-point g sum; a second word holds q_h when h is not constant.
template <int LPR, bool HV>
__global__ __launch_bounds__(512) void k_hist_row(Shared S, Task* tasks, int root, int rows_per_block) {
    constexpr int FB = 16 * LPR;
    extern __shared__ unsigned long long lds[];
    unsigned long long* lA = lds;
    unsigned long long* lB = lds + FB * NBIN;
    const Task& T = tasks[blockIdx.z];
    int node, start, count;
    if (!hist_target(T, root != 0, node, start, count)) return;
    const int r0 = blockIdx.x * rows_per_block;
    if (r0 >= count) return;
    const int r1 = min(count, r0 + rows_per_block);
    const int f0 = blockIdx.y * FB;
    for (int j = threadIdx.x; j < FB * NBIN * (HV ? 2 : 1); j += blockDim.x) lds[j] = 0;
    __syncthreads();
    const int sub = threadIdx.x % LPR;
    const int fcol = f0 + sub * 16;
    const bool contig = T.contiguous && root;
    for (int k = r0 + (int)threadIdx.x / LPR; k < r1; k += blockDim.x / LPR) {
        const int i = contig ? k : T.idx[start + k];
        if (fcol < S.fp) {
            const uint4 b = *(const uint4*)(S.bins_row + (size_t)i * S.fp + fcol);
            const unsigned long long p = T.pk[i];
            const unsigned long long q = HV ? T.qh[i] : 0ull;
            const unsigned int w4[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int bin = (w4[t >> 2] >> ((t & 3) * 8)) & 255;
                const int slot = (sub * 16 + t) * NBIN + bin;
                atomicAdd(&lA[slot], p);
                if (HV) atomicAdd(&lB[slot], q);
            }
        }
    }
    __syncthreads();
    hist_flush<HV>(T, lA, lB, FB, f0, S.n_features, node);
}

// Column-major kernel for contiguous rows: each lane loads 16 consecutive rows of one
// feature column (128-bit load); the block covers FB features and a long run of rows, so the
// per-block partial histograms stay small next to the bytes read.
template <int FB, bool HV>
__global__ __launch_bounds__(256) void k_hist_col(Shared S, Task* tasks, int rows_per_block) {
    __shared__ unsigned long long lA[FB * NBIN];
    __shared__ unsigned long long lB[HV ? FB * NBIN : 1];
    const Task& T = tasks[blockIdx.z];
    const int count = T.n_sample;
    const int r0 = blockIdx.x * rows_per_block;
    if (r0 >= count) return;
    const int r1 = min(count, r0 + rows_per_block);
    const int f0 = blockIdx.y * FB;
    for (int j = threadIdx.x; j < FB * NBIN; j += blockDim.x) {
        lA[j] = 0;
        if (HV) lB[j] = 0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    for (int base = r0 + wave * 1024; base < r1; base += nwave * 1024) {
        const int row = base + lane * 16;
        if (row >= r1) continue;
        unsigned long long p[16], q[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const bool ok = row + t < r1;
            p[t] = ok ? T.pk[row + t] : 0ull;
            q[t] = (HV && ok) ? T.qh[row + t] : 0ull;
        }
        for (int fl = 0; fl < FB; ++fl) {
            const int f = f0 + fl;
            if (f >= S.n_features) break;
            const uint4 b = *(const uint4*)(S.bins_col + (size_t)f * S.n_pad + row);
            const unsigned int w4[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                if (!p[t]) continue;
                const int bin = (w4[t >> 2] >> ((t & 3) * 8)) & 255;
                atomicAdd(&lA[fl * NBIN + bin], p[t]);
                if (HV) atomicAdd(&lB[fl * NBIN + bin], q[t]);
            }
        }
    }
    __syncthreads();
    hist_flush<HV>(T, lA, lB, FB, f0, S.n_features, 0);
}

void launch_hist(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, bool root,
                 bool contiguous_root, bool hconst, int max_count, int variant) {
    if (max_count <= 0) return;
    if (root && contiguous_root && variant != 2) {
        constexpr int FB = 8;
        const int rpb = std::min(MAX_FLUSH_ROWS, std::max(4096, ((max_count / 64) + 1023) / 1024 * 1024));
        dim3 grid((max_count + rpb - 1) / rpb, (S.n_features + FB - 1) / FB, n_tasks);
        if (hconst) hipLaunchKernelGGL((k_hist_col<FB, false>), grid, dim3(256), 0, st, S, tasks, rpb);
        else hipLaunchKernelGGL((k_hist_col<FB, true>), grid, dim3(256), 0, st, S, tasks, rpb);
        return;
    }
    // Row-major: LPR = 2 (32 features per block).
    constexpr int LPR = 2, FB = 32;
    const int rpb = std::min(MAX_FLUSH_ROWS, std::max(1024, ((max_count / 128) + 255) / 256 * 256));
    dim3 grid((max_count + rpb - 1) / rpb, (S.fp + FB - 1) / FB, n_tasks);
    const size_t lds = (size_t)FB * NBIN * 8 * (hconst ? 1 : 2);
    static bool lds_attr = false;
    if (!lds_attr) {  // dynamic LDS above 64 KiB needs the per-kernel opt-in
        (void)hipFuncSetAttribute((const void*)k_hist_row<LPR, false>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)k_hist_row<LPR, true>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        lds_attr = true;
    }
    if (hconst) hipLaunchKernelGGL((k_hist_row<LPR, false>), grid, dim3(512), lds, st, S, tasks, root ? 1 : 0, rpb);
    else hipLaunchKernelGGL((k_hist_row<LPR, true>), grid, dim3(512), lds, st, S, tasks, root ? 1 : 0, rpb);
}

// =================================================================== split scan
// Best numeric split of every feature of the scanned node(s) (reference best_numeric_missing,
// histogram.cpp:144-193): totals over all bins, bin 0 = missing, thresholds b in [1, nb-2],
// empty bins skipped, missing-right (ml=0) before missing-left, first maximum kept.  One wave per
// feature, 4 bins per lane; prefix sums are exact int64, so the scan order cannot change them.
// For the larger child the histogram is first formed as parent - smaller (exact).
constexpr int SCAN_FPB = 8;   // features per block (2 per wave)

// Lane j holds bins 4j..4j+3 (qg/c/qh preloaded; bins >= nb must be zero).
__device__ inline void scan_feature_wave(const long long* qg, const int* c, const long long* qh,
                                         bool hconst, double inv_g, double inv_h, int nb,
                                         double lam, double gamma, int min_leaf, double min_hess,
                                         double l1, double& out_gain, int& out_thr, int& out_ml,
                                         long long& totQ, long long& totQh, long long& totC) {
    const int lane = threadIdx.x & 63;
    long long lg = 0, lh = 0;
    long long lc = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        lg += qg[t];
        lh += qh[t];
        lc += c[t];
    }
    totQ = wave_sum_ll(lg);
    totQh = wave_sum_ll(lh);
    totC = wave_sum_ll(lc);
    const long long missQ = __shfl(qg[0], 0, 64), missQh = __shfl(qh[0], 0, 64);
    const long long missC = __shfl((long long)c[0], 0, 64);
    // prefix over bins 1..b (bin 0 excluded)
    long long pg[4], ph[4], pc[4];
    long long rg = 0, rh = 0, rc = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int b = lane * 4 + t;
        if (b >= 1) {
            rg += qg[t];
            rh += qh[t];
            rc += c[t];
        }
        pg[t] = rg;
        ph[t] = rh;
        pc[t] = rc;
    }
    const long long eg = wave_incl_ll(rg) - rg, eh = wave_incl_ll(rh) - rh, ec = wave_incl_ll(rc) - rc;
    const double tG = (double)totQ * inv_g;
    const double tH = hconst ? (double)totC : (double)totQh * inv_h;
    const double parent = leaf_score(tG, tH, lam, l1);
    double best = -INFINITY;
    int bkey = 0x7fffffff;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int b = lane * 4 + t;
        if (b < 1 || b > nb - 2 || c[t] == 0) continue;
        const long long QL = eg + pg[t], QhL = eh + ph[t], CL = ec + pc[t];
        for (int ml = 0; ml <= 1; ++ml) {
            const long long ql = QL + (ml ? missQ : 0), qhl = QhL + (ml ? missQh : 0);
            const long long cl = CL + (ml ? missC : 0);
            const long long cr = totC - cl;
            if (cl < min_leaf || cr < min_leaf) continue;
            const double gl = (double)ql * inv_g, gr = (double)(totQ - ql) * inv_g;
            const double hl = hconst ? (double)cl : (double)qhl * inv_h;
            const double hr = hconst ? (double)cr : (double)(totQh - qhl) * inv_h;
            if (hl < min_hess || hr < min_hess) continue;
            const double gain = 0.5 * (leaf_score(gl, hl, lam, l1) + leaf_score(gr, hr, lam, l1) - parent) - gamma;
            const int key = b * 2 + ml;
            if (cand_better(gain, key, best, bkey)) {
                best = gain;
                bkey = key;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double og = __shfl_xor(best, o, 64);
        const int ok = __shfl_xor(bkey, o, 64);
        if (cand_better(og, ok, best, bkey)) {
            best = og;
            bkey = ok;
        }
    }
    out_gain = best;
    out_thr = (bkey == 0x7fffffff) ? -1 : bkey >> 1;
    out_ml = (bkey == 0x7fffffff) ? 1 : (bkey & 1);
}

// mode 0: scan node 0 (root); mode 1: scan both children of the current split.
__global__ __launch_bounds__(256) void k_scan(Shared S, Task* tasks, GrowP P, int mode) {
    Task& T = tasks[blockIdx.z];
    int node, small = -1, parent = -1;
    if (mode == 0) {
        if (blockIdx.y != 0) return;
        node = 0;
    } else {
        if (T.done) return;
        parent = T.cur_leaf;
        const int L = T.nodes[parent].v[N_LEFT], R = T.nodes[parent].v[N_RIGHT];
        small = (T.nodes[L].seg_count <= T.nodes[R].seg_count) ? L : R;
        node = blockIdx.y == 0 ? L : R;
    }
    const int F = S.n_features;
    const size_t per = (size_t)F * NBIN;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int fl = wave; fl < SCAN_FPB; fl += 4) {
        const int f = blockIdx.x * SCAN_FPB + fl;
        if (f >= F) break;
        const size_t off = (size_t)node * per + (size_t)f * NBIN;
        long long qg[4], qh[4];
        int c[4];
        const bool sub = (mode == 1 && node != small);
        const size_t poff = sub ? (size_t)parent * per + (size_t)f * NBIN : 0;
        const size_t soff = sub ? (size_t)small * per + (size_t)f * NBIN : 0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int b = lane * 4 + t;
            if (sub) {
                // larger child = parent - smaller, exact in fixed point (reference
                // subtract_hist_inplace, histogram.cpp:108-117); stored for later use
                qg[t] = T.hg[poff + b] - T.hg[soff + b];
                c[t] = T.hc[poff + b] - T.hc[soff + b];
                qh[t] = T.hconst ? 0 : T.hh[poff + b] - T.hh[soff + b];
                T.hg[off + b] = qg[t];
                T.hc[off + b] = c[t];
                if (!T.hconst) T.hh[off + b] = qh[t];
            } else {
                qg[t] = T.hg[off + b];
                c[t] = T.hc[off + b];
                qh[t] = T.hconst ? 0 : T.hh[off + b];
            }
            if (b >= S.n_bins) {
                qg[t] = 0;
                c[t] = 0;
                qh[t] = 0;
            }
        }
        double gain;
        int thr, ml;
        long long tq, tqh, tc;
        const double inv_g = ldexp(1.0, -T.s_g), inv_h = ldexp(1.0, -T.s_h);
        scan_feature_wave(qg, c, qh, T.hconst, inv_g, inv_h, S.n_bins, P.l2, P.gamma, P.min_leaf,
                          P.min_hess, P.l1, gain, thr, ml, tq, tqh, tc);
        if (lane == 0) {
            FCand& o = T.fcand[(size_t)node * F + f];
            o.gain = gain;
            o.thr = thr;
            o.ml = ml;
            if (f == 0) {  // node value from its own sums (tree.cpp:323-328, 407-416)
                const double G = (double)tq * inv_g;
                const double H = T.hconst ? (double)tc : (double)tqh * inv_h;
                T.nodes[node].value = leaf_value(G, H, P.l2, P.l1, P.mds);
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && blockIdx.y == 0) {
        if (mode == 0) {
            T.pending0 = 0;
            T.pending1 = -1;
        } else {
            T.pending0 = T.nodes[parent].v[N_LEFT];
            T.pending1 = T.nodes[parent].v[N_RIGHT];
        }
    }
}

void launch_scan(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, const GrowP& P, int mode) {
    dim3 grid((S.n_features + SCAN_FPB - 1) / SCAN_FPB, mode == 0 ? 1 : 2, n_tasks);
    hipLaunchKernelGGL(k_scan, grid, dim3(256), 0, st, S, tasks, P, mode);
}

// =================================================================== best-first selection
// Reduce the per-feature results of freshly scanned nodes to one cached candidate each
// (tree.cpp:160-166: ascending features, skip invalid or gain < min_split_gain, strict '>'),
// then pick the leaf to split (tree.cpp:343-383: lowest node id among equal scores) and apply
// the split bookkeeping (tree.cpp:385-392, 445-456).  mode 1 applies a host-made decision.
__global__ __launch_bounds__(256) void k_select(Task* tasks, int F, GrowP P, int mode) {
    Task& T = tasks[blockIdx.x];
    if (T.done) return;
    __shared__ double sg[4];
    __shared__ int sk[4];
    __shared__ int s_leaf;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (mode == 0) {
        for (int pi = 0; pi < 2; ++pi) {
            const int nd = pi == 0 ? T.pending0 : T.pending1;
            if (nd < 0) continue;
            double best = -INFINITY;
            int bk = 0x7fffffff;
            for (int f = threadIdx.x; f < F; f += blockDim.x) {
                const FCand c = T.fcand[(size_t)nd * F + f];
                if (c.thr < 0 || c.gain < P.gamma) continue;
                if (cand_better(c.gain, f, best, bk)) {
                    best = c.gain;
                    bk = f;
                }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double og = __shfl_xor(best, o, 64);
                const int ok = __shfl_xor(bk, o, 64);
                if (cand_better(og, ok, best, bk)) {
                    best = og;
                    bk = ok;
                }
            }
            if (lane == 0) {
                sg[wave] = best;
                sk[wave] = bk;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
                    if (cand_better(sg[w], sk[w], best, bk)) {
                        best = sg[w];
                        bk = sk[w];
                    }
                DCand& c = T.cands[nd];
                const DNode& n = T.nodes[nd];
                const bool elig = n.v[N_DEPTH] < P.max_depth && n.v[N_NS] >= 2 * P.min_leaf;
                c.valid = (bk != 0x7fffffff) && elig;
                c.gain = best;
                c.feature = bk == 0x7fffffff ? -1 : bk;
                c.feature2 = -1;
                c.type = 0;
                if (bk != 0x7fffffff) {
                    const FCand fc = T.fcand[(size_t)nd * F + bk];
                    c.thr = fc.thr;
                    c.mgl = fc.ml;
                } else {
                    c.thr = -1;
                    c.mgl = 1;
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x != 0) return;
    if (T.n_nodes + 2 > MAX_NODES || T.n_nodes + 2 > 2 * P.max_leaves - 1) {
        T.done = 1;
        return;
    }
    int leaf = -1;
    if (mode == 0) {
        double best_score = P.gamma;
        for (int lid = 0; lid < T.n_nodes; ++lid) {
            const DNode& n = T.nodes[lid];
            if (!n.v[N_LEAF]) continue;
            const DCand& c = T.cands[lid];
            if (!c.valid || c.feature < 0 || !(c.gain >= P.gamma)) continue;
            if (c.gain > best_score) {
                best_score = c.gain;
                leaf = lid;
            }
        }
        if (leaf < 0) {
            T.done = 1;
            return;
        }
        const DCand& c = T.cands[leaf];
        T.cur_leaf = leaf;
        T.cur_type = c.type;
        T.cur_feat = c.feature;
        T.cur_feat2 = c.feature2;
        T.cur_thr = c.thr;
        T.cur_mgl = c.mgl;
        T.cur_gain = c.gain;
    } else {
        leaf = T.cur_leaf;
        const bool bad = leaf < 0 || leaf >= T.n_nodes || !T.nodes[leaf].v[N_LEAF] || T.cur_feat < 0 ||
                         T.cur_feat >= F || (T.cur_type == 1 && (T.cur_feat2 < 0 || T.cur_feat2 >= F)) ||
                         (T.cur_type != 0 && T.cur_type != 1);
        if (bad) {  // a host decision that does not name a splittable leaf stops the tree
            T.done = 1;
            return;
        }
    }
    (void)s_leaf;
    DNode& p = T.nodes[leaf];
    const int L = T.n_nodes, R = T.n_nodes + 1;
    p.v[N_LEAF] = 0;
    p.v[N_TYPE] = T.cur_type;
    p.v[N_FEAT] = T.cur_feat;
    p.v[N_FEAT2] = T.cur_feat2;
    p.v[N_THR] = T.cur_thr;
    p.v[N_MGL] = T.cur_mgl;
    p.v[N_LEFT] = L;
    p.v[N_RIGHT] = R;
    p.gain = T.cur_gain;
    for (int side = 0; side < 2; ++side) {
        DNode& c = T.nodes[side == 0 ? L : R];
        for (int k = 0; k < 10; ++k) c.v[k] = -1;
        c.v[N_LEAF] = 1;
        c.v[N_TYPE] = 0;
        c.v[N_MGL] = 1;
        c.v[N_NS] = 0;
        c.v[N_DEPTH] = p.v[N_DEPTH] + 1;
        c.seg_start = p.seg_start;
        c.seg_count = 0;
        c.value = 0.0;
        c.gain = 0.0;
        T.cands[side == 0 ? L : R].valid = 0;
    }
    T.n_nodes += 2;
}

void launch_select(hipStream_t st, Task* tasks, int n_tasks, int F, const GrowP& P, int mode) {
    hipLaunchKernelGGL(k_select, dim3(n_tasks), dim3(256), 0, st, tasks, F, P, mode);
}

// =================================================================== stable partition
// Reference split_indices (tree.cpp:68-91): rows of the leaf, in ascending order, go left when
// Axis: bin 0 -> missing_go_left else bin <= threshold; Pair: either bin 0 -> missing_go_left
// else pair_hash <= threshold.  Three passes: per-chunk left counts, stable scatter into the
// scratch index, copy back (block 0 records the children's row ranges).
constexpr int PART_CHUNK = 4096;   // rows per block (16 per thread)

__device__ inline bool row_goes_left(const Shared& S, const Task& T, int i) {
    const int b1 = S.bins_col[(size_t)T.cur_feat * S.n_pad + i];
    if (T.cur_type == 1) {
        const int b2 = S.bins_col[(size_t)T.cur_feat2 * S.n_pad + i];
        if (b1 == 0 || b2 == 0) return T.cur_mgl != 0;
        const unsigned z = ((unsigned)b1 * 2654435761u + (unsigned)b2 * 40503u) % (unsigned)S.n_bins;
        return (int)z <= T.cur_thr;
    }
    if (b1 == 0) return T.cur_mgl != 0;
    return b1 <= T.cur_thr;
}

__device__ inline int block_sum_int(int v, int* red) {
    v = wave_sum_i(v);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[wave] = v;
    __syncthreads();
    int s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    return s;
}

__global__ __launch_bounds__(256) void k_part_count(Shared S, Task* tasks) {
    Task& T = tasks[blockIdx.y];
    if (T.done) return;
    const DNode& p = T.nodes[T.cur_leaf];
    const int c0 = blockIdx.x * PART_CHUNK;
    if (c0 >= p.seg_count) return;
    __shared__ int red[4];
    int left = 0;
    const int k0 = c0 + threadIdx.x * 16;
    for (int k = k0; k < k0 + 16 && k < p.seg_count; ++k) left += row_goes_left(S, T, T.idx[p.seg_start + k]);
    const int tot = block_sum_int(left, red);
    if (threadIdx.x == 0) T.chunk_cnt[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_part_scatter(Shared S, Task* tasks) {
    Task& T = tasks[blockIdx.y];
    if (T.done) return;
    const DNode& p = T.nodes[T.cur_leaf];
    const int cnt = p.seg_count, st = p.seg_start;
    const int c0 = blockIdx.x * PART_CHUNK;
    if (c0 >= cnt) return;
    __shared__ int red[4];
    __shared__ int wsum[4];
    const int nch = (cnt + PART_CHUNK - 1) / PART_CHUNK;
    int before = 0, total = 0;
    for (int c = threadIdx.x; c < nch; c += blockDim.x) {
        const int v = T.chunk_cnt[c];
        total += v;
        if (c < (int)blockIdx.x) before += v;
    }
    before = block_sum_int(before, red);
    total = block_sum_int(total, red);
    // thread-local flags for 16 consecutive rows
    const int k0 = c0 + threadIdx.x * 16;
    unsigned flags = 0;
    int rows[16];
    int nl = 0;
    for (int t = 0; t < 16; ++t) {
        const int k = k0 + t;
        rows[t] = -1;
        if (k < cnt) {
            rows[t] = T.idx[st + k];
            if (row_goes_left(S, T, rows[t])) {
                flags |= 1u << t;
                ++nl;
            }
        }
    }
    // block-wide exclusive scan of nl
    const int incl = wave_incl_i(nl);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[wave] = incl;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wave; ++w) woff += wsum[w];
    int lrank = before + woff + incl - nl;                       // lefts before my first row
    int rrank = (k0 - c0) + (c0 - before) - (woff + incl - nl);  // rights before my first row
    // (k0 - c0) rows of this chunk precede me; (c0 - before) rights precede this chunk
    for (int t = 0; t < 16; ++t) {
        if (rows[t] < 0) break;
        if (flags & (1u << t)) T.idx_tmp[st + lrank++] = rows[t];
        else T.idx_tmp[st + total + rrank++] = rows[t];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) T.n_left = total;
}

__global__ __launch_bounds__(256) void k_part_copy(Task* tasks) {
    Task& T = tasks[blockIdx.y];
    if (T.done) return;
    const DNode& p = T.nodes[T.cur_leaf];
    const int cnt = p.seg_count, st = p.seg_start;
    const int c0 = blockIdx.x * PART_CHUNK;
    if (c0 >= cnt) return;
    for (int k = c0 + threadIdx.x; k < c0 + PART_CHUNK && k < cnt; k += blockDim.x) T.idx[st + k] = T.idx_tmp[st + k];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int nl = T.n_left;
        DNode& L = T.nodes[p.v[N_LEFT]];
        DNode& R = T.nodes[p.v[N_RIGHT]];
        L.seg_start = st;
        L.seg_count = nl;
        L.v[N_NS] = nl;
        R.seg_start = st + nl;
        R.seg_count = cnt - nl;
        R.v[N_NS] = cnt - nl;
    }
}

void launch_partition(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, int max_count) {
    const int nch = std::max(1, (max_count + PART_CHUNK - 1) / PART_CHUNK);
    dim3 grid(nch, n_tasks);
    hipLaunchKernelGGL(k_part_count, grid, dim3(256), 0, st, S, tasks);
    hipLaunchKernelGGL(k_part_scatter, grid, dim3(256), 0, st, S, tasks);
    hipLaunchKernelGGL(k_part_copy, grid, dim3(256), 0, st, tasks);
}

// =================================================================== host-driven helpers
// Fixed-point sum of g_use^2 over a node's rows: the leaf "energy" of the PAIR trigger
// (tree.cpp:356-357).
__global__ __launch_bounds__(256) void k_energy(Task* tasks, int task, int node) {
    Task& T = tasks[task];
    const DNode& n = T.nodes[node];
    long long acc = 0;
    const double sc = T.goss_scale;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n.seg_count; k += gridDim.x * blockDim.x) {
        const int i = T.idx[n.seg_start + k];
        double g = T.g[i];
        if (T.goss && T.scaled && T.scaled[i]) g *= sc;
        acc += __double2ll_rn(ldexp(g * g, T.s_e));
    }
    acc = wave_sum_ll(acc);
    if ((threadIdx.x & 63) == 0 && acc) atomicAdd((unsigned long long*)&T.energy[node], (unsigned long long)acc);
}

void launch_energy(hipStream_t st, Task* tasks, int task, int node, int count) {
    const int blocks = std::max(1, std::min(1024, (count + 255) / 256));
    hipLaunchKernelGGL(k_energy, dim3(blocks), dim3(256), 0, st, tasks, task, node);
}

// PAIR histograms: one hashed histogram per feature pair over a node's rows (the fused
// hash+accumulate pass of tree.cpp:243-267), then one scan per pair with lambda = pair_l2.
constexpr int MAX_PAIRS = 32;
__global__ __launch_bounds__(256) void k_pair_hist(Shared S, Task* tasks, int task, int node,
                                                   const int* __restrict__ pf, int n_pairs,
                                                   long long* ph_g, int* ph_c, long long* ph_h,
                                                   int rows_per_block) {
    extern __shared__ unsigned long long lds[];
    const Task& T = tasks[task];
    const bool hv = !T.hconst;
    unsigned long long* lA = lds;
    unsigned long long* lB = lds + (size_t)n_pairs * NBIN;
    const DNode& n = T.nodes[node];
    const int r0 = blockIdx.x * rows_per_block;
    if (r0 >= n.seg_count) return;
    const int r1 = min(n.seg_count, r0 + rows_per_block);
    for (int j = threadIdx.x; j < n_pairs * NBIN * (hv ? 2 : 1); j += blockDim.x) lds[j] = 0;
    __syncthreads();
    for (int k = r0 + threadIdx.x; k < r1; k += blockDim.x) {
        const int i = T.idx[n.seg_start + k];
        const unsigned char* row = S.bins_row + (size_t)i * S.fp;
        const unsigned long long p = T.pk[i];
        const unsigned long long q = hv ? T.qh[i] : 0ull;
        for (int pp = 0; pp < n_pairs; ++pp) {
            const unsigned b1 = row[pf[2 * pp]], b2 = row[pf[2 * pp + 1]];
            const unsigned z = (b1 * 2654435761u + b2 * 40503u) % (unsigned)S.n_bins;
            atomicAdd(&lA[pp * NBIN + z], p);
            if (hv) atomicAdd(&lB[pp * NBIN + z], q);
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n_pairs * NBIN; j += blockDim.x) {
        const unsigned long long v = lA[j];
        if (!v) continue;
        atomicAdd((unsigned long long*)&ph_g[j], (unsigned long long)unpack_q(v));
        atomicAdd(&ph_c[j], (int)unpack_c(v));
        if (hv) atomicAdd((unsigned long long*)&ph_h[j], lB[j]);
    }
}

__global__ __launch_bounds__(64) void k_pair_scan(Shared S, Task* tasks, int task, const long long* ph_g,
                                                  const int* ph_c, const long long* ph_h, GrowP P,
                                                  FCand* out) {
    const Task& T = tasks[task];
    const int pp = blockIdx.x;
    double gain;
    int thr, ml;
    long long tq, tqh, tc;
    const int lane = threadIdx.x & 63;
    long long qg[4], qh[4];
    int c[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int b = lane * 4 + t;
        const bool ok = b < S.n_bins;
        qg[t] = ok ? ph_g[(size_t)pp * NBIN + b] : 0;
        c[t] = ok ? ph_c[(size_t)pp * NBIN + b] : 0;
        qh[t] = (ok && !T.hconst) ? ph_h[(size_t)pp * NBIN + b] : 0;
    }
    scan_feature_wave(qg, c, qh, T.hconst, ldexp(1.0, -T.s_g),
                      ldexp(1.0, -T.s_h), S.n_bins, P.pair_l2, P.gamma, P.min_leaf, P.min_hess,
                      P.l1, gain, thr, ml, tq, tqh, tc);
    if (threadIdx.x == 0) {
        out[pp].gain = gain;
        out[pp].thr = thr;
        out[pp].ml = ml;
    }
}

void launch_pair(hipStream_t st, const Shared& S, Task* tasks, int task, int node, int count,
                 const int* pf, int n_pairs, long long* ph_g, int* ph_c, long long* ph_h,
                 bool hconst, const GrowP& P, FCand* out) {
    const int rpb = 4096;
    const int blocks = std::max(1, (count + rpb - 1) / rpb);
    const size_t lds = (size_t)n_pairs * NBIN * 8 * (hconst ? 1 : 2);
    static bool lds_attr = false;
    if (!lds_attr) {
        (void)hipFuncSetAttribute((const void*)k_pair_hist, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        lds_attr = true;
    }
    hipLaunchKernelGGL(k_pair_hist, dim3(blocks), dim3(256), lds, st, S, tasks, task, node, pf,
                       n_pairs, ph_g, ph_c, ph_h, rpb);
    hipLaunchKernelGGL(k_pair_scan, dim3(n_pairs), dim3(64), 0, st, S, tasks, task, ph_g, ph_c,
                       ph_h, P, out);
}

// =================================================================== tree end
// Leaf id of every sampled row from the final partition (all-rows samples), or by walking the
// tree over the bins (GOSS samples) — reference apply_all (tree.cpp:43-55).
__global__ __launch_bounds__(256) void k_nid_segments(Task* tasks) {
    Task& T = tasks[blockIdx.y];
    __shared__ int st[MAX_NODES], cn[MAX_NODES], id[MAX_NODES];
    __shared__ int nl;
    if (threadIdx.x == 0) nl = 0;
    __syncthreads();
    for (int k = threadIdx.x; k < T.n_nodes; k += blockDim.x) {
        const DNode& n = T.nodes[k];
        if (n.v[N_LEAF] && n.seg_count > 0) {
            // rank by segment start (starts of non-empty leaves are distinct)
            int r = 0;
            for (int j = 0; j < T.n_nodes; ++j) {
                const DNode& m = T.nodes[j];
                if (m.v[N_LEAF] && m.seg_count > 0 && m.seg_start < n.seg_start) ++r;
            }
            st[r] = n.seg_start;
            cn[r] = n.seg_count;
            id[r] = k;
            atomicAdd(&nl, 1);
        }
    }
    __syncthreads();
    const int L = nl;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < T.n_sample; k += gridDim.x * blockDim.x) {
        int lo = 0, hi = L - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (st[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        T.nid[T.idx[k]] = id[lo];
    }
}

__global__ __launch_bounds__(256) void k_nid_walk(Shared S, Task* tasks) {
    Task& T = tasks[blockIdx.y];
    __shared__ int nf[MAX_NODES], nf2[MAX_NODES], nt[MAX_NODES], nlft[MAX_NODES], nrgt[MAX_NODES], nflag[MAX_NODES];
    for (int k = threadIdx.x; k < T.n_nodes; k += blockDim.x) {
        const DNode& n = T.nodes[k];
        nf[k] = n.v[N_FEAT];
        nf2[k] = n.v[N_FEAT2];
        nt[k] = n.v[N_THR];
        nlft[k] = n.v[N_LEFT];
        nrgt[k] = n.v[N_RIGHT];
        nflag[k] = n.v[N_LEAF] | (n.v[N_TYPE] << 1) | (n.v[N_MGL] << 3);
    }
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < T.n_rows; i += gridDim.x * blockDim.x) {
        const unsigned char* row = S.bins_row + (size_t)i * S.fp;
        int k = 0;
        while (!(nflag[k] & 1)) {
            const int fl = nflag[k];
            const int b1 = row[nf[k]];
            bool left;
            if (((fl >> 1) & 3) == 1) {
                const int b2 = row[nf2[k]];
                if (b1 == 0 || b2 == 0) left = (fl >> 3) & 1;
                else left = (int)(((unsigned)b1 * 2654435761u + (unsigned)b2 * 40503u) % (unsigned)S.n_bins) <= nt[k];
            } else {
                left = b1 == 0 ? ((fl >> 3) & 1) : (b1 <= nt[k]);
            }
            k = left ? nlft[k] : nrgt[k];
        }
        T.nid[i] = k;
    }
}

// TORB sweep sums per (sweep segment, node) over ALL rows (booster.cpp:589-637): rows are time
// sorted, so segment s is the row range [cut_{s-1}, cut_s) (s = n_buckets: rows past the last
// cut, which only enter the full-train refit).  Each block folds a run of one segment into LDS.

__global__ __launch_bounds__(256) void k_torb_accum(Task* tasks, const TorbChunk* const* chunks,
                                                    const int* n_chunks) {
    Task& T = tasks[blockIdx.y];
    if ((int)blockIdx.x >= n_chunks[blockIdx.y]) return;
    const TorbChunk c = chunks[blockIdx.y][blockIdx.x];
    __shared__ long long ag[MAX_NODES], ah[MAX_NODES];
    __shared__ int ac[MAX_NODES];
    for (int k = threadIdx.x; k < MAX_NODES; k += blockDim.x) {
        ag[k] = 0;
        ah[k] = 0;
        ac[k] = 0;
    }
    __syncthreads();
    const int sr = T.s_r, srh = T.s_rh;
    for (int i = c.r0 + threadIdx.x; i < c.r1; i += blockDim.x) {
        const int k = T.nid[i];
        atomicAdd((unsigned long long*)&ag[k], (unsigned long long)__double2ll_rn(ldexp(T.g[i], sr)));
        if (!T.hconst) atomicAdd((unsigned long long*)&ah[k], (unsigned long long)__double2ll_rn(ldexp(T.h[i], srh)));
        atomicAdd(&ac[k], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < T.n_nodes; k += blockDim.x) {
        if (!ac[k]) continue;
        const size_t o = (size_t)c.s * MAX_NODES + k;
        atomicAdd((unsigned long long*)&T.tg[o], (unsigned long long)ag[k]);
        if (!T.hconst) atomicAdd((unsigned long long*)&T.th[o], (unsigned long long)ah[k]);
        atomicAdd(&T.tc[o], ac[k]);
    }
}

// Leaf refit and TORB prefix values (booster.cpp:591-648); one thread per node.
__global__ __launch_bounds__(256) void k_torb_final(Task* tasks, GrowP P, double min_cover, int iid) {
    Task& T = tasks[blockIdx.x];
    const int nn = T.n_nodes;
    const int k = threadIdx.x;
    if (k >= nn) return;
    const int B = T.n_buckets;
    const double inv = ldexp(1.0, -T.s_r), invh = ldexp(1.0, -T.s_rh);
    long long fg = 0, fh = 0, fc = 0;
    for (int s = 0; s <= B; ++s) {
        const size_t o = (size_t)s * MAX_NODES + k;
        fg += T.tg[o];
        fh += T.hconst ? 0 : T.th[o];
        fc += T.tc[o];
    }
    const double Hf = T.hconst ? (double)fc : (double)fh * invh;
    const double full = leaf_value((double)fg * inv, Hf, P.l2, P.l1, P.mds);
    T.values[k] = full;
    T.cover[k] = (int)fc;
    if (iid) {
        for (int b = 0; b < B; ++b) T.prefix[(size_t)b * nn + k] = full;
        return;
    }
    long long pg = 0, ph = 0, pc = 0;
    for (int b = 0; b < B; ++b) {
        const size_t o = (size_t)b * MAX_NODES + k;
        pg += T.tg[o];
        ph += T.hconst ? 0 : T.th[o];
        pc += T.tc[o];
        const double cover = (double)T.cuts[b];
        const double alpha = cover <= 0.0 ? 0.0 : fmin(1.0, cover / fmax(min_cover, 1.0));
        const double Hp = T.hconst ? (double)pc : (double)ph * invh;
        const double pref = cover > 0.0 ? leaf_value((double)pg * inv, Hp, P.l2, P.l1, P.mds) : full;
        T.prefix[(size_t)b * nn + k] = alpha * pref + (1.0 - alpha) * full;
    }
}

// pred += lr * prefix[bucket, leaf] (booster.cpp:655-664).
__global__ __launch_bounds__(256) void k_pred_update(Task* tasks, double lr) {
    Task& T = tasks[blockIdx.y];
    const int nn = T.n_nodes;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < T.n_rows; i += gridDim.x * blockDim.x)
        T.pred[i] += lr * T.prefix[(size_t)T.bucket[i] * nn + T.nid[i]];
}

__global__ void k_zero_torb(Task* tasks, int words) {
    Task& T = tasks[blockIdx.y];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < words; j += gridDim.x * blockDim.x) {
        T.tg[j] = 0;
        if (!T.hconst) T.th[j] = 0;
        T.tc[j] = 0;
    }
}

void launch_tree_end(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, bool walk,
                     int max_rows, const TorbChunk* const* chunks, const int* n_chunks,
                     int max_chunks, const GrowP& P, double min_cover, int iid, double lr) {
    const int rb = std::max(1, std::min(4096, (max_rows + 255) / 256));
    if (walk) hipLaunchKernelGGL(k_nid_walk, dim3(rb, n_tasks), dim3(256), 0, st, S, tasks);
    else hipLaunchKernelGGL(k_nid_segments, dim3(rb, n_tasks), dim3(256), 0, st, tasks);
    const int words = (MAX_BUCKETS + 1) * MAX_NODES;
    hipLaunchKernelGGL(k_zero_torb, dim3((words + 255) / 256, n_tasks), dim3(256), 0, st, tasks, words);
    hipLaunchKernelGGL(k_torb_accum, dim3(std::max(1, max_chunks), n_tasks), dim3(256), 0, st, tasks,
                       chunks, n_chunks);
    hipLaunchKernelGGL(k_torb_final, dim3(n_tasks), dim3(256), 0, st, tasks, P, min_cover, iid);
    hipLaunchKernelGGL(k_pred_update, dim3(rb, n_tasks), dim3(256), 0, st, tasks, lr);
}

void launch_tree_start(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, int max_rows) {
    const int rb = std::max(1, std::min(4096, (max_rows + 255) / 256));
    hipLaunchKernelGGL(k_tree_init, dim3(n_tasks), dim3(64), 0, st, tasks, S.n_features);
    hipLaunchKernelGGL(k_pack, dim3(rb, n_tasks), dim3(256), 0, st, tasks);
    hipLaunchKernelGGL(k_iota, dim3(rb, n_tasks), dim3(256), 0, st, tasks);
}

void launch_zero_slots(hipStream_t st, Task* tasks, int n_tasks, int F, int first, int count) {
    const size_t total = (size_t)F * NBIN * count;
    const int blocks = (int)std::min<size_t>(1024, (total + 255) / 256);
    hipLaunchKernelGGL(k_zero_slots, dim3(blocks, n_tasks), dim3(256), 0, st, tasks, F, first, count);
}

}  // namespace tgbm
