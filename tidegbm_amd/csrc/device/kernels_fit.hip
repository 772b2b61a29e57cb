// gfx950 kernels of the histogram-GBDT fit path.  Every kernel takes the task table and works
// on task blockIdx.z (or .y); kernels of the best-first loop read the current split from the
// task state, so the whole loop can run without host round trips.
//
// Built with -ffp-contract=off: every floating-point expression that mirrors a reference
// formula must round like the reference's x86-64 build (no fused multiply-add).
#include <hip/hip_ext.h>

#include <stdexcept>
#include <string>

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
        e[fl][k] = (f < F && k < (n_edges[f] & (CAT_EDGE_FLAG - 1))) ? edges[(size_t)f * NBIN + k] : 0.0;
    }
    if (threadIdx.x < FB) ne[threadIdx.x] = (f0 + (int)threadIdx.x < F) ? n_edges[f0 + threadIdx.x] : 0;
    __syncthreads();
    const int fl = threadIdx.x % FB;
    const int f = f0 + fl;
    if (f >= F) return;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(n, r0 + rows_per_block);
    const int cnt = ne[fl] & (CAT_EDGE_FLAG - 1);
    const bool cat = (ne[fl] & CAT_EDGE_FLAG) != 0;
    for (int i = r0 + (int)threadIdx.x / FB; i < r1; i += blockDim.x / FB) {
        const double x = X[(size_t)i * F + f];
        int b = 0;
        if (x == x) {
            int lo = 0, hi = cnt;
            if (cat) {  // lower_bound, then an exact match (binning.cpp:107-112)
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (e[fl][mid] < x) lo = mid + 1;
                    else hi = mid;
                }
                b = (lo < cnt && e[fl][lo] == x) ? 1 + lo : 0;
            } else {   // upper_bound (binning.cpp:113-116)
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (e[fl][mid] <= x) lo = mid + 1;
                    else hi = mid;
                }
                b = 1 + lo;
            }
        }
        bins_row[(size_t)i * fp + f] = (unsigned char)b;
        if (bins_col) bins_col[(size_t)f * n_pad + i] = (unsigned char)b;
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

// =================================================================== reductions
// Block maximum of two values, written to the task's per-block slots (no global atomics).
__device__ inline void block_max2_store(unsigned long long a, unsigned long long b, const Task& T) {
    __shared__ unsigned long long sa[16], sb[16];
    a = wave_max_u64(a);
    b = wave_max_u64(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            a = sa[k] > a ? sa[k] : a;
            b = sb[k] > b ? sb[k] : b;
        }
        T.red[blockIdx.x] = a;
        T.red[MAX_RED_BLOCKS + blockIdx.x] = b;
    }
}

// =================================================================== gradients
// MSE (objective.cpp:105-119, cost terms off): g = w (pred - y), h = w.
__global__ __launch_bounds__(256) void k_grad_mse(Task* tasks) {
    const Task T = tasks[blockIdx.y];   // a copy: the stores below cannot alias its fields
    unsigned long long mg = 0, mh = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < T.n_rows; i += gridDim.x * blockDim.x) {
        const double wi = T.w ? T.w[i] : 1.0;
        const double g = wi * (T.pred[i] - T.y[i]);
        T.g[i] = g;
        if (!T.hconst) T.h[i] = wi;   // h == 1 for every row otherwise (no weights), never stored
        const unsigned long long bg = (unsigned long long)__double_as_longlong(fabs(g));
        const unsigned long long bh = (unsigned long long)__double_as_longlong(fabs(wi));
        mg = bg > mg ? bg : mg;
        mh = bh > mh ? bh : mh;
    }
    block_max2_store(mg, mh, T);
    if (blockIdx.x == 0 && threadIdx.x == 0) tasks[blockIdx.y].n_red = gridDim.x;
}

// Within-group pairwise gradient (objective.cpp:179-219).  One block per group; each row
// replays, in the reference's order, the pairs that touch it: first as the later element b of
// pairs (a, b) for ascending a, then as the earlier element a for ascending b.  Row values of
// the group are staged in LDS (groups larger than the LDS stage read global memory).
constexpr int PAIR_MAX_GROUP = 4096;

template <bool LDS>
__device__ inline void pairwise_rows(Task& T, int g0, int m, const int* gr, const double* sy,
                                     const double* sp, unsigned long long& mg, unsigned long long& mh) {
    const int step = m > 48 ? max(1, m / 24) : 1;
    auto Y = [&](int q) { return LDS ? sy[q] : T.y[gr ? gr[q] : g0 + q]; };
    auto Pd = [&](int q) { return LDS ? sp[q] : T.pred[gr ? gr[q] : g0 + q]; };
    for (int p = threadIdx.x; p < m; p += blockDim.x) {
        const int rp = gr ? gr[p] : g0 + p;
        const double wp = T.w ? T.w[rp] : 1.0;
        double g = 0.0, h = 1e-6;
        if (m >= 2) {
            auto pair = [&](int a, int b) {  // a < b, positions within the group
                const double ya = Y(a), yb = Y(b);
                if (ya == yb) return;
                const bool b_hi = yb > ya;
                const int hi = b_hi ? b : a, lo = b_hi ? a : b;
                const double s = Pd(hi) - Pd(lo);
                const double sc = s < -30.0 ? -30.0 : (s > 30.0 ? 30.0 : s);
                const double e = exp(sc);
                const double dyw = fmax(fabs(Y(hi) - Y(lo)), 1e-6);
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

__global__ __launch_bounds__(256) void k_grad_pairwise(Task* tasks) {
    Task& T = tasks[blockIdx.y];
    __shared__ double sy[PAIR_MAX_GROUP], sp[PAIR_MAX_GROUP];
    unsigned long long mg = 0, mh = 0;
    for (int grp = blockIdx.x; grp < T.n_groups; grp += gridDim.x) {
        const int g0 = T.group_start[grp], m = T.group_start[grp + 1] - g0;
        const int* gr = T.group_rows ? T.group_rows + g0 : nullptr;  // null: rows g0.. in order
        if (m > PAIR_MAX_GROUP) {
            pairwise_rows<false>(T, g0, m, gr, sy, sp, mg, mh);
            continue;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < m; k += blockDim.x) {
            const int r = gr ? gr[k] : g0 + k;
            sy[k] = T.y[r];
            sp[k] = T.pred[r];
        }
        __syncthreads();
        pairwise_rows<true>(T, g0, m, gr, sy, sp, mg, mh);
    }
    // rows outside every group keep g = 0, h = 1e-6 (set by k_grad_fill): include 1e-6 in max h
    const unsigned long long fill_h = (unsigned long long)__double_as_longlong(1e-6);
    mh = mh > fill_h ? mh : fill_h;
    block_max2_store(mg, mh, T);
    if (blockIdx.x == 0 && threadIdx.x == 0) T.n_red = gridDim.x;
}

// Rows outside any group keep g = 0, h = 1e-6 (objective.cpp:180-181).
__global__ void k_grad_fill(Task* tasks) {
    Task& T = tasks[blockIdx.y];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < T.n_rows; i += gridDim.x * blockDim.x) {
        T.g[i] = 0.0;
        T.h[i] = 1e-6;
    }
}

void launch_grad(hipStream_t st, Task* tasks, int n_tasks, int objective, int max_rows,
                 int max_groups) {
    if (objective == 5) {
        hipLaunchKernelGGL(k_grad_fill, dim3(std::max(1, std::min(1024, (max_rows + 1023) / 1024)), n_tasks),
                           dim3(256), 0, st, tasks);
        hipLaunchKernelGGL(k_grad_pairwise, dim3(std::max(1, std::min(max_groups, MAX_RED_BLOCKS)), n_tasks),
                           dim3(256), 0, st, tasks);
    } else {
        hipLaunchKernelGGL(k_grad_mse, dim3(std::max(1, std::min(1024, (max_rows + 1023) / 1024)), n_tasks),
                           dim3(256), 0, st, tasks);
    }
}

// =================================================================== tree start
// Quantization exponents from the gradient maxima, root node over the sample.
__global__ __launch_bounds__(256) void k_tree_init(Task* tasks) {
    Task& T = tasks[blockIdx.x];
    unsigned long long a = 0, b = 0;
    for (int k = threadIdx.x; k < T.n_red; k += blockDim.x) {
        a = T.red[k] > a ? T.red[k] : a;
        b = T.red[MAX_RED_BLOCKS + k] > b ? T.red[MAX_RED_BLOCKS + k] : b;
    }
    __shared__ unsigned long long sa[4], sb[4];
    a = wave_max_u64(a);
    b = wave_max_u64(b);
    if ((threadIdx.x & 63) == 0) {
        sa[threadIdx.x >> 6] = a;
        sb[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int k = 1; k < 4; ++k) {
        a = sa[k] > a ? sa[k] : a;
        b = sb[k] > b ? sb[k] : b;
    }
    T.max_g_bits = a;
    T.max_h_bits = b;
    int eg = mag_exponent(a);
    int eh = mag_exponent(b);
    if (T.goss && T.goss_scale > 1.0) {
        int es;
        frexp(T.goss_scale, &es);
        eg += es;
        eh += es;
    }
    // bits per row shrink with the row count, so no int64 sum over the sample (histograms,
    // energy) or over all rows (refit, TORB) can wrap
    const int ns = max(T.n_sample, 1), nr = max(T.n_rows, 1);
    T.s_g = sum_bits(QG_BITS, ns) - eg;
    T.s_h = sum_bits(QH_BITS, ns) - eh;
    T.s_r = sum_bits(QR_BITS, nr) - mag_exponent(a);
    T.s_rh = sum_bits(QR_BITS, nr) - mag_exponent(b);
    T.s_e = sum_bits(QE_BITS, ns) - 2 * eg;
    T.n_nodes = 1;
    T.done = 0;
    T.err = 0;
    T.cur_nl = -1;
    T.pending0 = 0;
    T.pending1 = -1;
    T.n_slots = 0;
    T.spec_steps = 0;
    T.n_tnodes = 0;
    T.skip_end = 0;
    T.rs_valid = 0;
    T.spec_gb = 0;
    T.end_done = 0;
    T.cur_leaf = -1;
    T.scan_ticket = 0;
    T.part_lr = 0;
    DNode& r = T.nodes[0];
    for (int k = 0; k < 10; ++k) r.v[k] = -1;
    r.v[N_LEAF] = 1;
    r.v[N_TYPE] = 0;
    r.v[N_MGL] = 1;
    r.v[N_NS] = T.n_sample;
    r.v[N_DEPTH] = 0;
    r.seg_start = 0;
    r.seg_count = T.n_sample;
    r.buf = 0;
    r.value = 0.0;
    r.gain = 0.0;
}

// Packed per-row histogram operands: pk = 2^CNT_SHIFT + rint(g_use 2^s_g), qh = rint(h_use 2^s_h).
__global__ __launch_bounds__(256) void k_pack(Task* tasks) {
    const Task T = tasks[blockIdx.y];   // a copy: the stores below cannot alias its fields
    const int sg = T.s_g, sh = T.s_h;
    const bool gs = T.goss && T.scaled;
    const double sc = T.goss_scale;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < T.n_rows; i += gridDim.x * blockDim.x) {
        double g = T.g[i], h = T.hconst ? 1.0 : T.h[i];
        if (gs && T.scaled[i]) {
            g *= sc;
            h *= sc;
        }
        const long long q = __double2ll_rn(ldexp(g, sg));
        T.pk[i] = CNT_ONE + (unsigned long long)q;
        if (!T.hconst) T.qh[i] = (unsigned long long)__double2ll_rn(ldexp(h, sh));
        if (T.contiguous && i < T.n_sample) T.idx[i] = i;
    }
}

void launch_tree_start(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, int max_rows,
                       int max_nodes) {
    const int rb = std::max(1, std::min(2048, (max_rows + 1023) / 1024));
    hipLaunchKernelGGL(k_tree_init, dim3(n_tasks), dim3(256), 0, st, tasks);
    hipLaunchKernelGGL(k_pack, dim3(rb, n_tasks), dim3(256), 0, st, tasks);
    (void)S;
    (void)max_nodes;
}

// =================================================================== histogram build
__device__ void apply_nodes(const Task& T, Task* Tw, int nl);
__device__ void apply_nodes_d(const Task& T, Task* Tw, int leaf, int type, int feat, int feat2, int thr, int mgl,
                              double gain, int L, int start, int count, int buf, int nl, bool set_count);

// Which node a histogram launch fills: the root (node 0) or the smaller child of the current
// split (reference tree.cpp:418-428: left when n_left <= n_right), with its row range.  The
// children's ranges follow from the partition's left count; block (0,0) of the child launch
// records them in the node table and the step descriptor for the scan and the tree end.
__device__ inline bool hist_target(const Task& T, Task* Tw, bool root, bool leader, int& node, int& start,
                                   int& count, int& buf) {
    if (root) {
        node = 0;
        start = 0;
        count = T.n_sample;
        buf = 0;
        return true;
    }
    if (T.done) return false;
    const int nl = (int)(T.part_lr & 0xffffffffull), cnt = T.cur_count, st = T.cur_start;
    const int L = T.cur_left, R = T.cur_left + 1;
    const bool left_small = nl <= cnt - nl;
    node = left_small ? L : R;
    start = left_small ? st : st + nl;
    count = left_small ? nl : cnt - nl;
    buf = 1 - T.cur_buf;
    if (leader && threadIdx.x == 0) {
        // the partition must send exactly the rows the chosen split's histogram counts left
        if (T.cur_nl >= 0 && T.cur_nl != nl) Tw->err = 1;
        apply_nodes(T, Tw, nl);
        Tw->n_left = nl;
        Tw->small_node = node;
        Tw->big_node = left_small ? R : L;
        Tw->small_start = start;
        Tw->small_count = count;
        Tw->small_buf = buf;
    }
    return true;
}

// Multi-leaf step: block x of the launch builds row block rb of slot j's smaller child, where the
// blocks of slot j are blk0[j] .. blk0[j] + gx[j] - 1.  The slot's first block applies the node
// bookkeeping of its split (the selection assigned the children's ids and the node count) and
// records the smaller child and the partials' form for the scan.
__device__ inline bool spec_target(const Task& T, Task* Tw, const Slot* slots, int bx, bool by0, int& sj, int& rb,
                                   int& node, int& start, int& count, int& buf, int& per) {
    if (T.done) return false;
    sj = -1;
    for (int q = 0; q < T.n_slots; ++q)
        if (bx >= slots[q].blk0) sj = q;
    if (sj < 0) return false;
    const Slot& sl = slots[sj];
    Slot& slw = T.slots[sj];   // the slot in global memory, for what the scan reads
    rb = bx - sl.blk0;
    if (rb >= sl.gx) return false;
    const int nl = (int)(sl.part_lr & 0xffffffffull), cnt = sl.count, st = sl.start;
    const int L = sl.left, R = sl.left + 1;
    const bool left_small = nl <= cnt - nl;
    node = left_small ? L : R;
    start = left_small ? st : st + nl;
    count = left_small ? nl : cnt - nl;
    buf = 1 - sl.buf;
    per = sl.per;
    if (rb == 0 && by0 && threadIdx.x == 0) {
        // the partition must send exactly the rows the chosen split's histogram counts left
        if (sl.nl != nl) Tw->err = 1;
        apply_nodes_d(T, Tw, sl.leaf, sl.type, sl.feat, sl.feat2, sl.thr, sl.mgl, sl.gain, L, st, cnt, sl.buf, nl,
                      false);
        slw.small_node = node;
        slw.big_node = left_small ? R : L;
        slw.small_start = start;
        slw.small_count = count;
        slw.small_buf = buf;
        slw.hist_parts = (count + per - 1) / per;
        slw.hist_wide = per > WORD_ROWS ? 1 : 0;
    }
    return true;
}

// Store a block's LDS histogram as partial number `slot` of the launch (plain stores; the scan
// sums the partials exactly, so no global atomics are needed and nothing has to be zeroed).
// lA holds KC packed copies per (feature, bin) in consecutive words, summed here (a block of at
// most WORD_ROWS rows: their sum is a valid packed word); lB the h sums when h varies.
template <bool HV, int KC>
__device__ inline void hist_store_partial(const Task& T, const unsigned long long* lA, const unsigned long long* lB,
                                          int FB, int f0, int F, int slot) {
    const size_t base = (size_t)slot * F * NBIN + (size_t)f0 * NBIN;
    const int words = min(FB, F - f0) * NBIN;
    auto part = gbl(T.part);
    auto part_h = gbl(T.part_h);
    for (int j = threadIdx.x; j < words; j += blockDim.x) {
        unsigned long long v = 0;
#pragma unroll
        for (int c = 0; c < KC; ++c) v += lA[j * KC + c];
        part[base + j] = v;
        if (HV) part_h[base + j] = lB[j];
    }
}

// Row kernel (partitioned rows, or a sampled root): each lane loads 16 consecutive feature bins
// of one row (128-bit load); LPR lanes cover the block's FB = 16*LPR features of a row.  The
// block of NT threads walks its rows in steps of RPL rows per lane (8 at 512 threads, 4 at 1024):
// the index loads of those rows, then their bin and gradient loads, then the LDS atomics of the
// packed operands, so one step costs two memory round trips.  The block's row range is sized from
// the node's actual row count, so small nodes still spread over many blocks.
//
// The gathers bound this kernel: a wave's 128-bit loads touch 64 / LPR rows' 128-byte lines
// per instruction, and every feature block of a row block gathers the same rows again, so the
// cost per row falls as LPR grows (1M gathered rows of 128 features, loads alone: 170 us at
// LPR = 1, 86 us at 2, 44 us at 4, 29 us at 8).  LPR = 4 holds 64 features x 256 bins of one
// copy in 128 KB of LDS (one 1024-thread block per CU).
//
// The histogram has KC copies per (feature, bin) in consecutive words; thread t adds into copy
// (t / LPR) % KC, so the rows of a round alternate between the copies of every feature and a
// wave's atomics spread over the LDS banks.  A packed word may take at most WORD_ROWS rows.  A
// block of at most that many rows stores the sum of its copies as its partial.  A larger block
// (the "wide" form, the same choice in every block of a launch) adds a step's rows in rounds of
// KC * WORD_ROWS rows (WORD_ROWS per copy); after each round every thread folds the copies of its
// bins into int64 g sums and counts held in registers and clears them (one LDS exchange per word
// reads and clears it), and the partial is stored from those registers (g sums in part, counts in
// part_c).  The h sums (h varying) never come near a word's limit and stay in LDS.
// LDS: KC*FB*NBIN words for lA, then (HV) FB*NBIN words for lB.
//
// One block's rows [rb * per, (rb + 1) * per) of a node's range (start, count in index buffer
// buf, or rows start + k directly when contig) for features f0 .. f0 + 16 * LPR - 1, stored as
// partial pidx.  Every thread of the block calls it (it synchronizes the block).
template <int LPR, bool HV, int KC, int NT>
__device__ __forceinline__ void hist_row_block(const Task& T, const Shared& S, unsigned long long* lds, int start,
                                               int count, int buf, int per, int rb, int pidx, int f0, bool contig) {
    constexpr int FB = 16 * LPR;
    constexpr int W = FB * NBIN;
    constexpr int RPL = NT >= 1024 ? 4 : 8;       // rows per lane per step
    constexpr int STRIDE = NT / LPR;              // rows per lane-row of a step
    constexpr int UPR = KC * WORD_ROWS / STRIDE;  // lane-rows per round (WORD_ROWS rows per copy)
    constexpr int OWN = W / NT;                   // bins each thread folds in the wide form
    static_assert(RPL % UPR == 0 && W % NT == 0 && NT % KC == 0, "rounds, copies and bin ownership");
    unsigned long long* lA = lds;
    unsigned long long* lB = lds + (size_t)W * KC;
    const bool wide = per > WORD_ROWS;
    const int r0 = rb * per;
    if (r0 >= count) return;
    const int r1 = min(count, r0 + per);
    for (int j = threadIdx.x; j < W * (KC + (HV ? 1 : 0)); j += NT) lds[j] = 0;
    const int sub = threadIdx.x % LPR;
    // copies alternate by row (lane-row), so each copy takes half of a round's rows of every feature
    const int cp = ((int)threadIdx.x / LPR) % KC;
    const int fcol = f0 + sub * 16;
    const bool active = fcol < S.fp;
    const auto rows = gbl((const int*)(buf ? T.idx_tmp : T.idx) + start);
    const auto pk = gbl((const unsigned long long*)T.pk);
    const auto qhp = gbl((const unsigned long long*)T.qh);
    // folded sums of the bins this thread owns: int64 g sums, and counts two to a register (a
    // block folds at most MAX_BLOCK_ROWS < 2^16 rows, so a count never carries into its neighbour)
    static_assert(MAX_BLOCK_ROWS < 65536, "16-bit folded counts");
    constexpr int ACW = (OWN + 1) / 2;
    long long aq[OWN];
    unsigned int ac[ACW];
#pragma unroll
    for (int o = 0; o < OWN; ++o) aq[o] = 0;
#pragma unroll
    for (int o = 0; o < ACW; ++o) ac[o] = 0;
    __syncthreads();
    // every thread takes the same number of steps (the rounds synchronize the block)
    for (int s0 = r0; s0 < r1; s0 += STRIDE * RPL) {
        const int k = s0 + (int)threadIdx.x / LPR;
        int i[RPL];
#pragma unroll
        for (int u = 0; u < RPL; ++u) {
            const int kk = k + u * STRIDE;
            i[u] = (kk < r1 && active) ? (contig ? kk : rows[kk]) : -1;
        }
        uint4 b[RPL];
        unsigned long long p[RPL], q[RPL];
#pragma unroll
        for (int u = 0; u < RPL; ++u) {
            const int ii = i[u] < 0 ? 0 : i[u];
            b[u] = active ? *(const uint4*)(S.bins_row + (size_t)(T.row0 + ii) * S.fp + fcol) : make_uint4(0, 0, 0, 0);
            p[u] = pk[ii];
            q[u] = HV ? qhp[ii] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < RPL; ++u) {
            if (i[u] >= 0) {
                const unsigned int w4[4] = {b[u].x, b[u].y, b[u].z, b[u].w};
#pragma unroll
                for (int t = 0; t < 16; ++t) {
                    const int bin = (w4[t >> 2] >> ((t & 3) * 8)) & 255;
                    const int slot = (sub * 16 + t) * NBIN + bin;
                    atomicAdd(&lA[slot * KC + cp], p[u]);
                    if (HV) atomicAdd(&lB[slot], q[u]);
                }
            }
            if (wide && (u % UPR) == UPR - 1) {
                // a round is in (WORD_ROWS rows per copy): fold the copies into registers; thread
                // w reads its bins' copies starting at copy (w / 8) % KC, so the lanes of an LDS
                // cycle touch different banks
                __syncthreads();
#pragma unroll
                for (int o = 0; o < OWN; ++o) {
                    const int j = (threadIdx.x + o * NT) * KC;
#pragma unroll
                    for (int c = 0; c < KC; ++c) {
                        const int cc = (c + (int)threadIdx.x / 8) % KC;
                        const unsigned long long v = atomicExch(&lA[j + cc], 0ull);
                        aq[o] += unpack_q(v);
                        ac[o >> 1] += unpack_c(v) << (16 * (o & 1));
                    }
                }
                __syncthreads();
            }
        }
    }
    __syncthreads();
    if (!wide) {
        hist_store_partial<HV, KC>(T, lA, lB, FB, f0, S.n_features, pidx);
        return;
    }
    const size_t base = (size_t)pidx * S.n_features * NBIN + (size_t)f0 * NBIN;
    const int words = min(FB, S.n_features - f0) * NBIN;
    auto part = gbl(T.part);
    auto part_c = gbl(T.part_c);
    auto part_h = gbl(T.part_h);
#pragma unroll
    for (int o = 0; o < OWN; ++o) {
        const int j = threadIdx.x + o * NT;
        if (j < words) {
            part[base + j] = (unsigned long long)aq[o];
            part_c[base + j] = (ac[o >> 1] >> (16 * (o & 1))) & 0xffffu;
            if (HV) part_h[base + j] = lB[j];
        }
    }
}

// Exclusive prefix over v[0 .. n) into pre[0 .. n] (pre[n] = total) by a block of NT threads
// (512 or 1024), 1024 / NT consecutive entries per thread (n <= 1024); v is read by each thread
// for its own entries only.
template <int NT, class GetV>
__device__ inline void block_prefix_1024(int n, GetV v, int* pre) {
    constexpr int PER = 1024 / NT, NW = NT / 64;
    static_assert(PER * NT == 1024, "entries per thread");
    __shared__ int wtot[NW];
    const int t = threadIdx.x, wave = t >> 6;
    int e[PER], sum = 0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        e[u] = PER * t + u < n ? v(PER * t + u) : 0;
        sum += e[u];
    }
    const int incl = wave_incl_i(sum);
    if ((t & 63) == 63) wtot[wave] = incl;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; ++w) base += wtot[w];
    int ex = base + incl - sum;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        if (PER * t + u < n) pre[PER * t + u] = ex;
        ex += e[u];
    }
    if (t == 0) {
        int tot = 0;
        for (int w = 0; w < NW; ++w) tot += wtot[w];
        pre[n] = tot;
    }
    __syncthreads();
}

// root: 0 the smaller child of T.cur_*, 1 a sampled or whole-window root (grid: row blocks x
// feature blocks x tasks).
template <int LPR, bool HV, int KC>
__global__ __launch_bounds__(512) void k_hist_row(Task* tasks, Shared S, int root, int min_rows) {
    constexpr int FB = 16 * LPR;
    extern __shared__ unsigned long long lds[];
    const Task T = tasks[blockIdx.z];
    int node, start, count, buf;
    const bool leader = blockIdx.x == 0 && blockIdx.y == 0;
    if (!hist_target(T, tasks + blockIdx.z, root != 0, leader, node, start, count, buf)) return;
    (void)node;
    int per = (count + gridDim.x - 1) / gridDim.x;
    per = max(per, min_rows);
    per = min((per + 63) & ~63, MAX_BLOCK_ROWS);
    if (leader && threadIdx.x == 0) {
        tasks[blockIdx.z].hist_parts = (count + per - 1) / per;
        tasks[blockIdx.z].hist_wide = per > WORD_ROWS ? 1 : 0;
    }
    // the whole-window root reads rows directly
    hist_row_block<LPR, HV, KC, 512>(T, S, lds, start, count, buf, per, blockIdx.x, blockIdx.x, blockIdx.y * FB,
                                     T.contiguous && root == 1);
}

// Child builds of a multi-leaf step of n_tasks tasks: a fixed grid of blocks takes the (task, row
// block, feature block) items of all tasks.  h == 1: 1024 threads (one block per CU), 32 features
// per item (2 lanes per row), two copies per bin (128 KB of LDS).  Wider rows per item gather
// fewer cache lines per row: 1M x 128 fit, child builds 0.242 ms per tree against 0.271 for 16
// features per item in two 512-thread blocks per CU, and 0.263 with 64-feature items (4 lanes
// per row, one copy per bin: its LDS atomics collide more on the panel's bin distributions).
// h varies: 512 threads, 16 features per item with g and h planes (64 KB).  Each task's items are
// numbered row block * feature blocks + feature block within groups of 8 row blocks, and each
// task's first item is a multiple of 8, so row block gb's feature blocks all have item % 8 ==
// gb % 8; blocks go out round-robin over the 8 XCDs, and the blocks of XCD x take exactly the
// items with item % 8 == x: the 128-byte rows a row block's feature blocks share come into that
// XCD's L2 once.
constexpr int SPEC_ITEM_ROWS = 8192;   // rows of one multi-leaf child-build item at most

template <bool HV>
__global__ __launch_bounds__(HV ? 512 : 1024) void k_hist_spec(Task* tasks, Shared S, int n_tasks) {
    constexpr int NT = HV ? 512 : 1024, FB = HV ? 16 : 32;
    extern __shared__ unsigned long long lds[];
    const int fbs = (S.fp + FB - 1) / FB;
    auto task_items = [&](const Task& Ti) { return Ti.done ? 0 : (Ti.spec_gb + 7) / 8 * 8 * fbs; };
    __shared__ int ipre[1025];
    if (n_tasks == 1) {   // one task: no prefix to form
        if (threadIdx.x == 0) {
            ipre[0] = 0;
            ipre[1] = task_items(tasks[0]);
        }
        __syncthreads();
    } else {
        block_prefix_1024<NT>(n_tasks, [&](int i) { return task_items(tasks[i]); }, ipre);
    }
    const int items = ipre[n_tasks];
    __shared__ Slot slots[SPEC_W];
    __shared__ int s_item;
    int cur = -1;
    // Items i with i % 8 == x go to the blocks of XCD x (block id % 8): block b takes item b
    // first; items past the grid go out in turn from XCD x's counter (when there are any), so
    // large and small items balance within each XCD, and every block stops once its counter
    // passes the last item.
    const int xq = blockIdx.x % 8;
    for (int item = blockIdx.x; item < items;) {
        int lo = 0, hi = n_tasks - 1;   // the task: the last with ipre[t] <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ipre[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        const int tk = lo;
        const Task T = tasks[tk];
        if (tk != cur) {   // the task's slot table, into LDS in one round trip
            const int nw = (int)(sizeof(Slot) / 4) * T.n_slots;
            if ((int)threadIdx.x < nw) ((int*)slots)[threadIdx.x] = ((const int*)T.slots)[threadIdx.x];
            __syncthreads();
            cur = tk;
        }
        const int li = item - ipre[tk];
        const int x = li % 8, j = li / 8;
        const int fb = j % fbs, gb = (j / fbs) * 8 + x;
        int sj, rb, node, start, count, buf, per;
        if (gb < T.spec_gb && spec_target(T, tasks + tk, slots, gb, fb == 0, sj, rb, node, start, count, buf, per)) {
            if (HV)
                hist_row_block<1, true, 1, NT>(T, S, lds, start, count, buf, per, rb, slots[sj].blk0 + rb, fb * FB,
                                               false);
            else
                hist_row_block<2, false, 2, NT>(T, S, lds, start, count, buf, per, rb, slots[sj].blk0 + rb, fb * FB,
                                                false);
        }
        __syncthreads();   // the next item clears the LDS this one read (and may reload the slots)
        if (items <= (int)gridDim.x) break;
        if (threadIdx.x == 0) s_item = 8 * ((int)gridDim.x / 8 + atomicAdd(&S.spec_ctr[xq], 1)) + xq;
        __syncthreads();
        item = s_item;
        __syncthreads();
    }
}

// Root build over the contiguous rows, one block per 16384 rows x a range of features.  Each
// lane keeps the packed gradients of its 16 rows in registers and walks the block's features in
// passes of FB = 64 / K features (32 / K when h varies): per pass it adds into a K-copy LDS
// histogram (lane l updates copy l % K, and the K copies of a bin sit in consecutive words, so a
// wave's adds spread over the LDS banks whatever their bins) while the next pass's bin columns
// load.  A copy takes 1024 rows per pass (64 lanes x 16 rows), within a packed word's WORD_ROWS.
// After a pass each thread unpacks the copies of one bin into an int64 g sum and a count and
// clears them for the next pass (a copy's 1024 rows of one pass fit a packed word; the 16384 rows
// of all copies would not, so each copy is unpacked on its own).  The block's partial is wide
// (g sums and counts in two planes), or, once the task's root counts are cached, raw (below).
// Every lane takes the same path through the passes, so the next pass's loads stay in flight
// during this pass's adds.
constexpr int ROOT_RPL = 16;
constexpr int ROOT_THREADS = 1024;
constexpr int ROOT_ROWS = ROOT_THREADS * ROOT_RPL;

// Bins of 16 rows starting at `row` for features f .. f+FB-1, clamped to the block's last feature
// fl1 - 1 (so every call issues FB loads; the passes' load counts match and the adds of one pass
// need not wait for the next pass's loads).  `row` is always a readable row: the bin columns hold
// at least 64 rows past the last row.
template <int FB>
__device__ inline void root_load_bins(uint4 (&b)[FB], const unsigned char* colb, size_t n_pad, int f, int fl1,
                                      int row) {
#pragma unroll
    for (int fl = 0; fl < FB; ++fl) b[fl] = *(const uint4*)(colb + (size_t)min(f + fl, fl1 - 1) * n_pad + row);
}

// The passes of one root block (k_hist_root); RAW: packed words added modulo 2^64 (h == 1, the
// task's root counts cached).  Rows are counted from the 16-row boundary at or below the task's
// first row: the first `head` of the `count` rows (and any past the last) add zero words.
template <int K, bool HV, bool RAW>
__device__ __forceinline__ void hist_root_block(const Task& T, const Shared& S, unsigned long long* lds, int count,
                                                int head, int bx, int r0, int fr0, int fr1) {
    constexpr int FB = (HV ? 32 : 64) / K;
    constexpr int WORDS = FB * NBIN;
    constexpr int SPT = WORDS * (HV ? 2 : 1) / ROOT_THREADS;
    unsigned long long* lA = lds;
    unsigned long long* lB = lds + WORDS * K;
    for (int j = threadIdx.x; j < WORDS * K * (HV ? 2 : 1); j += ROOT_THREADS) lds[j] = 0;
    const int lane = threadIdx.x & 63;
    const int rep = lane % K;
    const int row = r0 + threadIdx.x * ROOT_RPL;
    const bool full = head == 0 && row + ROOT_RPL <= count;
    // lanes wholly past the end read the last row group's bins and add zeros; a lane across the
    // end reads past it (inside the column padding) and adds zeros for those rows
    const int ld_row = row < count ? row : (count - 1) & ~(ROOT_RPL - 1);
    const unsigned char* colb = S.bins_col + (T.row0 - head);   // 16-byte aligned column loads
    unsigned long long pg[ROOT_RPL], ph[HV ? ROOT_RPL : 1];
    if (full) {
        const uint4* pp = (const uint4*)(T.pk + row);
#pragma unroll
        for (int t = 0; t < ROOT_RPL / 2; ++t) {
            const uint4 v = ldg(pp + t);
            pg[2 * t] = (unsigned long long)v.y << 32 | v.x;
            pg[2 * t + 1] = (unsigned long long)v.w << 32 | v.z;
        }
        if (HV) {
            const uint4* qq = (const uint4*)(T.qh + row);
#pragma unroll
            for (int t = 0; t < ROOT_RPL / 2; ++t) {
                const uint4 v = ldg(qq + t);
                ph[2 * t] = (unsigned long long)v.y << 32 | v.x;
                ph[2 * t + 1] = (unsigned long long)v.w << 32 | v.z;
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < ROOT_RPL; ++t) {
            const bool in = row + t >= head && row + t < count;
            pg[t] = in ? gbl(T.pk)[row + t - head] : 0ull;
            if (HV) ph[t] = in ? gbl(T.qh)[row + t - head] : 0ull;
        }
    }
    uint4 bc[FB], bn[FB];
    root_load_bins<FB>(bc, colb, S.n_pad, fr0, fr1, ld_row);
    // Sums: thread t owns word w = t of the pass: bin w of the g copies when w < WORDS, else bin
    // w - WORDS of the h copies.  Stores go through a global-segment
    // pointer, so they count only against the vector memory counter.  A pass's sums are stored at
    // the start of the next pass, ahead of that pass's loads: the wait for those loads then never
    // waits on a store issued after them.
    static_assert(SPT == 1, "one word per thread and pass");
    const auto outA = gbl(T.part);
    const auto outC = gbl(T.part_c);
    const auto outB = gbl(T.part_h);
    const size_t pbase = (size_t)bx * S.n_features * NBIN;
    const int w = (int)threadIdx.x;
    const bool gword = w < WORDS;
    const int ws = gword ? w : w - WORDS;   // bin of the pass
    const int pr0 = (w / 2) % (K / 2);
    long long pend_q = 0;
    unsigned int pend_c = 0;
    int pend_f = -1;
    unsigned long long raw_prev = 0;   // raw builds: the previous pass's copy sum of this bin
    __syncthreads();
    auto flush = [&]() {
        if (pend_f < 0) return;
        const int nw = min(FB, fr1 - pend_f) * NBIN;
        if (ws >= nw) return;
        const size_t o = pbase + (size_t)pend_f * NBIN + ws;
        if (gword) {
            outA[o] = (unsigned long long)pend_q;
            if (!RAW) outC[o] = pend_c;
        } else {
            outB[o] = (unsigned long long)pend_q;
        }
    };
    auto pass = [&](const uint4(&b)[FB], int f) {
        const int nf = min(FB, fr1 - f);
#pragma unroll
        for (int fl = 0; fl < FB; ++fl) {
            if (fl >= nf) continue;
            const unsigned w4[4] = {b[fl].x, b[fl].y, b[fl].z, b[fl].w};
#pragma unroll
            for (int t = 0; t < ROOT_RPL; ++t) {
                const int bin = (w4[t >> 2] >> ((t & 3) * 8)) & 255;
                atomicAdd(&lA[(fl * NBIN + bin) * K + rep], pg[t]);
                if (HV) atomicAdd(&lB[(fl * NBIN + bin) * K + rep], ph[t]);
            }
        }
        __syncthreads();
        // The bin's copies in pairs (16-byte LDS accesses), read, cleared and unpacked in batches
        // (all eight pairs at once when h == 1; one at a time when h varies, which holds more
        // registers).  Thread w starts at pair (w / 2) % 8, so the 16 lanes an LDS cycle serves
        // touch 16 different 16-byte bank groups.
        constexpr int BATCH = HV ? 2 : 8;
        long long q = 0;
        unsigned int c = 0;
        uint4* cp = (uint4*)(lds + (size_t)w * K);
        if (RAW) {
            unsigned long long s = 0;
#pragma unroll
            for (int j0 = 0; j0 < K / 2; j0 += 4) {
                uint4 v[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) v[j] = cp[(pr0 + j0 + j) % (K / 2)];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    s += ((unsigned long long)v[j].y << 32 | v[j].x) + ((unsigned long long)v[j].w << 32 | v[j].z);
            }
            pend_q = (long long)(s - raw_prev);
            raw_prev = s;
            pend_f = f;
            __syncthreads();
            return;
        }
#pragma unroll
        for (int j0 = 0; j0 < K / 2; j0 += BATCH) {
            uint4 v[BATCH];
#pragma unroll
            for (int j = 0; j < BATCH; ++j) v[j] = cp[(pr0 + j0 + j) % (K / 2)];
#pragma unroll
            for (int j = 0; j < BATCH; ++j) cp[(pr0 + j0 + j) % (K / 2)] = make_uint4(0, 0, 0, 0);
#pragma unroll
            for (int j = 0; j < BATCH; ++j) {
                const unsigned long long a = (unsigned long long)v[j].y << 32 | v[j].x;
                const unsigned long long b = (unsigned long long)v[j].w << 32 | v[j].z;
                if (gword) {
                    q += unpack_q(a) + unpack_q(b);
                    c += unpack_c(a) + unpack_c(b);
                } else {
                    q += (long long)(a + b);   // h copies hold plain sums
                }
            }
        }
        pend_q = q;
        pend_c = c;
        pend_f = f;
        __syncthreads();
    };
    // the first pass's bins have arrived before the loop (vmcnt 0): inside it the adds then wait
    // only for loads issued in the previous pass
    __builtin_amdgcn_s_waitcnt(0x0F70);
    int f = fr0;
    for (; f + FB < fr1; f += FB) {
        flush();
        root_load_bins<FB>(bn, colb, S.n_pad, f + FB, fr1, ld_row);
        pass(bc, f);
#pragma unroll
        for (int fl = 0; fl < FB; ++fl) bc[fl] = bn[fl];
    }
    flush();
    pass(bc, f);
    flush();
}

template <int K, bool HV>
__global__ __launch_bounds__(ROOT_THREADS) void k_hist_root(Task* tasks, Shared S, int feats_per_block,
                                                             int xcd_major) {
    static_assert(K == 16, "copies per bin: the sums' bank spread and the register budget assume 16");
    constexpr int FB = (HV ? 32 : 64) / K;        // 128 KB of LDS
    constexpr int WORDS = FB * NBIN;              // bins of one pass (per sum)
    static_assert(WORDS * (HV ? 2 : 1) == ROOT_THREADS, "one word per thread and pass");
    __shared__ unsigned long long lds[WORDS * K * (HV ? 2 : 1)];
    const Task T = tasks[blockIdx.z];
    const int count = T.n_sample;
    int bx = blockIdx.x, by = blockIdx.y;
    if (xcd_major) {
        // Blocks go out round-robin over the 8 XCDs (id % 8).  Each XCD gets a contiguous run of
        // (row range, feature range) pairs, feature range fastest, so the blocks that read one
        // row range's gradients share that XCD's L2.
        const int nb = gridDim.x * gridDim.y, id = blockIdx.x + blockIdx.y * gridDim.x;
        const int x = id % 8, l = id / 8;
        const int base = nb / 8, rem = nb % 8;
        const int j = x * base + min(x, rem) + l;
        bx = j / gridDim.y;
        by = j % gridDim.y;
    }
    // With h == 1 the per-bin counts of the task's window are the same in every tree: once the
    // scan has recorded them (root_cnt), the build adds the packed words modulo 2^64 instead
    // ("raw": sum = count * 2^CNT_SHIFT + q-sum, the q-sum recovered by the scan from the count).
    // A raw pass needs no clearing: each pass stores the difference of its bin's copy sums from
    // the previous pass's.
    const bool raw = !HV && T.cnt_n == count && T.cnt_row0 == T.row0;
    // rows from the 16-row boundary at or below row0: the column loads stay 16-byte aligned
    const int head = T.row0 & (ROOT_RPL - 1);
    const int vcount = count + head;
    if (bx == 0 && by == 0 && threadIdx.x == 0) {
        tasks[blockIdx.z].hist_parts = (vcount + ROOT_ROWS - 1) / ROOT_ROWS;
        tasks[blockIdx.z].hist_wide = HV ? 1 : raw ? 2 : 3;
    }
    const int r0 = bx * ROOT_ROWS;
    const int fr0 = by * feats_per_block;
    const int fr1 = min(S.n_features, fr0 + feats_per_block);
    if (r0 >= vcount || fr0 >= fr1) return;
    if (raw) hist_root_block<K, false, true>(T, S, lds, vcount, head, bx, r0, fr0, fr1);
    else hist_root_block<K, HV, false>(T, S, lds, vcount, head, bx, r0, fr0, fr1);
}

// Launch with optional start/stop events recorded at the kernel's own boundaries (AMD extended
// launch), so timed runs measure the kernel and not the dispatch gap around it.
#define TIMED_LAUNCH(kernel, grid, block, lds, st, e0, e1, ...)                                      \
    do {                                                                                            \
        if (e0) hipExtLaunchKernelGGL(kernel, grid, block, lds, st, e0, e1, 0, __VA_ARGS__);        \
        else hipLaunchKernelGGL(kernel, grid, block, lds, st, __VA_ARGS__);                         \
    } while (0)

// Child-build tiles of a histogram variant.  Row kernel: LPR lanes per row (16 features each),
// 512 threads.  Default LPR = 1 (16 features per block: more blocks per node at the same partial
// count) and 1024-row minimum tiles, the fastest setting over the node sizes of a 1M x 128,
// 31-leaf fit.  Variants >= 16 (low byte) pick LPR = 1 << ((v - 16) / 8) and min_rows =
// 256 << ((v - 16) % 8).  At most 32 features per block (the wide form holds 8 * LPR bins per
// thread in registers), 16 when h varies.  Blocks per feature block: no block folds more than
// MAX_BLOCK_ROWS rows, and there are at most pcap partials per build: 32 by default, since more
// blocks speed the build up less than the extra partials slow the scan's sums down (whole-tree
// sweep at 1M x 128: 32 -> 836 trees/s, 64 -> 787, 16 -> 745); bits 12-13 of the variant pick
// 32, 64, 16 or 8.  (Fewer blocks per task when many tasks share the launch was slower too:
// batched walk-forward 2.27 s at 4096 blocks per launch vs 1.88 s at 32 partials per task.)
void child_tiles(int variant, bool hconst, int& lpr, int& min_rows, int& pcap) {
    lpr = 1;
    min_rows = 1024;
    const int cv = variant & 0xff;
    if (cv >= 16) {
        lpr = 1 << std::min(2, (cv - 16) / 8);
        min_rows = 256 << std::min(7, (cv - 16) % 8);
    }
    if (!hconst) lpr = 1;
    if (lpr > 2) lpr = 2;
    constexpr int pcaps[4] = {32, 64, 16, 8};
    pcap = pcaps[(variant >> 12) & 3];
}

void launch_hist(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, bool root,
                 bool contiguous_root, bool hconst, int max_count, int variant, hipEvent_t e0,
                 hipEvent_t e1) {
    if (max_count <= 0) return;
    if (root && contiguous_root && (variant & 0xff) != 2) {
        // Root variants (variant >> 20; the child variants use bits 0-15), all k_hist_root with
        // 16384 rows per block.  0: features per block sized so the grid about fills the GPU once,
        // XCD-major placement.  Otherwise 16 << bits 7-8 features per block and bit 3 XCD-major
        // placement.
        const int rv = (variant >> 20) & 0x3ff;
        int fpb = 16 << ((rv >> 7) & 3);
        int xcd = (rv >> 3) & 1;
        if (rv == 0) {
            // smallest power of two in [4, 64] with at most ~one block per CU (256 CUs)
            const long long want = (long long)((max_count + ROOT_RPL - 1 + ROOT_ROWS - 1) / ROOT_ROWS) * n_tasks * S.n_features;
            fpb = 4;
            while (fpb < 64 && want > 256LL * fpb) fpb *= 2;
            xcd = 1;
        }
        // a window's rows count from the 16-row boundary at or below its first row (up to 15 more)
        const int vmax = max_count + ROOT_RPL - 1;
        dim3 grid((vmax + ROOT_ROWS - 1) / ROOT_ROWS, (S.n_features + fpb - 1) / fpb, n_tasks);
        if (hconst) TIMED_LAUNCH((k_hist_root<16, false>), grid, dim3(ROOT_THREADS), 0, st, e0, e1, tasks, S, fpb, xcd);
        else TIMED_LAUNCH((k_hist_root<16, true>), grid, dim3(ROOT_THREADS), 0, st, e0, e1, tasks, S, fpb, xcd);
        return;
    }
    int lpr, min_rows, pcap;
    child_tiles(variant, hconst, lpr, min_rows, pcap);
    const int FB = 16 * lpr;
    const int fbs = (S.fp + FB - 1) / FB;
    int gx = std::max(1, std::min(pcap, (max_count + min_rows - 1) / min_rows));
    gx = std::max(gx, (max_count + MAX_BLOCK_ROWS - 1) / MAX_BLOCK_ROWS);
    gx = std::min(gx, S.part_cap);
    dim3 grid(gx, fbs, n_tasks);
    // LDS copies per bin: 4 for one task with h == 1 and 16 features per block (the launch has at
    // most 32 x 8 blocks, one per CU: 128 KB costs no occupancy), 2 when several tasks share the
    // launch (two 512-thread blocks per CU are resident by registers anyway), 1 otherwise
    const int kc = (!hconst || lpr > 1) ? 1 : (n_tasks == 1 ? 4 : 2);
    const size_t lds = (size_t)FB * NBIN * 8 * (kc + (hconst ? 0 : 1));
    const int rt = root ? 1 : 0;
    if (!hconst) TIMED_LAUNCH((k_hist_row<1, true, 1>), grid, dim3(512), lds, st, e0, e1, tasks, S, rt, min_rows);
    else if (lpr > 1) TIMED_LAUNCH((k_hist_row<2, false, 1>), grid, dim3(512), lds, st, e0, e1, tasks, S, rt, min_rows);
    else if (kc == 4) TIMED_LAUNCH((k_hist_row<1, false, 4>), grid, dim3(512), lds, st, e0, e1, tasks, S, rt, min_rows);
    else TIMED_LAUNCH((k_hist_row<1, false, 2>), grid, dim3(512), lds, st, e0, e1, tasks, S, rt, min_rows);
}

// Child builds of a multi-leaf step: every slot's smaller child in one launch of a fixed grid whose
// blocks take the step's (row block, feature block) items in turn; row block rb of slot j stores
// partial blk0_j + rb (the slots' row blocks numbered in turn).  h == 1: one 1024-thread block per
// CU, items of 32 features; h varies: two 512-thread blocks per CU, items of 16 features.
void spec_child_shape(bool hconst, int n_cu, int fp, int& blocks, int& fbs) {
    blocks = ((hconst ? 1 : 2) * n_cu + 7) / 8 * 8;   // a multiple of the 8 XCDs
    fbs = hconst ? (fp + 31) / 32 : (fp + 15) / 16;
}

void launch_hist_spec(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, bool hconst, int blocks,
                      hipEvent_t e0, hipEvent_t e1) {
    dim3 grid(std::max(1, blocks), 1, 1);   // the tasks' items share the grid
    const size_t lds = (size_t)32 * NBIN * 8 * 2;   // 32 features x 2 copies, or 16 features x (g, h)
    if (!hconst) TIMED_LAUNCH((k_hist_spec<true>), grid, dim3(512), lds / 2, st, e0, e1, tasks, S, n_tasks);
    else TIMED_LAUNCH((k_hist_spec<false>), grid, dim3(1024), lds, st, e0, e1, tasks, S, n_tasks);
}

// =================================================================== split scan
// Best numeric split of every feature of the scanned node(s) (reference best_numeric_missing,
// histogram.cpp:144-193): totals over all bins, bin 0 = missing, thresholds b in [1, nb-2],
// empty bins skipped, missing-right (ml=0) before missing-left, first maximum kept.  One wave per
// feature, 4 bins per lane; prefix sums are exact int64, so the scan order cannot change them.
// Lane j holds bins 4j..4j+3 (bins >= nb must be zero).
template <bool hconst, int T0, int T1, int TP = T1 - T0>
__device__ inline void scan_feature_wave_t(const long long* qg, const int* c, const long long* qh,
                                           double inv_g, double inv_h, int nb, double lam, double gamma,
                                           int min_leaf, double min_hess, double l1, int mono, double& out_gain,
                                           int& out_thr, int& out_ml, int& out_cl, long long& totQ,
                                           long long& totQh, long long& totC) {
    const int lane = threadIdx.x & 63;
    // bin 0 (missing) sits in lane 0
    const long long missQ = (long long)read_lane_u64((unsigned long long)qg[0], 0);
    const long long missQh = hconst ? 0 : (long long)read_lane_u64((unsigned long long)qh[0], 0);
    const long long missC = __builtin_amdgcn_readlane(c[0], 0);
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
    const long long ig = wave_incl_ll(rg), ic = wave_incl_ll(rc), ih = hconst ? 0 : wave_incl_ll(rh);
    const long long eg = ig - rg, eh = ih - (hconst ? 0 : rh), ec = ic - rc;
    // totals over all bins: bins 1.. from lane 63's inclusive prefix, plus the missing bin
    totQ = (long long)read_lane_u64((unsigned long long)ig, 63) + missQ;
    totQh = hconst ? 0 : (long long)read_lane_u64((unsigned long long)ih, 63) + missQh;
    totC = (long long)read_lane_u64((unsigned long long)ic, 63) + missC;
    const double tG = (double)totQ * inv_g;
    const double tH = hconst ? (double)totC : (double)totQh * inv_h;
    const double parent = leaf_score(tG, tH, lam, l1);
    double best = -INFINITY;
    int bkey = 0x7fffffff;
    int bcl = 0;
    // The lane's candidates are evaluated as straight-line code, TP bins (2 * TP candidates) per
    // pass (an invalid one is computed and then ignored), so their divisions are independent and
    // overlap; they are then taken in key order with the same strict comparison, so the result
    // equals a sequential scan's.  Only bins lane*4+t with T0 <= t < T1 are candidates of this call.
    static_assert((T1 - T0) % TP == 0, "passes of TP bins");
#pragma unroll
    for (int p0 = T0; p0 < T1; p0 += TP) {
        double cg[2 * TP];
        int cclv[2 * TP];
        bool cok[2 * TP];
#pragma unroll
        for (int u = 0; u < TP; ++u) {
            const int t = p0 + u;
            const int b = lane * 4 + t;
            const bool bin_ok = (b >= 1) & (b <= nb - 2) & (c[t] != 0);
            const long long QL = eg + pg[t], QhL = eh + ph[t], CL = ec + pc[t];
#pragma unroll
            for (int ml = 0; ml <= 1; ++ml) {
                const long long ql = QL + (ml ? missQ : 0), qhl = QhL + (ml ? missQh : 0);
                const long long cl = CL + (ml ? missC : 0);
                const long long cr = totC - cl;
                const double gl = (double)ql * inv_g, gr = (double)(totQ - ql) * inv_g;
                const double hl = hconst ? (double)cl : (double)qhl * inv_h;
                const double hr = hconst ? (double)cr : (double)(totQh - qhl) * inv_h;
                bool ok = bin_ok & (cl >= min_leaf) & (cr >= min_leaf) & (hl >= min_hess) & (hr >= min_hess);
                if (mono != 0) {   // wave-uniform: the feature's monotone constraint (histogram.cpp:177-180)
                    const double wl = weight_of(gl, hl, lam, l1), wr = weight_of(gr, hr, lam, l1);
                    ok &= mono > 0 ? !(wl < wr) : !(wl > wr);
                }
                cok[u * 2 + ml] = ok;
                cg[u * 2 + ml] = 0.5 * (leaf_score(gl, hl, lam, l1) + leaf_score(gr, hr, lam, l1) - parent) - gamma;
                cclv[u * 2 + ml] = (int)cl;
            }
        }
#pragma unroll
        for (int j = 0; j < 2 * TP; ++j) {
            const int key = (lane * 4 + p0 + (j >> 1)) * 2 + (j & 1);
            const bool take = cok[j] & ((cg[j] > best) | ((cg[j] == best) & (key < bkey)));
            best = take ? cg[j] : best;
            bkey = take ? key : bkey;
            bcl = take ? cclv[j] : bcl;
        }
    }
    // lanes hold increasing keys, so the wave's first maximum is the best lane's best
    const int wl = wave_best_lane(best, bkey);
    best = __longlong_as_double((long long)read_lane_u64((unsigned long long)__double_as_longlong(best), wl));
    bkey = __builtin_amdgcn_readlane(bkey, wl);
    bcl = __builtin_amdgcn_readlane(bcl, wl);
    out_gain = best;
    out_thr = (bkey == 0x7fffffff) ? -1 : bkey >> 1;
    out_ml = (bkey == 0x7fffffff) ? 1 : (bkey & 1);
    out_cl = (bkey == 0x7fffffff) ? 0 : bcl;
}

__device__ inline void scan_feature_wave(const long long* qg, const int* c, const long long* qh, bool hconst,
                                         double inv_g, double inv_h, int nb, double lam, double gamma,
                                         int min_leaf, double min_hess, double l1, double& out_gain,
                                         int& out_thr, int& out_ml, int& out_cl, long long& totQ,
                                         long long& totQh, long long& totC) {
    // (PAIR splits: the reference scans the pair histogram without a monotone constraint,
    // tree.cpp:268-269)
    if (hconst)
        scan_feature_wave_t<true, 0, 4>(qg, c, qh, inv_g, inv_h, nb, lam, gamma, min_leaf, min_hess, l1, 0,
                                        out_gain, out_thr, out_ml, out_cl, totQ, totQh, totC);
    else
        scan_feature_wave_t<false, 0, 4>(qg, c, qh, inv_g, inv_h, nb, lam, gamma, min_leaf, min_hess, l1, 0,
                                         out_gain, out_thr, out_ml, out_cl, totQ, totQh, totC);
}
// Half of the candidates (bins lane*4 + {0,1} for half 0, lane*4 + {2,3} for half 1): two waves
// holding the same sums split one feature's scan.
__device__ inline void scan_feature_half(const long long* qg, const int* c, const long long* qh, bool hconst,
                                         int half, double inv_g, double inv_h, int nb, double lam, double gamma,
                                         int min_leaf, double min_hess, double l1, int mono, double& out_gain, int& out_thr,
                                         int& out_ml, int& out_cl, long long& totQ, long long& totQh,
                                         long long& totC) {
    if (hconst) {
        if (half == 0)
            scan_feature_wave_t<true, 0, 2>(qg, c, qh, inv_g, inv_h, nb, lam, gamma, min_leaf, min_hess, l1,
                                            mono, out_gain, out_thr, out_ml, out_cl, totQ, totQh, totC);
        else
            scan_feature_wave_t<true, 2, 4>(qg, c, qh, inv_g, inv_h, nb, lam, gamma, min_leaf, min_hess, l1,
                                            mono, out_gain, out_thr, out_ml, out_cl, totQ, totQh, totC);
    } else {
        if (half == 0)
            scan_feature_wave_t<false, 0, 2>(qg, c, qh, inv_g, inv_h, nb, lam, gamma, min_leaf, min_hess, l1,
                                             mono, out_gain, out_thr, out_ml, out_cl, totQ, totQh, totC);
        else
            scan_feature_wave_t<false, 2, 4>(qg, c, qh, inv_g, inv_h, nb, lam, gamma, min_leaf, min_hess, l1,
                                             mono, out_gain, out_thr, out_ml, out_cl, totQ, totQh, totC);
    }
}

// One split decision (the node to split and its predicate).
struct Decision {
    int leaf;              // -1: the tree stops growing
    int type, feat, feat2, thr, mgl;
    int nl;                // rows the split sends left (from the node's histogram; -1: unknown)
    int start, count, buf; // the leaf's rows: range in the index buffer `buf`
    double gain;
};

// Best-first selection of the next split (tree.cpp:343-392, 445-456), computed by a whole
// 256-thread block: reduce the last scan's per-feature results of the two pending nodes to their
// leaf candidates (tree.cpp:160-166: ascending features, skip invalid or gain < min_split_gain,
// first maximum kept), then pick the leaf with the best candidate (lowest node id on ties; the
// gain must exceed min_split_gain).  All inputs are loaded up front so their latencies overlap.
// Every block that calls it reaches the same decision; only a block with `writer` set stores the
// pending candidates in the candidate cache.
__device__ void select_decide(const Task& T, int F, const GrowP& P, bool writer, Decision& out) {
    __shared__ double sg[2][16];
    __shared__ int sk[2][16], st[2][16], sm[2][16], scl[2][16];
    __shared__ DCand pc[2];
    __shared__ double lg[16];
    __shared__ int lk[16];
    // every node's cached candidate and row range, so the final choice reads no global memory
    __shared__ DCand ncand[MAX_NODES];
    __shared__ int nstart[MAX_NODES], ncount[MAX_NODES], nbuf[MAX_NODES];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int pend[2] = {T.pending0, T.pending1};
    const int nn = T.n_nodes;
    // inputs of the leaf choice that do not depend on this step's scan (one node per thread;
    // the block has at least MAX_NODES threads)
    const int lid = threadIdx.x;
    DCand lc;
    lc.valid = 0;
    bool lleaf = false;
    if (lid < nn) {
        lc = ldg(T.cands + lid);
        const DNode nd = ldg(T.nodes + lid);
        lleaf = nd.v[N_LEAF] != 0;
        ncand[lid] = lc;
        nstart[lid] = nd.seg_start;
        ncount[lid] = nd.seg_count;
        nbuf[lid] = nd.buf;
    }
    int ndepth[2] = {0, 0}, nns[2] = {0, 0};
    if (threadIdx.x == 0) {
#pragma unroll
        for (int pi = 0; pi < 2; ++pi)
            if (pend[pi] >= 0) {
                ndepth[pi] = T.nodes[pend[pi]].v[N_DEPTH];
                nns[pi] = T.nodes[pend[pi]].v[N_NS];
            }
    }
    double best[2] = {-INFINITY, -INFINITY};
    int bk[2] = {0x7fffffff, 0x7fffffff}, bt[2] = {-1, -1}, bm[2] = {1, 1}, bcl[2] = {0, 0};
#pragma unroll
    for (int pi = 0; pi < 2; ++pi) {
        const int nd = pend[pi];
        if (nd < 0) continue;
        for (int f = threadIdx.x; f < F; f += blockDim.x) {
            const FCand c = ldg(T.fcand + (size_t)nd * F + f);
            if (c.thr < 0 || c.gain < P.gamma) continue;
            if (cand_better(c.gain, f, best[pi], bk[pi])) {
                best[pi] = c.gain;
                bk[pi] = f;
                bt[pi] = c.thr;
                bm[pi] = c.ml;
                bcl[pi] = c.cl;
            }
        }
    }
#pragma unroll
    for (int pi = 0; pi < 2; ++pi) {
        const int wl = wave_best_lane(best[pi], bk[pi]);
        best[pi] = __longlong_as_double((long long)read_lane_u64((unsigned long long)__double_as_longlong(best[pi]), wl));
        bk[pi] = __builtin_amdgcn_readlane(bk[pi], wl);
        bt[pi] = __builtin_amdgcn_readlane(bt[pi], wl);
        bm[pi] = __builtin_amdgcn_readlane(bm[pi], wl);
        bcl[pi] = __builtin_amdgcn_readlane(bcl[pi], wl);
        if (lane == 0) {
            sg[pi][wave] = best[pi];
            sk[pi][wave] = bk[pi];
            st[pi][wave] = bt[pi];
            sm[pi][wave] = bm[pi];
            scl[pi][wave] = bcl[pi];
        }
    }
    __syncthreads();
    if (wave == 0) {
        // the waves' results side by side: lanes 0-15 for pending node 0, lanes 16-31 for node 1
        const int pl = lane >> 4, wv = lane & 15;
        const bool have = pl < 2 && wv < nw;
        const double gv = have ? sg[pl][wv] : -INFINITY;
        const int kv = have ? sk[pl][wv] : 0x7fffffff;
        const int tv = have ? st[pl][wv] : -1, mv = have ? sm[pl][wv] : 1, cv = have ? scl[pl][wv] : 0;
#pragma unroll
        for (int pi = 0; pi < 2; ++pi) {
            const bool mine = pl == pi;
            const double gm = mine ? gv : -INFINITY;
            const int km = mine ? kv : 0x7fffffff;
            const int wl = wave_best_lane(gm, km);
            const double b = __longlong_as_double((long long)read_lane_u64((unsigned long long)__double_as_longlong(gm), wl));
            const int k = __builtin_amdgcn_readlane(km, wl);
            const int t = __builtin_amdgcn_readlane(mine ? tv : -1, wl);
            const int m = __builtin_amdgcn_readlane(mine ? mv : 1, wl);
            const int cl = __builtin_amdgcn_readlane(mine ? cv : 0, wl);
            const int nd = pend[pi];
            if (lane != 0 || nd < 0) continue;
            DCand c;
            const bool elig = ndepth[pi] < P.max_depth && nns[pi] >= 2 * P.min_leaf;
            c.valid = (k != 0x7fffffff) && elig;
            c.gain = b;
            c.feature = k == 0x7fffffff ? -1 : k;
            c.feature2 = -1;
            c.type = 0;
            c.thr = k == 0x7fffffff ? -1 : t;
            c.mgl = k == 0x7fffffff ? 1 : m;
            c.cl = k == 0x7fffffff ? 0 : cl;
            c.pad = 0;
            pc[pi] = c;
            if (writer) T.cands[nd] = c;
        }
    }
    __syncthreads();
    // leaf choice: strict '>' over ascending node ids == (gain desc, id asc)
    double lb = -INFINITY;
    int lbid = 0x7fffffff;
    if (lid < nn && lleaf) {
        const DCand c = lid == pend[0] ? pc[0] : (lid == pend[1] ? pc[1] : lc);
        if (c.valid && c.feature >= 0 && c.gain > P.gamma) {
            lb = c.gain;
            lbid = lid;
        }
    }
    {
        const int wl = wave_best_lane(lb, lbid);
        lb = __longlong_as_double((long long)read_lane_u64((unsigned long long)__double_as_longlong(lb), wl));
        lbid = __builtin_amdgcn_readlane(lbid, wl);
    }
    if (lane == 0) {
        lg[wave] = lb;
        lk[wave] = lbid;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nw; ++w)
            if (cand_better(lg[w], lk[w], lb, lbid)) {
                lb = lg[w];
                lbid = lk[w];
            }
        if (T.n_nodes + 2 > MAX_NODES || T.n_nodes + 2 > 2 * P.max_leaves - 1 || lbid == 0x7fffffff) {
            out.leaf = -1;
        } else {
            const DCand c = lbid == pend[0] ? pc[0] : (lbid == pend[1] ? pc[1] : ncand[lbid]);
            out.leaf = lbid;
            out.start = nstart[lbid];
            out.count = ncount[lbid];
            out.buf = nbuf[lbid];
            out.type = c.type;
            out.feat = c.feature;
            out.feat2 = c.feature2;
            out.thr = c.thr;
            out.mgl = c.mgl;
            out.nl = c.cl;
            out.gain = c.gain;
        }
    }
    __syncthreads();
}

// Node bookkeeping of the split in T.cur_* (tree.cpp:385-392, 445-456): the split node becomes
// internal, its children get ids L = cur_left and L + 1 with their row ranges; one thread.
__device__ void apply_nodes_d(const Task& T, Task* Tw, int leaf, int type, int feat, int feat2, int thr, int mgl,
                              double gain, int L, int start, int count, int buf, int nl, bool set_count) {
    const int R = L + 1;
    DNode& p = T.nodes[leaf];
    p.v[N_LEAF] = 0;
    p.v[N_TYPE] = type;
    p.v[N_FEAT] = feat;
    p.v[N_FEAT2] = feat2;
    p.v[N_THR] = thr;
    p.v[N_MGL] = mgl;
    p.v[N_LEFT] = L;
    p.v[N_RIGHT] = R;
    p.gain = gain;
    for (int side = 0; side < 2; ++side) {
        DNode& c = T.nodes[side == 0 ? L : R];
        for (int k = 0; k < 10; ++k) c.v[k] = -1;
        c.v[N_LEAF] = 1;
        c.v[N_TYPE] = 0;
        c.v[N_MGL] = 1;
        c.v[N_NS] = side == 0 ? nl : count - nl;
        c.v[N_DEPTH] = p.v[N_DEPTH] + 1;
        c.seg_start = side == 0 ? start : start + nl;
        c.seg_count = side == 0 ? nl : count - nl;
        c.buf = 1 - buf;
        c.value = 0.0;
        c.gain = 0.0;
        T.cands[side == 0 ? L : R].valid = 0;
    }
    if (set_count) Tw->n_nodes = L + 2;
}

__device__ void apply_nodes(const Task& T, Task* Tw, int nl) {
    apply_nodes_d(T, Tw, T.cur_leaf, T.cur_type, T.cur_feat, T.cur_feat2, T.cur_thr, T.cur_mgl, T.cur_gain,
                  T.cur_left, T.cur_start, T.cur_count, T.cur_buf, nl, true);
}

// The partial histograms of one histogram build: packed words or (wide) int64 g sums with
// counts in part_c, or (2) raw packed-word sums; h sums in part_h when h varies.
struct Parts {
    const unsigned long long* part;
    const unsigned int* part_c;
    const unsigned long long* part_h;
    int wide, hconst;
};

// Exact sums, over partials k = k0, k0 + dk, ... < np of one histogram build, of bins
// lane*4 .. lane*4+3 of feature f.  PR partials (4 or 2) are loaded per round so their latencies
// overlap.
template <int PR = 4>
__device__ inline void sum_partials(const Parts& T, int F, int f, int k0, int dk, int np, long long sq[4],
                                    long long sc[4], long long sh[4]) {
    const int lane = threadIdx.x & 63;
    const size_t stride = (size_t)F * NBIN;
    const size_t off = (size_t)f * NBIN + lane * 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        sq[t] = 0;
        sc[t] = 0;
        sh[t] = 0;
    }
    int k = k0;
    if (T.wide == 2) {
        // raw packed-word sums modulo 2^64 (the counts come from the task's root counts)
        unsigned long long uq[4] = {0, 0, 0, 0};
        for (; k + (PR - 1) * dk < np; k += PR * dk) {
            uint4 v[2 * PR];
#pragma unroll
            for (int u = 0; u < PR; ++u) {
                const uint4* a = (const uint4*)(T.part + (size_t)(k + u * dk) * stride + off);
                v[2 * u] = ldg(a);
                v[2 * u + 1] = ldg(a + 1);
            }
#pragma unroll
            for (int u = 0; u < 2 * PR; ++u) {
                const int t = (u & 1) * 2;
                uq[t] += (unsigned long long)v[u].y << 32 | v[u].x;
                uq[t + 1] += (unsigned long long)v[u].w << 32 | v[u].z;
            }
        }
        for (; k < np; k += dk) {
            const auto w = gbl((const unsigned long long*)T.part + (size_t)k * stride + off);
#pragma unroll
            for (int t = 0; t < 4; ++t) uq[t] += w[t];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) sq[t] = (long long)uq[t];
    } else if (T.wide) {
        for (; k + (PR - 1) * dk < np; k += PR * dk) {
            uint4 v[2 * PR], c[PR];
#pragma unroll
            for (int u = 0; u < PR; ++u) {
                const size_t o = (size_t)(k + u * dk) * stride + off;
                const uint4* a = (const uint4*)(T.part + o);
                v[2 * u] = ldg(a);
                v[2 * u + 1] = ldg(a + 1);
                c[u] = ldg((const uint4*)(T.part_c + o));
            }
#pragma unroll
            for (int u = 0; u < PR; ++u) {
                sq[0] += (long long)((unsigned long long)v[2 * u].y << 32 | v[2 * u].x);
                sq[1] += (long long)((unsigned long long)v[2 * u].w << 32 | v[2 * u].z);
                sq[2] += (long long)((unsigned long long)v[2 * u + 1].y << 32 | v[2 * u + 1].x);
                sq[3] += (long long)((unsigned long long)v[2 * u + 1].w << 32 | v[2 * u + 1].z);
                sc[0] += c[u].x;
                sc[1] += c[u].y;
                sc[2] += c[u].z;
                sc[3] += c[u].w;
            }
        }
        for (; k < np; k += dk) {
            const size_t o = (size_t)k * stride + off;
            const auto w = gbl((const long long*)T.part + o);
            const auto c = gbl((const unsigned int*)T.part_c + o);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                sq[t] += w[t];
                sc[t] += c[t];
            }
        }
    } else {
        for (; k + (PR - 1) * dk < np; k += PR * dk) {
            uint4 v[2 * PR];
#pragma unroll
            for (int u = 0; u < PR; ++u) {
                const uint4* a = (const uint4*)(T.part + (size_t)(k + u * dk) * stride + off);
                v[2 * u] = ldg(a);
                v[2 * u + 1] = ldg(a + 1);
            }
#pragma unroll
            for (int u = 0; u < 2 * PR; ++u) {
                const unsigned long long w0 = (unsigned long long)v[u].y << 32 | v[u].x;
                const unsigned long long w1 = (unsigned long long)v[u].w << 32 | v[u].z;
                const int t = (u & 1) * 2;
                sq[t] += unpack_q(w0);
                sc[t] += unpack_c(w0);
                sq[t + 1] += unpack_q(w1);
                sc[t + 1] += unpack_c(w1);
            }
        }
        for (; k < np; k += dk) {
            const auto w = gbl((const unsigned long long*)T.part + (size_t)k * stride + off);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                sq[t] += unpack_q(w[t]);
                sc[t] += unpack_c(w[t]);
            }
        }
    }
    if (!T.hconst) {
        for (int j = k0; j < np; j += dk) {
            const auto w = gbl((const unsigned long long*)T.part_h + (size_t)j * stride + off);
#pragma unroll
            for (int t = 0; t < 4; ++t) sh[t] += (long long)w[t];
        }
    }
}

// Best categorical split of one feature (reference best_categorical, histogram.cpp:195-244) for
// the node of every wave with `mine` set (one node per wave).  The bins holding rows, missing
// (bin 0) included, are ranked by G / (H + cat_smooth); equal ratios rank by bin index (the
// reference's std::sort is not stable, so it may order equal ratios differently).  The scan over
// that order uses exact int64 prefix sums; every position but the last is a candidate and the
// first maximum wins.  The left set is the bins ranked up to the best position.  Every wave of
// the block calls this (it synchronizes the block); kq/kh/kc are the block's [4][NBIN] LDS rows,
// whose contents every wave has consumed on entry (row `wave` is this wave's scratch).
__device__ void scan_cat(bool mine, const long long* qg, const int* c, const long long* qh, bool hconst,
                         double inv_g, double inv_h, const GrowP& P, long long (*kq)[NBIN], long long (*kh)[NBIN],
                         int (*kc)[NBIN], double& out_gain, int& out_pos, int& out_ml, int& out_cl,
                         unsigned long long (&mask)[4], long long& totQ, long long& totQh, long long& totC) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double sm = P.cat_smooth > 0.0 ? P.cat_smooth : 0.0;
    double key[4];
    int rank[4] = {0, 0, 0, 0};
    __syncthreads();
    if (mine) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const double g = (double)qg[t] * inv_g;
            const double h = hconst ? (double)c[t] : (double)qh[t] * inv_h;
            key[t] = g / (h + sm);
            kq[wave][lane * 4 + t] = __double_as_longlong(key[t]);
            kc[wave][lane * 4 + t] = c[t];
        }
    }
    __syncthreads();
    if (mine) {
        for (int j = 0; j < NBIN; ++j) {
            const double kj = __longlong_as_double(kq[wave][j]);
            const bool vj = kc[wave][j] > 0;
#pragma unroll
            for (int t = 0; t < 4; ++t)
                rank[t] += (vj && (kj < key[t] || (kj == key[t] && j < lane * 4 + t))) ? 1 : 0;
        }
    }
    __syncthreads();
    int m = 0;
    if (mine) {
        m = wave_sum_i((c[0] > 0) + (c[1] > 0) + (c[2] > 0) + (c[3] > 0));
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (c[t] > 0) {
                kq[wave][rank[t]] = qg[t];
                kh[wave][rank[t]] = hconst ? 0 : qh[t];
                kc[wave][rank[t]] = c[t];
            }
    }
    __syncthreads();
    out_gain = -INFINITY;
    out_pos = -1;
    out_ml = 1;
    out_cl = 0;
    mask[0] = mask[1] = mask[2] = mask[3] = 0;
    if (!mine) return;
    // exact prefix sums over ranked positions lane*4 .. lane*4+3
    long long pq[4], ph[4], pc[4];
    long long rq = 0, rh = 0, rc = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int p = lane * 4 + t;
        if (p < m) {
            rq += kq[wave][p];
            rh += kh[wave][p];
            rc += kc[wave][p];
        }
        pq[t] = rq;
        ph[t] = rh;
        pc[t] = rc;
    }
    const long long iq = wave_incl_ll(rq), ih = hconst ? 0 : wave_incl_ll(rh), ic = wave_incl_ll(rc);
    const long long eq = iq - rq, eh = ih - (hconst ? 0 : rh), ec = ic - rc;
    totQ = (long long)read_lane_u64((unsigned long long)iq, 63);
    totQh = hconst ? 0 : (long long)read_lane_u64((unsigned long long)ih, 63);
    totC = (long long)read_lane_u64((unsigned long long)ic, 63);
    const double tG = (double)totQ * inv_g;
    const double tH = hconst ? (double)totC : (double)totQh * inv_h;
    const double parent = leaf_score(tG, tH, P.l2, P.l1);
    double best = -INFINITY;
    int bkey = 0x7fffffff, bcl = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int p = lane * 4 + t;
        const long long ql = eq + pq[t], qhl = eh + ph[t], cl = ec + pc[t];
        const long long cr = totC - cl;
        const double gl = (double)ql * inv_g, gr = (double)(totQ - ql) * inv_g;
        const double hl = hconst ? (double)cl : (double)qhl * inv_h;
        const double hr = hconst ? (double)cr : (double)(totQh - qhl) * inv_h;
        const bool ok = (p + 1 < m) & (cl >= P.min_leaf) & (cr >= P.min_leaf) & (hl >= P.min_hess) & (hr >= P.min_hess);
        const double gain = 0.5 * (leaf_score(gl, hl, P.l2, P.l1) + leaf_score(gr, hr, P.l2, P.l1) - parent) - P.gamma;
        const bool take = ok & (gain > best);
        best = take ? gain : best;
        bkey = take ? p : bkey;
        bcl = take ? (int)cl : bcl;
    }
    const int wl = wave_best_lane(best, bkey);
    best = __longlong_as_double((long long)read_lane_u64((unsigned long long)__double_as_longlong(best), wl));
    bkey = __builtin_amdgcn_readlane(bkey, wl);
    bcl = __builtin_amdgcn_readlane(bcl, wl);
    if (bkey == 0x7fffffff) return;
    // left set: bin lane*4+t is bit (lane & 15)*4 + t of word lane >> 4
    unsigned long long bal[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) bal[t] = __ballot(c[t] > 0 && rank[t] <= bkey);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        unsigned long long word = 0;
        for (int j = 0; j < 16; ++j)
#pragma unroll
            for (int t = 0; t < 4; ++t) word |= ((bal[t] >> (16 * w + j)) & 1ull) << (4 * j + t);
        mask[w] = word;
    }
    out_gain = best;
    out_pos = bkey;
    out_cl = bcl;
    out_ml = (int)(mask[0] & 1ull);
}

// mode 0: the root (node 0) from the root launch's partials.  mode 1: both children of the
// current split: the smaller from the child launch's partials, the larger as parent - smaller
// (exact in fixed point; reference subtract_hist_inplace, histogram.cpp:108-117).  A block owns
// one feature: its four waves sum interleaved quarters of the partials, then waves 0 and 2 scan
// the smaller child and waves 1 and 3 the larger (half of the candidates each); wave 0 / 1 stores
// the child's histogram, kept for later use as a parent, and its best split of the feature.
// One scan of feature f: task `task`'s root (mode 0), the children of its current split (mode
// 1), or the children of slot sj of its multi-leaf step (mode 2).  Every thread of the block
// calls it; it synchronizes the block.
__device__ __forceinline__ void scan_one(Task* tasks, const Shared& S, const GrowP& P, int mode, int task,
                                         int f) {
    const Task T = tasks[task];
    Task& Tw = tasks[task];
    if (T.done) return;
    const int F = S.n_features;
    const size_t per = (size_t)F * NBIN;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // the build this scan reads: its parent, children and partials
    const int parent = T.cur_leaf, small_node = T.small_node, big_node = T.big_node;
    Parts pp;
    pp.part = T.part;
    pp.part_c = T.part_c;
    pp.part_h = T.part_h;
    pp.wide = T.hist_wide;
    pp.hconst = T.hconst;
    const int n_parts = T.hist_parts;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Tw.pending0 = mode == 0 ? 0 : T.cur_left;
        Tw.pending1 = mode == 0 ? -1 : T.cur_left + 1;
        Tw.part_lr = 0;   // the histogram launch has consumed it; the next partition counts anew
        if (mode == 0) {
            Tw.pend_first = 0;   // multi-leaf growth: the root's candidate is reduced next
            Tw.n_pend = 1;
        }
    }
    __shared__ long long rq[4][NBIN], rh[4][NBIN];
    __shared__ int rc[4][NBIN];
    // waves 1 and 3 of a child scan derive the larger child from the parent's sums: load them
    // now, so their latency overlaps the partial loads
    long long par_q[4] = {0, 0, 0, 0}, par_h[4] = {0, 0, 0, 0};
    int par_c[4] = {0, 0, 0, 0};
    if (mode != 0 && (wave & 1)) {
        const size_t po = (size_t)parent * per + (size_t)f * NBIN + lane * 4;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            par_q[t] = gbl(T.hg)[po + t];
            par_c[t] = gbl(T.hc)[po + t];
            if (!T.hconst) par_h[t] = gbl(T.hh)[po + t];
        }
    }
    {
        long long sq[4], sc[4], sh[4];
        sum_partials(pp, F, f, wave, 4, n_parts, sq, sc, sh);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            rq[wave][lane * 4 + t] = sq[t];
            rc[wave][lane * 4 + t] = (int)sc[t];
            rh[wave][lane * 4 + t] = sh[t];
        }
    }
    __syncthreads();
    // waves 0 and 2 scan the smaller child (the root in mode 0), waves 1 and 3 the larger; the
    // two waves of a node hold the same sums and each evaluates half of its candidates
    const int side = wave & 1, half = wave >> 1;
    const bool active = mode != 0 || side == 0;
    long long qg[4], qh[4];
    int c[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int b = lane * 4 + t;
        // modulo 2^64 (raw sums wrap; the other forms never do)
        qg[t] = (long long)((unsigned long long)rq[0][b] + (unsigned long long)rq[1][b] +
                            (unsigned long long)rq[2][b] + (unsigned long long)rq[3][b]);
        c[t] = rc[0][b] + rc[1][b] + rc[2][b] + rc[3][b];
        qh[t] = rh[0][b] + rh[1][b] + rh[2][b] + rh[3][b];
    }
    if (mode == 0 && T.hist_wide == 2) {
        // raw root sums: count * 2^CNT_SHIFT + q-sum modulo 2^64, |q-sum| < 2^62
        const auto rc_in = gbl((const int*)T.root_cnt);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            c[t] = rc_in[(size_t)f * NBIN + lane * 4 + t];
            qg[t] = (long long)((unsigned long long)qg[t] - ((unsigned long long)(unsigned)c[t] << CNT_SHIFT));
        }
    } else if (mode == 0 && T.hist_wide == 3 && wave == 0) {
        // the task's whole-window root: keep its counts for the raw builds of later trees
        auto rc_out = gbl(T.root_cnt);
#pragma unroll
        for (int t = 0; t < 4; ++t) rc_out[(size_t)f * NBIN + lane * 4 + t] = c[t];
        if (f == 0 && lane == 0) {
            Tw.cnt_row0 = T.row0;
            Tw.cnt_n = T.n_sample;
        }
    }
    int node = -1;
    double gain = -INFINITY;
    int thr = -1, ml = 1, cl = 0;
    long long tq = 0, tqh = 0, tc = 0;
    const bool catf = S.is_cat != nullptr && S.is_cat[f] != 0;
    const double inv_g = ldexp(1.0, -T.s_g), inv_h = ldexp(1.0, -T.s_h);
    if (active) {
        if (side == 0) {
            node = mode == 0 ? 0 : small_node;
        } else {
            node = big_node;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                qg[t] = par_q[t] - qg[t];
                c[t] = par_c[t] - c[t];
                qh[t] = T.hconst ? 0 : par_h[t] - qh[t];
            }
        }
        if (half == 0) {
            // both children's histograms are kept for later use as parents
            const size_t o = (size_t)node * per + (size_t)f * NBIN + lane * 4;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                gbl(T.hg)[o + t] = qg[t];
                gbl(T.hc)[o + t] = c[t];
                if (!T.hconst) gbl(T.hh)[o + t] = qh[t];
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (lane * 4 + t >= S.n_bins) {
                qg[t] = 0;
                c[t] = 0;
                qh[t] = 0;
            }
        if (!catf)
            scan_feature_half(qg, c, qh, T.hconst, half, inv_g, inv_h, S.n_bins, P.l2, P.gamma, P.min_leaf,
                              P.min_hess, P.l1, S.monotone ? (int)S.monotone[f] : 0, gain, thr, ml, cl, tq, tqh, tc);
    }
    unsigned long long cmask[4] = {0, 0, 0, 0};
    if (catf)   // block-uniform: every wave takes the categorical scan's barriers
        scan_cat(active && half == 0, qg, c, qh, T.hconst, inv_g, inv_h, P, rq, rh, rc, gain, thr, ml, cl, cmask,
                 tq, tqh, tc);
    __shared__ double xg[4];
    __shared__ int xt[4], xm[4], xc[4];
    if (lane == 0) {
        xg[wave] = gain;
        xt[wave] = thr;
        xm[wave] = ml;
        xc[wave] = cl;
    }
    __syncthreads();
    if (!active || half != 0 || lane != 0) return;
    // the other half's best (keys interleave across the halves: key = bin * 2 + missing-left)
    const int o = wave + 2;
    const int key = thr < 0 ? 0x7fffffff : thr * 2 + ml;
    const int okey = xt[o] < 0 ? 0x7fffffff : xt[o] * 2 + xm[o];
    if (cand_better(xg[o], okey, gain, key)) {
        gain = xg[o];
        thr = xt[o];
        ml = xm[o];
        cl = xc[o];
    }
    FCand fc;
    fc.gain = gain;
    fc.thr = thr;
    fc.ml = ml;
    fc.cl = cl;
    fc.pad = 0;
    stg(T.fcand + (size_t)node * S.n_features + f, fc);
    if (catf) {
        const auto fm = gbl(T.fmask + ((size_t)node * S.n_features + f) * 4);
#pragma unroll
        for (int w = 0; w < 4; ++w) fm[w] = cmask[w];
    }
    if (f == 0) {
        // node value from the node's own sums (tree.cpp:323-328, 407-416)
        const double G = (double)tq * inv_g;
        const double H = T.hconst ? (double)tc : (double)tqh * inv_h;
        T.nodes[node].value = leaf_value(G, H, P.l2, P.l1, P.mds);
    }
}

// The split scan of the root (mode 0) or of a split's two children (mode 1): grid (features, 1,
// tasks).
__global__ __launch_bounds__(256) void k_scan(Task* tasks, Shared S, GrowP P, int mode) {
    scan_one(tasks, S, P, mode, blockIdx.z, blockIdx.x);
}

// The split scan of a multi-leaf step (numeric features: categorical features grow one split at a
// time on the host).  One wave per (task slot, feature, child), with no block barrier between
// slots, so the waves of a CU overlap one another's loads: wave w of block x scans feature
// 2x + (w >> 1) of child w & 1 (0 the built smaller child, 1 the larger = parent - smaller; both
// waves sum the slot's partials, the second from the cache) for the active (task, slot) pairs
// q = z, z + gridDim.z, ... of all tasks.  Each wave stores its child's histogram row (the
// parent of a later split), the feature's best split, and for feature 0 the child's value.
template <bool HC>   // h == 1 for every task of the session
__global__ __launch_bounds__(256, 4) void k_scan_spec(Task* tasks, Shared S, GrowP P, int n_tasks) {
    __shared__ int pre[1025];
    {   // exclusive prefix of the tasks' slot counts (256 threads, four entries each; <= 1024 tasks)
        __shared__ int wtot[4];
        const int tt = threadIdx.x, wave = tt >> 6;
        int v[4], sum = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = 4 * tt + u;
            v[u] = i < n_tasks && !tasks[i].done ? tasks[i].n_slots : 0;
            sum += v[u];
        }
        const int incl = wave_incl_i(sum);
        if ((tt & 63) == 63) wtot[wave] = incl;
        __syncthreads();
        int base = 0;
        for (int w = 0; w < wave; ++w) base += wtot[w];
        int ex = base + incl - sum;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = 4 * tt + u;
            if (i < n_tasks) pre[i] = ex;
            ex += v[u];
        }
        if (tt == 0) pre[n_tasks] = wtot[0] + wtot[1] + wtot[2] + wtot[3];
        __syncthreads();
    }
    const int pairs = pre[n_tasks];
    const int F = S.n_features;
    const size_t per_node = (size_t)F * NBIN;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int f = 2 * blockIdx.x + (wave >> 1), side = wave & 1;
    if (f >= F) return;   // no barrier follows
    for (int q = blockIdx.z; q < pairs; q += gridDim.z) {
        int lo = 0, hi = n_tasks - 1;   // the task: the last with pre[t] <= q
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= q) lo = mid;
            else hi = mid - 1;
        }
        const Task T = tasks[lo];
        const Slot sl = ldg(tasks[lo].slot_tab + (q - pre[lo]));   // issued beside the task's loads
        Parts pp;
        pp.part = T.part + (size_t)sl.blk0 * per_node;
        pp.part_c = T.part_c + (size_t)sl.blk0 * per_node;
        pp.part_h = T.part_h ? T.part_h + (size_t)sl.blk0 * per_node : nullptr;
        pp.wide = sl.hist_wide;
        pp.hconst = HC;
        const size_t fo = (size_t)f * NBIN + lane * 4;
        long long par_q[4] = {0, 0, 0, 0}, par_h[4] = {0, 0, 0, 0};
        int par_c[4] = {0, 0, 0, 0};
        if (side) {   // the parent's sums, loaded before the partials so the latencies overlap
            const size_t po = (size_t)sl.leaf * per_node + fo;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                par_q[t] = gbl(T.hg)[po + t];
                par_c[t] = gbl(T.hc)[po + t];
                if (!HC) par_h[t] = gbl(T.hh)[po + t];
            }
        }
        long long sq[4], sc[4], sh[4];
        sum_partials<2>(pp, F, f, 0, 1, sl.hist_parts, sq, sc, sh);   // two per round: fewer live registers
        long long qg[4], qh[4];
        int c[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            qg[t] = side ? par_q[t] - sq[t] : sq[t];
            c[t] = side ? par_c[t] - (int)sc[t] : (int)sc[t];
            qh[t] = HC ? 0 : (side ? par_h[t] - sh[t] : sh[t]);
        }
        const int node = side ? sl.big_node : sl.small_node;
        const size_t o = (size_t)node * per_node + fo;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            gbl(T.hg)[o + t] = qg[t];
            gbl(T.hc)[o + t] = c[t];
            if (!HC) gbl(T.hh)[o + t] = qh[t];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (lane * 4 + t >= S.n_bins) {
                qg[t] = 0;
                c[t] = 0;
                qh[t] = 0;
            }
        const double inv_g = ldexp(1.0, -T.s_g), inv_h = ldexp(1.0, -T.s_h);
        const int mono = S.monotone ? (int)S.monotone[f] : 0;
        double gain;
        int thr, ml, cl;
        long long tq, tqh, tc;
        // the lane's 8 candidates in two passes of 4 (fewer live registers: more resident waves)
        scan_feature_wave_t<HC, 0, 4, 2>(qg, c, qh, inv_g, inv_h, S.n_bins, P.l2, P.gamma, P.min_leaf, P.min_hess,
                                         P.l1, mono, gain, thr, ml, cl, tq, tqh, tc);
        if (lane == 0) {
            FCand fc;
            fc.gain = gain;
            fc.thr = thr;
            fc.ml = ml;
            fc.cl = cl;
            fc.pad = 0;
            stg(T.fcand + (size_t)node * F + f, fc);
            if (f == 0) {
                // node value from the node's own sums (tree.cpp:323-328, 407-416)
                const double G = (double)tq * inv_g;
                const double H = HC ? (double)tc : (double)tqh * inv_h;
                T.nodes[node].value = leaf_value(G, H, P.l2, P.l1, P.mds);
            }
        }
    }
}

void launch_scan(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, const GrowP& P, int mode, bool hconst) {
    if (mode == 2) {
        // as many z-blocks as one task's slots, or up to 64 pairs in flight at a time over the tasks
        const int z = std::min(64, std::max(SPEC_W, n_tasks));
        const dim3 grid((S.n_features + 1) / 2, 1, z);
        if (hconst) hipLaunchKernelGGL(k_scan_spec<true>, grid, dim3(256), 0, st, tasks, S, P, n_tasks);
        else hipLaunchKernelGGL(k_scan_spec<false>, grid, dim3(256), 0, st, tasks, S, P, n_tasks);
        return;
    }
    hipLaunchKernelGGL(k_scan, dim3(S.n_features, 1, n_tasks), dim3(256), 0, st, tasks, S, P, mode);
}

// =================================================================== partition
// Reference split_indices (tree.cpp:68-91) decides which rows of the leaf go left: Axis: bin 0
// -> missing_go_left else bin <= threshold; Pair: either bin 0 -> missing_go_left else
// pair_hash <= threshold.  The rows move into the other index buffer (lefts at the front of the
// leaf's range, rights at the back); the reference keeps them in ascending order, this kernel
// does not need to (see k_partition).
// Rows per block (4 per thread): smaller chunks start the gathers of a leaf's rows on more
// CUs at once (1M x 128 fit: 0.979 ms/tree at 1024 rows, 1.00 at 2048, 1.03 at 512; the 64-fit
// walk-forward batch: 122 ms of partitions at 1024 rows, 159 ms at 4096).
constexpr int PART_CHUNK = 1024;
constexpr int PART_RPT = 4;

__device__ inline bool row_goes_left(const Shared& S, const Task& T, const Decision& d, int i) {
    const int b1 = S.bins_col[(size_t)d.feat * S.n_pad + T.row0 + i];
    if (d.type >= 2) {   // Cat / Regime, cat_goes_left (tree.hpp:27-30): the bin's bit in the leaf's mask
        const auto fm = gbl((const unsigned long long*)T.fmask + ((size_t)d.leaf * S.n_features + d.feat) * 4);
        return (fm[b1 >> 6] >> (b1 & 63)) & 1ull;
    }
    if (d.type == 1) {
        const int b2 = S.bins_col[(size_t)d.feat2 * S.n_pad + T.row0 + i];
        if (b1 == 0 || b2 == 0) return d.mgl != 0;
        const unsigned z = ((unsigned)b1 * 2654435761u + (unsigned)b2 * 40503u) % (unsigned)S.n_bins;
        return (int)z <= d.thr;
    }
    if (b1 == 0) return d.mgl != 0;
    return b1 <= d.thr;
}

// Decide the split and partition the split leaf's rows.  mode 0: every block runs the same
// best-first selection (select_decide), so no separate selection launch is needed; mode 1: the
// host wrote the decision into T.cur_* (validated here).  Block 0 records the decision and the
// leaf's row range in the step descriptor; the node table itself is updated by the next
// (histogram) launch, so no block of this launch can observe a half-applied split.
//
// Each block places its rows with one atomic reservation per side: lefts fill the leaf's range
// from the front, rights from the back.  The order of rows inside a child's range therefore
// depends on block timing, but nothing downstream depends on it: every histogram, leaf and TORB
// sum is an exact integer sum over the set of rows.
__global__ __launch_bounds__(256) void k_partition(Task* tasks, Shared S, GrowP P, int mode) {
    const Task T = tasks[blockIdx.y];
    Task& Tw = tasks[blockIdx.y];
    if (T.done) return;
    __shared__ Decision sd;
    __shared__ int s_slot, s_chunk;
    __shared__ Slot pslots[SPEC_W];
    if (mode == 2) {
        // multi-leaf step: the blocks of slot j are chunk0[j] .. chunk0[j + 1) - 1; the slot
        // table comes into LDS in one round trip
        const int nw = (int)(sizeof(Slot) / 4) * T.n_slots;
        if ((int)threadIdx.x < nw) ((int*)pslots)[threadIdx.x] = ((const int*)T.slots)[threadIdx.x];
        __syncthreads();
        if (threadIdx.x == 0) {
            int j = -1;
            for (int q = 0; q < T.n_slots; ++q)
                if ((int)blockIdx.x >= pslots[q].chunk0) j = q;
            s_slot = j;
            sd.leaf = -1;
            if (j >= 0) {
                const Slot& sl = pslots[j];
                s_chunk = (int)blockIdx.x - sl.chunk0;
                sd.leaf = s_chunk * PART_CHUNK < sl.count ? sl.leaf : -1;
                sd.type = sl.type;
                sd.feat = sl.feat;
                sd.feat2 = sl.feat2;
                sd.thr = sl.thr;
                sd.mgl = sl.mgl;
                sd.nl = sl.nl;
                sd.gain = sl.gain;
                sd.start = sl.start;
                sd.count = sl.count;
                sd.buf = sl.buf;
            }
        }
        __syncthreads();
        if (sd.leaf < 0) return;   // past the step's last chunk (the selection decides when the tree ends)
    } else if (mode == 0) {
        select_decide(T, S.n_features, P, blockIdx.x == 0, sd);
    } else {
        if (threadIdx.x == 0) {
            const int leaf = T.cur_leaf, F = S.n_features;
            const bool bad = leaf < 0 || leaf >= T.n_nodes || !T.nodes[leaf].v[N_LEAF] || T.cur_feat < 0 ||
                             T.cur_feat >= F || (T.cur_type == 1 && (T.cur_feat2 < 0 || T.cur_feat2 >= F)) ||
                             T.cur_type < 0 || T.cur_type > 3 || (T.cur_type >= 2 && !T.fmask) ||
                             (T.cur_type >= 2 && !(S.is_cat && S.is_cat[T.cur_feat])) || T.n_nodes + 2 > MAX_NODES ||
                             T.n_nodes + 2 > 2 * P.max_leaves - 1;
            sd.leaf = bad ? -1 : leaf;   // a decision that does not name a splittable leaf stops the tree
            sd.type = T.cur_type;
            sd.feat = T.cur_feat;
            sd.feat2 = T.cur_feat2;
            sd.thr = T.cur_thr;
            sd.mgl = T.cur_mgl;
            sd.nl = T.cur_nl;   // known for decisions of k_select, -1 for the host's
            sd.gain = T.cur_gain;
            if (!bad) {
                const DNode nd = ldg(T.nodes + leaf);
                sd.start = nd.seg_start;
                sd.count = nd.seg_count;
                sd.buf = nd.buf;
            }
        }
        __syncthreads();
    }
    const Decision d = sd;
    if (d.leaf < 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) Tw.done = 1;
        return;
    }
    const int cnt = d.count, st = d.start, lbuf = d.buf;
    const int chunk = mode == 2 ? s_chunk : (int)blockIdx.x;
    unsigned long long* lr = mode == 2 ? &T.slots[s_slot].part_lr : &Tw.part_lr;
    if (mode != 2 && blockIdx.x == 0 && threadIdx.x == 0) {
        Tw.cur_leaf = d.leaf;
        Tw.cur_type = d.type;
        Tw.cur_feat = d.feat;
        Tw.cur_feat2 = d.feat2;
        Tw.cur_thr = d.type >= 2 ? -1 : d.thr;   // a categorical node has no threshold bin
        if (d.type >= 2) {   // the split node keeps its mask for the exported tree
            const auto fm = gbl((const unsigned long long*)T.fmask + ((size_t)d.leaf * S.n_features + d.feat) * 4);
            const auto nm = gbl(T.nmask + (size_t)d.leaf * 4);
#pragma unroll
            for (int w = 0; w < 4; ++w) nm[w] = fm[w];
        }
        Tw.cur_mgl = d.mgl;
        Tw.cur_gain = d.gain;
        Tw.cur_start = st;
        Tw.cur_count = cnt;
        Tw.cur_buf = lbuf;
        Tw.cur_left = T.n_nodes;
        Tw.cur_nl = d.nl;
    }
    const int c0 = chunk * PART_CHUNK;
    if (c0 >= cnt) return;
    __shared__ int wsum[4], wsumr[4];
    __shared__ int baseL, baseR;
    const auto src = gbl((const int*)(lbuf ? T.idx_tmp : T.idx) + st);
    const auto dst = gbl((lbuf ? T.idx : T.idx_tmp) + st);
    const int k0 = c0 + threadIdx.x * PART_RPT;
    unsigned flags = 0;
    int rows[PART_RPT];
    int nl = 0, nr = 0;
#pragma unroll
    for (int t = 0; t < PART_RPT; ++t) rows[t] = (k0 + t < cnt) ? src[k0 + t] : -1;
#pragma unroll
    for (int t = 0; t < PART_RPT; ++t) {
        if (rows[t] < 0) continue;
        if (row_goes_left(S, T, d, rows[t])) {
            flags |= 1u << t;
            ++nl;
        } else {
            ++nr;
        }
    }
    const int incl = wave_incl_i(nl), inclr = wave_incl_i(nr);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) {
        wsum[wave] = incl;
        wsumr[wave] = inclr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int tl = 0, tr = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            tl += wsum[w];
            tr += wsumr[w];
        }
        // one reservation per block for both sides: same-address device atomics serialize
        const unsigned long long old = atomicAdd(lr, (unsigned long long)tr << 32 | (unsigned)tl);
        baseL = (int)(old & 0xffffffffull);
        baseR = (int)(old >> 32);
    }
    __syncthreads();
    int woff = 0, woffr = 0;
    for (int w = 0; w < wave; ++w) {
        woff += wsum[w];
        woffr += wsumr[w];
    }
    int lpos = baseL + woff + incl - nl;
    int rpos = baseR + woffr + inclr - nr;
#pragma unroll
    for (int t = 0; t < PART_RPT; ++t) {
        if (rows[t] < 0) break;
        if (flags & (1u << t)) dst[lpos++] = rows[t];
        else dst[cnt - 1 - rpos++] = rows[t];
    }
}

// Standalone selection (one block per task) for launches that grow many tasks at once: the
// partition then reads the decision instead of every partition block recomputing it.
__global__ __launch_bounds__(256) void k_select(Task* tasks, Shared S, GrowP P) {
    const Task T = tasks[blockIdx.x];
    Task& Tw = tasks[blockIdx.x];
    if (T.done) return;
    __shared__ Decision sd;
    select_decide(T, S.n_features, P, true, sd);
    if (threadIdx.x != 0) return;
    if (sd.leaf < 0) {
        Tw.done = 1;
        return;
    }
    Tw.cur_leaf = sd.leaf;
    Tw.cur_type = sd.type;
    Tw.cur_feat = sd.feat;
    Tw.cur_feat2 = sd.feat2;
    Tw.cur_thr = sd.thr;
    Tw.cur_mgl = sd.mgl;
    Tw.cur_gain = sd.gain;
    Tw.cur_nl = sd.nl;
}

// =================================================================== multi-leaf growth
// Best-first growth (tree.cpp:343-392, 445-456) splits one leaf at a time, always the leaf with
// the best candidate.  A leaf's candidate depends only on its rows (PAIR splits off), so leaves
// may be split ahead of their turn: each step below splits the leaf best-first takes next plus
// the next-best leaves of the frontier (up to SPEC_W), and the tree is the best-first replay over
// the splits made.  A split the replay does not reach is dropped at the end (its leaf stays a
// leaf; only the work is lost), so the tree, its node ids and every value are the one-split-
// per-step tree's.

// A grown node's candidate as a replay key: its gain if best-first may split it, else -inf.
__device__ inline double spec_key(const DCand& c, const GrowP& P) {
    return (c.valid && c.feature >= 0 && c.gain > P.gamma) ? c.gain : -INFINITY;
}

// The best-first replay over the grown nodes, by one whole wave: lane i holds frontier entry i
// (node, replay id, key).  It takes the best entry (largest key, then lowest replay id, as
// select_decide) while that node has been split (left[] >= 0) and the leaf budget allows,
// replacing it by its children.  Returns 1 when the replay ends (no candidate or the budget is
// spent: the tree is complete), 0 when the best entry is a leaf not split yet.  fid[] (if given)
// receives the replay ids, fsplit[] marks the replayed splits.
// With resume set, the replay continues from the frontier in fnode / fsim / n_front / n_final
// (a stop of an earlier replay over the same tree: splits are only ever added, so continuing
// from there equals replaying from the root).
__device__ inline int spec_replay(const double* key, const int* left, const GrowP& P, int& fnode, int& fsim,
                                  double& fg, int& n_front, int& n_final, int* fid, unsigned char* fsplit,
                                  bool resume = false) {
    const int lane = threadIdx.x & 63;
    if (resume) {
        fg = fnode >= 0 ? key[fnode] : -INFINITY;
    } else {
        fnode = lane == 0 ? 0 : -1;
        fsim = 0;
        fg = lane == 0 ? key[0] : -INFINITY;
        n_front = 1;
        n_final = 1;
        if (fid && lane == 0) fid[0] = 0;
    }
    const int budget = min(2 * P.max_leaves - 1, MAX_NODES);
    while (true) {
        const double g = fnode >= 0 ? fg : -INFINITY;
        const int wl = wave_best_lane(g, fnode >= 0 ? fsim : 0x7fffffff);
        const double bg = read_lane_f64(g, wl);
        const int bx = __builtin_amdgcn_readlane(fnode, wl);
        if (!(bg > -INFINITY) || n_final + 2 > budget) return 1;
        const int L = left[bx];
        if (L < 0) return 0;
        if (lane == wl) {
            fnode = L;
            fsim = n_final;
            fg = key[L];
        }
        if (lane == n_front) {
            fnode = L + 1;
            fsim = n_final + 1;
            fg = key[L + 1];
        }
        if (lane == 0) {
            if (fid) {
                fid[L] = n_final;
                fid[L + 1] = n_final + 1;
            }
            if (fsplit) fsplit[bx] = 1;
        }
        ++n_front;
        n_final += 2;
    }
}

// Candidates (tree.cpp:160-166: ascending features, first maximum kept) of the nodes the last
// scan produced, [pend_first, pend_first + n_pend), into T.cands, key[] and pc[]; a 256-thread
// block, 16 threads per pending node (n_pend <= 2 * SPEC_W = 16), each over features sub,
// sub + 16, ... with its loads issued together, the 16 partial bests combined in feature order.
// Recomputing them gives the same values, so the step's selection and the tree's final numbering
// may both call it.
__device__ void spec_pending(const Task& T, int F, const GrowP& P, double* key, DCand* pc) {
    const int p0 = T.pend_first, np = T.n_pend;
    __shared__ double rg[256];
    __shared__ int rk[256], rt[256], rm[256], rcl[256];
    {
        const int pi = threadIdx.x >> 4, sub = threadIdx.x & 15;
        double best = -INFINITY;
        int bk = 0x7fffffff, bt = -1, bm = 1, bcl = 0;
        if (pi < np) {
            const int nd = p0 + pi;
            for (int f0 = sub; f0 < F; f0 += 16 * 8) {
                FCand c[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int f = f0 + 16 * u;
                    if (f < F) c[u] = ldg(T.fcand + (size_t)nd * F + f);
                    else c[u].thr = -1;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int f = f0 + 16 * u;
                    if (c[u].thr < 0 || c[u].gain < P.gamma) continue;
                    if (cand_better(c[u].gain, f, best, bk)) {
                        best = c[u].gain;
                        bk = f;
                        bt = c[u].thr;
                        bm = c[u].ml;
                        bcl = c[u].cl;
                    }
                }
            }
        }
        rg[threadIdx.x] = best;
        rk[threadIdx.x] = bk;
        rt[threadIdx.x] = bt;
        rm[threadIdx.x] = bm;
        rcl[threadIdx.x] = bcl;
        __syncthreads();
        if (sub == 0 && pi < np) {
            for (int j = 1; j < 16; ++j) {
                const int o = threadIdx.x + j;
                if (cand_better(rg[o], rk[o], best, bk)) {
                    best = rg[o];
                    bk = rk[o];
                    bt = rt[o];
                    bm = rm[o];
                    bcl = rcl[o];
                }
            }
            const int nd = p0 + pi;
            const DNode d = ldg(T.nodes + nd);
            DCand c;
            c.valid = bk != 0x7fffffff && d.v[N_DEPTH] < P.max_depth && d.v[N_NS] >= 2 * P.min_leaf;
            c.gain = best;
            c.feature = bk == 0x7fffffff ? -1 : bk;
            c.feature2 = -1;
            c.type = 0;
            c.thr = bk == 0x7fffffff ? -1 : bt;
            c.mgl = bk == 0x7fffffff ? 1 : bm;
            c.cl = bk == 0x7fffffff ? 0 : bcl;
            c.pad = 0;
            pc[pi] = c;
            T.cands[nd] = c;
            key[nd] = spec_key(c, P);
        }
    }
}

// Selection of a multi-leaf step (one block per task): reduce the candidates of the nodes the last
// scan produced, replay best-first over all grown nodes (resuming from the last step's stop), and
// split the leaf the replay stops at together with the next-best unsplit leaves of its frontier:
// at most SPEC_W, at most the splits the leaf budget still allows, and never so many that the
// remaining splits could not get node ids.  Writes the slots (with each slot's partition blocks
// and histogram tiles), the children's ids and the node count; sets done when the replay
// completes the tree.  Every global load is issued before the first use (one node and its
// cached candidate per thread, the pending nodes' per-feature results, the stored stop), so the
// kernel waits for memory about once.
__global__ __launch_bounds__(256) void k_spec_select(Task* tasks, Shared S, GrowP P, int width, int min_per,
                                                     int rb_target, int step_parts) {
    Task& Tw = tasks[blockIdx.x];
    const Task T = tasks[blockIdx.x];
    // the step's child builds hand out their items from these counters (the first task's block
    // resets them, done or not)
    if (blockIdx.x == 0 && threadIdx.x < 8 && S.spec_ctr) S.spec_ctr[threadIdx.x] = 0;
    if (T.done) return;
    const int F = S.n_features;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    static_assert(MAX_NODES <= 256, "one node per thread");
    __shared__ double key[MAX_NODES];
    __shared__ int left[MAX_NODES], nst[MAX_NODES], ncnt[MAX_NODES], nbuf[MAX_NODES], ndep[MAX_NODES], nns[MAX_NODES];
    __shared__ DCand scand[MAX_NODES];
    __shared__ double rg[256];
    __shared__ int rk[256], rt[256], rm[256], rcl[256];
    const int nn = T.n_nodes;
    const int p0 = T.pend_first, np = T.n_pend;
    // ---- loads
    const int k = threadIdx.x;
    DNode nd;
    DCand cd;
    if (k < nn) {
        nd = ldg(T.nodes + k);
        cd = ldg(T.cands + k);
    }
    const int pi = threadIdx.x >> 4, sub = threadIdx.x & 15;   // 16 threads per pending node
    FCand fc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int f = sub + 16 * u;
        if (pi < np && f < F) fc[u] = ldg(T.fcand + (size_t)(p0 + pi) * F + f);
        else fc[u].thr = -1;
    }
    int fnode = -1, fsim = 0;
    const bool resume = T.rs_valid != 0;
    if (wave == 0 && resume) {
        fnode = T.rs_node[lane];
        fsim = T.rs_sim[lane];
    }
    // ---- the grown nodes into LDS
    if (k < nn) {
        left[k] = nd.v[N_LEAF] ? -1 : nd.v[N_LEFT];
        nst[k] = nd.seg_start;
        ncnt[k] = nd.seg_count;
        nbuf[k] = nd.buf;
        ndep[k] = nd.v[N_DEPTH];
        nns[k] = nd.v[N_NS];
        scand[k] = cd;
        key[k] = spec_key(cd, P);
    }
    // ---- the pending nodes' candidates (tree.cpp:160-166): ascending features, first maximum
    // kept; each thread's features sub, sub + 16, ..., combined in feature order
    {
        double best = -INFINITY;
        int bk = 0x7fffffff, bt = -1, bm = 1, bcl = 0;
        for (int f0 = 0;; f0 += 128) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int f = f0 + sub + 16 * u;
                if (fc[u].thr < 0 || fc[u].gain < P.gamma) continue;
                if (cand_better(fc[u].gain, f, best, bk)) {
                    best = fc[u].gain;
                    bk = f;
                    bt = fc[u].thr;
                    bm = fc[u].ml;
                    bcl = fc[u].cl;
                }
            }
            if (f0 + 128 >= F) break;
#pragma unroll
            for (int u = 0; u < 8; ++u) {   // more than 128 features: the next 128
                const int f = f0 + 128 + sub + 16 * u;
                if (pi < np && f < F) fc[u] = ldg(T.fcand + (size_t)(p0 + pi) * F + f);
                else fc[u].thr = -1;
            }
        }
        rg[threadIdx.x] = best;
        rk[threadIdx.x] = bk;
        rt[threadIdx.x] = bt;
        rm[threadIdx.x] = bm;
        rcl[threadIdx.x] = bcl;
        __syncthreads();
        if (sub == 0 && pi < np) {
            for (int j = 1; j < 16; ++j) {
                const int o = threadIdx.x + j;
                if (cand_better(rg[o], rk[o], best, bk)) {
                    best = rg[o];
                    bk = rk[o];
                    bt = rt[o];
                    bm = rm[o];
                    bcl = rcl[o];
                }
            }
            const int node = p0 + pi;
            DCand c;
            c.valid = bk != 0x7fffffff && ndep[node] < P.max_depth && nns[node] >= 2 * P.min_leaf;
            c.gain = best;
            c.feature = bk == 0x7fffffff ? -1 : bk;
            c.feature2 = -1;
            c.type = 0;
            c.thr = bk == 0x7fffffff ? -1 : bt;
            c.mgl = bk == 0x7fffffff ? 1 : bm;
            c.cl = bk == 0x7fffffff ? 0 : bcl;
            c.pad = 0;
            scand[node] = c;
            T.cands[node] = c;
            key[node] = spec_key(c, P);
        }
    }
    __syncthreads();
    if (wave != 0) return;
    // ---- the replay (wave 0)
    int n_front = resume ? T.rs_front : 1, n_final = resume ? T.rs_final : 1;
    double fg;
    const int complete = spec_replay(key, left, P, fnode, fsim, fg, n_front, n_final, nullptr, nullptr, resume);
    // keep the stop for the next step's selection
    T.rs_node[lane] = fnode;
    T.rs_sim[lane] = fsim;
    if (lane == 0) {
        Tw.rs_front = n_front;
        Tw.rs_final = n_final;
        Tw.rs_valid = 1;
    }
    if (complete) {
        if (lane == 0) {
            Tw.done = 1;
            Tw.n_slots = 0;
            Tw.spec_gb = 0;
        }
        return;
    }
    // the replay stopped at its best entry, an unsplit leaf: it and the next-best unsplit leaves
    const int remaining = (min(2 * P.max_leaves - 1, MAX_NODES) - n_final) / 2;
    int kmax = min(min(width, SPEC_W), remaining);
    kmax = min(kmax, (T.spec_cap - nn) / 2 - (remaining - 1));
    kmax = max(kmax, 1);
    double g = (fnode >= 0 && left[fnode] < 0) ? fg : -INFINITY;
    const int kk = fnode >= 0 ? fsim : 0x7fffffff;
    int chosen = -1, n_chosen = 0;   // lane j < n_chosen holds the j-th chosen leaf
    for (int j = 0; j < kmax; ++j) {
        const int wl = wave_best_lane(g, kk);
        if (!(read_lane_f64(g, wl) > -INFINITY)) break;
        const int leaf = __builtin_amdgcn_readlane(fnode, wl);
        if (lane == j) chosen = leaf;
        if (lane == wl) g = -INFINITY;
        ++n_chosen;
    }
    // lane j < n_chosen writes slot j; chunk and tile offsets are exclusive prefix sums
    int chunks = 0, gx = 0, small = 0;
    Slot sl;
    if (lane < n_chosen) {
        const DCand c = scand[chosen];
        sl.leaf = chosen;
        sl.left = nn + 2 * lane;
        sl.start = nst[chosen];
        sl.count = ncnt[chosen];
        sl.buf = nbuf[chosen];
        sl.type = c.type;
        sl.feat = c.feature;
        sl.feat2 = c.feature2;
        sl.thr = c.thr;
        sl.mgl = c.mgl;
        sl.nl = c.cl;
        sl.gain = c.gain;
        sl.part_lr = 0;
        sl.small_node = sl.big_node = sl.small_start = sl.small_count = sl.small_buf = -1;
        sl.hist_parts = 0;
        sl.hist_wide = 0;
        sl.pad = 0;
        chunks = (sl.count + PART_CHUNK - 1) / PART_CHUNK;
        small = min(c.cl, sl.count - c.cl);
    }
    // rows per histogram tile, the same for every slot: the step's smaller children in at most
    // rb_target row blocks (so the step's items fit the resident blocks: each slot rounds its
    // row blocks up, so the tile is sized for rb_target - n_chosen of them), at least min_per
    // rows and at most SPEC_ITEM_ROWS (a launch's last items then end close together: 64-fit
    // walk-forward batch 1.25 s at 8,192 rows, 1.32 s at 4,096 and at 16,384, 1.50 s at 32,768),
    // yet enough for at most step_parts blocks over all slots, and at most MAX_BLOCK_ROWS
    const int total_small = wave_sum_i(small);
    const int rb = max(1, rb_target - n_chosen);
    int per = max(min_per, (total_small + rb - 1) / rb);
    per = min(per, SPEC_ITEM_ROWS);
    per = max(per, (total_small + step_parts - n_chosen - 1) / (step_parts - n_chosen));
    per = min((per + 63) & ~63, MAX_BLOCK_ROWS);
    if (lane < n_chosen) gx = max(1, (small + per - 1) / per);
    const int c_incl = wave_incl_i(chunks), g_incl = wave_incl_i(gx);
    if (nn + 2 * n_chosen > T.spec_cap) {   // cannot happen while spec_cap >= 2 * max_leaves - 1
        if (lane == 0) {
            Tw.err = 3;
            Tw.done = 1;
            Tw.n_slots = 0;
            Tw.spec_gb = 0;
        }
        return;
    }
    if (lane < n_chosen) {
        sl.chunk0 = c_incl - chunks;
        sl.blk0 = g_incl - gx;
        sl.gx = gx;
        sl.per = per;
        T.slots[lane] = sl;
    }
    const int n_gb = __builtin_amdgcn_readlane(g_incl, 63);
    if (lane == 0) {
        Tw.n_slots = n_chosen;
        Tw.n_nodes = nn + 2 * n_chosen;
        Tw.pend_first = nn;
        Tw.n_pend = 2 * n_chosen;
        Tw.spec_steps = T.spec_steps + 1;
        Tw.spec_gb = n_gb;
    }
}

// The grown tree in best-first order (one block per task, after the last step): replay
// best-first over all grown nodes to the end, number the replayed nodes by the replay (the ids
// one-split-per-step growth assigns), and rewrite the node table in that numbering.  A grown
// split the replay did not take leaves its node a leaf.  The grown nodes are kept (tnodes) with,
// for each grown leaf, the final leaf that holds its rows (tmap): the leaf-id pass walks their
// row ranges.  An unfinished replay (more steps are needed) sets skip_end, so the tree-end
// kernels enqueued behind it change nothing.
__global__ __launch_bounds__(256) void k_spec_finalize(Task* tasks, Shared S, GrowP P) {
    Task& Tw = tasks[blockIdx.x];
    const Task T = tasks[blockIdx.x];
    if (T.end_done) {   // a later end of a tree that has ended (another task needed more steps)
        if (threadIdx.x == 0) Tw.skip_end = 1;
        return;
    }
    __shared__ DNode sn[MAX_NODES];
    __shared__ double key[MAX_NODES];
    __shared__ int left[MAX_NODES], fid[MAX_NODES], par[MAX_NODES];
    __shared__ unsigned char fsplit[MAX_NODES];
    __shared__ DCand pc[2 * SPEC_W];
    __shared__ int s_final, s_ok;
    const int nn = T.n_nodes;
    for (int k = threadIdx.x; k < nn; k += blockDim.x) {
        sn[k] = T.nodes[k];
        key[k] = spec_key(T.cands[k], P);
        fid[k] = -1;
        par[k] = -1;
        fsplit[k] = 0;
    }
    __syncthreads();
    // the last step's children may not have their candidates yet (their selection has not run)
    if (!T.done) spec_pending(T, S.n_features, P, key, pc);
    __syncthreads();
    for (int k = threadIdx.x; k < nn; k += blockDim.x) {
        left[k] = sn[k].v[N_LEAF] ? -1 : sn[k].v[N_LEFT];
        if (!sn[k].v[N_LEAF]) {
            par[sn[k].v[N_LEFT]] = k;
            par[sn[k].v[N_RIGHT]] = k;
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        int fnode, fsim, n_front, n_final;
        double fg;
        const int complete = spec_replay(key, left, P, fnode, fsim, fg, n_front, n_final, fid, fsplit);
        if (threadIdx.x == 0) {
            s_final = n_final;
            s_ok = complete;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Tw.skip_end = s_ok ? 0 : 1;
        if (s_ok) Tw.end_done = 1;
        if (T.host_flag) {   // the host waits for this: {1 finished | 2 more steps needed, steps}
            T.host_flag[1] = T.spec_steps;
            __threadfence_system();
            T.host_flag[0] = s_ok ? 1 : 2;
        }
    }
    if (!s_ok) return;   // more steps are needed: the host runs them and enqueues the tree end again
    for (int k = threadIdx.x; k < nn; k += blockDim.x) {
        DNode d = sn[k];
        T.tnodes[k] = d;
        // the final leaf holding a grown leaf's rows: its deepest replayed ancestor-or-self
        int a = k;
        while (fid[a] < 0) a = par[a];
        T.tmap[k] = d.v[N_LEAF] ? fid[a] : -1;
        if (fid[k] < 0) continue;
        if (fsplit[k]) {
            d.v[N_LEFT] = fid[sn[k].v[N_LEFT]];
            d.v[N_RIGHT] = fid[sn[k].v[N_RIGHT]];
        } else if (!d.v[N_LEAF]) {
            // a grown split best-first does not take: the node is a leaf of the tree
            d.v[N_LEAF] = 1;
            d.v[N_TYPE] = 0;
            d.v[N_FEAT] = -1;
            d.v[N_FEAT2] = -1;
            d.v[N_THR] = -1;
            d.v[N_LEFT] = -1;
            d.v[N_RIGHT] = -1;
            d.v[N_MGL] = 1;
            d.gain = 0.0;
        }
        T.nodes[fid[k]] = d;
    }
    if (threadIdx.x == 0) {
        Tw.n_tnodes = nn;
        Tw.n_nodes = s_final;
    }
}

// Growth state of the active tasks for the host: out[0] = 1 when every task's growth is done,
// out[1] = the most steps a task took (one block).
__global__ __launch_bounds__(256) void k_spec_status(const Task* tasks, int n_tasks, int* out) {
    __shared__ int s_all, s_max;
    if (threadIdx.x == 0) {
        s_all = 1;
        s_max = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n_tasks; i += blockDim.x) {
        if (!tasks[i].done) atomicAnd(&s_all, 0);
        atomicMax(&s_max, tasks[i].spec_steps);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[1] = s_max;
        out[0] = s_all;
    }
}

void launch_spec_status(hipStream_t st, const Task* tasks, int n_tasks, int* out) {
    hipLaunchKernelGGL(k_spec_status, dim3(1), dim3(256), 0, st, tasks, n_tasks, out);
}

void launch_spec_select(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, const GrowP& P, int width,
                        int min_per, int rb_target, int step_parts) {
    hipLaunchKernelGGL(k_spec_select, dim3(n_tasks), dim3(256), 0, st, tasks, S, P, width, min_per, rb_target,
                       step_parts);
}

void launch_spec_finalize(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, const GrowP& P) {
    hipLaunchKernelGGL(k_spec_finalize, dim3(n_tasks), dim3(256), 0, st, tasks, S, P);
}

void launch_select(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, const GrowP& P) {
    hipLaunchKernelGGL(k_select, dim3(n_tasks), dim3(256), 0, st, tasks, S, P);
}

void launch_partition(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, int max_count, const GrowP& P,
                      int mode) {
    // mode 2: the slots' leaves are disjoint, so their chunks number at most rows / chunk + slots
    const int nch = std::max(1, (max_count + PART_CHUNK - 1) / PART_CHUNK) + (mode == 2 ? SPEC_W : 0);
    dim3 grid(nch, n_tasks);
    hipLaunchKernelGGL(k_partition, grid, dim3(256), 0, st, tasks, S, P, mode);
}

// =================================================================== host-driven helpers
// Fixed-point sum of g_use^2 over a node's rows: the leaf "energy" of the PAIR trigger
// (tree.cpp:356-357).
__global__ __launch_bounds__(256) void k_energy(Task* tasks, int task, int node) {
    Task& T = tasks[task];
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
        const long long s = sacc[0] + sacc[1] + sacc[2] + sacc[3];
        if (s) atomicAdd((unsigned long long*)&T.energy[node], (unsigned long long)s);
    }
}

void launch_energy(hipStream_t st, Task* tasks, int task, int node, int count) {
    const int blocks = std::max(1, std::min(256, (count + 1023) / 1024));
    hipLaunchKernelGGL(k_energy, dim3(blocks), dim3(256), 0, st, tasks, task, node);
}

// PAIR histograms: one hashed histogram per feature pair over a node's rows (the fused
// hash+accumulate pass of tree.cpp:243-267), then one scan per pair with lambda = pair_l2.
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
    const int* rows = node_rows(T, n) + n.seg_start;
    for (int j = threadIdx.x; j < n_pairs * NBIN * (hv ? 2 : 1); j += blockDim.x) lds[j] = 0;
    __syncthreads();
    for (int k = r0 + threadIdx.x; k < r1; k += blockDim.x) {
        const int i = rows[k];
        const unsigned char* row = S.bins_row + (size_t)(T.row0 + i) * S.fp;
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
    int thr, ml, cl;
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
                      P.l1, gain, thr, ml, cl, tq, tqh, tc);
    if (threadIdx.x == 0) {
        out[pp].gain = gain;
        out[pp].thr = thr;
        out[pp].ml = ml;
        out[pp].cl = cl;
    }
}

void launch_pair(hipStream_t st, const Shared& S, Task* tasks, int task, int node, int count,
                 const int* pf, int n_pairs, long long* ph_g, int* ph_c, long long* ph_h,
                 bool hconst, const GrowP& P, FCand* out) {
    const int rpb = WORD_ROWS;   // a block's packed words then take at most WORD_ROWS rows
    const int blocks = std::max(1, (count + rpb - 1) / rpb);
    const size_t lds = (size_t)n_pairs * NBIN * 8 * (hconst ? 1 : 2);
    hipLaunchKernelGGL(k_pair_hist, dim3(blocks), dim3(256), lds, st, S, tasks, task, node, pf,
                       n_pairs, ph_g, ph_c, ph_h, rpb);
    hipLaunchKernelGGL(k_pair_scan, dim3(n_pairs), dim3(64), 0, st, S, tasks, task, ph_g, ph_c,
                       ph_h, P, out);
}

// =================================================================== tree end
// Leaves sorted by the start of their row range (one block per task).
// Leaf id of every sampled row from the final partition (all-rows samples) — reference
// apply_all (tree.cpp:43-55) restricted to the rows the tree was grown on.  Every block first
// orders the non-empty leaves by their segment start in LDS (at most MAX_NODES nodes), then
// finds each of its rows' segment by binary search.
__global__ __launch_bounds__(256) void k_nid_segments(Task* tasks) {
    const Task T = tasks[blockIdx.y];   // a copy: the stores below cannot alias its fields
    if (T.skip_end) return;
    __shared__ int st[MAX_NODES];        // segment start of non-empty leaves, -1 otherwise
    __shared__ int tab[MAX_NODES * 4];   // {seg_start, seg_count, node, buf} by start
    __shared__ int nl;
    if (threadIdx.x == 0) nl = 0;
    // multi-leaf growth: the grown leaves' row ranges, each mapped to the final leaf holding it
    const bool grown = T.n_tnodes > 0;
    const DNode* nodes = grown ? T.tnodes : T.nodes;
    const int nn = grown ? T.n_tnodes : T.n_nodes;
    for (int k = threadIdx.x; k < MAX_NODES; k += blockDim.x) {
        st[k] = -1;
        if (k < nn) {
            const DNode& n = nodes[k];
            if (n.v[N_LEAF] && n.seg_count > 0) st[k] = n.seg_start;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nn; k += blockDim.x) {
        const int s0 = st[k];
        if (s0 < 0) continue;
        int r = 0;  // starts of non-empty leaves are distinct
        for (int j = 0; j < nn; ++j) r += (st[j] >= 0 && st[j] < s0);
        const DNode& n = nodes[k];
        tab[4 * r] = s0;
        tab[4 * r + 1] = n.seg_count;
        tab[4 * r + 2] = grown ? T.tmap[k] : k;
        tab[4 * r + 3] = n.buf;
        atomicAdd(&nl, 1);
    }
    __syncthreads();
    const int L = nl;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < T.n_sample; k += gridDim.x * blockDim.x) {
        int lo = 0, hi = L - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tab[4 * mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        const int* rows = tab[4 * lo + 3] ? T.idx_tmp : T.idx;
        T.nid[rows[k]] = tab[4 * lo + 2];
    }
}

// Tree walk over the bins for every row (GOSS samples).
__global__ __launch_bounds__(256) void k_nid_walk(Shared S, Task* tasks) {
    Task& T = tasks[blockIdx.y];
    if (T.skip_end) return;
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
        const unsigned char* row = S.bins_row + (size_t)(T.row0 + i) * S.fp;
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
    if (T.skip_end || (int)blockIdx.x >= n_chunks[blockIdx.y]) return;
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

// Leaf refit and TORB prefix values (booster.cpp:591-648).  One wave per node (block x = node):
// lanes hold sweep segments, an exact int64 wave scan gives the prefix sums at every bucket cut.
__global__ __launch_bounds__(64) void k_torb_final(Task* tasks, GrowP P, double min_cover, int iid) {
    const Task T = tasks[blockIdx.y];
    if (T.skip_end) return;
    const int nn = T.n_nodes;
    const int B = T.n_buckets;
    const int lane = threadIdx.x & 63;
    const double inv = ldexp(1.0, -T.s_r), invh = ldexp(1.0, -T.s_rh);
    for (int k = blockIdx.x; k < nn; k += gridDim.x) {
        // full sums over segments 0..B
        long long fg = 0, fh = 0, fc = 0;
        for (int s = lane; s <= B; s += 64) {
            const size_t o = (size_t)s * MAX_NODES + k;
            fg += T.tg[o];
            fh += T.hconst ? 0 : T.th[o];
            fc += T.tc[o];
        }
        fg = wave_sum_ll(fg);
        fh = wave_sum_ll(fh);
        fc = wave_sum_ll(fc);
        const double Hf = T.hconst ? (double)fc : (double)fh * invh;
        const double full = leaf_value((double)fg * inv, Hf, P.l2, P.l1, P.mds);
        if (lane == 0) {
            T.values[k] = full;
            T.cover[k] = (int)fc;
        }
        if (iid) {
            for (int b = lane; b < B; b += 64) T.prefix[(size_t)b * nn + k] = full;
        }
        long long cg = 0, ch = 0, cc = 0;  // carry of the prefix across 64-segment chunks
        for (int s0 = 0; s0 < (iid ? 0 : B); s0 += 64) {
            const int b = s0 + lane;
            long long xg = 0, xh = 0, xc = 0;
            if (b < B) {
                const size_t o = (size_t)b * MAX_NODES + k;
                xg = T.tg[o];
                xh = T.hconst ? 0 : T.th[o];
                xc = T.tc[o];
            }
            const long long pg = cg + wave_incl_ll(xg), ph = ch + wave_incl_ll(xh), pc = cc + wave_incl_ll(xc);
            if (b < B) {
                const double cover = (double)T.cuts[b];
                const double alpha = cover <= 0.0 ? 0.0 : fmin(1.0, cover / fmax(min_cover, 1.0));
                const double Hp = T.hconst ? (double)pc : (double)ph * invh;
                const double pref = cover > 0.0 ? leaf_value((double)pg * inv, Hp, P.l2, P.l1, P.mds) : full;
                T.prefix[(size_t)b * nn + k] = alpha * pref + (1.0 - alpha) * full;
            }
            cg = __shfl(pg, 63, 64);
            ch = __shfl(ph, 63, 64);
            cc = __shfl(pc, 63, 64);
        }
        // the node's sums are consumed: clear them for the next tree's k_torb_accum (the
        // session starts them at zero)
        for (int s = lane; s <= B; s += 64) {
            const size_t o = (size_t)s * MAX_NODES + k;
            T.tg[o] = 0;
            if (!T.hconst) T.th[o] = 0;
            T.tc[o] = 0;
        }
    }
}

// pred += lr * prefix[bucket, leaf] (booster.cpp:655-664), one chunk of PRED_CHUNK rows per
// block.  The same pass forms the chunk's share of the training loss recorded after the tree
// (objective.cpp:323-328: sum of w (pred - y)^2 and of w; booster.cpp:704-706), except for the
// pairwise objective with groups, whose loss k_loss_groups computes.  Thread sums, then waves,
// then the block's waves in order: a fixed order that depends only on the task's own rows.
__global__ __launch_bounds__(256) void k_pred_update(Task* tasks, double lr, int pairwise) {
    const Task T = tasks[blockIdx.y];   // a copy: the stores below cannot alias its fields
    const int r0 = blockIdx.x * PRED_CHUNK;
    if (T.skip_end || r0 >= T.n_rows) return;
    const int r1 = min(T.n_rows, r0 + PRED_CHUNK);
    const int nn = T.n_nodes;
    const bool mse = !(pairwise && T.has_groups);
    double s = 0.0, sw = 0.0;
    for (int i = r0 + (int)threadIdx.x; i < r1; i += blockDim.x) {
        const double p = T.pred[i] + lr * T.prefix[(size_t)T.bucket[i] * nn + T.nid[i]];
        T.pred[i] = p;
        if (mse) {
            const double wi = T.w ? T.w[i] : 1.0;
            const double d = p - T.y[i];
            s += wi * d * d;
            sw += wi;
        }
    }
    if (!mse) return;
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_down(s, o, 64);
        sw += __shfl_down(sw, o, 64);
    }
    __shared__ double ss[4], sws[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        ss[wave] = s;
        sws[wave] = sw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        T.loss_part[2 * blockIdx.x] = ((ss[0] + ss[1]) + ss[2]) + ss[3];
        T.loss_part[2 * blockIdx.x + 1] = ((sws[0] + sws[1]) + sws[2]) + sws[3];
    }
}

void launch_tree_end(hipStream_t st, const Shared& S, Task* tasks, int n_tasks, bool walk,
                     int max_rows, const TorbChunk* const* chunks, const int* n_chunks,
                     int max_chunks, const GrowP& P, double min_cover, int iid, double lr, int pairwise) {
    const int rb = std::max(1, std::min(2048, (max_rows + 1023) / 1024));
    if (walk) {
        hipLaunchKernelGGL(k_nid_walk, dim3(rb, n_tasks), dim3(256), 0, st, S, tasks);
    } else {
        hipLaunchKernelGGL(k_nid_segments, dim3(rb, n_tasks), dim3(256), 0, st, tasks);
    }
    hipLaunchKernelGGL(k_torb_accum, dim3(std::max(1, max_chunks), n_tasks), dim3(256), 0, st, tasks,
                       chunks, n_chunks);
    hipLaunchKernelGGL(k_torb_final, dim3(MAX_NODES, n_tasks), dim3(64), 0, st, tasks, P, min_cover, iid);
    hipLaunchKernelGGL(k_pred_update, dim3(std::max(1, loss_chunks(max_rows)), n_tasks), dim3(256), 0, st, tasks, lr,
                       pairwise);
}

// =================================================================== training loss
// Pairwise objective with groups (objective.cpp:291-321): one wave per group computes the
// group's Pearson correlation with the reference's sequential sums.  The wave stages the group's
// rows in LDS (their loads all in flight), then adds them in row order (every lane holds the
// same running sums), so the sums match the reference's loop.  k_export averages the groups in
// order.  The mean squared error of the other objectives comes from k_pred_update.
__global__ __launch_bounds__(64) void k_loss_groups(Task* tasks) {
    const Task& T = tasks[blockIdx.y];
    const int grp = blockIdx.x;
    if (T.skip_end || grp >= T.n_groups) return;
    constexpr int STAGE = 1024;   // rows staged in LDS at a time (all their loads in flight)
    __shared__ double ys[STAGE], ps[STAGE];
    const int g0 = T.group_start[grp], m = T.group_start[grp + 1] - g0;
    const int* gr = T.group_rows ? T.group_rows + g0 : nullptr;
    double val = __longlong_as_double(0x7ff8000000000000ll);   // NaN: the group does not count
    if (m >= 3) {
        // a lane's STAGE / 64 rows are loaded together: every row index first, then every value,
        // so each stage costs two memory round trips instead of two per 64 rows
        auto stage = [&](int base, int cnt) {
            constexpr int PER_LANE = STAGE / 64;
            __syncthreads();
            int r[PER_LANE];
#pragma unroll
            for (int u = 0; u < PER_LANE; ++u) {
                const int k = threadIdx.x + u * 64;
                r[u] = k < cnt ? (gr ? gr[base + k] : g0 + base + k) : -1;
            }
            double yv[PER_LANE], pv[PER_LANE];
#pragma unroll
            for (int u = 0; u < PER_LANE; ++u) {
                yv[u] = r[u] >= 0 ? T.y[r[u]] : 0.0;
                pv[u] = r[u] >= 0 ? T.pred[r[u]] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < PER_LANE; ++u) {
                const int k = threadIdx.x + u * 64;
                if (k < cnt) {
                    ys[k] = yv[u];
                    ps[k] = pv[u];
                }
            }
            __syncthreads();
        };
        double my = 0.0, mp = 0.0;
        for (int base = 0; base < m; base += STAGE) {
            const int cnt = min(STAGE, m - base);
            stage(base, cnt);
            for (int k = 0; k < cnt; ++k) {
                my += ys[k];
                mp += ps[k];
            }
        }
        my /= m;
        mp /= m;
        double vy = 0.0, vp = 0.0, c = 0.0;
        for (int base = 0; base < m; base += STAGE) {
            const int cnt = min(STAGE, m - base);
            if (base > 0 || m > STAGE) stage(base, cnt);   // a group of <= STAGE rows is still staged
            for (int k = 0; k < cnt; ++k) {
                const double dy = ys[k] - my, dp = ps[k] - mp;
                vy += dy * dy;
                vp += dp * dp;
                c += dy * dp;
            }
        }
        if (vy > 1e-18 && vp > 1e-18) val = c / sqrt(vy * vp);
    }
    if (threadIdx.x == 0) T.loss_part[2 * loss_chunks(T.cap_rows) + grp] = val;
}

void launch_loss(hipStream_t st, Task* tasks, int n_tasks, int max_groups) {
    hipLaunchKernelGGL(k_loss_groups, dim3(std::max(1, max_groups), n_tasks), dim3(64), 0, st, tasks);
}

// Finished trees of all active tasks packed into one staging buffer (one copy to the host):
// task i at out + i*L.stride holds n_nodes, nodes[0..nn), values[0..nn), cover[0..nn) and
// prefix[0..nn*n_buckets) at the offsets of L.
// The header also carries the tree's training loss: the in-order mean of the group correlations
// (pairwise with groups; minus their mean, as objective.cpp:320-321 reports it), else the chunk
// sums of k_pred_update added in a fixed order.
__global__ __launch_bounds__(256) void k_export(Task* tasks, ExportLayout L, char* out, int pairwise) {
    Task& T = tasks[blockIdx.x];
    if (T.skip_end) return;
    char* base = out + (size_t)blockIdx.x * L.stride;
    const int nn = T.n_nodes;
    double loss = 0.0;
    if (pairwise && T.has_groups) {
        // the block stages 2048 group values at a time in LDS (all loads in flight at once);
        // thread 0 adds them in group order
        constexpr int STAGE = 2048;
        __shared__ double gvs[STAGE];
        const double* gv = T.loss_part + 2 * loss_chunks(T.cap_rows);
        const int ngr = T.n_groups;
        double acc = 0.0;
        int ng = 0;
        constexpr int PER_THREAD = STAGE / 256;   // launched with 256 threads
        for (int base = 0; base < ngr; base += STAGE) {
            const int cnt = min(STAGE, ngr - base);
            double v[PER_THREAD];   // a thread's loads issued together, then stored
#pragma unroll
            for (int u = 0; u < PER_THREAD; ++u) {
                const int j = threadIdx.x + u * 256;
                v[u] = j < cnt ? gv[base + j] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < PER_THREAD; ++u) {
                const int j = threadIdx.x + u * 256;
                if (j < cnt) gvs[j] = v[u];
            }
            __syncthreads();
            if (threadIdx.x == 0)
                for (int j = 0; j < cnt; ++j) {
                    const double v = gvs[j];
                    if (v == v) {
                        acc += v;
                        ng += 1;
                    }
                }
            __syncthreads();
        }
        loss = ng > 0 ? -acc / ng : 0.0;
    } else {
        double a = 0.0, b = 0.0;
        const int nb = loss_chunks(T.n_rows);
        for (int j = threadIdx.x; j < nb; j += blockDim.x) {
            a += T.loss_part[2 * j];
            b += T.loss_part[2 * j + 1];
        }
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_down(a, o, 64);
            b += __shfl_down(b, o, 64);
        }
        __shared__ double sa[4], sb[4];
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) {
            sa[wave] = a;
            sb[wave] = b;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const double s = ((sa[0] + sa[1]) + sa[2]) + sa[3], sw = ((sb[0] + sb[1]) + sb[2]) + sb[3];
            loss = sw > 0 ? s / sw : 0.0;
        }
    }
    if (threadIdx.x == 0) {
        T.train_loss = loss;
        *(int*)base = nn;
        *((int*)base + 1) = T.err;
        *((double*)base + 1) = loss;   // header bytes 8..15
    }
    const int* src = (const int*)T.nodes;
    int* dst = (int*)(base + L.off_nodes);
    const int words = nn * (int)(sizeof(DNode) / sizeof(int));
    for (int j = threadIdx.x; j < words; j += blockDim.x) dst[j] = src[j];
    double* vals = (double*)(base + L.off_values);
    int* cov = (int*)(base + L.off_cover);
    for (int j = threadIdx.x; j < nn; j += blockDim.x) {
        vals[j] = T.values[j];
        cov[j] = T.cover[j];
    }
    double* pre = (double*)(base + L.off_prefix);
    const int np = nn * T.n_buckets;
    for (int j = threadIdx.x; j < np; j += blockDim.x) pre[j] = T.prefix[j];
}

void launch_export(hipStream_t st, Task* tasks, int n_tasks, const ExportLayout& L, char* out, int pairwise) {
    hipLaunchKernelGGL(k_export, dim3(n_tasks), dim3(256), 0, st, tasks, L, out, pairwise);
}

// Dynamic LDS above 64 KiB needs a per-kernel opt-in; done once, outside any stream capture.
void init_kernel_attributes() {
    static bool done = false;
    if (done) return;
    // dynamic LDS opt-ins: what each kernel is launched with at most (the multi-leaf child
    // kernels also hold their slot table in static LDS, and static plus dynamic must fit 160 KB)
    auto opt_in = [](const void* f, int bytes) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) throw std::runtime_error(std::string("kernel LDS attribute: ") + hipGetErrorString(e));
    };
    opt_in((const void*)k_hist_row<1, false, 4>, 128 * 1024);
    opt_in((const void*)k_hist_row<1, false, 2>, 128 * 1024);
    opt_in((const void*)k_hist_row<1, true, 1>, 128 * 1024);
    opt_in((const void*)k_hist_row<2, false, 1>, 128 * 1024);
    opt_in((const void*)k_hist_spec<false>, 128 * 1024);
    opt_in((const void*)k_hist_spec<true>, 64 * 1024);
    opt_in((const void*)k_pair_hist, 160 * 1024);
    done = true;
}

}  // namespace tgbm
