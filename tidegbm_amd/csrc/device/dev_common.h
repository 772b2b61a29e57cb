// Shared device-side definitions for the gfx950 histogram-GBDT kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tgbm {

// ---------------------------------------------------------------- sizes
constexpr int NBIN = 256;          // histogram stride per feature (uint8 bins: 0 = missing)
constexpr int MAX_NODES = 256;     // per tree (max_leaves <= 128)
// n_edges[f] with this bit set: feature f is categorical and its "edges" are its sorted category
// values; a value bins to 1 + its index, an unlisted value or NaN to 0 (binning.cpp:107-112).
constexpr int CAT_EDGE_FLAG = 1 << 16;
constexpr int MAX_FEATURES = 1024;
constexpr int MAX_BUCKETS = 256;   // TORB buckets (+1 sweep segment for rows past the last cut)
constexpr int MAX_RED_BLOCKS = 2048;

// ---------------------------------------------------------------- fixed point
// Histogram sums are exact integer sums of per-row quantized values, so they are independent
// of summation order (deterministic under any atomic schedule) and sibling subtraction is
// exact.  g is quantized as q = rint(g * 2^s_g) with |q| <= 2^qg, qg = min(QG_BITS, 62 -
// ceil(log2 rows)), so a node's int64 total cannot wrap; h (when not identically 1) likewise
// with QH_BITS.  In LDS one 64-bit word per (feature, bin) packs count << CNT_SHIFT | q-sum
// (53-bit two's complement), valid while the word takes at most WORD_ROWS rows: |sum| <=
// 2^10 * 2^41 < 2^52 and count <= 1024 < 2^11.  Blocks that fold more rows keep an int64 g word
// and a 32-bit count word per bin instead (the "wide" form).
constexpr int CNT_SHIFT = 53;
constexpr int QG_BITS = 41;
constexpr int WORD_ROWS = 1024;
constexpr int QH_BITS = 39;
constexpr int QR_BITS = 39;        // refit / TORB sums over all rows (plain int64)
constexpr int QE_BITS = 39;        // leaf energy sum of g^2
constexpr int MAX_BLOCK_ROWS = 32768;   // rows one histogram block may fold (h sums: 2^15 * 2^39)
constexpr int PRED_CHUNK = 1024;   // rows per block of the prediction update and its loss partial
__host__ __device__ inline int loss_chunks(int rows) { return (rows + PRED_CHUNK - 1) / PRED_CHUNK; }
constexpr unsigned long long CNT_ONE = 1ull << CNT_SHIFT;
static_assert((1LL << (QG_BITS + 10)) < (1LL << (CNT_SHIFT - 1)) && WORD_ROWS < (1 << (64 - CNT_SHIFT)),
              "packed word: q-sum and count of WORD_ROWS rows must fit their fields");

// Bits per row of a fixed-point quantity summed over `rows` rows in int64: at most `cap`, and
// rows * 2^bits <= 2^62.
__host__ __device__ inline int sum_bits(int cap, long long rows) {
    int l = 0;
    while ((1LL << l) < rows) ++l;
    return cap < 62 - l ? cap : 62 - l;
}

__host__ __device__ inline long long unpack_q(unsigned long long v) {
    return (long long)(v << (64 - CNT_SHIFT)) >> (64 - CNT_SHIFT);
}
__host__ __device__ inline unsigned int unpack_c(unsigned long long v) {
    const long long q = unpack_q(v);
    return (unsigned int)((v - (unsigned long long)q) >> CNT_SHIFT);
}

// Exponent e with |x| < 2^e for the largest magnitude x (bits of a non-negative double).
__host__ __device__ inline int mag_exponent(unsigned long long max_bits) {
    const double m = __longlong_as_double((long long)max_bits);
    if (!(m > 0.0)) return -1074;
    int e;
    frexp(m, &e);  // m = f * 2^e, f in [0.5, 1)  =>  m < 2^e
    return e;
}

// ---------------------------------------------------------------- tree records
// Node layout: v[] follows the export order {is_leaf, split_type, feature, feature2, thr,
// left, right, missing_go_left, n_samples, depth} of reference Node (tree.hpp:11-25).
struct DNode {
    int v[10];
    int seg_start, seg_count;   // rows of this node in the task's partitioned index array
    int buf;                    // which of the two index buffers holds the node's rows
    int pad1;
    double value, gain;
};
enum { N_LEAF = 0, N_TYPE, N_FEAT, N_FEAT2, N_THR, N_LEFT, N_RIGHT, N_MGL, N_NS, N_DEPTH };

// One leaf split of a multi-leaf step (speculative best-first growth, one task per session):
// written by the step's selection, then read by its partition, histogram and scan launches.
constexpr int SPEC_W = 8;   // leaves split per step at most
struct Slot {
    int leaf, left;             // the split leaf; its children get ids left and left + 1
    int start, count, buf;      // the leaf's rows: range in index buffer `buf`
    int type, feat, feat2, thr, mgl;
    int nl;                     // rows the split sends left (from the leaf's histogram counts)
    int chunk0;                 // the slot's first partition block in the step's launch
    int blk0, gx, per;          // first histogram row block, row blocks, rows per block
    int small_node, big_node, small_start, small_count, small_buf;
    int hist_parts, hist_wide;  // partials of the slot's histogram blocks and their form
    int pad;
    double gain;
    unsigned long long part_lr; // partition reservations: right << 32 | left
};
// kernels copy the slot table into LDS as 4-byte words, one per thread of a 256-thread block
static_assert(sizeof(Slot) % 8 == 0 && sizeof(Slot) / 4 * SPEC_W <= 256, "slot table copy");

struct FCand {      // best split of one feature of one node
    double gain;
    int thr, ml;
    int cl, pad;    // rows the split sends left (exact, from the node's histogram counts)
};

struct DCand {      // cached leaf candidate
    double gain;
    int feature, feature2, thr, type, mgl, valid;
    int cl, pad;    // rows the split sends left
};

// Byte offsets of one task's finished tree in the export staging buffer.
struct ExportLayout {
    size_t off_nodes, off_values, off_cover, off_prefix, stride;
};
__host__ inline ExportLayout export_layout(int max_nodes, int max_buckets) {
    ExportLayout L;
    auto up = [](size_t x, size_t a) { return (x + a - 1) / a * a; };
    L.off_nodes = 16;
    L.off_values = up(L.off_nodes + sizeof(DNode) * max_nodes, 8);
    L.off_cover = L.off_values + sizeof(double) * max_nodes;
    L.off_prefix = up(L.off_cover + sizeof(int) * max_nodes, 8);
    L.stride = up(L.off_prefix + sizeof(double) * max_nodes * max_buckets, 256);
    return L;
}

// Per-task state and pointers, resident in device memory (one entry per task).
struct Task {
    // ---- row window (set at creation or by a later bind)
    int n_rows;            // this task uses rows [row0, row0 + n_rows) of the shared matrix
    int row0;              // local row i of every per-row array is shared-matrix row row0 + i
    int cap_rows;          // capacity of the per-row arrays
    int pad_r;
    int n_buckets;
    int hconst;            // h == 1 for every row
    int has_groups;
    // ---- per tree (host resets at tree begin)
    int n_sample;
    int contiguous;        // sample == rows [0, n_sample)
    int n_nodes;
    int done;
    int s_g, s_h;          // quantization exponents for histogram g/h
    int s_r, s_e;          // refit/TORB (g) and energy exponents
    int s_rh, pad_s;       // refit/TORB exponent for h
    double goss_scale;
    int goss;              // 1 when some rows use scaled g/h
    int pending0, pending1; // nodes scanned in the last step (-1 = none) awaiting reduction
    int pad_p;
    // ---- current split
    int cur_leaf, cur_type, cur_feat, cur_feat2, cur_thr, cur_mgl;
    double cur_gain;
    int n_left;
    int pad_a;
    // ---- step descriptor: the split leaf's rows (set with the decision) and the children
    // (set by the partition), so kernels read one cache line instead of chasing nodes[]
    int cur_start, cur_count, cur_buf, cur_left;
    int small_node, big_node, small_start, small_count, small_buf, pad_c;
    unsigned long long part_lr; // rows placed by the running partition: right << 32 | left
    int cur_nl;                 // rows the current split sends left (-1: not known in advance)
    int err;                    // nonzero: an internal consistency check failed in this tree
    int hist_parts;            // partial histograms written by the last histogram launch
    int hist_wide;             // how those partials hold the sums: 0 packed words; 1 int64 g sums in
                               // part with counts in part_c; 2 the packed words' sums modulo 2^64
                               // in part, counts from root_cnt; 3 as 1, and the build was the
                               // task's whole-window root, so the scan records its counts
    int cnt_row0, cnt_n;       // the row window root_cnt holds counts of (cnt_n = -1: none)
    double train_loss;         // training loss after the last finished tree (exported)
    double* loss_part;         // [2 * loss_chunks(cap_rows)] per-chunk (sum w d^2, sum w), then
                               // [n_groups] per-group correlations (pairwise)
    // ---- reductions
    unsigned long long max_g_bits, max_h_bits;
    unsigned long long* red;   // per-block maxima of |g| and |h| [2][MAX_RED_BLOCKS]
    int n_red;                 // blocks that wrote red[] in the last gradient launch
    int scan_ticket;           // completion counter of the scan launch (last block selects)
    int pad_t;
    // ---- per-row arrays (length n_rows)
    const double* y;
    const double* w;       // may be null (all 1)
    double* pred;
    double* g;
    double* h;
    unsigned long long* pk;   // packed (1 << CNT_SHIFT) + q_g
    unsigned long long* qh;   // q_h (only when !hconst)
    const int* bucket;
    const unsigned char* seg; // TORB sweep segment per row
    const unsigned char* scaled; // GOSS flag per row (may be null)
    int* nid;
    int* idx;              // partitioned sample rows (buffer 0)
    int* idx_tmp;          // buffer 1: a split moves the parent's rows into the other buffer
    const int* group_start; // pairwise groups: positions [gs[k], gs[k+1]) of group_rows
    const int* group_rows;  // rows in stable group order (null: identity, groups contiguous)
    int n_groups;
    int pad_b;
    // ---- partition scratch
    int* chunk_cnt;
    // ---- tree state
    DNode* nodes;          // [MAX_NODES]
    DCand* cands;          // [MAX_NODES]
    FCand* fcand;          // [MAX_NODES][n_features]
    long long* hg;         // [MAX_NODES][F][NBIN]
    int* hc;               // [MAX_NODES][F][NBIN]
    long long* hh;         // [MAX_NODES][F][NBIN] (only when !hconst)
    long long* energy;     // [MAX_NODES] fixed-point sum of g^2 (host-driven mode)
    unsigned long long* part;    // [part_cap][F][NBIN] per-block partial histograms (packed, or
                                 // int64 g sums when the launch was wide)
    unsigned int* part_c;        // [part_cap][F][NBIN] partial counts of wide launches
    unsigned long long* part_h;  // [part_cap][F][NBIN] partial h sums (only when !hconst)
    int* root_cnt;               // [F][NBIN] per-bin row counts of the task's whole-window root
    // ---- TORB
    const int* cuts;       // [n_buckets]
    long long* tg;         // [(MAX_BUCKETS+1) * MAX_NODES]
    long long* th;
    int* tc;
    double* values;        // [MAX_NODES]
    double* prefix;        // [n_buckets * MAX_NODES]
    int* cover;            // [MAX_NODES]
    // ---- categorical splits (null when the session has no categorical feature)
    unsigned long long* fmask;  // [MAX_NODES][n_features][4] bins a feature's best split sends left
    unsigned long long* nmask;  // [MAX_NODES][4] the split node's mask (exported with the tree)
    // ---- multi-leaf growth (null slots: one split per step)
    Slot* slots;           // [SPEC_W] the current step's splits (this task's slot_tab)
    DNode* tnodes;         // [MAX_NODES] the grown nodes in creation order (kept by the finalize)
    int* tmap;             // [MAX_NODES] the final leaf holding each grown leaf's rows
    int n_slots;           // splits of the current step
    int spec_steps;        // steps taken in this tree
    int pend_first, n_pend; // nodes [pend_first, pend_first + n_pend) scanned, not yet reduced
    int n_tnodes;          // grown nodes (0: the tree grew one split per step)
    int spec_cap;          // nodes the task's histogram buffers hold
    int* host_flag;        // mapped host memory: {done, steps} written when a tree's growth ends
    int skip_end;          // the tree's growth was not finished: the tree-end kernels do nothing
    int rs_valid;          // the replay's frontier below is this tree's (0: replay from the root)
    int rs_front, rs_final; // frontier entries and replay ids used at the last stop
    int spec_gb;           // histogram row blocks of the current step (all slots)
    int end_done;          // this tree's end has run (a later end of the same tree does nothing)
    int* rs_node;          // [64] frontier nodes at the last stop (-1: empty entry)
    int* rs_sim;           // [64] their replay ids
    // The step's slot table inside the descriptor: a kernel that knows the task's index reaches a
    // slot at a fixed offset from it, without waiting for the slots pointer first.
    Slot slot_tab[SPEC_W];
};

struct Shared {            // the session-wide matrix
    int n_rows, n_features, fp, n_pad, n_bins;
    int part_cap;          // per-task capacity of per-block partial histograms
    const unsigned char* bins_row;   // [n_rows][fp]
    const unsigned char* bins_col;   // [n_features][n_pad]
    const unsigned char* is_cat;     // [n_features] categorical flags, or null
    const signed char* monotone;     // [n_features] monotone constraints (+1 / -1 / 0), or null
    int* spec_ctr;                   // [8] item counters of multi-leaf child builds, one per XCD
};

// A run of rows inside one TORB sweep segment (rows are time sorted, segment s = rows
// [cut_{s-1}, cut_s), s = n_buckets for rows past the last cut).
struct TorbChunk {
    int r0, r1, s, pad;
};

// Grow parameters copied to kernels by value.
struct GrowP {
    int max_depth, max_leaves, min_leaf;
    double min_hess, l2, l1, mds, gamma, pair_l2, cat_smooth;
};

// ---------------------------------------------------------------- leaf formulas
// Reference tree.hpp:43-67 and histogram.cpp:119-142 (no FMA: -ffp-contract=off).
__device__ inline double soft_thr(double G, double l1) {
    // selects rather than branches, so candidate evaluations stay straight-line code
    const double t = G > l1 ? G - l1 : (G < -l1 ? G + l1 : 0.0);
    return l1 > 0 ? t : G;
}
__device__ inline double leaf_score(double G, double H, double lam, double l1) {
    const double t = soft_thr(G, l1);
    return (t * t) / (H + lam);
}
// Leaf weight used by the monotone check of the split scan (histogram.cpp:136-142; no
// max_delta_step clip there).
__device__ inline double weight_of(double G, double H, double lam, double l1) {
    const double den = H + lam;
    if (l1 <= 0) return -G / den;
    if (G > l1) return -(G - l1) / den;
    if (G < -l1) return -(G + l1) / den;
    return 0.0;
}
__device__ inline double leaf_value(double G, double H, double l2, double l1, double mds) {
    double v;
    if (l1 > 0) {
        if (G > l1) v = -(G - l1) / (H + l2);
        else if (G < -l1) v = -(G + l1) / (H + l2);
        else v = 0.0;
    } else {
        v = -G / (H + l2);
    }
    if (mds > 0) {
        if (v > mds) v = mds;
        if (v < -mds) v = -mds;
    }
    return v;
}

// ---------------------------------------------------------------- wave helpers (wave64)
// Lane moves through DPP (data-parallel primitives of the VALU, a few cycles each) instead of
// LDS-routed shuffles.  Every helper must be called by all 64 lanes of the wave.  Controls
// (GFX9 encoding): row_shr:n = 0x110 + n within rows of 16 lanes; row_bcast:15 = 0x142 (lane 15
// of each row into the next row); row_bcast:31 = 0x143 (lane 31 into rows 2 and 3).  A lane
// whose source is outside its row, or whose row is masked off, receives `id`.
template <int CTRL, int ROW_MASK>
__device__ inline int dpp_i(int v, int id) {
    return __builtin_amdgcn_update_dpp(id, v, CTRL, ROW_MASK, 0xf, false);
}
template <int CTRL, int ROW_MASK>
__device__ inline unsigned long long dpp_u64(unsigned long long v, unsigned long long id) {
    const unsigned lo = (unsigned)dpp_i<CTRL, ROW_MASK>((int)(unsigned)v, (int)(unsigned)id);
    const unsigned hi = (unsigned)dpp_i<CTRL, ROW_MASK>((int)(unsigned)(v >> 32), (int)(unsigned)(id >> 32));
    return (unsigned long long)hi << 32 | lo;
}
__device__ inline unsigned long long read_lane_u64(unsigned long long v, int lane) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), lane);
    return (unsigned long long)hi << 32 | lo;
}
// Lane `lane`'s double, in every lane (lane must be wave-uniform).
__device__ inline double read_lane_f64(double v, int lane) {
    return __longlong_as_double((long long)read_lane_u64((unsigned long long)__double_as_longlong(v), lane));
}
// Inclusive scan over the lanes with an associative operation (Hillis-Steele in each row, then
// the row totals): lane 63 ends with the whole wave's result.
template <class Op>
__device__ inline unsigned long long wave_scan_u64(unsigned long long x, unsigned long long id, Op op) {
    x = op(x, dpp_u64<0x111, 0xf>(x, id));
    x = op(x, dpp_u64<0x112, 0xf>(x, id));
    x = op(x, dpp_u64<0x114, 0xf>(x, id));
    x = op(x, dpp_u64<0x118, 0xf>(x, id));
    x = op(x, dpp_u64<0x142, 0xa>(x, id));
    x = op(x, dpp_u64<0x143, 0xc>(x, id));
    return x;
}
template <class Op>
__device__ inline int wave_scan_i(int x, int id, Op op) {
    x = op(x, dpp_i<0x111, 0xf>(x, id));
    x = op(x, dpp_i<0x112, 0xf>(x, id));
    x = op(x, dpp_i<0x114, 0xf>(x, id));
    x = op(x, dpp_i<0x118, 0xf>(x, id));
    x = op(x, dpp_i<0x142, 0xa>(x, id));
    x = op(x, dpp_i<0x143, 0xc>(x, id));
    return x;
}
struct OpAdd64 {
    __device__ unsigned long long operator()(unsigned long long a, unsigned long long b) const { return a + b; }
};
struct OpMax64 {
    __device__ unsigned long long operator()(unsigned long long a, unsigned long long b) const { return a > b ? a : b; }
};
struct OpAddI {
    __device__ int operator()(int a, int b) const { return a + b; }
};
struct OpMinI {
    __device__ int operator()(int a, int b) const { return a < b ? a : b; }
};
// Inclusive prefix over lanes (two's complement sums: signed and unsigned adds agree).
__device__ inline long long wave_incl_ll(long long v) {
    return (long long)wave_scan_u64((unsigned long long)v, 0ull, OpAdd64());
}
__device__ inline int wave_incl_i(int v) { return wave_scan_i(v, 0, OpAddI()); }
// Totals, the same in every lane.
__device__ inline long long wave_sum_ll(long long v) { return (long long)read_lane_u64((unsigned long long)wave_incl_ll(v), 63); }
__device__ inline int wave_sum_i(int v) { return __builtin_amdgcn_readlane(wave_incl_i(v), 63); }
__device__ inline unsigned long long wave_max_u64(unsigned long long v) {
    return read_lane_u64(wave_scan_u64(v, 0ull, OpMax64()), 63);
}
__device__ inline int wave_min_i(int v) {
    return __builtin_amdgcn_readlane(wave_scan_i(v, 0x7fffffff, OpMinI()), 63);
}
// Order-preserving 64-bit key of a double (larger double, larger key; -0 and +0 equal); NaN
// maps to 0, below every number.
__device__ inline unsigned long long gain_key(double d) {
    d = d + 0.0;
    if (d != d) return 0ull;
    const unsigned long long u = (unsigned long long)__double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
// The lane holding the wave's best candidate: largest gain, then smallest key.
__device__ inline int wave_best_lane(double gain, int key) {
    const unsigned long long gk = gain_key(gain);
    const unsigned long long m = wave_max_u64(gk);
    const int kmin = wave_min_i(gk == m ? key : 0x7fffffff);
    return (int)__builtin_ctzll(__ballot(gk == m && key == kmin));
}

// Global-segment view of a device pointer.  Pointers read from memory (the Task descriptors)
// are generic to the compiler, and generic accesses also count against the LDS/scalar counter,
// so every LDS wait would wait for them as well; hot kernels access through this view.
template <class T>
__device__ inline __attribute__((address_space(1))) T* gbl(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

// Load / store of a whole struct or vector through the global segment (4-byte pieces, which the
// compiler merges into wide accesses when the type's alignment allows).
template <class T>
__device__ inline T ldg(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "4-byte pieces");
    T v;
    const auto s = (const __attribute__((address_space(1))) unsigned*)p;
    unsigned* d = (unsigned*)&v;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = s[i];
    return v;
}
template <class T>
__device__ inline void stg(T* p, const T& v) {
    static_assert(sizeof(T) % 4 == 0, "4-byte pieces");
    const auto d = (__attribute__((address_space(1))) unsigned*)p;
    const unsigned* s = (const unsigned*)&v;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = s[i];
}

// Candidate ordering: larger gain first, then smaller key (feature/bin/missing order).
__device__ inline bool cand_better(double ga, int ka, double gb, int kb) {
    return ga > gb || (ga == gb && ka < kb);
}

__device__ inline const int* node_rows(const Task& T, const DNode& n) {
    return n.buf ? T.idx_tmp : T.idx;
}

}  // namespace tgbm
