// GOSS row selection on the device (reference booster.cpp:554-585): rows ordered by |g|
// descending, the first ceil(top * S) kept, then ceil(other * S) rows drawn from the rest in the
// order std::shuffle leaves them, with g and h of the drawn rows scaled.
//
// The reference orders with std::sort, which leaves the order of equal |g| unspecified, and
// draws with std::shuffle on its mt19937.  Here the rows are ordered by a radix sort of the |g|
// bit patterns (non-negative doubles order like their bits) and the draw is the permutation that
// the same std::shuffle call applies to the rest (computed on the host from the same generator
// state: it depends on the rest's size only).  When no two rows share a |g| value, the sorted
// order is unique and the selection is the reference's exactly; the launch reports ties so the
// caller can fall back to the host path, which uses the same std::sort.
//
// The sort is a least-significant-digit radix sort of (key, row) pairs, 8 bits per pass, each
// pass stable: per 4096-key tile a digit count, one exclusive scan over (digit, tile) in digit
// order, then a scatter in which every wave ranks its keys in order (lanes with the same digit
// found by ballots over the digit's bits) and the tile, ordered by digit in LDS, is written out
// in runs of consecutive destinations.  Digits are taken from the complemented key, so the
// ascending sort puts |g| in descending order and equal keys stay in row order.  A pass whose
// digit is the same for every key is skipped; which passes run, and so which buffer holds the
// result, follows from the OR and AND of all keys, computed on the device first.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dev_common.h"
#include "launch.h"

namespace tgbm {

namespace {

constexpr int RS_TILE = 4096;      // keys per block of the count and scatter kernels
constexpr int RS_THREADS = 256;    // 4 waves, 1024 keys each
constexpr int RS_PASSES = 8;

__device__ inline int rs_digit(unsigned long long key, int shift) { return 255 - (int)((key >> shift) & 255ull); }

// Number of passes before pass `pass` (0 .. RS_PASSES) whose digit varies over the keys; m holds
// the OR and the AND of all keys.
__device__ inline int rs_passes_before(const unsigned long long* m, int pass) {
    const unsigned long long v = m[0] ^ m[1];
    int before = 0;
    for (int q = 0; q < pass; ++q) before += ((v >> (8 * q)) & 255ull) != 0;
    return before;
}
// Whether pass `pass` (< RS_PASSES) runs, and how many passes run before it.
__device__ inline bool rs_pass_on(const unsigned long long* m, int pass, int& before) {
    before = rs_passes_before(m, pass);
    return (((m[0] ^ m[1]) >> (8 * pass)) & 255ull) != 0;
}

constexpr int MAX_KEY_BLOCKS = 4096;

// The lanes of the wave holding a key (ok) with the same 8-bit digit as this lane: the AND of one
// ballot per digit bit.  Every lane of the wave must call it.
__device__ inline unsigned long long rs_match(bool ok, int dg) {
    unsigned long long same = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (dg >> b) & 1;
        const unsigned long long bb = __ballot(ok && bit);
        same &= bit ? bb : ~bb;
    }
    return same;
}

}  // namespace

__global__ __launch_bounds__(256) void k_goss_keys(const double* __restrict__ g, int n,
                                                   unsigned long long* __restrict__ keys, int* __restrict__ vals,
                                                   unsigned char* __restrict__ keep,
                                                   unsigned char* __restrict__ scaled, unsigned long long* part) {
    unsigned long long o = 0, a = ~0ull;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned long long k = (unsigned long long)__double_as_longlong(fabs(g[i]));
        keys[i] = k;
        vals[i] = i;
        keep[i] = 0;
        scaled[i] = 0;
        o |= k;
        a &= k;
    }
    // the block's OR and AND of its keys: part[block], part[MAX_KEY_BLOCKS + block]
    o = read_lane_u64(wave_scan_u64(o, 0ull, [](unsigned long long x, unsigned long long y) { return x | y; }), 63);
    a = read_lane_u64(wave_scan_u64(a, ~0ull, [](unsigned long long x, unsigned long long y) { return x & y; }), 63);
    __shared__ unsigned long long so[4], sa[4];
    if ((threadIdx.x & 63) == 0) {
        so[threadIdx.x >> 6] = o;
        sa[threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = so[0] | so[1] | so[2] | so[3];
        part[MAX_KEY_BLOCKS + blockIdx.x] = sa[0] & sa[1] & sa[2] & sa[3];
    }
}

// mask[0] = OR, mask[1] = AND of the blocks' partial results.
__global__ __launch_bounds__(1024) void k_goss_mask(const unsigned long long* part, int blocks,
                                                    unsigned long long* mask) {
    unsigned long long o = 0, a = ~0ull;
    for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
        o |= part[b];
        a &= part[MAX_KEY_BLOCKS + b];
    }
    o = read_lane_u64(wave_scan_u64(o, 0ull, [](unsigned long long x, unsigned long long y) { return x | y; }), 63);
    a = read_lane_u64(wave_scan_u64(a, ~0ull, [](unsigned long long x, unsigned long long y) { return x & y; }), 63);
    __shared__ unsigned long long so[16], sa[16];
    if ((threadIdx.x & 63) == 0) {
        so[threadIdx.x >> 6] = o;
        sa[threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 16; ++w) {
            o |= so[w];
            a &= sa[w];
        }
        mask[0] = o;
        mask[1] = a;
    }
}

// Digit counts of one tile: counts[digit * n_tiles + tile].  Each lane loads its 16 keys first
// (one memory round trip), then the waves count them.
__global__ __launch_bounds__(RS_THREADS) void k_rs_count(const unsigned long long* __restrict__ ka,
                                                         const unsigned long long* __restrict__ kb, int n, int pass,
                                                         const unsigned long long* mask, unsigned int* counts,
                                                         int n_tiles) {
    int before;
    if (!rs_pass_on(mask, pass, before)) return;
    const unsigned long long* src = (before & 1) ? kb : ka;
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    const int t0 = blockIdx.x * RS_TILE;
    constexpr int KPT = RS_TILE / RS_THREADS;
    unsigned long long key[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        const int i = t0 + j * RS_THREADS + (int)threadIdx.x;
        key[j] = i < n ? src[i] : 0ull;
    }
    __syncthreads();
    const unsigned long long below = (1ull << (threadIdx.x & 63)) - 1ull;
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        // one atomic per group of lanes with the same digit (digits cluster in the high bits)
        const bool ok = t0 + j * RS_THREADS + (int)threadIdx.x < n;
        const int dg = ok ? rs_digit(key[j], 8 * pass) : 0;
        const unsigned long long same = rs_match(ok, dg);
        if (ok && (same & below) == 0) atomicAdd(&h[dg], (unsigned int)__popcll(same));
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

// Per digit d (one block each): exclusive prefix sums over the tiles of counts[d * n_tiles + t] in
// place, and the digit's total in totals[d].
__global__ __launch_bounds__(1024) void k_rs_scan_digits(unsigned int* counts, int n_tiles, int pass,
                                                         const unsigned long long* mask, unsigned int* totals) {
    int before;
    if (!rs_pass_on(mask, pass, before)) return;
    unsigned int* c = counts + (size_t)blockIdx.x * n_tiles;
    __shared__ unsigned int wt[16];
    const int wave = threadIdx.x >> 6;
    unsigned int carry = 0;
    for (int c0 = 0; c0 < n_tiles; c0 += 1024) {
        const int j = c0 + (int)threadIdx.x;
        const unsigned int v = j < n_tiles ? c[j] : 0u;
        const int incl = wave_incl_i((int)v);
        if ((threadIdx.x & 63) == 63) wt[wave] = (unsigned int)incl;
        __syncthreads();
        unsigned int before_w = 0, chunk = 0;
        for (int w = 0; w < 16; ++w) {
            before_w += w < wave ? wt[w] : 0u;
            chunk += wt[w];
        }
        if (j < n_tiles) c[j] = carry + before_w + (unsigned int)incl - v;
        carry += chunk;
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Exclusive prefix sums of c[0 .. m) in place by one block (pass < 0: always; else only when the
// pass runs); *total, when given, receives the sum.
__global__ __launch_bounds__(1024) void k_rs_scan(unsigned int* c, int m, int pass, const unsigned long long* mask,
                                                  int* total) {
    if (pass >= 0) {
        int before;
        if (!rs_pass_on(mask, pass, before)) return;
    }
    const int per = (m + 1023) / 1024;
    const int j0 = min(m, (int)threadIdx.x * per), j1 = min(m, j0 + per);
    unsigned int s = 0;
    for (int j = j0; j < j1; ++j) s += c[j];
    // exclusive scan of the per-thread sums: waves, then the 16 wave totals
    const int incl = wave_incl_i((int)s);
    __shared__ int wt[16];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wt[wave] = incl;
    __syncthreads();
    if (threadIdx.x < 64) {
        const int v = threadIdx.x < 16 ? wt[threadIdx.x] : 0;
        const int iv = wave_incl_i(v);
        if (threadIdx.x < 16) wt[threadIdx.x] = iv - v;
        if (threadIdx.x == 15 && total) *total = iv;
    }
    __syncthreads();
    unsigned int run = (unsigned int)(wt[wave] + incl) - s;
    for (int j = j0; j < j1; ++j) {
        const unsigned int v = c[j];
        c[j] = run;
        run += v;
    }
}

// Stable scatter of one tile.  Wave w holds keys w*1024 .. w*1024+1023 of the tile (16 per lane,
// loaded first) and ranks them in order: a key's place among the wave's keys of its digit is the
// count of that digit in the wave's earlier steps plus its rank among the lanes of its step.  The
// keys are placed in tile order by digit in LDS, then written out by consecutive threads, so the
// keys of one digit (a run of consecutive destinations) are stored together.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(unsigned long long* ka, int* va, unsigned long long* kb,
                                                           int* vb, int n, int pass, const unsigned long long* mask,
                                                           const unsigned int* __restrict__ offs,
                                                           const unsigned int* __restrict__ totals, int n_tiles) {
    int before;
    if (!rs_pass_on(mask, pass, before)) return;
    const unsigned long long* ks = (before & 1) ? kb : ka;
    const int* vs = (before & 1) ? vb : va;
    unsigned long long* kd = (before & 1) ? ka : kb;
    int* vd = (before & 1) ? va : vb;
    const int shift = 8 * pass;
    constexpr int STEPS = RS_TILE / RS_THREADS;   // 16 keys per lane
    __shared__ unsigned int wh[4][256];   // per-wave digit counts, then each wave's first tile position
    __shared__ unsigned int dstart[256];  // the digit's first position in the tile
    __shared__ unsigned int gbase[256];   // the digit's first destination of this tile's keys
    __shared__ unsigned long long sk[RS_TILE];
    __shared__ int sv[RS_TILE];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t0 = blockIdx.x * RS_TILE;
    const int w0 = t0 + wave * 1024;
    unsigned long long key[STEPS];
    int val[STEPS];
#pragma unroll
    for (int j = 0; j < STEPS; ++j) {
        const int i = w0 + j * 64 + lane;
        key[j] = i < n ? ks[i] : 0ull;
        val[j] = i < n ? vs[i] : 0;
    }
    for (int d = threadIdx.x; d < 4 * 256; d += RS_THREADS) (&wh[0][0])[d] = 0;
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    int pos[STEPS];
#pragma unroll
    for (int j = 0; j < STEPS; ++j) {
        const bool ok = w0 + j * 64 + lane < n;
        const int dg = ok ? rs_digit(key[j], shift) : 0;
        const unsigned long long same = rs_match(ok, dg);
        // every lane reads the count before the group's first lane adds to it (in-order LDS
        // accesses of one wave)
        pos[j] = ok ? (int)wh[wave][dg] + __popcll(same & below) : 0;
        if (ok && (same & below) == 0) wh[wave][dg] += (unsigned int)__popcll(same);
    }
    __syncthreads();
    {
        // thread d: digit d's count in the tile, its first tile position (an exclusive scan over
        // the digits), its first destination (the smaller digits' totals plus this tile's offset
        // among the tiles), and each wave's first tile position within the digit
        const int d = threadIdx.x;
        const unsigned int c0 = wh[0][d], c1 = wh[1][d], c2 = wh[2][d], c3 = wh[3][d];
        const unsigned int cnt = c0 + c1 + c2 + c3;
        const unsigned int tot = totals[d];
        __shared__ unsigned int dw[4], tw[4];
        const int incl_c = wave_incl_i((int)cnt);
        const int incl_t = wave_incl_i((int)tot);
        if ((threadIdx.x & 63) == 63) {
            dw[wave] = (unsigned int)incl_c;
            tw[wave] = (unsigned int)incl_t;
        }
        __syncthreads();
        unsigned int ls = (unsigned int)incl_c - cnt, gs = (unsigned int)incl_t - tot;
        for (int w = 0; w < wave; ++w) {
            ls += dw[w];
            gs += tw[w];
        }
        dstart[d] = ls;
        gbase[d] = gs + offs[(size_t)d * n_tiles + blockIdx.x];
        wh[0][d] = ls;
        wh[1][d] = ls + c0;
        wh[2][d] = ls + c0 + c1;
        wh[3][d] = ls + c0 + c1 + c2;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < STEPS; ++j) {
        if (w0 + j * 64 + lane < n) {
            const int p = (int)wh[wave][rs_digit(key[j], shift)] + pos[j];
            sk[p] = key[j];
            sv[p] = val[j];
        }
    }
    __syncthreads();
    const int m = min(RS_TILE, n - t0);
    for (int p = threadIdx.x; p < m; p += RS_THREADS) {
        const unsigned long long k = sk[p];
        const int d = rs_digit(k, shift);
        const unsigned int g = gbase[d] + (unsigned int)p - dstart[d];
        kd[g] = k;
        vd[g] = sv[p];
    }
}

// Keep the top rows, mark the drawn rest rows (rest[perm[k]], k < n_other), flag equal keys
// whose order decides the sample.  The sorted pairs are in the buffer the last pass wrote.
__global__ __launch_bounds__(256) void k_goss_mark(const unsigned long long* ka, const int* va,
                                                   const unsigned long long* kb, const int* vb, int n, int n_top,
                                                   const int* __restrict__ perm, int n_other,
                                                   unsigned char* __restrict__ keep,
                                                   unsigned char* __restrict__ scaled, int* tie,
                                                   const unsigned long long* mask) {
    const int passes = rs_passes_before(mask, RS_PASSES);
    const unsigned long long* keys = (passes & 1) ? kb : ka;
    const int* order = (passes & 1) ? vb : va;
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        // equal keys both among the kept top rows leave the sample unchanged
        if (j > 0 && j >= n_top && keys[j] == keys[j - 1]) *tie = 1;
        if (j < n_top) keep[order[j]] = 1;
    }
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n_other; k += stride) {
        const int row = order[n_top + perm[k]];
        keep[row] = 1;
        scaled[row] = 1;
    }
}

// Rows with keep set, per 4096-row tile: counts[tile].
__global__ __launch_bounds__(RS_THREADS) void k_flag_count(const unsigned char* __restrict__ keep, int n,
                                                           unsigned int* counts) {
    const int t0 = blockIdx.x * RS_TILE;
    int c = 0;
    for (int j = threadIdx.x; j < RS_TILE; j += RS_THREADS) {
        const int i = t0 + j;
        c += (i < n && keep[i]) ? 1 : 0;
    }
    c = wave_sum_i(c);
    __shared__ int wt[4];
    if ((threadIdx.x & 63) == 0) wt[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = (unsigned int)(wt[0] + wt[1] + wt[2] + wt[3]);
}

// The kept rows in ascending order: thread t of a tile writes its 16 consecutive rows' kept ones.
__global__ __launch_bounds__(RS_THREADS) void k_flag_scatter(const unsigned char* __restrict__ keep, int n,
                                                             const unsigned int* __restrict__ offs, int* out) {
    const int r0 = blockIdx.x * RS_TILE + threadIdx.x * 16;
    int c = 0;
#pragma unroll
    for (int u = 0; u < 16; ++u) c += (r0 + u < n && keep[r0 + u]) ? 1 : 0;
    const int incl = wave_incl_i(c);
    __shared__ int wt[4];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wt[wave] = incl;
    __syncthreads();
    int pos = (int)offs[blockIdx.x] + incl - c;
    for (int w = 0; w < wave; ++w) pos += wt[w];
#pragma unroll
    for (int u = 0; u < 16; ++u)
        if (r0 + u < n && keep[r0 + u]) out[pos++] = r0 + u;
}

// Scratch of one GOSS launch: the keys' OR / AND (256 bytes), the blocks' partial OR / AND, the
// digit totals, then the digit counts per tile.
size_t goss_temp_bytes(int n) {
    const size_t tiles = (size_t)(std::max(n, 1) + RS_TILE - 1) / RS_TILE;
    return 256 + sizeof(unsigned long long) * 2 * MAX_KEY_BLOCKS + sizeof(unsigned int) * 256 * (tiles + 1);
}

hipError_t launch_goss(hipStream_t st, const double* g, int n, int n_top, int n_other, const int* perm,
                       unsigned long long* k0, unsigned long long* k1, int* v0, int* v1, unsigned char* keep,
                       unsigned char* scaled, int* sample_out, int* n_sample_out, int* tie, void* tmp,
                       size_t tmp_bytes) {
    if (tmp_bytes < goss_temp_bytes(n)) return hipErrorInvalidValue;
    unsigned long long* mask = (unsigned long long*)tmp;
    unsigned long long* part = (unsigned long long*)((char*)tmp + 256);
    unsigned int* totals = (unsigned int*)(part + 2 * MAX_KEY_BLOCKS);
    unsigned int* counts = totals + 256;
    const int tiles = (std::max(n, 1) + RS_TILE - 1) / RS_TILE;
    const int blocks = std::max(1, std::min(MAX_KEY_BLOCKS, (n + 255) / 256));
    hipLaunchKernelGGL(k_goss_keys, dim3(blocks), dim3(256), 0, st, g, n, k0, v0, keep, scaled, part);
    hipLaunchKernelGGL(k_goss_mask, dim3(1), dim3(1024), 0, st, part, blocks, mask);
    for (int pass = 0; pass < RS_PASSES; ++pass) {
        hipLaunchKernelGGL(k_rs_count, dim3(tiles), dim3(RS_THREADS), 0, st, k0, k1, n, pass, mask, counts, tiles);
        hipLaunchKernelGGL(k_rs_scan_digits, dim3(256), dim3(1024), 0, st, counts, tiles, pass, mask, totals);
        hipLaunchKernelGGL(k_rs_scatter, dim3(tiles), dim3(RS_THREADS), 0, st, k0, v0, k1, v1, n, pass, mask, counts,
                           totals, tiles);
    }
    hipLaunchKernelGGL(k_goss_mark, dim3(blocks), dim3(256), 0, st, k0, v0, k1, v1, n, n_top, perm, n_other, keep,
                       scaled, tie, mask);
    hipLaunchKernelGGL(k_flag_count, dim3(tiles), dim3(RS_THREADS), 0, st, keep, n, counts);
    hipLaunchKernelGGL(k_rs_scan, dim3(1), dim3(1024), 0, st, counts, tiles, -1, mask, n_sample_out);
    hipLaunchKernelGGL(k_flag_scatter, dim3(tiles), dim3(RS_THREADS), 0, st, keep, n, counts, sample_out);
    return hipGetLastError();
}

}  // namespace tgbm
