// The first k entries of std::shuffle over n values, with the same generator draws, without
// shuffling the n values.
//
// The reference draws its GOSS rows and its bin sub-samples with std::shuffle over a whole array
// and then keeps a prefix (reference booster.cpp:572-576, binning.cpp:94-96).  libstdc++'s
// std::shuffle, when the generator's range is below n^2, runs Fisher-Yates forwards: for
// i = 1 .. n-1 it swaps positions i and j = uniform_int_distribution<unsigned long>(0, i)(g).
// Position i holds its original value until step i, so the prefix [0, k) after all steps is:
// the same swaps run on [0, k) for i < k, then for i >= k the value v[i] written to j whenever
// j < k (the value leaving for position i never returns below k).  That touches a k-entry array
// instead of n random positions, and the draws are made in blocks from an MT19937 that produces
// the same stream as the std::mt19937 it was taken from.  For small n (generator range >= n^2)
// libstdc++ takes another path; there the whole array is shuffled with std::shuffle itself.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <random>
#include <vector>

namespace tide {

namespace mt_detail {

constexpr int N = 624, M = 397;
constexpr uint32_t kMatrix = 0x9908b0dfu, kB = 0x9d2c5680u, kC = 0xefc60000u;

inline uint32_t temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & kB;
    y ^= (y << 15) & kC;
    y ^= y >> 18;
    return y;
}
// x[k + N] = x[k + M] ^ twist(upper bit of x[k], lower 31 bits of x[k + 1])
inline uint32_t next_term(uint32_t xk, uint32_t xk1, uint32_t xkm) {
    const uint32_t y = (xk & 0x80000000u) | (xk1 & 0x7fffffffu);
    return xkm ^ (y >> 1) ^ ((0u - (y & 1u)) & kMatrix);
}
// The next N terms of the recurrence in place (x[0] the oldest) and their tempered outputs.
__attribute__((always_inline)) inline void refill_body(uint32_t* __restrict x, uint32_t* __restrict out) {
    int i = 0;
    for (; i < N - M; ++i) x[i] = next_term(x[i], x[i + 1], x[i + M]);
    for (; i < N - 1; ++i) x[i] = next_term(x[i], x[i + 1], x[i + M - N]);
    x[N - 1] = next_term(x[N - 1], x[0], x[M - 1]);
    for (int k = 0; k < N; ++k) out[k] = temper(x[k]);
}
// Lemire products of one run: js[c] = high word of x[c] * (r0 + c); returns nonzero when some
// low word is below its range (a possible rejection).
__attribute__((always_inline)) inline uint32_t run_body(const uint32_t* __restrict x, uint32_t r0, uint32_t run,
                                                        unsigned long* __restrict js) {
    uint32_t low_below = 0;
    for (uint32_t c = 0; c < run; ++c) {
        const uint32_t r = r0 + c;
        const uint64_t prod = (uint64_t)x[c] * r;
        js[c] = (unsigned long)(prod >> 32);
        low_below |= (uint32_t)((uint32_t)prod < r);
    }
    return low_below;
}

// The same loops compiled for wider vector units; the widest the CPU supports is picked once.
inline void refill_base(uint32_t* x, uint32_t* out) { refill_body(x, out); }
__attribute__((target("avx2"))) inline void refill_avx2(uint32_t* x, uint32_t* out) { refill_body(x, out); }
__attribute__((target("avx512f,avx512bw"))) inline void refill_avx512(uint32_t* x, uint32_t* out) {
    refill_body(x, out);
}
inline uint32_t run_base(const uint32_t* x, uint32_t r0, uint32_t run, unsigned long* js) {
    return run_body(x, r0, run, js);
}
__attribute__((target("avx2"))) inline uint32_t run_avx2(const uint32_t* x, uint32_t r0, uint32_t run,
                                                          unsigned long* js) {
    return run_body(x, r0, run, js);
}
__attribute__((target("avx512f,avx512bw"))) inline uint32_t run_avx512(const uint32_t* x, uint32_t r0, uint32_t run,
                                                                      unsigned long* js) {
    return run_body(x, r0, run, js);
}

struct Kernels {
    void (*refill)(uint32_t*, uint32_t*);
    uint32_t (*run)(const uint32_t*, uint32_t, uint32_t, unsigned long*);
};
enum Isa { kBase = 0, kAvx2 = 1, kAvx512 = 2 };
inline bool isa_supported(int isa) {
    __builtin_cpu_init();
    if (isa == kAvx512) return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw");
    if (isa == kAvx2) return __builtin_cpu_supports("avx2");
    return true;
}
inline Kernels kernels(int isa) {
    if (isa == kAvx512) return {refill_avx512, run_avx512};
    if (isa == kAvx2) return {refill_avx2, run_avx2};
    return {refill_base, run_base};
}
inline int best_isa() { return isa_supported(kAvx512) ? kAvx512 : (isa_supported(kAvx2) ? kAvx2 : kBase); }

}  // namespace mt_detail

// MT19937 (Matsumoto and Nishimura, ACM TOMACS 8(1), 1998) producing the outputs of a
// std::mt19937 from its current position on.  The state is recovered from 624 outputs drawn from
// the source (tempering is invertible, and 624 consecutive untempered outputs are the recurrence's
// last 624 terms); those outputs are served first.  The source is left 624 draws further on.
// Outputs are generated 624 at a time in plain loops the compiler vectorizes (built for the
// widest vector unit the CPU has).
class Mt19937Stream {
  public:
    using result_type = std::mt19937::result_type;
    static constexpr result_type min() { return std::mt19937::min(); }
    static constexpr result_type max() { return std::mt19937::max(); }

    // isa: mt_detail::kBase / kAvx2 / kAvx512 variant of the block loops (-1: the widest the CPU
    // supports; a variant it does not support falls back to kBase).
    explicit Mt19937Stream(std::mt19937& src, int isa = -1) {
        if (isa < 0 || !mt_detail::isa_supported(isa)) isa = isa < 0 ? mt_detail::best_isa() : mt_detail::kBase;
        k_ = mt_detail::kernels(isa);
        for (int i = 0; i < N; ++i) out_[i] = (uint32_t)src();
        for (int i = 0; i < N; ++i) x_[i] = untemper(out_[i]);
        p_ = 0;
    }
    result_type operator()() {
        if (p_ == N) refill();
        return out_[p_++];
    }
    // js[t] = uniform_int_distribution<unsigned long>(0, i0 + t)(*this) for t = 0 .. m-1, in
    // order, as libstdc++ draws it from a 32-bit generator (Lemire's method, D. Lemire, "Fast
    // random integer generation in an interval", ACM TOMACS 29(1), 2019): with r = i0 + t + 1 and
    // x the next output, the 64-bit product x * r is kept when its low word is at least
    // (2^32 - r) mod r, and x is redrawn otherwise; the draw is the high word.  A low word at or
    // above r is always kept, so runs of 64 draws are computed together and a run that holds a
    // low word below r is redone one draw at a time.  Requires i0 + m <= 2^32 - 1.
    void positions(size_t i0, size_t m, unsigned long* js) {
        size_t t = 0;
        while (t < m) {
            if (p_ == N) refill();
            const size_t run = std::min<size_t>({kRun, m - t, (size_t)(N - p_)});
            if (!k_.run(out_ + p_, (uint32_t)(i0 + t + 1), (uint32_t)run, js + t)) {
                p_ += (int)run;
                t += run;
                continue;
            }
            for (size_t c = 0; c < run; ++c, ++t) js[t] = bounded((uint32_t)(i0 + t + 1));
        }
    }

  private:
    static constexpr int N = mt_detail::N;
    static constexpr size_t kRun = 64;

    // One draw in [0, r) by Lemire's method (see positions()).
    unsigned long bounded(uint32_t r) {
        uint64_t prod = (uint64_t)(*this)() * r;
        uint32_t low = (uint32_t)prod;
        if (low < r) {
            const uint32_t threshold = (0u - r) % r;
            while (low < threshold) {
                prod = (uint64_t)(*this)() * r;
                low = (uint32_t)prod;
            }
        }
        return (unsigned long)(prod >> 32);
    }
    // Inverse of y ^= y >> s: each pass fixes s more high bits.
    static uint32_t unshift_right(uint32_t y, int s) {
        uint32_t r = y;
        for (int k = 0; k < 32 / s + 1; ++k) r = y ^ (r >> s);
        return r;
    }
    // Inverse of y ^= (y << s) & mask: each pass fixes s more low bits.
    static uint32_t unshift_left(uint32_t y, int s, uint32_t mask) {
        uint32_t r = y;
        for (int k = 0; k < 32 / s + 1; ++k) r = y ^ ((r << s) & mask);
        return r;
    }
    static uint32_t untemper(uint32_t y) {
        y = unshift_right(y, 18);
        y = unshift_left(y, 15, mt_detail::kC);
        y = unshift_left(y, 7, mt_detail::kB);
        return unshift_right(y, 11);
    }
    void refill() {
        k_.refill(x_, out_);
        p_ = 0;
    }

    mt_detail::Kernels k_;
    uint32_t x_[N];
    uint32_t out_[N];
    int p_;
};

// head[0 .. min(k, n)) = the first entries of std::shuffle(a, a + n, g) where a[i] = value(i),
// drawing from g exactly as that call does.
template <class T, class Value>
void shuffle_head(Mt19937Stream& g, size_t n, size_t k, Value value, std::vector<T>& head) {
    k = std::min(k, n);
    head.clear();
    if (n == 0) return;
    const unsigned long range = (unsigned long)(Mt19937Stream::max() - Mt19937Stream::min());
    if (range / n >= n) {   // libstdc++ draws two positions per call here: shuffle the whole array
        std::vector<T> a(n);
        for (size_t i = 0; i < n; ++i) a[i] = value(i);
        std::shuffle(a.begin(), a.end(), g);
        head.assign(a.begin(), a.begin() + (long)k);
        return;
    }
    head.resize(k + 1);   // head[k] takes the writes of draws that land at or past k
    for (size_t i = 0; i < k; ++i) head[i] = value(i);
    constexpr size_t kChunk = 2048;
    unsigned long js[kChunk];
    for (size_t i = 1; i < n; i += kChunk) {
        const size_t m = std::min(kChunk, n - i);
        g.positions(i, m, js);
        size_t t = 0;
        for (; t < m && i + t < k; ++t) std::swap(head[i + t], head[js[t]]);
        for (; t < m; ++t) head[js[t] < k ? js[t] : k] = value(i + t);
    }
    head.resize(k);
}

}  // namespace tide
