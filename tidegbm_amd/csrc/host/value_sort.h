// Ascending sort of NaN-free doubles, giving the same sequence as std::sort.
//
// Two doubles that compare equal have the same bits unless they are -0.0 and +0.0, so an
// ascending NaN-free sequence is unique apart from the order of zeros of both signs.  When both
// signs of zero are present, std::sort's own order is kept by calling it; otherwise an LSD radix
// sort over the order-preserving 64-bit keys (sign bit flipped for positives, all bits flipped for
// negatives) gives the same values in the same order.  Digit passes in which every key holds the
// same byte are skipped: doubles converted from float keep their low 29 mantissa bits zero.
// The radix sort holds two 8-byte keys per value; above 4M values (bin sub-sampling off on a
// large column) std::sort sorts in place instead.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace tide {

inline void sort_values(std::vector<double>& v) {
    const size_t n = v.size();
    if (n < 2048 || n > ((size_t)1 << 22)) {
        std::sort(v.begin(), v.end());
        return;
    }
    constexpr uint64_t kSign = 0x8000000000000000ull;
    std::vector<uint64_t> a(n), b(n);
    bool pos_zero = false, neg_zero = false;
    for (size_t i = 0; i < n; ++i) {
        uint64_t u;
        std::memcpy(&u, &v[i], sizeof u);
        pos_zero |= u == 0;
        neg_zero |= u == kSign;
        a[i] = (u & kSign) ? ~u : (u | kSign);
    }
    if (pos_zero && neg_zero) {
        std::sort(v.begin(), v.end());
        return;
    }
    std::vector<size_t> count(8 * 256, 0);
    for (size_t i = 0; i < n; ++i)
        for (int d = 0; d < 8; ++d) ++count[d * 256 + ((a[i] >> (8 * d)) & 0xff)];
    for (int d = 0; d < 8; ++d) {
        size_t* c = &count[d * 256];
        if (*std::max_element(c, c + 256) == n) continue;   // one byte value: the pass moves nothing
        size_t sum = 0;
        for (int k = 0; k < 256; ++k) {
            const size_t t = c[k];
            c[k] = sum;
            sum += t;
        }
        for (size_t i = 0; i < n; ++i) b[c[(a[i] >> (8 * d)) & 0xff]++] = a[i];
        a.swap(b);
    }
    for (size_t i = 0; i < n; ++i) {
        const uint64_t k = a[i];
        const uint64_t u = (k & kSign) ? (k & ~kSign) : ~k;
        std::memcpy(&v[i], &u, sizeof u);
    }
}

}  // namespace tide
