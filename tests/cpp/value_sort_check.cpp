// Checks tidegbm_amd/csrc/host/value_sort.h against std::sort: the sorted sequences must be
// bitwise identical for NaN-free inputs of many shapes (float-converted values, heavy duplicates,
// negatives, infinities, denormals, one sign of zero), and inputs holding both signs of zero must
// come back in std::sort's own order.  Prints one line per failure and "ok <checks>" at the end.
#include "value_sort.h"

#include <cmath>
#include <cstdio>
#include <limits>
#include <random>

using tide::sort_values;

static int failures = 0, checks = 0;

static bool same_bits(const std::vector<double>& a, const std::vector<double>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) == 0);
}

static void check(const char* what, std::vector<double> v) {
    std::vector<double> ref = v;
    std::sort(ref.begin(), ref.end());
    sort_values(v);
    ++checks;
    if (!same_bits(v, ref)) {
        ++failures;
        std::printf("FAIL %s (n=%zu)\n", what, v.size());
    }
}

int main() {
    std::mt19937_64 g(11);
    std::normal_distribution<double> nd;
    for (size_t n : {0ul, 1ul, 100ul, 2047ul, 2048ul, 5000ul, 262144ul}) {
        std::vector<double> v(n);
        for (auto& x : v) x = nd(g);
        check("normal", v);
        for (auto& x : v) x = (double)(float)nd(g);
        check("float-converted", v);
        for (auto& x : v) x = (double)(long)(nd(g) * 3.0);   // heavy duplicates, zeros of one sign
        check("integers", v);
        for (auto& x : v) x = -std::fabs(nd(g)) * 1e-300;   // negatives reaching the denormals
        check("tiny negatives", v);
        for (size_t i = 0; i < n; i += 17) v[i] = (i & 1) ? std::numeric_limits<double>::infinity()
                                                           : -std::numeric_limits<double>::infinity();
        check("infinities", v);
        for (auto& x : v) x = std::ldexp(1.0, (int)(g() % 2000) - 1000) * ((g() & 1) ? 1 : -1);
        check("powers of two", v);
        for (auto& x : v) x = (g() % 3 == 0) ? -0.0 : nd(g);
        check("negative zeros", v);
        for (auto& x : v) x = (g() % 3 == 0) ? ((g() & 1) ? -0.0 : 0.0) : nd(g);
        check("zeros of both signs", v);
    }
    std::printf("%s %d\n", failures ? "failed" : "ok", checks);
    return failures ? 1 : 0;
}
