// Checks tidegbm_amd/csrc/host/shuffle_head.h against the standard library it replaces:
// Mt19937Stream's outputs against std::mt19937's, and shuffle_head's prefix and the generator's
// next output against std::shuffle over the whole array, for sizes on both sides of libstdc++'s
// two-draws-per-call threshold (n^2 <= 2^32 - 1, i.e. n <= 65535), prefixes of every length class,
// and integer and double values.  Prints one line per failure and "ok <checks>" at the end.
#include "shuffle_head.h"

#include <cstdio>
#include <numeric>

using tide::Mt19937Stream;
using tide::shuffle_head;

static int failures = 0, checks = 0;

static void expect(bool ok, const char* what, long a, long b) {
    ++checks;
    if (!ok) {
        ++failures;
        std::printf("FAIL %s (%ld, %ld)\n", what, a, b);
    }
}

int main() {
    // the stream, from generators at several positions within and across 624-word blocks, with
    // every block-loop variant this CPU can run
    for (int isa = tide::mt_detail::kBase; isa <= tide::mt_detail::kAvx512; ++isa) {
        if (!tide::mt_detail::isa_supported(isa)) {
            std::printf("note: vector variant %d not supported here, not checked\n", isa);
            continue;
        }
        for (int skip : {0, 1, 311, 623, 624, 625, 5000}) {
            std::mt19937 a(1234u + skip), b(1234u + skip);
            for (int i = 0; i < skip; ++i) {
                a();
                b();
            }
            Mt19937Stream s(b, isa);
            bool same = true;
            for (int i = 0; i < 200000 && same; ++i) same = a() == s();
            expect(same, "stream", skip, isa);
        }
        // the bounded draws of a large shuffle with this variant
        std::mt19937 a(99), b(99);
        std::vector<int> p(300000);
        std::iota(p.begin(), p.end(), 0);
        std::shuffle(p.begin(), p.end(), a);
        Mt19937Stream s(b, isa);
        std::vector<int> head;
        shuffle_head<int>(s, p.size(), 30000, [](size_t i) { return (int)i; }, head);
        expect(std::equal(head.begin(), head.end(), p.begin()) && a() == s(), "variant prefix", isa, 0);
    }
    // the standard's own check value: the 10000th output of a default-seeded mt19937
    {
        std::mt19937 src;
        Mt19937Stream s(src);
        unsigned long v = 0;
        for (int i = 0; i < 10000; ++i) v = s();
        expect(v == 4123659995ul, "10000th output", (long)v, 4123659995l);
    }
    const long sizes[] = {0, 1, 2, 3, 64, 1000, 65535, 65536, 65537, 100000, 250001, 1200000};
    for (long n : sizes) {
        for (long k : {0l, 1l, n / 10, n / 2, n - 1, n, n + 5}) {
            if (k < 0) continue;
            for (unsigned seed : {0u, 7u}) {
                std::mt19937 a(seed), b(seed);
                std::vector<int> p(n);
                std::iota(p.begin(), p.end(), 0);
                std::shuffle(p.begin(), p.end(), a);
                Mt19937Stream s(b);
                std::vector<int> head;
                shuffle_head<int>(s, (size_t)n, (size_t)k, [](size_t i) { return (int)i; }, head);
                const size_t kk = std::min<size_t>(k, n);
                bool same = head.size() == kk && std::equal(head.begin(), head.end(), p.begin());
                expect(same, "int prefix", n, k);
                expect(a() == s(), "next draw after int prefix", n, k);
            }
        }
    }
    // values (the bin sub-sample) rather than positions, with the reference's per-feature seeds
    for (long n : {70000l, 300000l}) {
        std::vector<double> v(n);
        for (long i = 0; i < n; ++i) v[i] = 0.5 * (double)((i * 7919) % 1000003) - 3.0;
        for (int f = 0; f < 3; ++f) {
            const unsigned seed = 42u + (unsigned)f * 2654435761u;
            std::mt19937 a(seed), b(seed);
            std::vector<double> w = v;
            std::shuffle(w.begin(), w.end(), a);
            Mt19937Stream s(b);
            std::vector<double> head;
            shuffle_head<double>(s, (size_t)n, 65536, [&](size_t i) { return v[i]; }, head);
            expect(std::equal(head.begin(), head.end(), w.begin()) && head.size() == 65536, "double prefix", n, f);
        }
    }
    std::printf("%s %d\n", failures ? "failed" : "ok", checks);
    return failures ? 1 : 0;
}
