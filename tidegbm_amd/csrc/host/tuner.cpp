// Online histogram-variant tuner (see tuner.h).
#include "tuner.h"

#include <limits>
#include <map>
#include <mutex>

namespace tide {
namespace {

std::mutex g_mu;
std::map<std::string, int>& cache() {
    static std::map<std::string, int> m;
    return m;
}

// Child-build tiles (bits 0-15 of the variant): defaults (16 features and >= 1024 rows per
// block, at most 32 partials), 512- or 2048-row minimum tiles, 32 features per block (only when
// h = 1; with h varying it means 16).
const std::vector<int> kChild = {0, 17, 19, 26};
// Root-build variants (bits 20-29): defaults (features per block sized to the grid), and the
// multi-pass root kernel with 16, 32 or 64 features per block, XCD-major placement.
const int kRootFpb16 = ((6 << 4) | (0 << 7) | 8) << 20;
const int kRootFpb32 = ((6 << 4) | (1 << 7) | 8) << 20;
const int kRootFpb64 = ((6 << 4) | (2 << 7) | 8) << 20;

}  // namespace

HistTuner::HistTuner(std::string key, int n_trees, bool enabled) : key_(std::move(key)) {
    if (!enabled) return;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = cache().find(key_);
        if (it != cache().end()) {
            best_ = it->second;
            return;
        }
    }
    if (n_trees < kMinTrees) return;
    active_ = true;
    c1_ = kChild;
    t1_.assign(c1_.size(), std::numeric_limits<double>::infinity());
    stage2_start_ = kRounds * (int)c1_.size();
    end_ = stage2_start_ + kRounds * 4;
}

const std::vector<int>& HistTuner::cands(int tree, int& base) const {
    if (tree < stage2_start_) {
        base = 0;
        return c1_;
    }
    base = stage2_start_;
    return c2_;
}

int HistTuner::mode(int tree) const {
    if (!active_ || tree >= end_) return best_;
    int base;
    const std::vector<int>& c = cands(tree, base);
    return c[(tree - base) % c.size()];
}

void HistTuner::record(int tree, double seconds) {
    if (!active_ || tree >= end_) return;
    int base;
    const std::vector<int>& c = cands(tree, base);
    const int round = (tree - base) / (int)c.size(), i = (tree - base) % (int)c.size();
    std::vector<double>& t = tree < stage2_start_ ? t1_ : t2_;
    if (round > 0 && seconds < t[i]) t[i] = seconds;
    // the fastest candidate, if it beats the stage's first candidate by the margin
    auto pick = [](const std::vector<double>& t) {
        size_t b = 0;
        for (size_t k = 1; k < t.size(); ++k)
            if (t[k] < t[b]) b = k;
        return t[b] < (1.0 - kMargin) * t[0] ? b : 0;
    };
    if (tree == stage2_start_ - 1) {
        // stage 1 done: the best child tiles, then the root variants with them
        const int child = c1_[pick(t1_)];
        c2_ = {child, child | kRootFpb16, child | kRootFpb32, child | kRootFpb64};
        t2_.assign(c2_.size(), std::numeric_limits<double>::infinity());
        best_ = child;
    } else if (tree == end_ - 1) {
        best_ = c2_[pick(t2_)];
        std::lock_guard<std::mutex> lk(g_mu);
        cache()[key_] = best_;
    }
}

}  // namespace tide
