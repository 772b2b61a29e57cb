// Online choice of the histogram kernel variant (tgbm_grow_params::hist_mode) for one fit.
#pragma once

#include <string>
#include <vector>

namespace tide {

// Every variant yields bit-identical trees (the histogram sums are exact integers), so the first
// trees of a fit are grown with different candidates: each candidate grows one tree that captures
// its launch graph, then two timed trees, in rounds over the candidates (so their timed trees sit
// at similar places of the boosting sequence); the candidate with the least timed tree is kept
// for the remaining trees when it beats the stage's first candidate (the defaults) by kMargin,
// since tree times vary by a few percent from tree to tree.  Two stages: child-build tiles with
// the default root build, then root-build variants with the best child tiles.  A choice is
// remembered per workload shape for later fits of the process.
class HistTuner {
  public:
    // key: the workload shape; n_trees: trees of the fit; enabled: tuning allowed (the caller's
    // configuration did not force a variant)
    HistTuner(std::string key, int n_trees, bool enabled);
    int mode(int tree) const;                  // variant for tree `tree`
    void record(int tree, double seconds);     // wall time of tree `tree`
    int chosen() const { return best_; }
    // the variant of tree `tree` no longer depends on timings still to be recorded
    bool settled(int tree) const { return !active_ || tree >= end_; }
    // trees a tuned fit needs (shorter fits keep the defaults)
    static constexpr int kMinTrees = 54;

  private:
    static constexpr int kRounds = 4;          // one capture round, three timed rounds
    static constexpr double kMargin = 0.03;    // a candidate must be this much faster than the first
    std::string key_;
    bool active_ = false;
    int best_ = 0;
    int stage2_start_ = 0;
    int end_ = 0;
    std::vector<int> c1_, c2_;                 // candidates of the stages
    std::vector<double> t1_, t2_;              // least timed seconds per candidate
    const std::vector<int>& cands(int tree, int& base) const;
};

}  // namespace tide
