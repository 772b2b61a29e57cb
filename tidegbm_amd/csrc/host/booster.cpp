// Host orchestration of a fit.  Integer preparation (row order, TORB buckets and cuts, bin edges,
// GOSS selection) is done here with the same library calls and call order as the reference so
// that the rows, bins and samples handed to the device are identical; every per-row and
// per-histogram computation runs on the GPU.
#include "booster.h"
#include "shuffle_head.h"
#include "tuner.h"
#include "value_sort.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <iterator>
#include <numeric>
#include <thread>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace tide {

void dev_check(int rc) {
    if (rc != 0) throw Error(tgbm_last_error());
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

namespace {

struct SessionGuard {
    tgbm_session* s = nullptr;
    ~SessionGuard() {
        if (s) tgbm_session_destroy(s);
    }
};

// Quantile right-edges of one feature (reference quantile_edges, binning.cpp:14-32).
std::vector<double> quantile_edges(std::vector<double>& v, int nb) {
    sort_values(v);   // std::sort's sequence
    v.erase(std::unique(v.begin(), v.end()), v.end());
    if (v.empty()) return {0.0};
    if ((int)v.size() <= nb) return v;
    std::vector<double> e;
    e.reserve(nb - 1);
    const size_t last = v.size() - 1;
    for (int i = 1; i < nb; ++i) {
        const double q = (double)i / nb;
        const double pos = q * last;
        const size_t lo = (size_t)pos;
        const size_t hi = std::min(lo + 1, last);
        const double fr = pos - lo;
        e.push_back(v[lo] * (1 - fr) + v[hi] * fr);
    }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    return e;
}

}  // namespace

// Per-feature edges over non-NaN values with the reference's sub-sampling
// (Binner::fit numeric branch, binning.cpp:84-99).  Groups of up to 8 adjacent features run in
// parallel: one pass over the row-major rows gathers a group's columns (8 doubles share a cache
// line), with at most 32M gathered values held per thread.
void fit_edges(const double* X, int n, int F, int max_bin, int bin_sample, uint32_t seed,
               std::vector<std::vector<double>>& edges) {
    edges.assign(F, {});
    if (F <= 0) return;
    const int threads = std::max(1, omp_get_max_threads());
    long long gs = std::min(8, std::max(1, F / threads));
    gs = std::max(1ll, std::min(gs, (32ll << 20) / std::max(n, 1)));
    const int n_groups = (int)((F + gs - 1) / gs);
#pragma omp parallel for schedule(dynamic) if (n_groups > 1)
    for (int gi = 0; gi < n_groups; ++gi) {
        const int f0 = (int)(gi * gs), nf = std::min(F - f0, (int)gs);
        std::vector<std::vector<double>> cols(nf);
        for (auto& v : cols) v.reserve(n);
        for (int i = 0; i < n; ++i) {
            const double* r = X + (size_t)i * F + f0;
            for (int q = 0; q < nf; ++q) {
                const double x = r[q];
                if (x == x) cols[q].push_back(x);
            }
        }
        for (int q = 0; q < nf; ++q) {
            const int f = f0 + q;
            std::vector<double>& v = cols[q];
            if (bin_sample > 0 && (int)v.size() > bin_sample) {
                std::mt19937 rng(seed + (uint32_t)f * 2654435761u);
                Mt19937Stream g(rng);
                std::vector<double> head;
                shuffle_head<double>(g, v.size(), (size_t)bin_sample, [&](size_t i) { return v[i]; }, head);
                v.swap(head);
            }
            edges[f] = quantile_edges(v, max_bin - 1);
            std::vector<double>().swap(v);
        }
    }
}

namespace {

double leaf_value_host(double G, double H, const Config& c) {  // tree.hpp:53-67
    double v;
    if (c.l1 > 0) {
        if (G > c.l1) v = -(G - c.l1) / (H + c.l2);
        else if (G < -c.l1) v = -(G + c.l1) / (H + c.l2);
        else v = 0.0;
    } else {
        v = -G / (H + c.l2);
    }
    if (c.max_delta_step > 0) v = std::min(std::max(v, -c.max_delta_step), c.max_delta_step);
    return v;
}

}  // namespace

// The power-of-two row count the device launches are sized for (workload shape key of the tuner).
long long launch_rows_of(long long n) {
    long long r = 1;
    while (r < n) r <<= 1;
    return r;
}

tgbm_grow_params grow_params(const Config& c) {
    tgbm_grow_params p;
    std::memset(&p, 0, sizeof p);
    p.max_depth = c.max_depth;
    p.max_leaves = c.max_leaves;
    p.min_samples_leaf = c.min_samples_leaf;
    p.min_child_weight = c.min_child_weight;
    p.l2 = c.l2;
    p.l1 = c.l1;
    p.max_delta_step = c.max_delta_step;
    p.min_split_gain = c.min_split_gain;
    p.pair_l2 = c.pair_l2;
    p.cat_smooth = c.cat_smooth;
    p.hist_mode = c.hist_mode;
    p.host_driven = (c.pair_splits || c.force_host_driver) ? 1 : 0;
    return p;
}

// Host mirror of one node during host-driven (PAIR-capable) growth.
namespace {

struct LeafInfo {
    std::vector<double> gain;
    std::vector<int> thr, ml;
    double energy = 0.0;
    int rows = 0, depth = 0;
    bool leaf = true;
    bool cached = false;
    tgbm_cand cand{};
};

void fetch_leaf(tgbm_session* s, int node, int F, LeafInfo& li) {
    li.gain.resize(F);
    li.thr.resize(F);
    li.ml.resize(F);
    dev_check(tgbm_leaf_scan_results(s, 0, node, li.gain.data(), li.thr.data(), li.ml.data(), &li.energy, &li.rows));
}

// Candidate of one leaf (reference best_for_leaf, tree.cpp:106-281, with no interaction-group,
// regime or HALO branches): argmax over ascending features of their axis or categorical
// (is_cat) best splits, then the PAIR search on weak leaves over the top-k features ordered by
// std::sort.
// Features a node may split on (tree.cpp:293-317): all, or under interaction group ixg >= 0 the
// unconstrained ones and those of the group, with the regime feature appended when missing.
std::vector<int> node_features(const Config& c, int F, int ixg) {
    std::vector<int> all(F);
    std::iota(all.begin(), all.end(), 0);
    if (c.regime_feature >= 0 && c.regime_feature < F &&
        std::find(all.begin(), all.end(), c.regime_feature) == all.end())
        all.push_back(c.regime_feature);
    if (ixg < 0 || c.interaction_group.empty()) return all;
    std::vector<int> out;
    for (int f : all) {
        const int g = f < (int)c.interaction_group.size() ? c.interaction_group[f] : -1;
        if (g < 0 || g == ixg) out.push_back(f);
    }
    if (c.regime_feature >= 0 && std::find(out.begin(), out.end(), c.regime_feature) == out.end())
        out.push_back(c.regime_feature);
    return out.empty() ? all : out;
}

tgbm_cand leaf_candidate(tgbm_session* s, int node, LeafInfo& li, const Config& c, int F,
                         int pair_budget, const tgbm_grow_params& gp, const std::vector<int>& is_cat,
                         const std::vector<int>& group_of, const std::vector<int>& active) {
    tgbm_cand best{};
    best.gain = -1e300;
    best.feature = -1;
    best.feature2 = -1;
    best.threshold = -1;
    best.type = 0;
    best.missing_go_left = 1;
    auto gain_of = [&](int f) { return li.thr[f] >= 0 ? li.gain[f] : -1e300; };
    auto cand_of = [&](int f) {
        tgbm_cand r{};
        r.gain = li.gain[f];
        r.feature = f;
        r.feature2 = -1;
        r.threshold = li.thr[f];   // categorical: the best position of the ranked bins
        r.missing_go_left = li.ml[f];
        r.type = (!is_cat.empty() && is_cat[f]) ? 2 : 0;
        return r;
    };
    // one feature per feature group: its best-gain member, first in `active` order on ties;
    // groups numbered below 256 are listed once (tree.cpp:135-158)
    std::map<int, int> group_best;
    for (int f : active) {
        const int grp = group_of.empty() ? -1 : group_of[f];
        if (grp < 0) continue;
        auto it = group_best.find(grp);
        if (it == group_best.end()) group_best[grp] = f;
        else if (gain_of(f) > gain_of(it->second)) it->second = f;
    }
    std::vector<int> allowed;
    std::vector<char> seen(256, 0);
    for (int f : active) {
        const int grp = group_of.empty() ? -1 : group_of[f];
        if (grp >= 0) {
            if (grp < 256 && seen[grp]) continue;
            if (grp < 256) seen[grp] = 1;
            f = group_best[grp];
        }
        allowed.push_back(f);
    }
    for (int f : allowed) {
        if (li.thr[f] < 0 || li.gain[f] < c.min_split_gain) continue;
        if (li.gain[f] > best.gain) best = cand_of(f);
    }
    // regime priority (tree.cpp:168-176)
    if (c.regime_priority && c.regime_feature >= 0 && c.regime_feature < F) {
        const int rf = c.regime_feature;
        if (li.thr[rf] >= 0 && li.gain[rf] > 0 && (best.feature < 0 || li.gain[rf] >= 0.5 * std::max(best.gain, 0.0))) {
            best = cand_of(rf);
            best.type = (!is_cat.empty() && is_cat[rf]) ? 3 : 0;
        }
    }
    const bool weak = (best.gain < c.pair_trigger * std::max(li.energy * 0.01, 1e-12)) || best.gain <= 0;
    if (c.pair_splits && pair_budget > 0 && allowed.size() >= 2 && weak) {
        std::vector<int> top = allowed;
        std::sort(top.begin(), top.end(), [&](int a, int b) { return gain_of(a) > gain_of(b); });
        const int k = std::min(c.pair_top_k, (int)top.size());
        std::vector<int> f1, f2;
        for (int a = 0; a < k; ++a)
            for (int b = a + 1; b < k; ++b) {
                f1.push_back(top[a]);
                f2.push_back(top[b]);
            }
        std::vector<tgbm_cand> pc(f1.size());
        for (size_t off = 0; off < f1.size(); off += 32) {
            const int m = (int)std::min<size_t>(32, f1.size() - off);
            dev_check(tgbm_pair_search(s, 0, node, f1.data() + off, f2.data() + off, m, &gp, pc.data() + off));
        }
        for (auto& p : pc) {
            if (!p.valid) continue;
            if (p.gain > best.gain) best = p;
        }
    }
    best.valid = best.feature >= 0;
    return best;
}

// Host-driven best-first loop (reference grow_tree, tree.cpp:343-463).
void grow_host_driven(tgbm_session* s, const Config& c, int F, const tgbm_grow_params& gp,
                      const std::vector<int>& is_cat, const std::vector<int>& group_of) {
    std::vector<LeafInfo> nodes(1);
    fetch_leaf(s, 0, F, nodes[0]);
    nodes[0].depth = 0;
    std::vector<int> ixg(1, -1);   // interaction group each node's path is bound to (tree.cpp:436-445)
    int pair_used = 0;
    std::vector<double> ch_gain;    // a split's two children's scan results (left's F, then right's)
    std::vector<int> ch_thr, ch_ml;
    for (int it = 0; it < c.max_leaves - 1; ++it) {
        int best_leaf = -1;
        tgbm_cand best{};
        double best_score = c.min_split_gain;
        for (int lid = 0; lid < (int)nodes.size(); ++lid) {
            LeafInfo& li = nodes[lid];
            if (!li.leaf || li.depth >= c.max_depth) continue;
            if (li.rows < 2 * c.min_samples_leaf) continue;
            if (!li.cached) {
                li.cand = leaf_candidate(s, lid, li, c, F, c.max_pair_splits_per_tree - pair_used, gp, is_cat,
                                         group_of, node_features(c, F, ixg[lid]));
                li.cached = true;
            }
            const tgbm_cand& r = li.cand;
            if (r.feature >= 0 && r.gain >= c.min_split_gain && r.gain > best_score) {
                best_score = r.gain;
                best_leaf = lid;
                best = r;
            }
        }
        if (best_leaf < 0) break;
        if (best.type == 1) {
            ++pair_used;
            for (auto& li : nodes) li.cached = false;
        }
        nodes[best_leaf].cached = false;
        nodes[best_leaf].leaf = false;
        tgbm_decision d{best_leaf, best.type, best.feature, best.type == 1 ? best.feature2 : -1,
                        best.threshold, best.missing_go_left, best.gain};
        int left = -1;
        double ch_energy[2];
        int ch_rows[2];
        ch_gain.resize(2 * (size_t)F);
        ch_thr.resize(2 * (size_t)F);
        ch_ml.resize(2 * (size_t)F);
        dev_check(tgbm_apply_split_fetch(s, 0, &d, &gp, &left, ch_gain.data(), ch_thr.data(), ch_ml.data(), ch_energy,
                                         ch_rows));
        if (left < 0) break;
        nodes.resize(left + 2);
        int child_ixg = ixg[best_leaf];
        if (child_ixg < 0 && best.feature >= 0 && best.feature < (int)c.interaction_group.size() &&
            c.interaction_group[best.feature] >= 0)
            child_ixg = c.interaction_group[best.feature];
        ixg.resize(left + 2, -1);
        ixg[left] = ixg[left + 1] = child_ixg;
        for (int side = 0; side < 2; ++side) {
            LeafInfo& ch = nodes[left + side];
            ch.gain.assign(ch_gain.begin() + (size_t)side * F, ch_gain.begin() + (size_t)(side + 1) * F);
            ch.thr.assign(ch_thr.begin() + (size_t)side * F, ch_thr.begin() + (size_t)(side + 1) * F);
            ch.ml.assign(ch_ml.begin() + (size_t)side * F, ch_ml.begin() + (size_t)(side + 1) * F);
            ch.energy = ch_energy[side];
            ch.rows = ch_rows[side];
            ch.depth = nodes[best_leaf].depth + 1;
            ch.leaf = true;
            ch.cached = false;
        }
    }
}

// GOSS row counts of one tree (reference booster.cpp:564-570).
struct GossCounts {
    int n_top, n_rest, n_other;
};
GossCounts goss_counts(int S, const Config& c) {
    GossCounts k;
    k.n_top = std::min(S, std::max(1, (int)std::ceil(c.goss_top * S)));
    k.n_rest = S - k.n_top;
    k.n_other = std::min((int)std::ceil(c.goss_other * S), k.n_rest);
    return k;
}

// The generator draws of a GOSS fit, produced ahead of the trees.  Per tree the reference makes
// one std::shuffle of the rest (booster.cpp:572) and one tree seed draw (:588).  The shuffle's
// swaps depend on the generator and the rest's size only, never on the values shuffled, so the
// rest after the shuffle is rest[P[i]] with P the same call applied to 0..n_rest-1.  A thread
// computes the first n_other entries of P for each tree in order (shuffle_head: the same draws,
// without shuffling n_rest entries) and queues them (the drawn positions).  The thread draws from
// a stream taken from the generator, which is not to be used again by the caller.
class GossDraws {
  public:
    GossDraws(std::mt19937& rng, int n_rest, int n_other, int n_trees)
        : rng_(rng), n_rest_(n_rest), n_other_(n_other), n_trees_(n_trees) {
        th_ = std::thread([this] { run(); });
    }
    ~GossDraws() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
    }
    // The drawn positions of the next tree (valid until the following call).
    const std::vector<int>& next() {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !q_.empty(); });
        cur_ = std::move(q_.front());
        q_.pop_front();
        cv_.notify_all();
        return cur_;
    }

  private:
    static constexpr size_t kAhead = 4;
    void run() {
        Mt19937Stream g(rng_);
        for (int m = 0; m < n_trees_; ++m) {
            std::vector<int> head;
            shuffle_head<int>(g, (size_t)n_rest_, (size_t)n_other_, [](size_t i) { return (int)i; }, head);
            (void)g();  // the tree's own seed; only colsample would use it
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || q_.size() < kAhead; });
            if (stop_) return;
            q_.push_back(std::move(head));
            cv_.notify_all();
        }
    }
    std::mt19937& rng_;
    const int n_rest_, n_other_, n_trees_;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::vector<int>> q_;
    std::vector<int> cur_;
    bool stop_ = false;
};

// GOSS sample on the host (reference booster.cpp:557-583) with the drawn rest positions of the
// tree.  std::sort with the reference's comparator orders equal |g| in an unspecified but fixed
// way; sorting (|g|, row) records by |g| performs the same comparisons and moves as sorting row
// indices through g, so it yields the same order, with contiguous reads.
void goss_sample_host(const std::vector<double>& g, int n, const Config& c, const std::vector<int>& drawn,
                      std::vector<int>& sample, std::vector<unsigned char>& scaled, double& scale) {
    struct Rec {
        double a;
        int row;
    };
    std::vector<Rec> order(n);
    for (int i = 0; i < n; ++i) order[i] = Rec{std::abs(g[i]), i};
    std::sort(order.begin(), order.end(), [](const Rec& x, const Rec& y) { return x.a > y.a; });
    const GossCounts k = goss_counts(n, c);
    if ((int)drawn.size() != k.n_other) throw Error("internal: GOSS draw size");
    std::vector<char> keep(n, 0);
    scaled.assign(n, 0);
    for (int i = 0; i < k.n_top; ++i) keep[order[i].row] = 1;
    for (int i = 0; i < k.n_other; ++i) {
        const int row = order[k.n_top + drawn[i]].row;
        keep[row] = 1;
        scaled[row] = 1;
    }
    scale = (1.0 - c.goss_top) / std::max(c.goss_other, 1e-12);
    sample.clear();
    for (int i = 0; i < n; ++i)
        if (keep[i]) sample.push_back(i);
}

}  // namespace

// TORB time buckets, causal cuts and initial scores of one time-sorted fit
// (booster.cpp:36-72 buckets, :311-364 cuts, :399-457 init_prefix / init pred).
TimeLayout time_layout(const Config& cfg, const long long* ts, const double* y, const double* ws, int n) {
    TimeLayout tl;
    std::vector<int>& bucket = tl.bucket;
    std::vector<int>& cuts = tl.cuts;
    bucket.assign(n, 0);
    int n_buckets = 1;
    long long embargo = cfg.embargo;
    if (!cfg.iid_mode) {
        if (!ts) throw Error("timestamp is required unless iid_mode");
        std::vector<long long> u(ts, ts + n);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        const int uniq_n = std::max(2, (int)u.size());
        int want = cfg.n_time_buckets;
        if (want <= 0) {
            long long step = 1;
            if (u.size() >= 2) {
                std::vector<long long> d;
                for (size_t k = 1; k < u.size(); ++k)
                    if (u[k] - u[k - 1] > 0) d.push_back(u[k] - u[k - 1]);
                if (!d.empty()) {
                    std::nth_element(d.begin(), d.begin() + d.size() / 2, d.end());
                    step = std::max<long long>(d[d.size() / 2], 1);
                }
            }
            const int h_steps = (int)(std::max<long long>(cfg.horizon, 0) / step);
            want = std::max(2, std::min(h_steps >= 8 ? 64 : 32, std::max(uniq_n, 2)));
        }
        n_buckets = cfg.finer_torb ? std::max(2, std::min(want, uniq_n)) : std::max(2, want);
        if (n_buckets > 256) throw Error("the GPU path supports at most 256 time buckets");
        std::vector<int> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ts[a] < ts[b]; });
        for (int r = 0; r < n; ++r) {
            int b = (int)((double)r / std::max(n, 1) * n_buckets);
            if (b >= n_buckets) b = n_buckets - 1;
            bucket[order[r]] = b;
        }
        embargo += std::max<long long>(cfg.horizon, 0);
        bool sorted = true;
        for (int i = 1; i < n && sorted; ++i) sorted = ts[i] >= ts[i - 1];
        if (!sorted) throw Error("internal: rows not time-sorted");
        std::vector<int> first(n_buckets, -1);
        for (int i = 0; i < n; ++i)
            if (first[bucket[i]] < 0) first[bucket[i]] = i;
        cuts.assign(n_buckets, 0);
        int run = 0;
        for (int b = 0; b < n_buckets; ++b) {
            int c = 0;
            if (first[b] >= 0) c = (int)(std::lower_bound(ts, ts + n, ts[first[b]] - embargo) - ts);
            run = std::max(run, c);
            cuts[b] = run;
        }
    }
    tl.n_buckets = n_buckets;
    auto init_of = [&](int m) -> double {  // Objective::init_score (objective.cpp:80-100)
        if (m <= 0) return 0;
        if (cfg.objective == 5) return 0;
        double sw = 0, sy = 0;
        for (int i = 0; i < m; ++i) {
            sw += ws[i];
            sy += ws[i] * y[i];
        }
        return sw > 0 ? sy / sw : 0;
    };
    std::vector<double>& pred = tl.pred;
    pred.assign(n, 0.0);
    tl.init_prefix.assign(n_buckets, 0.0);
    if (cfg.iid_mode) {
        tl.init_score = init_of(n);
        std::fill(pred.begin(), pred.end(), tl.init_score);
    } else {
        // running weighted sums snapshot at each cut: identical to refolding [0, cut) per bucket
        std::vector<double> at(n_buckets, NAN);
        if (cfg.objective == 5) {
            for (int b = 0; b < n_buckets; ++b) at[b] = cuts[b] > 0 ? 0.0 : NAN;
        } else {
            double sw = 0, sy = 0;
            int i = 0;
            for (int b = 0; b < n_buckets; ++b) {
                for (; i < cuts[b]; ++i) {
                    sw += ws[i];
                    sy += ws[i] * y[i];
                }
                if (cuts[b] > 0) at[b] = sw > 0 ? sy / sw : 0;
            }
        }
        tl.init_prefix = at;
        tl.init_score = init_of(n);
        for (int b = 0; b < n_buckets; ++b)
            if (!std::isfinite(tl.init_prefix[b])) tl.init_prefix[b] = tl.init_score;
        for (int i = 0; i < n; ++i) pred[i] = tl.init_prefix[bucket[i]];
    }
    return tl;
}

void bin_on_device(tgbm_session* s, const double* X, const std::vector<std::vector<double>>& edges,
                   const std::vector<int>& is_cat, const std::vector<std::vector<double>>& cats) {
    std::vector<double> E;
    std::vector<int> ne;
    flat_edges(edges, E, ne, is_cat, cats);
    dev_check(tgbm_session_bin(s, X, E.data(), ne.data()));
}

// Binner::fit's categorical branch (binning.cpp:48-82): count each distinct non-NaN value, keep
// those seen at least cat_min_count times (else the single most frequent), keep the
// max_bin - 1 most frequent with std::partial_sort when more remain, and sort the survivors.
std::vector<double> category_values(const double* X, int n, int F, int f, int max_bin, int cat_min_count) {
    std::vector<double> vals;
    vals.reserve(n);
    for (int i = 0; i < n; ++i) {
        const double x = X[(size_t)i * F + f];
        if (!std::isnan(x)) vals.push_back(x);
    }
    std::sort(vals.begin(), vals.end());
    std::vector<std::pair<int, double>> counts;   // (occurrences, value)
    for (size_t i = 0; i < vals.size();) {
        size_t j = i;
        while (j < vals.size() && vals[j] == vals[i]) ++j;
        counts.emplace_back((int)(j - i), vals[i]);
        i = j;
    }
    const int need = std::max(1, cat_min_count);
    std::vector<std::pair<int, double>> keep;
    std::copy_if(counts.begin(), counts.end(), std::back_inserter(keep),
                 [&](const std::pair<int, double>& q) { return q.first >= need; });
    if (keep.empty() && !counts.empty())
        keep.push_back(*std::max_element(counts.begin(), counts.end(),
                                         [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
                                             return a.first < b.first;
                                         }));
    const size_t cap = (size_t)std::max(0, max_bin - 1);
    if (keep.size() > cap) {
        std::partial_sort(keep.begin(), keep.begin() + cap, keep.end(),
                          [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
                              return a.first > b.first;
                          });
        keep.resize(cap);
    }
    std::vector<double> out(keep.size());
    std::transform(keep.begin(), keep.end(), out.begin(), [](const std::pair<int, double>& q) { return q.second; });
    std::sort(out.begin(), out.end());
    return out;
}

std::vector<unsigned long long> tree_cat_masks(const Tree& t) {
    bool any = false;
    for (const Node& d : t.nodes) any |= !d.is_leaf && d.split_type >= 2;
    if (!any) return {};
    std::vector<unsigned long long> m(t.nodes.size() * 4, 0ull);
    for (size_t k = 0; k < t.nodes.size(); ++k) std::copy(t.nodes[k].cat_left, t.nodes[k].cat_left + 4, &m[4 * k]);
    return m;
}

Tree unpack_tree(int nn, const int* ints, const double* dbl, const double* vals, const int* cov, const double* pref,
                 int n_buckets) {
    Tree t;
    t.nodes.resize(nn);
    for (int k = 0; k < nn; ++k) {
        const int* v = &ints[(size_t)k * TGBM_NODE_INTS];
        Node& d = t.nodes[k];
        d.is_leaf = v[0];
        d.split_type = v[1];
        d.feature = v[2];
        d.feature2 = v[3];
        d.threshold = v[4];
        d.left = v[5];
        d.right = v[6];
        d.missing_go_left = v[7];
        d.n_samples = v[8];
        d.depth = v[9];
        d.value = dbl[2 * k];
        d.gain = dbl[2 * k + 1];
    }
    t.values.assign(vals, vals + nn);
    t.cover.assign(cov, cov + nn);
    t.n_buckets = n_buckets;
    t.prefix_values.assign(pref, pref + (size_t)nn * n_buckets);
    return t;
}

void Config::validate() const {  // reference config.hpp:78-97
    if (n_estimators < 1) throw Error("n_estimators must be >= 1");
    if (learning_rate <= 0 || learning_rate > 1.5) throw Error("learning_rate out of range");
    if (max_bin < 8 || max_bin > 256) throw Error("max_bin must be in [8, 256]");
    if (!iid_mode && n_time_buckets == 1) throw Error("n_time_buckets must be 0 (auto) or >= 2");
    if (!iid_mode && n_time_buckets < 0) throw Error("n_time_buckets must be 0 (auto) or >= 2");
    if (!iid_mode && horizon < 0)
        throw Error("horizon is required in time mode (label length in timestamp units). "
                    "Overlapping returns without horizon leak into TORB.");
    if (device != "cpu" && device != "gpu") throw Error("device must be cpu or gpu");
    if (device == "cpu")
        throw Error("device=cpu is not provided by this library: it implements the GPU path only");
    // limits of the device path (each fails loudly rather than silently changing results)
    if (max_bin > 255) throw Error("the GPU path stores uint8 bins: max_bin must be <= 255");
    if (max_leaves < 2 || max_leaves > 128) throw Error("the GPU path supports max_leaves in [2, 128]");
    if (objective != 0 && objective != 5) throw Error("the GPU path supports objective mse or pairwise");
    if (halo) throw Error("halo is not supported on the GPU path");
    if (subsample < 1.0) throw Error("subsample < 1 is not supported on the GPU path");
    if (colsample < 1.0) throw Error("colsample < 1 is not supported on the GPU path");
    if (time_decay < 1.0) throw Error("time_decay < 1 is not supported on the GPU path");
    if (tree_decay < 1.0) throw Error("tree_decay < 1 is not supported on the GPU path");
    if (min_samples_leaf < 0) throw Error("min_samples_leaf must be >= 0");
}

struct Booster::EvalRun {
    tgbm_evalset* set = nullptr;
    bool ic = false;        // per-date correlation (higher is better) instead of squared error
    bool scored = false;
    int use = 0;            // trees summed into the device predictions
    double last = 0.0;      // metric of those predictions
    double best = 0.0;
    int stall = 0;
    ~EvalRun() {
        if (set) tgbm_eval_destroy(set);
    }
};

Booster::~Booster() { drop_forest(); }

void Booster::set_eval(const double* X, const double* y, const long long* ts, int n, int nf) {
    if (n < 0) throw Error("n_eval must be >= 0");
    if (n > 0 && (!X || !y || nf <= 0)) throw Error("an eval set needs X, y and n_features > 0");
    n_eval = n;
    eval_features = n > 0 ? nf : 0;
    eval_X.assign(X, X + (size_t)n * eval_features);
    eval_y.assign(y, y + n);
    if (n > 0 && ts) eval_ts.assign(ts, ts + n);
    else eval_ts.clear();
}

// The validation step after tree m (booster.cpp:707-732).  The reference scores the eval rows
// with predict_multi_binned, which sums the first min(trees, best_iteration) trees once an
// iteration has improved (all trees before that): after the first finite score the scored
// prefix stops growing and so does the metric.  The device predictions hold the running sum of
// the scored prefix; a step whose prefix is unchanged reuses the metric it already computed.
bool Booster::eval_step(int m) {
    EvalRun& ev = *eval_run_;
    const int use = std::min((int)trees.size(), best_iteration > 0 ? best_iteration : cfg.n_estimators);
    if (!ev.scored || use != ev.use) {
        int from = ev.use;
        if (!ev.scored || use < ev.use) {
            dev_check(tgbm_eval_reset(ev.set, init_score));
            from = 0;
        }
        std::vector<int> ints;
        std::vector<double> vals;
        for (int t = from; t < use; ++t) {
            ints.clear();
            vals.clear();
            tree_node_ints(trees[t], ints, vals);
            const std::vector<unsigned long long> cm = tree_cat_masks(trees[t]);
            dev_check(tgbm_eval_add_tree(ev.set, (int)trees[t].nodes.size(), ints.data(), vals.data(),
                                         cfg.learning_rate, cm.empty() ? nullptr : cm.data()));
        }
        dev_check(tgbm_eval_metric(ev.set, ev.ic ? 1 : 0, &ev.last));
        ev.use = use;
        ev.scored = true;
    }
    const double vl = ev.last;
    evals_valid.push_back(vl);
    const bool improved = ev.ic ? (vl > ev.best + 1e-12) : (vl + 1e-12 < ev.best);
    if (improved) {
        ev.best = vl;
        ev.stall = 0;
        best_iteration = m + 1;
        return false;
    }
    ev.stall += 1;
    if (cfg.early_stopping_rounds > 0 && ev.stall >= cfg.early_stopping_rounds) {
        const int keep = best_iteration;
        if (keep > 0 && keep < (int)trees.size()) trees.resize(keep);
        return true;
    }
    return false;
}

void Booster::drop_forest() {
    if (forest_) tgbm_forest_destroy(forest_);
    forest_ = nullptr;
    forest_trees_ = -1;
}

void tree_node_ints(const Tree& tr, std::vector<int>& ints, std::vector<double>& vals) {
    for (size_t k = 0; k < tr.nodes.size(); ++k) {
        const Node& d = tr.nodes[k];
        const int v[10] = {d.is_leaf, d.split_type, d.feature, d.feature2, d.threshold, d.left, d.right,
                           d.missing_go_left, d.n_samples, d.depth};
        ints.insert(ints.end(), v, v + 10);
        vals.push_back(tr.values.empty() ? d.value : tr.values[k]);
    }
}

void flat_edges(const std::vector<std::vector<double>>& edges, std::vector<double>& E, std::vector<int>& ne,
                const std::vector<int>& is_cat, const std::vector<std::vector<double>>& cats) {
    const int F = (int)edges.size();
    E.assign((size_t)F * 256, 0.0);
    ne.assign(F, 0);
    for (int f = 0; f < F; ++f) {
        const bool cat = (int)is_cat.size() == F && is_cat[f];
        const std::vector<double>& e = cat ? cats.at(f) : edges[f];
        if (e.size() > 254) throw Error("more than 254 bin edges or categories in a feature");
        ne[f] = (int)e.size() | (cat ? TGBM_CAT_EDGES : 0);
        std::copy(e.begin(), e.end(), E.begin() + (size_t)f * 256);
    }
}

tgbm_forest* make_forest(const Booster& b) {
    const int use = std::min((int)b.trees.size(), b.best_iteration > 0 ? b.best_iteration : b.cfg.n_estimators);
    const int F = b.n_features;
    std::vector<int> tn(use);
    std::vector<int> ints;
    std::vector<double> vals, lrs(use, b.cfg.learning_rate);
    std::vector<unsigned long long> cm;
    const bool cat = b.has_cat();
    for (int t = 0; t < use; ++t) {
        tn[t] = (int)b.trees[t].nodes.size();
        tree_node_ints(b.trees[t], ints, vals);
        if (cat)
            for (const Node& d : b.trees[t].nodes) cm.insert(cm.end(), d.cat_left, d.cat_left + 4);
    }
    std::vector<double> E;
    std::vector<int> ne;
    flat_edges(b.edges, E, ne, b.is_cat, b.cat_values);
    tgbm_forest* out = nullptr;
    dev_check(tgbm_forest_create(&out, F, b.n_bins, use, tn.data(), ints.data(), vals.data(), lrs.data(), b.init_score,
                                 E.data(), ne.data(), cat ? cm.data() : nullptr));
    return out;
}

std::shared_ptr<Booster> Booster::clone_model() const {
    auto m = std::make_shared<Booster>();
    m->cfg = cfg;
    m->n_features = n_features;
    m->n_bins = n_bins;
    m->n_time_buckets = n_time_buckets;
    m->best_iteration = best_iteration;
    m->init_score = init_score;
    m->init_prefix = init_prefix;
    m->edges = edges;
    m->is_cat = is_cat;
    m->group_of = group_of;
    m->feature_names = feature_names;
    m->cat_values = cat_values;
    m->trees = trees;
    m->fitted = fitted;
    return m;
}

void Booster::ensure_forest() {
    const int use = std::min((int)trees.size(), best_iteration > 0 ? best_iteration : cfg.n_estimators);
    if (forest_ && forest_trees_ == use) return;
    drop_forest();
    forest_ = make_forest(*this);
    forest_trees_ = use;
}

void Booster::predict_raw(const double* X, int n, int nf, double* out) {
    if (!fitted) throw Error("not fitted");
    if (nf != n_features) throw Error("n_features mismatch");
    if (n <= 0) return;
    ensure_forest();
    dev_check(tgbm_forest_predict(forest_, X, n, out));
}

void Booster::predict_binned_u16(const unsigned short* xb, int n, double* out) {
    if (!fitted) throw Error("not fitted");
    if (n <= 0) return;
    for (size_t j = 0; j < (size_t)n * n_features; ++j)
        if (xb[j] > 255) throw Error("bin index above 255 is not representable on the GPU path");
    ensure_forest();
    dev_check(tgbm_forest_predict_binned_u16(forest_, xb, n, out));
}

void Booster::fit(const double* X0, const double* y0, int n, int F, const long long* ts0, const double* w0,
                  const int* is_cat_in, const long long* ent0, const long long* rg0) {
    const double t0 = now_s();
    cfg.validate();
    if (n <= 0 || F <= 0) throw Error("empty input");
    if (!group_of.empty() && (int)group_of.size() != F) throw Error("feature groups need one entry per feature");
    if (!cfg.monotone.empty() && (int)cfg.monotone.size() != F)
        throw Error("monotone constraints need one entry per feature");
    feature_names.clear();   // the model file then names the features f0 .. f{F-1} (booster.cpp:244-245)
    is_cat.assign(F, 0);
    if (is_cat_in)
        for (int f = 0; f < F; ++f) is_cat[f] = is_cat_in[f] ? 1 : 0;
    if (!has_cat()) is_cat.clear();
    if (tgbm_device_count() < 1) throw Error("GPU device not available: no HIP device was found for device=gpu");
    if (n >= 1000000) cfg.pair_splits = false;  // booster.cpp:237-240
    drop_forest();
    fitted = false;
    n_features = F;
    // ---- stable sort by (timestamp, entity, index) (booster.cpp:264-309)
    std::vector<double> Xs, ys, ws(n, 1.0);
    std::vector<long long> tss, rgs;
    const double* X = X0;
    const double* y = y0;
    const long long* ts = ts0;
    const long long* rg = rg0;
    bool any_weight = false;
    if (w0) {
        std::copy(w0, w0 + n, ws.begin());
        for (int i = 0; i < n; ++i) any_weight |= w0[i] != 1.0;
    }
    if (!cfg.iid_mode && ts0 && n > 1) {
        std::vector<int> ord(n);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
            if (ts0[a] != ts0[b]) return ts0[a] < ts0[b];
            if (ent0) return ent0[a] < ent0[b];
            return a < b;
        });
        bool perm = false;
        for (int i = 0; i < n && !perm; ++i) perm = ord[i] != i;
        if (perm) {
            Xs.resize((size_t)n * F);
            ys.resize(n);
            tss.resize(n);
            if (rg0) rgs.resize(n);
            std::vector<double> w2(n);
            for (int i = 0; i < n; ++i) {
                const int s = ord[i];
                std::memcpy(&Xs[(size_t)i * F], X0 + (size_t)s * F, sizeof(double) * F);
                ys[i] = y0[s];
                w2[i] = ws[s];
                tss[i] = ts0[s];
                if (rg0) rgs[i] = rg0[s];
            }
            ws.swap(w2);
            X = Xs.data();
            y = ys.data();
            ts = tss.data();
            if (rg0) rg = rgs.data();
        }
    }
    // ---- TORB buckets, causal cuts and initial scores
    TimeLayout tl = time_layout(cfg, ts, y, ws.data(), n);
    n_time_buckets = tl.n_buckets;
    init_prefix = tl.init_prefix;
    init_score = tl.init_score;
    // ---- bin edges (booster.cpp:378-380) and bins on the device
    fit_edges(X, n, F, cfg.max_bin, cfg.bin_sample, (uint32_t)cfg.random_state, edges);
    cat_values.assign(F, {});
    for (int f = 0; f < F; ++f)
        if (!is_cat.empty() && is_cat[f]) {   // a categorical feature has a vocabulary, no edges
            edges[f].clear();
            cat_values[f] = category_values(X, n, F, f, cfg.max_bin, cfg.cat_min_count);
        }
    n_bins = cfg.max_bin;   // Binner::max_bins (binning.cpp:130-134)
    for (int f = 0; f < F; ++f)
        n_bins = std::max(n_bins, (!is_cat.empty() && is_cat[f]) ? (int)cat_values[f].size() + 1
                                                                  : (int)edges[f].size() + 2);
    // ---- eval set (booster.cpp:459-478): later than the training rows, binned with their edges
    std::unique_ptr<EvalRun> ev;
    evals_valid.clear();
    if (n_eval > 0) {
        if (eval_features != F) throw Error("eval set: n_features mismatch");
        if (!cfg.iid_mode) {
            if (eval_ts.empty()) throw Error("eval_timestamp is required when iid_mode=false");
            long long tmax = ts[0];
            for (int i = 1; i < n; ++i) tmax = std::max(tmax, ts[i]);
            long long tmin_e = eval_ts[0];
            for (int i = 1; i < n_eval; ++i) tmin_e = std::min(tmin_e, eval_ts[i]);
            const long long embargo = cfg.embargo + std::max<long long>(cfg.horizon, 0);
            if (tmin_e < tmax - embargo)
                throw Error(
                    "eval timestamps overlap the training future. "
                    "Validation must be strictly later than train (minus embargo/horizon).");
        }
        ev.reset(new EvalRun());
        ev->ic = cfg.eval_metric == "ic" || (cfg.eval_metric == "auto" && (cfg.objective == 5 || !cfg.iid_mode));
        ev->best = ev->ic ? -1e300 : 1e300;
        // dates of the correlation metric: grouped_ic orders the rows with std::sort by timestamp
        // (not stable; the order inside a date feeds the sums) and keeps dates of >= 3 rows
        std::vector<int> order, gstart{0};
        if (ev->ic && n_eval >= 3) {
            std::vector<int> o(n_eval);
            std::iota(o.begin(), o.end(), 0);
            if (!eval_ts.empty()) {
                const long long* te = eval_ts.data();
                std::sort(o.begin(), o.end(), [&](int a, int b) { return te[a] < te[b]; });
                for (int i = 0; i < n_eval;) {
                    int j = i + 1;
                    while (j < n_eval && te[o[j]] == te[o[i]]) ++j;
                    if (j - i >= 3) {
                        order.insert(order.end(), o.begin() + i, o.begin() + j);
                        gstart.push_back((int)order.size());
                    }
                    i = j;
                }
            } else {
                order = o;
                gstart.push_back(n_eval);
            }
        }
        std::vector<double> E;
        std::vector<int> ne;
        flat_edges(edges, E, ne, is_cat, cat_values);
        dev_check(tgbm_eval_create(&ev->set, F, n_bins, E.data(), ne.data(), eval_X.data(), eval_y.data(), n_eval,
                                   order.empty() ? nullptr : order.data(), gstart.data(), (int)gstart.size() - 1));
    }
    // ---- device session
    const bool use_goss = (cfg.goss > 0) || (cfg.goss < 0 && n >= 250000);
    const bool goss_active = use_goss && cfg.goss_top < 1.0 && n > 64;
    const bool hconst = cfg.objective == 0 && !any_weight && !goss_active;
    SessionGuard sg;
    dev_check(tgbm_session_create(&sg.s, n, F, n_bins, 1, &n, cfg.max_leaves, hconst ? 1 : 0));
    if (has_cat()) dev_check(tgbm_session_set_cat(sg.s, is_cat.data()));
    if (!cfg.monotone.empty()) dev_check(tgbm_session_set_monotone(sg.s, cfg.monotone.data()));
    dev_check(tgbm_session_set_spec(sg.s, cfg.spec_width > 1 ? cfg.spec_width : 0));
    bin_on_device(sg.s, X, edges, is_cat, cat_values);
    const long long* groups = (cfg.objective == 5) ? rg : nullptr;
    dev_check(tgbm_task_set_rows(sg.s, 0, y, any_weight ? ws.data() : nullptr, tl.pred.data(), tl.bucket.data(),
                                 nullptr, tl.n_buckets, cfg.iid_mode ? nullptr : tl.cuts.data(), groups));
    trees.clear();
    if (cfg.profile) dev_check(tgbm_timer_enable(sg.s, cfg.profile));
    t_setup = now_s() - t0;
    const double t1 = now_s();
    std::mt19937 rng((uint32_t)cfg.random_state);
    best_iteration = 0;   // booster.cpp:491
    eval_run_ = ev.get();
    try {
        grow_trees(sg.s, n, cfg.n_estimators, !cfg.iid_mode, tl.cuts, rng, goss_active, nullptr, false);
    } catch (...) {
        eval_run_ = nullptr;
        throw;
    }
    eval_run_ = nullptr;
    if (cfg.profile) {
        for (int k = 0; k < 8; ++k) {
            dev_check(tgbm_timer_read(sg.s, k, &prof_ms[k], &prof_n[k], &prof_bytes[k]));
            const int m = tgbm_timer_log(sg.s, k, nullptr, nullptr, 0);
            if (m < 0) dev_check(-1);
            prof_log_ms[k].resize(m);
            prof_log_rows[k].resize(m);
            if (m > 0) tgbm_timer_log(sg.s, k, prof_log_ms[k].data(), prof_log_rows[k].data(), m);
        }
    }
    if (best_iteration <= 0) best_iteration = cfg.n_estimators;   // booster.cpp:739
    fitted = true;
    t_boost = now_s() - t1;
}

// The boosting loop over the device session (booster.cpp:667-703 with grow_head :519-665).
void Booster::grow_trees(tgbm_session* s, int n, int n_trees, bool time_mode, const std::vector<int>& cuts,
                         std::mt19937& rng, bool goss, const double*, bool for_update) {
    (void)cuts;
    tgbm_grow_params gp = grow_params(cfg);
    const int F = n_features;
    // categorical splits are decided on the host (the device-driven selection handles axis splits)
    // (and so are feature groups, interaction groups and the regime feature)
    const bool host_driven = cfg.pair_splits || cfg.force_host_driver || has_cat() || !group_of.empty() ||
                             !cfg.interaction_group.empty() || cfg.regime_feature >= 0;
    gp.host_driven = host_driven ? 1 : 0;
    // the histogram kernel variant: forced, or chosen over the first trees (all give the same trees)
    const std::string shape = "fit:" + std::to_string(launch_rows_of(n)) + ":" + std::to_string(F) + ":" +
                              std::to_string(cfg.objective) + ":" + std::to_string(goss) + ":" +
                              std::to_string(cfg.max_leaves) + ":" + std::to_string(host_driven);
    HistTuner tuner(shape, n_trees, cfg.autotune != 0 && cfg.hist_mode == 0 && !for_update);
    const double min_cover =
        cfg.torb_min_cover > 0.0 ? cfg.torb_min_cover : (double)(2 * std::max(cfg.min_samples_leaf, 1));
    const int MN = 2 * cfg.max_leaves - 1;
    std::vector<int> ints((size_t)MN * TGBM_NODE_INTS);
    std::vector<double> dbl((size_t)MN * 2), vals(MN), pref((size_t)MN * 256);
    std::vector<int> cov(MN);
    std::vector<double> g;
    std::vector<int> sample;
    std::vector<unsigned char> scaled;
    // GOSS (booster.cpp:557-583) runs on the device; the generator's draws come from a thread
    // that works ahead, and a tree whose |g| has ties is sampled on the host with std::sort.
    std::unique_ptr<GossDraws> draws;
    const GossCounts gk = goss_counts(n, cfg);
    if (goss) draws.reset(new GossDraws(rng, gk.n_rest, gk.n_other, n_trees));
    goss_host_trees = 0;
    goss_wait_s = goss_sample_s = 0.0;
    const double t_loop = now_s();
    if (!for_update) {
        tree_end_times.clear();
        evals_train.clear();   // booster.cpp:481: a fit starts a new eval record
    }
    // Device-driven fits without GOSS or an eval set enqueue tree m + 1 before reading tree m back
    // (its gradients follow tree m's prediction update in stream order), so the device does not
    // wait for the host between trees.  Trees whose variant the tuner has yet to decide run one
    // at a time.
    const bool pipeline = !host_driven && !goss && !eval_run_ && !for_update;
    const int iid = (time_mode && !for_update) ? 0 : 1;
    std::vector<double> t_enq(n_trees, 0.0);
    auto enqueue = [&](int m) {
        dev_check(tgbm_task_grad(s, -1, cfg.objective));
        if (goss) {
            const double t0 = now_s();
            const std::vector<int>& drawn = draws->next();
            const double t1 = now_s();
            goss_wait_s += t1 - t0;
            int tie = 0;
            dev_check(tgbm_task_goss(s, 0, cfg.goss_top, cfg.goss_other, drawn.data(), (int)drawn.size(), &tie));
            if (tie) {
                ++goss_host_trees;
                g.resize(n);
                dev_check(tgbm_task_get_grad(s, 0, g.data(), nullptr));
                double scale = 1.0;
                goss_sample_host(g, n, cfg, drawn, sample, scaled, scale);
                dev_check(tgbm_task_set_sample(s, 0, sample.data(), (int)sample.size(), scaled.data(), scale));
            }
            goss_sample_s += now_s() - t1;
        } else {
            (void)rng();  // the tree's own seed (booster.cpp:586-588); only colsample would use it
        }
        if (cfg.hist_mode == 0) gp.hist_mode = tuner.mode(m);
        t_enq[m] = now_s();
        dev_check(tgbm_tree_begin(s, &gp));
        if (host_driven) grow_host_driven(s, cfg, F, gp, is_cat, group_of);
        else dev_check(tgbm_tree_grow_device(s, &gp));
        dev_check(tgbm_tree_end_async(s, cfg.learning_rate, min_cover, iid, &gp));
    };
    int enqueued = 0;
    for (int m = 0; m < n_trees; ++m) {
        if (enqueued == m) enqueue(enqueued++);
        if (pipeline && enqueued == m + 1 && m + 1 < n_trees && tuner.settled(m + 1)) enqueue(enqueued++);
        int nn = 0;
        dev_check(tgbm_tree_collect(s, &gp, &nn, ints.data(), dbl.data(), vals.data(), cov.data(), pref.data()));
        tuner.record(m, now_s() - t_enq[m]);   // the collect waited for tree m
        if (!for_update) {
            double loss = 0.0;
            dev_check(tgbm_task_train_loss(s, 0, &loss));
            evals_train.push_back(loss);   // booster.cpp:704-706
        }
        Tree t = unpack_tree(nn, ints.data(), dbl.data(), vals.data(), cov.data(), pref.data(), n_time_buckets);
        if (has_cat()) {
            std::vector<unsigned long long> cm((size_t)nn * 4);
            dev_check(tgbm_task_cat_masks(s, 0, nn, cm.data()));
            for (int k = 0; k < nn; ++k)
                if (!t.nodes[k].is_leaf && t.nodes[k].split_type >= 2) std::copy(&cm[4 * k], &cm[4 * k] + 4, t.nodes[k].cat_left);
        }
        if (for_update) {
            // Booster::update keeps grow_tree's cover (node sample counts) and one bucket
            for (int k = 0; k < nn; ++k) t.cover[k] = t.nodes[k].n_samples;
            t.n_buckets = 1;
            t.prefix_values = t.values;
        }
        trees.push_back(std::move(t));
        if (!for_update) tree_end_times.push_back(now_s() - t_loop);  // tree_end synchronized
        if (!for_update && eval_run_ && eval_step(m)) break;            // early stopping
    }
    hist_variant = gp.hist_mode;
}

void Booster::update(const double* X, const double* y, int n, int nf, const double* w, int extra_trees) {
    if (!fitted) throw Error("not fitted");
    if (nf != n_features) throw Error("n_features mismatch");
    if (n <= 0 || extra_trees <= 0) return;
    bool any_weight = false;
    if (w)
        for (int i = 0; i < n; ++i) any_weight |= w[i] != 1.0;
    std::vector<double> pred(n);
    predict_raw(X, n, nf, pred.data());  // booster.cpp:859: current model, frozen binner
    const bool hconst = cfg.objective == 0 && !any_weight;
    SessionGuard sg;
    dev_check(tgbm_session_create(&sg.s, n, nf, n_bins, 1, &n, cfg.max_leaves, hconst ? 1 : 0));
    if (has_cat()) dev_check(tgbm_session_set_cat(sg.s, is_cat.data()));
    if (!cfg.monotone.empty()) dev_check(tgbm_session_set_monotone(sg.s, cfg.monotone.data()));
    dev_check(tgbm_session_set_spec(sg.s, cfg.spec_width > 1 ? cfg.spec_width : 0));
    bin_on_device(sg.s, X, edges, is_cat, cat_values);
    std::vector<int> zeros(n, 0);
    dev_check(tgbm_task_set_rows(sg.s, 0, y, any_weight ? w : nullptr, pred.data(), zeros.data(), nullptr, 1, nullptr,
                                 nullptr));
    std::mt19937 rng((uint32_t)cfg.random_state + (uint32_t)trees.size());
    const int n_before = (int)trees.size();
    drop_forest();
    grow_trees(sg.s, n, extra_trees, false, {}, rng, false, nullptr, true);
    (void)n_before;
    best_iteration = (int)trees.size();
}

}  // namespace tide
