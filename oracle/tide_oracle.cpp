// ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker; the product never links it.
//
// A CPU restatement (scalar C++, no OpenMP) of the reference TideGBM histogram-GBDT path,
// restricted to what the device path covers: numeric features, the MSE and within-date
// pairwise objectives, TORB time buckets, GOSS, PAIR splits, best-first growth, leaf refit,
// predict and update.  Every function cites the reference file:line it restates.
// It is pinned bit-for-bit against the reference core itself (oracle/_ref, built from the
// reference sources by oracle/Makefile.ref) on seeded panels; see tests/test_oracle_pin.py.
//
// Compiled with the reference's flags (-O3 -std=c++17 -fno-math-errno, x86-64 baseline: no FMA
// contraction is possible) so that sequential FP64 sums and glibc exp match the reference.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

using i64 = long long;

struct Cfg {  // defaults: reference cpp/include/tidegbm/config.hpp:13-76, cpp/src/bindings.cpp:43-126
    int n_estimators = 100;
    double learning_rate = 0.1;
    int max_depth = 6;
    int max_leaves = 31;
    int min_samples_leaf = 20;
    double min_child_weight = 1e-3;
    double l2 = 1.0;
    double l1 = 0.0;
    double max_delta_step = 0.0;
    double min_split_gain = 0.0;
    int max_bin = 255;
    int n_time_buckets = 0;
    i64 horizon = -1;
    i64 embargo = 0;
    bool iid = false;
    bool pair_splits = true;
    double pair_trigger = 0.25;
    int pair_top_k = 6;
    int max_pair_splits = 3;
    double pair_l2 = 2.0;
    int objective = 0;  // 0 MSE, 5 pairwise (types.hpp:26-35)
    int random_state = 0;
    int goss = -1;
    double goss_top = 0.2;
    double goss_other = 0.1;
    int bin_sample = 262144;
    bool finer_torb = true;
    double torb_min_cover = 0.0;
    int cat_min_count = 5;          // config.hpp / bindings.cpp cat_min_count
    int regime_feature = -1;        // config.hpp:55-56
    bool regime_priority = true;
    std::vector<int> interaction_group;  // config.hpp:59, per feature, -1 = unconstrained
    double cat_smooth = 1.0;        // config.hpp:70
    int early_stopping_rounds = 0;  // config.hpp early_stopping_rounds (api.py:23-24: None -> 0)
    int eval_metric = 0;            // config.hpp:76 "auto" | "loss" | "ic" as 0 | 1 | 2
};

struct Node {  // tree.hpp:11-25
    int is_leaf = 1, split_type = 0, feature = -1, feature2 = -1, thr = -1, left = -1, right = -1,
        mgl = 1, n_samples = 0, depth = 0;
    double value = 0.0, gain = 0.0;
    uint64_t cat_left[4] = {0, 0, 0, 0};   // categorical split: bins that go left
};

struct Tree {  // tree.hpp:32-41
    std::vector<Node> nodes;
    std::vector<double> values, prefix;
    std::vector<int> cover;
    int n_buckets = 1;
};

// ---------------------------------------------------------------- leaf formulas (tree.hpp:43-67)
static double soft(double G, double l1) {
    if (l1 <= 0) return G;
    if (G > l1) return G - l1;
    if (G < -l1) return G + l1;
    return 0.0;
}
static double score_of(double G, double H, double lam, double l1) {  // histogram.cpp:126-129
    const double t = soft(G, l1);
    return (t * t) / (H + lam);
}
static double weight_of(double G, double H, double lam, double l1) {  // histogram.cpp:136-142
    const double den = H + lam;
    if (l1 <= 0) return -G / den;
    if (G > l1) return -(G - l1) / den;
    if (G < -l1) return -(G + l1) / den;
    return 0.0;
}
static double leaf_val(double G, double H, double l2, double l1, double mds) {  // tree.hpp:53-67
    double v;
    if (l1 > 0) {
        if (G > l1) v = -(G - l1) / (H + l2);
        else if (G < -l1) v = -(G + l1) / (H + l2);
        else v = 0.0;
    } else {
        v = -G / (H + l2);
    }
    if (mds > 0) v = std::min(std::max(v, -mds), mds);
    return v;
}

// ---------------------------------------------------------------- binning (binning.cpp:14-128)
// Quantile right-edges of one feature: sorted distinct sample values when they fit in
// `nb` slots, else nb-1 linearly interpolated quantiles (binning.cpp:14-32).
static std::vector<double> edges_of(std::vector<double>& v, int nb) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    if (v.empty()) return {0.0};
    if ((int)v.size() <= nb) return v;
    std::vector<double> e;
    const size_t top = v.size() - 1;
    for (int i = 1; i < nb; ++i) {
        const double pos = (double)i / nb * top;
        const size_t lo = (size_t)pos;
        const size_t hi = std::min(lo + 1, top);
        const double fr = pos - lo;
        e.push_back(v[lo] * (1 - fr) + v[hi] * fr);
    }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    return e;
}

// Per-feature edges over non-NaN values; subsample with mt19937(seed + f*2654435761) and
// std::shuffle when more than bin_sample values (binning.cpp:84-99).
static void fit_edges(const double* X, int n, int F, int max_bin, int bin_sample, uint32_t seed,
                      std::vector<std::vector<double>>& edges) {
    edges.assign(F, {});
    for (int f = 0; f < F; ++f) {
        std::vector<double> v;
        v.reserve(n);
        for (int i = 0; i < n; ++i) {
            const double x = X[(size_t)i * F + f];
            if (x == x) v.push_back(x);
        }
        if (bin_sample > 0 && (int)v.size() > bin_sample) {
            std::mt19937 rng(seed + (uint32_t)f * 2654435761u);
            std::shuffle(v.begin(), v.end(), rng);
            v.resize((size_t)bin_sample);
        }
        edges[f] = edges_of(v, max_bin - 1);
    }
}

// Category vocabulary of one categorical feature (binning.cpp:48-81): the distinct non-NaN
// values seen at least cat_min_count times (the most frequent one if none is), at most
// max_bin - 1 of them by frequency (std::partial_sort), in ascending order.
static std::vector<double> categories_of(const double* X, int n, int F, int f, int max_bin, int cat_min_count) {
    std::vector<double> raw;
    raw.reserve(n);
    for (int i = 0; i < n; ++i) {
        const double v = X[(size_t)i * F + f];
        if (v == v) raw.push_back(v);
    }
    std::sort(raw.begin(), raw.end());
    std::vector<std::pair<int, double>> freq;
    for (size_t i = 0; i < raw.size();) {
        size_t j = i + 1;
        while (j < raw.size() && raw[j] == raw[i]) ++j;
        freq.emplace_back((int)(j - i), raw[i]);
        i = j;
    }
    const int min_c = std::max(1, cat_min_count);
    std::vector<std::pair<int, double>> kept;
    for (const auto& q : freq)
        if (q.first >= min_c) kept.push_back(q);
    if (kept.empty() && !freq.empty())
        kept.push_back(*std::max_element(freq.begin(), freq.end(),
                                         [](const auto& a, const auto& b) { return a.first < b.first; }));
    const int cap = max_bin - 1;
    if ((int)kept.size() > cap) {
        std::partial_sort(kept.begin(), kept.begin() + cap, kept.end(),
                          [](const auto& a, const auto& b) { return a.first > b.first; });
        kept.resize(cap);
    }
    std::vector<double> u;
    for (const auto& q : kept) u.push_back(q.second);
    std::sort(u.begin(), u.end());
    return u;
}

// Bin of one value: 0 for NaN, else 1 + number of edges <= v (binning.cpp:110-123).
static int bin_of(const std::vector<double>& e, double x) {
    if (!(x == x)) return 0;
    return 1 + (int)(std::upper_bound(e.begin(), e.end(), x) - e.begin());
}

// Bin of a categorical value: 1 + its index in the vocabulary, 0 for NaN or an unseen value
// (binning.cpp:107-112).
static int cat_bin_of(const std::vector<double>& voc, double x) {
    if (!(x == x)) return 0;
    const auto it = std::lower_bound(voc.begin(), voc.end(), x);
    return (it != voc.end() && *it == x) ? (int)(it - voc.begin()) + 1 : 0;
}

// n_bins = max(max_bin, max over features of n_edges+2, or categories+1) (binning.cpp:82,99,130-134).
static int total_bins(const std::vector<std::vector<double>>& edges, const std::vector<int>& is_cat,
                      const std::vector<std::vector<double>>& cats, int max_bin) {
    int m = 2;
    for (size_t f = 0; f < edges.size(); ++f)
        m = std::max(m, (!is_cat.empty() && is_cat[f]) ? (int)cats[f].size() + 1 : (int)edges[f].size() + 2);
    return std::max(m, max_bin);
}

// ---------------------------------------------------------------- histograms
struct Hist {  // histogram.hpp:14-46, planar
    int F = 0, nb = 0;
    std::vector<double> g, h;
    std::vector<i64> c;
    void init(int F_, int nb_) {
        F = F_;
        nb = nb_;
        g.assign((size_t)F * nb, 0.0);
        h.assign((size_t)F * nb, 0.0);
        c.assign((size_t)F * nb, 0);
    }
};

// Per feature, rows in ascending idx order: fold g, h, count into the row's bin
// (histogram.cpp:44-76; row-major build_hist :24-42 gives identical sums).
static void build(const uint16_t* xb, int F, const int* idx, int m, const double* g,
                  const double* h, Hist& out, int nb) {
    out.init(F, nb);
    for (int f = 0; f < F; ++f) {
        double* G = &out.g[(size_t)f * nb];
        double* H = &out.h[(size_t)f * nb];
        i64* C = &out.c[(size_t)f * nb];
        for (int k = 0; k < m; ++k) {
            const int i = idx[k];
            const int b = xb[(size_t)i * F + f];
            G[b] += g[i];
            H[b] += h[i];
            C[b] += 1;
        }
    }
}

static void subtract_into(Hist& parent, const Hist& child) {  // histogram.cpp:108-117
    for (size_t i = 0; i < parent.g.size(); ++i) {
        parent.g[i] -= child.g[i];
        parent.h[i] -= child.h[i];
        parent.c[i] -= child.c[i];
    }
}

struct Cand {  // histogram.hpp:65-74
    double gain = -1e300;
    int feature = -1, feature2 = -1, thr = -1, type = 0, mgl = 1;
    uint64_t cat_left[4] = {0, 0, 0, 0};
};

// Best numeric split of feature f with bin 0 = missing, thresholds b in [1, nb-2], empty bins
// skipped, missing-right tried before missing-left, strict '>' keeps the first maximum
// (histogram.cpp:144-193).
static Cand scan_feature(const Hist& H, int f, double lam, double gamma, int min_leaf,
                         double min_hess, double l1, int mono = 0) {
    Cand best;
    const int nb = H.nb;
    const double* G = &H.g[(size_t)f * nb];
    const double* Hh = &H.h[(size_t)f * nb];
    const i64* C = &H.c[(size_t)f * nb];
    double tG = 0, tH = 0;
    i64 tC = 0;
    for (int b = 0; b < nb; ++b) {
        tG += G[b];
        tH += Hh[b];
        tC += C[b];
    }
    const double parent = score_of(tG, tH, lam, l1);
    double GL = 0, HL = 0;
    i64 CL = 0;
    for (int b = 1; b < nb - 1; ++b) {
        GL += G[b];
        HL += Hh[b];
        CL += C[b];
        if (C[b] == 0) continue;
        for (int ml = 0; ml <= 1; ++ml) {
            const double gl = GL + (ml ? G[0] : 0);
            const double hl = HL + (ml ? Hh[0] : 0);
            const i64 cl = CL + (ml ? C[0] : 0);
            const double gr = tG - gl, hr = tH - hl;
            const i64 cr = tC - cl;
            if (cl < min_leaf || cr < min_leaf) continue;
            if (hl < min_hess || hr < min_hess) continue;
            // monotone constraint of the feature (histogram.cpp:177-180): the leaf weights must
            // keep its order
            if (mono != 0) {
                const double wl = weight_of(gl, hl, lam, l1), wr = weight_of(gr, hr, lam, l1);
                if (mono > 0 && wl < wr) continue;
                if (mono < 0 && wl > wr) continue;
            }
            const double gain =
                0.5 * (score_of(gl, hl, lam, l1) + score_of(gr, hr, lam, l1) - parent) - gamma;
            if (gain > best.gain) {
                best.gain = gain;
                best.feature = f;
                best.thr = b;
                best.type = 0;
                best.mgl = ml;
            }
        }
    }
    return best;
}

// Best categorical split of feature f (histogram.cpp:195-244): the bins with rows, missing
// (bin 0) included, ordered by G / (H + cat_smooth) with std::sort (not stable: equal ratios
// take the library's order), then a left-to-right scan of that order with sequential FP64
// sums; the left set is the bin mask up to the best position, strict '>' keeps the first.
static Cand scan_categorical(const Hist& H, int f, double lam, double gamma, int min_leaf, double min_hess,
                             double cat_smooth, double l1) {
    Cand best;
    const int nb = H.nb;
    const double* G = &H.g[(size_t)f * nb];
    const double* Hh = &H.h[(size_t)f * nb];
    const i64* C = &H.c[(size_t)f * nb];
    std::vector<int> order;
    double tG = 0, tH = 0;
    i64 tC = 0;
    for (int b = 0; b < nb; ++b) {
        tG += G[b];
        tH += Hh[b];
        tC += C[b];
        if (C[b] > 0) order.push_back(b);
    }
    if (order.size() < 2) return best;
    const double sm = std::max(cat_smooth, 0.0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return G[a] / (Hh[a] + sm) < G[b] / (Hh[b] + sm); });
    const double parent = score_of(tG, tH, lam, l1);
    double GL = 0, HL = 0;
    i64 CL = 0;
    uint64_t mask[4] = {0, 0, 0, 0};
    for (size_t i = 0; i + 1 < order.size(); ++i) {
        const int b = order[i];
        GL += G[b];
        HL += Hh[b];
        CL += C[b];
        mask[b >> 6] |= uint64_t(1) << (b & 63);
        const double GR = tG - GL, HR = tH - HL;
        const i64 CR = tC - CL;
        if (CL < min_leaf || CR < min_leaf) continue;
        if (HL < min_hess || HR < min_hess) continue;
        const double gain = 0.5 * (score_of(GL, HL, lam, l1) + score_of(GR, HR, lam, l1) - parent) - gamma;
        if (gain > best.gain) {
            best.gain = gain;
            best.feature = f;
            best.type = 2;
            best.thr = -1;
            best.mgl = (mask[0] & 1ull) ? 1 : 0;
            for (int k = 0; k < 4; ++k) best.cat_left[k] = mask[k];
        }
    }
    return best;
}

// pair hash (histogram.hpp:81-85)
static inline int pair_hash(int b1, int b2, int nb) {
    return (int)(((uint32_t)b1 * 2654435761u + (uint32_t)b2 * 40503u) % (uint32_t)nb);
}

// Node routing of one bin row (tree.cpp:15-41 / 68-91): Axis: bin 0 -> missing direction,
// else bin <= threshold; Pair: either bin 0 -> missing direction, else hash <= threshold.
static bool goes_left(const Node& nd, const uint16_t* row, int nb) {
    if (nd.split_type == 2 || nd.split_type == 3) {  // cat_goes_left (tree.hpp:27-30)
        const int b = row[nd.feature];
        if (b >= 256) return nd.mgl != 0;
        return (nd.cat_left[b >> 6] >> (b & 63)) & 1ull;
    }
    if (nd.split_type == 1) {
        const int b1 = row[nd.feature], b2 = row[nd.feature2];
        if (b1 == 0 || b2 == 0) return nd.mgl != 0;
        return pair_hash(b1, b2, nb) <= nd.thr;
    }
    const int b = row[nd.feature];
    if (b == 0) return nd.mgl != 0;
    return b <= nd.thr;
}

static int leaf_of(const Tree& t, const uint16_t* row, int nb) {  // tree.cpp:15-41
    int k = 0;
    while (!t.nodes[k].is_leaf) k = goes_left(t.nodes[k], row, nb) ? t.nodes[k].left : t.nodes[k].right;
    return k;
}

// ---------------------------------------------------------------- leaf candidate (tree.cpp:106-281)
struct GrowCtx {
    const uint16_t* xb;
    int F, nb;
    const double *g, *h;
    const Cfg* cfg;
    const int* is_cat;   // per feature, or null
    const int* group_of; // per feature feature group (Booster::fit's group_of), or null
    const int* monotone; // per feature monotone constraint (Booster::fit's monotone), or null
};

// Features a node may split on (tree.cpp:304-317): every feature, or under an interaction group
// ixg >= 0 the unconstrained features and those of group ixg, with the regime feature appended
// when it is missing; an empty list falls back to every feature.
static std::vector<int> active_features(const Cfg& c, int F, int ixg) {
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

static Cand leaf_candidate(const GrowCtx& cx, const std::vector<int>& rows, const Hist& hist,
                           int pair_budget, double energy, const std::vector<int>& active) {
    const Cfg& c = *cx.cfg;
    const int F = cx.F;
    std::vector<Cand> per_f(F);
    for (int f : active)  // tree.cpp:126-133
        per_f[f] = (cx.is_cat && cx.is_cat[f])
                       ? scan_categorical(hist, f, c.l2, c.min_split_gain, c.min_samples_leaf, c.min_child_weight,
                                          c.cat_smooth, c.l1)
                       : scan_feature(hist, f, c.l2, c.min_split_gain, c.min_samples_leaf, c.min_child_weight,
                                      c.l1, cx.monotone ? cx.monotone[f] : 0);
    // feature groups (tree.cpp:135-158): a group is represented by its best-gain feature (the
    // first one in `active` order wins ties); groups numbered below 256 appear once
    std::map<int, int> group_best;
    for (int f : active) {
        const int grp = cx.group_of ? cx.group_of[f] : -1;
        if (grp < 0) continue;
        auto it = group_best.find(grp);
        if (it == group_best.end()) group_best[grp] = f;
        else if (per_f[f].gain > per_f[it->second].gain) it->second = f;
    }
    std::vector<int> allowed;
    std::vector<char> seen(256, 0);
    for (int f : active) {
        const int grp = cx.group_of ? cx.group_of[f] : -1;
        if (grp >= 0) {
            if (grp < 256 && seen[grp]) continue;
            if (grp < 256) seen[grp] = 1;
            f = group_best[grp];
        }
        allowed.push_back(f);
    }
    Cand best;
    for (int f : allowed) {  // strict '>' keeps the first maximum (tree.cpp:160-166)
        if (per_f[f].feature < 0 || per_f[f].gain < c.min_split_gain) continue;
        if (per_f[f].gain > best.gain) best = per_f[f];
    }
    // regime priority (tree.cpp:168-176): the regime feature's split wins when its gain is at
    // least half the best; on a categorical regime feature it becomes a Regime node (type 3)
    if (c.regime_priority && c.regime_feature >= 0 && c.regime_feature < F) {
        const int rf = c.regime_feature;
        if (per_f[rf].feature >= 0 && per_f[rf].gain > 0) {
            if (best.feature < 0 || per_f[rf].gain >= 0.5 * std::max(best.gain, 0.0)) {
                best = per_f[rf];
                best.type = (cx.is_cat && cx.is_cat[rf]) ? 3 : 0;
            }
        }
    }
    // PAIR search on weak leaves (tree.cpp:215-279)
    const bool weak =
        (best.gain < c.pair_trigger * std::max(energy * 0.01, 1e-12)) || best.gain <= 0;
    if (c.pair_splits && pair_budget > 0 && allowed.size() >= 2 && weak) {
        std::vector<int> top = allowed;
        std::sort(top.begin(), top.end(),
                  [&](int a, int b) { return per_f[a].gain > per_f[b].gain; });
        const int k = std::min(c.pair_top_k, (int)top.size());
        std::vector<std::pair<int, int>> pairs;
        for (int a = 0; a < k; ++a)
            for (int b = a + 1; b < k; ++b) pairs.emplace_back(top[a], top[b]);
        const int m = (int)rows.size();
        for (auto& pr : pairs) {
            const int f1 = pr.first, f2 = pr.second;
            if (f1 < 0 || f2 < 0 || f1 >= F || f2 >= F || f1 == f2) continue;
            Hist hz;
            hz.init(1, cx.nb);
            for (int q = 0; q < m; ++q) {
                const int i = rows[q];
                const uint16_t* row = cx.xb + (size_t)i * F;
                const int z = pair_hash(row[f1], row[f2], cx.nb);
                hz.g[z] += cx.g[i];
                hz.h[z] += cx.h[i];
                hz.c[z] += 1;
            }
            Cand pc = scan_feature(hz, 0, c.pair_l2, c.min_split_gain, c.min_samples_leaf,
                                   c.min_child_weight, c.l1);
            pc.type = 1;
            pc.feature = f1;
            pc.feature2 = f2;
            if (pc.thr < 0) continue;  // no valid split
            if (pc.gain > best.gain) best = pc;
        }
    }
    return best;
}

// ---------------------------------------------------------------- best-first growth (tree.cpp:283-472)
static Tree grow(const GrowCtx& cx, std::vector<int> sample) {
    const Cfg& c = *cx.cfg;
    Tree t;
    std::sort(sample.begin(), sample.end());
    Node root;
    root.n_samples = (int)sample.size();
    double sg = 0, sh = 0;
    for (int i : sample) {
        sg += cx.g[i];
        sh += cx.h[i];
    }
    root.value = leaf_val(sg, sh, c.l2, c.l1, c.max_delta_step);
    t.nodes.push_back(root);
    std::vector<std::vector<int>> rows(1, sample);
    std::vector<Hist> hists(1);
    build(cx.xb, cx.F, sample.data(), (int)sample.size(), cx.g, cx.h, hists[0], cx.nb);
    std::vector<Cand> cache(1);
    std::vector<char> cached(1, 0);
    std::vector<int> ixg(1, -1);   // interaction group a node's path is bound to (tree.cpp:440-445)
    int pair_used = 0;
    for (int it = 0; it < c.max_leaves - 1; ++it) {
        int best_leaf = -1;
        Cand best;
        double best_score = c.min_split_gain;
        for (int lid = 0; lid < (int)t.nodes.size(); ++lid) {
            if (!t.nodes[lid].is_leaf || t.nodes[lid].depth >= c.max_depth) continue;
            if ((int)rows[lid].size() < 2 * c.min_samples_leaf) continue;
            if (!cached[lid]) {
                double energy = 0;
                for (int i : rows[lid]) energy += cx.g[i] * cx.g[i];
                cache[lid] = leaf_candidate(cx, rows[lid], hists[lid], c.max_pair_splits - pair_used, energy,
                                            active_features(c, cx.F, ixg[lid]));
                cached[lid] = 1;
            }
            const Cand& r = cache[lid];
            if (r.feature >= 0 && r.gain >= c.min_split_gain && r.gain > best_score) {
                best_score = r.gain;
                best_leaf = lid;
                best = r;
            }
        }
        if (best_leaf < 0) break;
        Node& p = t.nodes[best_leaf];
        p.is_leaf = 0;
        p.feature = best.feature;
        p.feature2 = best.feature2;
        p.thr = best.thr;
        p.split_type = best.type;
        p.mgl = best.mgl;
        p.gain = best.gain;
        for (int k = 0; k < 4; ++k) p.cat_left[k] = best.cat_left[k];
        if (best.type == 1) {  // a PAIR split consumes budget and invalidates every cache
            ++pair_used;
            std::fill(cached.begin(), cached.end(), 0);
        }
        cached[best_leaf] = 0;
        std::vector<int> L, R;  // stable partition (tree.cpp:68-91)
        for (int i : rows[best_leaf])
            (goes_left(p, cx.xb + (size_t)i * cx.F, cx.nb) ? L : R).push_back(i);
        Node ln, rn;
        ln.depth = rn.depth = p.depth + 1;
        ln.n_samples = (int)L.size();
        rn.n_samples = (int)R.size();
        for (auto pr : {std::make_pair(&ln, &L), std::make_pair(&rn, &R)}) {
            double G = 0, H = 0;
            for (int i : *pr.second) {
                G += cx.g[i];
                H += cx.h[i];
            }
            pr.first->value = leaf_val(G, H, c.l2, c.l1, c.max_delta_step);
        }
        const bool left_small = L.size() <= R.size();
        Hist small;
        const std::vector<int>& sm = left_small ? L : R;
        build(cx.xb, cx.F, sm.data(), (int)sm.size(), cx.g, cx.h, small, cx.nb);
        Hist big = std::move(hists[best_leaf]);
        subtract_into(big, small);
        const int lid = (int)t.nodes.size();
        p.left = lid;
        p.right = lid + 1;
        t.nodes.push_back(ln);
        t.nodes.push_back(rn);
        rows.push_back(std::move(L));
        rows.push_back(std::move(R));
        hists.push_back(left_small ? std::move(small) : std::move(big));
        hists.push_back(left_small ? std::move(big) : std::move(small));
        hists[best_leaf] = Hist();
        rows[best_leaf].clear();
        cache.emplace_back();
        cache.emplace_back();
        cached.push_back(0);
        cached.push_back(0);
        int child_ixg = ixg[best_leaf];
        if (child_ixg < 0 && best.feature >= 0 && best.feature < (int)c.interaction_group.size() &&
            c.interaction_group[best.feature] >= 0)
            child_ixg = c.interaction_group[best.feature];
        ixg.push_back(child_ixg);
        ixg.push_back(child_ixg);
    }
    return t;
}

// ---------------------------------------------------------------- objectives (objective.cpp)
static void grad_mse(const double* y, const double* p, const double* w, int n, double* g,
                     double* h) {  // objective.cpp:105-119 (cost terms off)
    for (int i = 0; i < n; ++i) {
        g[i] = w[i] * (p[i] - y[i]);
        h[i] = w[i];
    }
}

// Within-group pairwise gradient (objective.cpp:179-219); groups are contiguous runs of equal
// keys when sorted, else a stable sort by key (group_runs, objective.cpp:18-39).
static void grad_pairwise(const double* y, const double* p, const double* w, const i64* grp,
                          int n, double* g, double* h) {
    std::fill(g, g + n, 0.0);
    std::fill(h, h + n, 1e-6);
    if (!grp) return;
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    bool sorted = true;
    for (int i = 1; i < n; ++i)
        if (grp[i] < grp[i - 1]) sorted = false;
    if (!sorted)
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return grp[a] < grp[b]; });
    int s0 = 0;
    while (s0 < n) {
        int s1 = s0 + 1;
        while (s1 < n && grp[ord[s1]] == grp[ord[s0]]) ++s1;
        const int m = s1 - s0;
        const int* ix = ord.data() + s0;
        if (m >= 2) {
            const int step = m > 48 ? std::max(1, m / 24) : 1;
            for (int a = 0; a < m; ++a)
                for (int b = a + step; b < m; b += step) {
                    const int i = ix[a], j = ix[b];
                    if (y[i] == y[j]) continue;
                    const int hi = y[j] > y[i] ? j : i;
                    const int lo = y[j] > y[i] ? i : j;
                    const double s = p[hi] - p[lo];
                    const double e = std::exp(std::min(std::max(s, -30.0), 30.0));
                    const double dyw = std::max(std::abs(y[hi] - y[lo]), 1e-6);
                    const double lam = -1.0 / (1.0 + e) * dyw;
                    const double d2 = e / ((1 + e) * (1 + e)) * dyw;
                    g[hi] += w[hi] * lam;
                    g[lo] -= w[lo] * lam;
                    h[hi] += w[hi] * d2;
                    h[lo] += w[lo] * d2;
                }
        }
        s0 = s1;
    }
}

// Training loss reported after every tree (objective.cpp:262-329, recorded at booster.cpp:704-706):
// the pairwise objective with groups reports minus the mean within-group Pearson correlation
// over groups of at least 3 rows whose variances exceed 1e-18; every other objective here the
// weighted mean squared error.  Sequential sums in row / group order.
static double loss_of(int objective, const double* y, const double* p, const double* w, const i64* grp,
                      int n) {
    if (objective == 5 && grp) {
        std::vector<int> ord(n);
        std::iota(ord.begin(), ord.end(), 0);
        bool sorted = true;
        for (int i = 1; i < n; ++i)
            if (grp[i] < grp[i - 1]) sorted = false;
        if (!sorted) std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return grp[a] < grp[b]; });
        double acc = 0;
        int ng = 0;
        int s0 = 0;
        while (s0 < n) {
            int s1 = s0 + 1;
            while (s1 < n && grp[ord[s1]] == grp[ord[s0]]) ++s1;
            const int m = s1 - s0;
            const int* ix = ord.data() + s0;
            s0 = s1;
            if (m < 3) continue;
            double my = 0, mp = 0;
            for (int k = 0; k < m; ++k) {
                my += y[ix[k]];
                mp += p[ix[k]];
            }
            my /= m;
            mp /= m;
            double vy = 0, vp = 0, c = 0;
            for (int k = 0; k < m; ++k) {
                const double dy = y[ix[k]] - my, dp = p[ix[k]] - mp;
                vy += dy * dy;
                vp += dp * dp;
                c += dy * dp;
            }
            if (vy > 1e-18 && vp > 1e-18) {
                acc += c / std::sqrt(vy * vp);
                ng += 1;
            }
        }
        return ng > 0 ? -acc / ng : 0.0;
    }
    double sum = 0, sw = 0;
    for (int i = 0; i < n; ++i) {
        const double d = p[i] - y[i];
        sum += w[i] * d * d;
        sw += w[i];
    }
    return sw > 0 ? sum / sw : 0;
}

// Validation metric helpers (booster.cpp:132-174).  pearson_ix: sequential means and sums over
// the listed rows; a group of fewer than 3 rows or with a variance below 1e-18 scores 0.
static double pearson_ix(const double* y, const double* p, const int* ix, int m) {
    if (m < 3) return 0.0;
    double my = 0, mp = 0;
    for (int k = 0; k < m; ++k) {
        my += y[ix[k]];
        mp += p[ix[k]];
    }
    my /= m;
    mp /= m;
    double vy = 0, vp = 0, c = 0;
    for (int k = 0; k < m; ++k) {
        const double dy = y[ix[k]] - my;
        const double dp = p[ix[k]] - mp;
        vy += dy * dy;
        vp += dp * dp;
        c += dy * dp;
    }
    if (vy < 1e-18 || vp < 1e-18) return 0.0;
    return c / std::sqrt(vy * vp);
}

// grouped_ic (booster.cpp:155-174): rows ordered by timestamp with std::sort (not stable: the
// order inside a date is the library's, and it feeds the sequential sums), the mean of the
// per-date correlations over dates of at least 3 rows; no timestamps = one correlation.
static double grouped_ic(const double* y, const double* pred, const i64* ts, int n) {
    if (n < 3) return 0.0;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    if (!ts) return pearson_ix(y, pred, order.data(), n);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return ts[a] < ts[b]; });
    double s = 0;
    int ng = 0;
    for (int i = 0; i < n;) {
        int j = i + 1;
        while (j < n && ts[order[j]] == ts[order[i]]) ++j;
        if (j - i >= 3) {
            s += pearson_ix(y, pred, order.data() + i, j - i);
            ng += 1;
        }
        i = j;
    }
    return ng > 0 ? s / ng : 0.0;
}

static double wmean(const double* y, const double* w, int n) {  // objective.cpp:41-49
    double sw = 0, sy = 0;
    for (int i = 0; i < n; ++i) {
        sw += w[i];
        sy += w[i] * y[i];
    }
    return sw > 0 ? sy / sw : 0;
}

// ---------------------------------------------------------------- model and fit (booster.cpp:231-742)
struct Model {
    Cfg cfg;
    int F = 0, nb = 255, n_buckets = 1, best_iteration = 0;
    std::vector<std::vector<double>> edges;
    double init_score = 0.0;
    std::vector<double> init_prefix;
    std::vector<Tree> trees;
    std::vector<double> evals_train;   // training loss after each tree of the last fit
    std::vector<double> evals_valid;   // validation metric after each tree of the last fit
    // Eval set of later fits (Booster::fit's X_eval, y_eval, ts_eval, n_eval, booster.cpp:231-234).
    std::vector<double> Xe, ye;
    std::vector<i64> te;
    int n_eval = 0;
    bool fitted = false;
    // Caller-supplied bin edges replace Binner::fit (checker mode for fits that share a binned
    // matrix, e.g. walk-forward folds binned once); empty = fit edges as the reference does.
    std::vector<std::vector<double>> fixed_edges;
    // Categorical features (Booster::fit's is_cat, booster.cpp:246-247) and their vocabularies.
    std::vector<int> is_cat;
    std::vector<std::vector<double>> cat_values;
    std::vector<int> group_of;   // Booster::fit's group_of (feature groups), empty = none
    std::vector<int> monotone;   // Booster::fit's monotone (booster.cpp:250-251), empty = none

    bool cat(int f) const { return !is_cat.empty() && is_cat[f] != 0; }

    void bins_of(const double* X, int n, std::vector<uint16_t>& xb) const {
        xb.resize((size_t)n * F);
        for (int i = 0; i < n; ++i)
            for (int f = 0; f < F; ++f) {
                const double x = X[(size_t)i * F + f];
                xb[(size_t)i * F + f] = (uint16_t)(cat(f) ? cat_bin_of(cat_values[f], x) : bin_of(edges[f], x));
            }
    }

    double init_of(const double* y, const double* w, int n) const {  // objective.cpp:80-100
        if (n <= 0) return 0;
        if (cfg.objective == 5) return 0;
        return wmean(y, w, n);
    }

    void grad(const double* y, const double* p, const double* w, const i64* grp, int n, double* g,
              double* h) const {
        if (cfg.objective == 5) grad_pairwise(y, p, w, grp, n, g, h);
        else grad_mse(y, p, w, n, g, h);
    }

    void fit(const double* X0, const double* y0, int n, int F_, const i64* ts0, const double* w0,
             const i64* ent0, const i64* rg0) {
        Cfg& c = cfg;
        if (!c.iid && c.horizon < 0) throw std::runtime_error("horizon is required in time mode");
        if (n >= 1000000) c.pair_splits = false;  // booster.cpp:237-240
        F = F_;
        std::vector<double> w(n, 1.0);
        if (w0) std::copy(w0, w0 + n, w.begin());
        // stable sort by (ts, entity, index) (booster.cpp:264-309)
        std::vector<int> ord(n);
        std::iota(ord.begin(), ord.end(), 0);
        if (!c.iid && ts0 && n > 1)
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
                if (ts0[a] != ts0[b]) return ts0[a] < ts0[b];
                if (ent0) return ent0[a] < ent0[b];
                return a < b;
            });
        std::vector<double> X((size_t)n * F), y(n), ws(n);
        std::vector<i64> ts(ts0 ? n : 0), rg(rg0 ? n : 0);
        for (int i = 0; i < n; ++i) {
            const int s = ord[i];
            std::memcpy(&X[(size_t)i * F], X0 + (size_t)s * F, sizeof(double) * F);
            y[i] = y0[s];
            ws[i] = w[s];
            if (ts0) ts[i] = ts0[s];
            if (rg0) rg[i] = rg0[s];
        }
        // TORB buckets and causal cuts (booster.cpp:36-72, 311-364)
        std::vector<int> bucket(n, 0);
        n_buckets = 1;
        std::vector<int> cut;
        i64 embargo = c.embargo;
        if (!c.iid) {
            if (!ts0) throw std::runtime_error("timestamp is required unless iid_mode");
            std::vector<i64> u(ts);
            std::sort(u.begin(), u.end());
            u.erase(std::unique(u.begin(), u.end()), u.end());
            const int un = std::max(2, (int)u.size());
            int want = c.n_time_buckets;
            if (want <= 0) {
                i64 step = 1;
                if (u.size() >= 2) {
                    std::vector<i64> d;
                    for (size_t k = 1; k < u.size(); ++k)
                        if (u[k] - u[k - 1] > 0) d.push_back(u[k] - u[k - 1]);
                    if (!d.empty()) {
                        std::nth_element(d.begin(), d.begin() + d.size() / 2, d.end());
                        step = std::max<i64>(d[d.size() / 2], 1);
                    }
                }
                const int hs = (int)(std::max<i64>(c.horizon, 0) / step);
                want = std::max(2, std::min(hs >= 8 ? 64 : 32, std::max(un, 2)));
            }
            n_buckets = c.finer_torb ? std::max(2, std::min(want, un)) : std::max(2, want);
            // rank by stable timestamp order; rows are already sorted so rank == index
            std::vector<int> o(n);
            std::iota(o.begin(), o.end(), 0);
            std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return ts[a] < ts[b]; });
            for (int r = 0; r < n; ++r) {
                int b = (int)((double)r / std::max(n, 1) * n_buckets);
                if (b >= n_buckets) b = n_buckets - 1;
                bucket[o[r]] = b;
            }
            embargo += std::max<i64>(c.horizon, 0);
            cut.assign(n_buckets, 0);
            std::vector<int> first(n_buckets, -1);
            for (int i = 0; i < n; ++i)
                if (first[bucket[i]] < 0) first[bucket[i]] = i;
            int run = 0;
            for (int b = 0; b < n_buckets; ++b) {
                int k = 0;
                if (first[b] >= 0)
                    k = (int)(std::lower_bound(ts.begin(), ts.end(), ts[first[b]] - embargo) - ts.begin());
                run = std::max(run, k);
                cut[b] = run;
            }
        }
        // bins (booster.cpp:378-390)
        if (!fixed_edges.empty()) {
            if ((int)fixed_edges.size() != F) throw std::runtime_error("fixed edges: feature count mismatch");
            edges = fixed_edges;
        } else {
            fit_edges(X.data(), n, F, c.max_bin, c.bin_sample, (uint32_t)c.random_state, edges);
        }
        if (!is_cat.empty() && (int)is_cat.size() != F) throw std::runtime_error("is_cat: feature count mismatch");
        if (!group_of.empty() && (int)group_of.size() != F) throw std::runtime_error("group_of: feature count mismatch");
        if (!monotone.empty() && (int)monotone.size() != F) throw std::runtime_error("monotone: feature count mismatch");
        cat_values.assign(F, {});
        for (int f = 0; f < F; ++f)
            if (cat(f)) {  // categorical features have no edges, only a vocabulary (binning.cpp:48-82)
                edges[f].clear();
                cat_values[f] = categories_of(X.data(), n, F, f, c.max_bin, c.cat_min_count);
            }
        nb = total_bins(edges, is_cat, cat_values, c.max_bin);
        std::vector<uint16_t> xb;
        bins_of(X.data(), n, xb);
        // initial scores (booster.cpp:399-457)
        std::vector<double> pred(n);
        init_prefix.assign(n_buckets, 0.0);
        if (c.iid) {
            init_score = init_of(y.data(), ws.data(), n);
            std::fill(pred.begin(), pred.end(), init_score);
        } else {
            for (int b = 0; b < n_buckets; ++b)
                init_prefix[b] = cut[b] > 0 ? init_of(y.data(), ws.data(), cut[b]) : NAN;
            init_score = init_of(y.data(), ws.data(), n);
            for (int b = 0; b < n_buckets; ++b)
                if (!std::isfinite(init_prefix[b])) init_prefix[b] = init_score;
            for (int i = 0; i < n; ++i) pred[i] = init_prefix[bucket[i]];
        }
        // eval set (booster.cpp:459-478): later than the training rows (minus the embargo), binned
        // with the training edges
        std::vector<uint16_t> xbe;
        if (n_eval > 0) {
            if (Xe.size() != (size_t)n_eval * F) throw std::runtime_error("eval set: feature count mismatch");
            if (!c.iid) {
                if (te.empty()) throw std::runtime_error("eval_timestamp is required when iid_mode=false");
                i64 tmax = ts[0];
                for (int i = 1; i < n; ++i) tmax = std::max(tmax, ts[i]);
                i64 tmin_e = te[0];
                for (int i = 1; i < n_eval; ++i) tmin_e = std::min(tmin_e, te[i]);
                if (tmin_e < tmax - embargo)
                    throw std::runtime_error(
                        "eval timestamps overlap the training future. "
                        "Validation must be strictly later than train (minus embargo/horizon).");
            }
            bins_of(Xe.data(), n_eval, xbe);
        }
        const bool eval_ic = n_eval > 0 && (c.eval_metric == 2 || (c.eval_metric == 0 && (c.objective == 5 || !c.iid)));
        double best_valid = eval_ic ? -1e300 : 1e300;
        int stall = 0;
        best_iteration = 0;
        evals_valid.clear();
        trees.clear();
        evals_train.clear();
        std::mt19937 rng((uint32_t)c.random_state);
        std::vector<double> g(n), h(n);
        const i64* grp = rg0 ? rg.data() : nullptr;
        const bool use_goss = (c.goss > 0) || (c.goss < 0 && n >= 250000);
        const double min_cover =
            c.torb_min_cover > 0.0 ? c.torb_min_cover : (double)(2 * std::max(c.min_samples_leaf, 1));
        for (int m = 0; m < c.n_estimators; ++m) {
            grad(y.data(), pred.data(), ws.data(), grp, n, g.data(), h.data());
            // grow_head (booster.cpp:519-665)
            std::vector<int> sample(n);
            std::iota(sample.begin(), sample.end(), 0);
            std::vector<double> gg, hh;
            const double* gu = g.data();
            const double* hu = h.data();
            if (use_goss && c.goss_top < 1.0 && sample.size() > 64) {  // booster.cpp:554-585
                gg = g;
                hh = h;
                std::vector<int> o = sample;
                std::sort(o.begin(), o.end(),
                          [&](int a, int b) { return std::abs(g[a]) > std::abs(g[b]); });
                const int S = (int)sample.size();
                const int n_top = std::min(S, std::max(1, (int)std::ceil(c.goss_top * S)));
                std::vector<char> keep(n, 0);
                for (int i = 0; i < n_top; ++i) keep[o[i]] = 1;
                std::vector<int> rest(o.begin() + n_top, o.end());
                int n_other = std::min((int)std::ceil(c.goss_other * S), (int)rest.size());
                std::shuffle(rest.begin(), rest.end(), rng);
                const double sc = (1.0 - c.goss_top) / std::max(c.goss_other, 1e-12);
                for (int i = 0; i < n_other; ++i) {
                    keep[rest[i]] = 1;
                    gg[rest[i]] *= sc;
                    hh[rest[i]] *= sc;
                }
                sample.clear();
                for (int i : o)
                    if (keep[i]) sample.push_back(i);
                std::sort(sample.begin(), sample.end());
                gu = gg.data();
                hu = hh.data();
            }
            GrowCtx cx{xb.data(), F, nb, gu, hu, &c, is_cat.empty() ? nullptr : is_cat.data(),
                       group_of.empty() ? nullptr : group_of.data(),
                       monotone.empty() ? nullptr : monotone.data()};
            (void)rng();  // the tree's own seed (booster.cpp:586-588); only colsample uses it
            Tree t = grow(cx, sample);
            const int nn = (int)t.nodes.size();
            std::vector<int> nid(n);
            for (int i = 0; i < n; ++i) nid[i] = leaf_of(t, &xb[(size_t)i * F], nb);
            std::vector<double> Gs(nn, 0), Hs(nn, 0);  // refit over all rows (booster.cpp:589-600)
            t.cover.assign(nn, 0);
            for (int i = 0; i < n; ++i) {
                Gs[nid[i]] += g[i];
                Hs[nid[i]] += h[i];
                t.cover[nid[i]] += 1;
            }
            t.values.resize(nn);
            for (int k = 0; k < nn; ++k) t.values[k] = leaf_val(Gs[k], Hs[k], c.l2, c.l1, c.max_delta_step);
            t.n_buckets = n_buckets;
            t.prefix.assign((size_t)n_buckets * nn, 0.0);
            if (c.iid) {
                for (int b = 0; b < n_buckets; ++b)
                    for (int k = 0; k < nn; ++k) t.prefix[(size_t)b * nn + k] = t.values[k];
            } else {  // TORB prefix sweep (booster.cpp:609-648)
                std::vector<double> pG(nn, 0), pH(nn, 0);
                int swept = 0;
                for (int b = 0; b < n_buckets; ++b) {
                    for (; swept < cut[b]; ++swept) {
                        pG[nid[swept]] += g[swept];
                        pH[nid[swept]] += h[swept];
                    }
                    const double cover = (double)cut[b];
                    const double alpha = cover <= 0.0 ? 0.0 : std::min(1.0, cover / std::max(min_cover, 1.0));
                    for (int k = 0; k < nn; ++k) {
                        const double full = t.values[k];
                        const double pref =
                            cover > 0.0 ? leaf_val(pG[k], pH[k], c.l2, c.l1, c.max_delta_step) : full;
                        t.prefix[(size_t)b * nn + k] = alpha * pref + (1.0 - alpha) * full;
                    }
                }
            }
            trees.push_back(std::move(t));
            const Tree& T = trees.back();
            const double lr = c.learning_rate;
            for (int i = 0; i < n; ++i) pred[i] += lr * T.prefix[(size_t)bucket[i] * nn + nid[i]];
            evals_train.push_back(loss_of(c.objective, y.data(), pred.data(), ws.data(), grp, n));
            if (n_eval > 0) {  // booster.cpp:707-732
                // predict_multi_binned (booster.cpp:775-799) scores the first best_iteration trees
                // once an iteration has improved (all grown trees before that)
                const int use = std::min((int)trees.size(), best_iteration > 0 ? best_iteration : c.n_estimators);
                std::vector<double> ev(n_eval);
                for (int i = 0; i < n_eval; ++i) {
                    double s = init_score;
                    for (int k = 0; k < use; ++k)
                        s += c.learning_rate * trees[k].values[leaf_of(trees[k], &xbe[(size_t)i * F], nb)];
                    ev[i] = s;
                }
                double vl;
                if (eval_ic) {
                    vl = grouped_ic(ye.data(), ev.data(), te.empty() ? nullptr : te.data(), n_eval);
                } else {  // Objective::loss with unit weights and no groups: mean squared error
                    double s = 0, sw = 0;
                    for (int i = 0; i < n_eval; ++i) {
                        const double d = ev[i] - ye[i];
                        s += 1.0 * d * d;
                        sw += 1.0;
                    }
                    vl = sw > 0 ? s / sw : 0;
                }
                evals_valid.push_back(vl);
                const bool improved = eval_ic ? (vl > best_valid + 1e-12) : (vl + 1e-12 < best_valid);
                if (improved) {
                    best_valid = vl;
                    stall = 0;
                    best_iteration = m + 1;
                } else {
                    stall += 1;
                    if (c.early_stopping_rounds > 0 && stall >= c.early_stopping_rounds) {
                        const int keep = best_iteration;
                        if (keep > 0 && keep < (int)trees.size()) trees.resize(keep);
                        break;
                    }
                }
            }
        }
        if (best_iteration <= 0) best_iteration = c.n_estimators;  // booster.cpp:739
        fitted = true;
    }

    // predict_raw_binned (booster.cpp:744-773): init_score + sum_t lr * values_t[leaf]
    void predict(const double* X, int n, double* out) const {
        if (!fitted) throw std::runtime_error("not fitted");
        std::vector<uint16_t> xb;
        bins_of(X, n, xb);
        const int use = std::min((int)trees.size(), best_iteration > 0 ? best_iteration : cfg.n_estimators);
        for (int i = 0; i < n; ++i) {
            double s = init_score;
            for (int t = 0; t < use; ++t) s += cfg.learning_rate * trees[t].values[leaf_of(trees[t], &xb[(size_t)i * F], nb)];
            out[i] = s;
        }
    }

    // update (booster.cpp:844-884): frozen bins, grow `extra` trees on all rows, no TORB.
    void update(const double* X, const double* y, int n, const double* w0, int extra) {
        if (!fitted) throw std::runtime_error("not fitted");
        std::vector<uint16_t> xb;
        bins_of(X, n, xb);
        std::vector<double> pred(n), g(n), h(n), w(n, 1.0);
        if (w0) std::copy(w0, w0 + n, w.begin());
        predict(X, n, pred.data());
        std::mt19937 rng((uint32_t)cfg.random_state + (uint32_t)trees.size());
        std::vector<int> all(n);
        std::iota(all.begin(), all.end(), 0);
        for (int m = 0; m < extra; ++m) {
            grad(y, pred.data(), w.data(), nullptr, n, g.data(), h.data());
            GrowCtx cx{xb.data(), F, nb, g.data(), h.data(), &cfg, is_cat.empty() ? nullptr : is_cat.data(),
                       group_of.empty() ? nullptr : group_of.data(),
                       monotone.empty() ? nullptr : monotone.data()};
            (void)rng();
            Tree t = grow(cx, all);
            const int nn = (int)t.nodes.size();
            std::vector<int> nid(n);
            for (int i = 0; i < n; ++i) nid[i] = leaf_of(t, &xb[(size_t)i * F], nb);
            std::vector<double> Gs(nn, 0), Hs(nn, 0);
            t.cover.assign(nn, 0);
            for (int i = 0; i < n; ++i) {
                Gs[nid[i]] += g[i];
                Hs[nid[i]] += h[i];
            }
            t.values.resize(nn);
            for (int k = 0; k < nn; ++k) t.values[k] = leaf_val(Gs[k], Hs[k], cfg.l2, cfg.l1, cfg.max_delta_step);
            for (int k = 0; k < nn; ++k) t.cover[k] = t.nodes[k].n_samples;  // grow_tree's cover
            t.n_buckets = 1;
            t.prefix = t.values;
            trees.push_back(std::move(t));
            const Tree& T = trees.back();
            for (int i = 0; i < n; ++i) pred[i] += cfg.learning_rate * T.values[nid[i]];
        }
        best_iteration = (int)trees.size();
    }
};

}  // namespace orc

// ================================================================= extern "C" checker API
using orc::Model;
static thread_local std::string g_orc_err;

template <class Fn>
static int orc_guard(Fn&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception& e) {
        g_orc_err = e.what();
        return -1;
    }
}

extern "C" {

const char* orc_last_error(void) { return g_orc_err.c_str(); }
void* orc_create(void) { return new Model(); }
void orc_free(void* h) { delete static_cast<Model*>(h); }

int orc_set(void* hd, const char* key, double v) {
    orc::Cfg& c = static_cast<Model*>(hd)->cfg;
    const std::string k(key);
    const int iv = (int)v;
    if (k == "n_estimators") c.n_estimators = iv;
    else if (k == "learning_rate") c.learning_rate = v;
    else if (k == "max_depth") c.max_depth = iv;
    else if (k == "max_leaves") c.max_leaves = iv;
    else if (k == "min_samples_leaf") c.min_samples_leaf = iv;
    else if (k == "min_child_weight") c.min_child_weight = v;
    else if (k == "l2") c.l2 = v;
    else if (k == "l1") c.l1 = v;
    else if (k == "max_delta_step") c.max_delta_step = v;
    else if (k == "min_split_gain") c.min_split_gain = v;
    else if (k == "max_bin") c.max_bin = iv;
    else if (k == "n_time_buckets") c.n_time_buckets = iv;
    else if (k == "embargo") c.embargo = (long long)v;
    else if (k == "pair_splits") c.pair_splits = iv != 0;
    else if (k == "pair_trigger") c.pair_trigger = v;
    else if (k == "pair_top_k") c.pair_top_k = iv;
    else if (k == "max_pair_splits_per_tree") c.max_pair_splits = iv;
    else if (k == "pair_l2") c.pair_l2 = v;
    else if (k == "objective") c.objective = iv;
    else if (k == "random_state") c.random_state = iv;
    else if (k == "goss") c.goss = iv;
    else if (k == "goss_top") c.goss_top = v;
    else if (k == "goss_other") c.goss_other = v;
    else if (k == "bin_sample") c.bin_sample = iv;
    else if (k == "finer_torb") c.finer_torb = iv != 0;
    else if (k == "torb_min_cover") c.torb_min_cover = v;
    else if (k == "early_stopping_rounds") c.early_stopping_rounds = iv;
    else if (k == "cat_min_count") c.cat_min_count = iv;
    else if (k == "regime_feature") c.regime_feature = iv;
    else if (k == "regime_priority") c.regime_priority = iv != 0;
    else if (k == "cat_smooth") c.cat_smooth = v;
    else if (k == "eval_metric") {
        if (iv < 0 || iv > 2) {
            g_orc_err = "eval_metric code must be 0 (auto), 1 (loss) or 2 (ic)";
            return -1;
        }
        c.eval_metric = iv;
    }
    else if (k == "n_jobs") { /* scalar oracle */ }
    else {
        g_orc_err = "unknown key " + k;
        return -1;
    }
    return 0;
}

int orc_fit(void* hd, const double* X, const double* y, int n, int F, const long long* ts,
            const double* w, const long long* entity, const long long* rank_groups, int iid,
            long long horizon) {
    Model* m = static_cast<Model*>(hd);
    return orc_guard([&] {
        m->cfg.iid = iid != 0;
        if (!m->cfg.iid) m->cfg.horizon = horizon;
        m->fit(X, y, n, F, ts, w, entity, rank_groups);
    });
}

// Fix the bin edges of later fits (edges[f*256 + k], k < n_edges[f]); n_features = 0 clears.
int orc_set_edges(void* hd, const double* edges, const int* n_edges, int n_features) {
    Model* m = static_cast<Model*>(hd);
    return orc_guard([&] {
        m->fixed_edges.assign(n_features, {});
        for (int f = 0; f < n_features; ++f) {
            if (n_edges[f] < 0 || n_edges[f] > 255) throw std::runtime_error("edge count out of range");
            m->fixed_edges[f].assign(edges + (size_t)f * 256, edges + (size_t)f * 256 + n_edges[f]);
        }
    });
}

int orc_update(void* hd, const double* X, const double* y, int n, int F, const double* w, int extra) {
    Model* m = static_cast<Model*>(hd);
    return orc_guard([&] {
        if (F != m->F) throw std::runtime_error("n_features mismatch");
        m->update(X, y, n, w, extra);
    });
}

int orc_predict(void* hd, const double* X, int n, int F, double* out) {
    Model* m = static_cast<Model*>(hd);
    return orc_guard([&] {
        if (F != m->F) throw std::runtime_error("n_features mismatch");
        m->predict(X, n, out);
    });
}

int orc_n_bins(void* hd) { return static_cast<Model*>(hd)->nb; }
int orc_n_time_buckets(void* hd) { return static_cast<Model*>(hd)->n_buckets; }
double orc_init_score(void* hd) { return static_cast<Model*>(hd)->init_score; }
int orc_init_prefix(void* hd, double* out, int cap) {
    const auto& v = static_cast<Model*>(hd)->init_prefix;
    for (int i = 0; i < (int)v.size() && i < cap; ++i) out[i] = v[i];
    return (int)v.size();
}
int orc_edges(void* hd, int f, double* out, int cap) {
    const auto& e = static_cast<Model*>(hd)->edges.at(f);
    for (int i = 0; i < (int)e.size() && i < cap; ++i) out[i] = e[i];
    return (int)e.size();
}
int orc_n_trees(void* hd) { return (int)static_cast<Model*>(hd)->trees.size(); }
// Eval set of later fits: X (n x n_features, row-major), y, ts (may be NULL in iid mode); n = 0 clears.
int orc_set_eval(void* hd, const double* X, const double* y, const long long* ts, int n, int n_features) {
    Model* m = static_cast<Model*>(hd);
    return orc_guard([&] {
        if (n < 0 || (n > 0 && (!X || !y || n_features <= 0))) throw std::runtime_error("bad eval set");
        m->n_eval = n;
        m->Xe.assign(X, X + (size_t)n * std::max(n_features, 0));
        m->ye.assign(y, y + n);
        if (ts) m->te.assign(ts, ts + n);
        else m->te.clear();
    });
}
int orc_best_iteration(void* hd) { return static_cast<Model*>(hd)->best_iteration; }
int orc_evals_valid(void* hd, double* out, int cap) {
    const auto& e = static_cast<Model*>(hd)->evals_valid;
    if (out) std::copy(e.begin(), e.begin() + std::min<size_t>(e.size(), (size_t)std::max(cap, 0)), out);
    return (int)e.size();
}
int orc_evals_train(void* hd, double* out, int cap) {
    const auto& e = static_cast<Model*>(hd)->evals_train;
    for (int i = 0; i < (int)e.size() && i < cap; ++i) out[i] = e[i];
    return (int)e.size();
}
int orc_tree_info(void* hd, int t, int* nn, int* nbk) {
    const orc::Tree& tr = static_cast<Model*>(hd)->trees.at(t);
    *nn = (int)tr.nodes.size();
    *nbk = tr.n_buckets;
    return 0;
}
int orc_tree_get(void* hd, int t, int* ints, double* dbl, double* values, int* cover, double* prefix) {
    const orc::Tree& tr = static_cast<Model*>(hd)->trees.at(t);
    for (size_t k = 0; k < tr.nodes.size(); ++k) {
        const orc::Node& d = tr.nodes[k];
        const int v[10] = {d.is_leaf, d.split_type, d.feature, d.feature2, d.thr, d.left, d.right,
                           d.mgl, d.n_samples, d.depth};
        std::memcpy(ints + 10 * k, v, sizeof v);
        dbl[2 * k] = d.value;
        dbl[2 * k + 1] = d.gain;
        values[k] = tr.values[k];
        cover[k] = tr.cover[k];
    }
    std::copy(tr.prefix.begin(), tr.prefix.end(), prefix);
    return 0;
}

// Categorical split masks of tree t: 4 words per node (bins that go left).
int orc_tree_cat(void* hd, int t, unsigned long long* out) {
    const orc::Tree& tr = static_cast<Model*>(hd)->trees.at(t);
    for (size_t k = 0; k < tr.nodes.size(); ++k)
        for (int w = 0; w < 4; ++w) out[4 * k + w] = tr.nodes[k].cat_left[w];
    return (int)tr.nodes.size();
}

// Feature groups of later fits (Booster::fit's group_of, -1 = none); n_features = 0 clears.
int orc_set_groups(void* hd, const int* group_of, int n_features) {
    static_cast<Model*>(hd)->group_of.assign(group_of, group_of + std::max(n_features, 0));
    return 0;
}

// Monotone constraints of later fits (Booster::fit's monotone: +1 / -1 / 0 per feature);
// n_features = 0 clears.
int orc_set_monotone(void* hd, const int* monotone, int n_features) {
    static_cast<Model*>(hd)->monotone.assign(monotone, monotone + std::max(n_features, 0));
    return 0;
}

// Interaction groups (Config::interaction_group, -1 = unconstrained); n = 0 clears.
int orc_set_interaction(void* hd, const int* groups, int n) {
    static_cast<Model*>(hd)->cfg.interaction_group.assign(groups, groups + std::max(n, 0));
    return 0;
}

// Categorical flags of later fits (is_cat[f] != 0); n_features = 0 clears.
int orc_set_cat(void* hd, const int* is_cat, int n_features) {
    Model* m = static_cast<Model*>(hd);
    m->is_cat.assign(is_cat, is_cat + std::max(n_features, 0));
    return 0;
}

// Category vocabulary of feature f of the last fit (ascending); returns its size.
int orc_cat_values(void* hd, int f, double* out, int cap) {
    const auto& v = static_cast<Model*>(hd)->cat_values.at(f);
    for (int i = 0; i < (int)v.size() && i < cap; ++i) out[i] = v[i];
    return (int)v.size();
}

// ---- per-function checkers over caller buffers ----

// Edges per feature into edges[f*256 + k], counts into n_edges[f]; returns n_bins.
int orc_fit_edges(const double* X, int n, int F, int max_bin, int bin_sample, unsigned seed,
                  double* edges, int* n_edges) {
    std::vector<std::vector<double>> e;
    orc::fit_edges(X, n, F, max_bin, bin_sample, seed, e);
    for (int f = 0; f < F; ++f) {
        n_edges[f] = (int)e[f].size();
        std::copy(e[f].begin(), e[f].end(), edges + (size_t)f * 256);
    }
    return orc::total_bins(e, {}, {}, max_bin);
}

// Row-major u8 bins from edges laid out as in orc_fit_edges.
int orc_bin_transform(const double* X, int n, int F, const double* edges, const int* n_edges,
                      unsigned char* out) {
    for (int f = 0; f < F; ++f) {
        std::vector<double> e(edges + (size_t)f * 256, edges + (size_t)f * 256 + n_edges[f]);
        for (int i = 0; i < n; ++i) out[(size_t)i * F + f] = (unsigned char)orc::bin_of(e, X[(size_t)i * F + f]);
    }
    return 0;
}

// FP64 histogram over rows idx (ascending order of the caller's list) of row-major u8 bins.
int orc_hist(const unsigned char* xb, int F, const int* idx, int m, const double* g,
             const double* h, int nb, double* G, double* H, long long* C) {
    std::vector<uint16_t> x16;  // widen to the reference's u16 storage
    int mx = 0;
    for (int k = 0; k < m; ++k) mx = std::max(mx, idx[k]);
    x16.resize((size_t)(mx + 1) * F);
    for (int k = 0; k < m; ++k)
        for (int f = 0; f < F; ++f) x16[(size_t)idx[k] * F + f] = xb[(size_t)idx[k] * F + f];
    orc::Hist out;
    orc::build(x16.data(), F, idx, m, g, h, out, nb);
    std::copy(out.g.begin(), out.g.end(), G);
    std::copy(out.h.begin(), out.h.end(), H);
    std::copy(out.c.begin(), out.c.end(), C);
    return 0;
}

// Per-feature best split (gain/bin/missing-left per feature) and the leaf's argmax.
int orc_scan(const double* G, const double* H, const long long* C, int F, int nb, double lam,
             double gamma, int min_leaf, double min_hess, double l1, double* gain_f, int* bin_f,
             int* ml_f, int* best_feature) {
    orc::Hist hist;
    hist.init(F, nb);
    std::copy(G, G + (size_t)F * nb, hist.g.begin());
    std::copy(H, H + (size_t)F * nb, hist.h.begin());
    std::copy(C, C + (size_t)F * nb, hist.c.begin());
    orc::Cand best;
    for (int f = 0; f < F; ++f) {
        orc::Cand c = orc::scan_feature(hist, f, lam, gamma, min_leaf, min_hess, l1);
        gain_f[f] = c.gain;
        bin_f[f] = c.thr;
        ml_f[f] = c.mgl;
        if (c.feature >= 0 && c.gain >= gamma && c.gain > best.gain) best = c;
    }
    *best_feature = best.feature;
    return 0;
}

// Best split of one feature with its monotone constraint (histogram.cpp:144-193), the same
// entry point as the reference driver's tref_best_numeric.
int orc_scan_feature(const double* G, const double* H, const long long* C, int F, int nb, int f, double lam,
                     double gamma, int min_leaf, double min_hess, int monotone, double l1, double* gain, int* bin,
                     int* miss_left) {
    orc::Hist hist;
    hist.init(F, nb);
    std::copy(G, G + (size_t)F * nb, hist.g.begin());
    std::copy(H, H + (size_t)F * nb, hist.h.begin());
    std::copy(C, C + (size_t)F * nb, hist.c.begin());
    const orc::Cand c = orc::scan_feature(hist, f, lam, gamma, min_leaf, min_hess, l1, monotone);
    *gain = c.gain;
    *bin = c.thr;
    *miss_left = c.mgl;
    return 0;
}

int orc_grad(int kind, const double* y, const double* p, const double* w, const long long* grp,
             int n, double* g, double* h) {
    std::vector<double> ones;
    if (!w) {
        ones.assign(n, 1.0);
        w = ones.data();
    }
    if (kind == 5) orc::grad_pairwise(y, p, w, grp, n, g, h);
    else orc::grad_mse(y, p, w, n, g, h);
    return 0;
}

}  // extern "C"
