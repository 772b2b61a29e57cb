// TGBM v3 model files and the JSON dump.
//
// The file layout is the reference's (writer and reader: cpp/src/serialize.cpp:65-255; magic and
// version: cpp/include/tidegbm/types.hpp:22-23; JSON dump: serialize.cpp:266-324), so that either
// library loads the other's files.  Here the layout is stated once, as a schema: `walk` visits every
// field of a `ModelFile` record in file order, and the two archives run it in each direction
// (`Emit` appends the fields, `Parse` reads them back with bounds checks).  A Booster converts to
// and from the record; the checks a file must pass before a Booster is built from it (feature
// counts, child links that point forward, supported options) sit in that conversion.
#include <cstring>
#include <sstream>
#include <type_traits>

#include "booster.h"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "TGBM files are little-endian, as is this host");

namespace tide {

namespace {

constexpr char kMagic[4] = {'T', 'G', 'B', 'M'};
constexpr uint32_t kVersion = 3;

// ------------------------------------------------------------------ the record
struct FileFeature {        // one feature of the binner
    int32_t is_cat = 0;
    std::vector<double> edges, cat_values;
    int32_t n_bins = 2;     // bins this feature uses: edges + 2, or categories + 1
};

struct FileNode {           // a node, fields in file order
    int32_t is_leaf = 1, split_type = 0, feature = -1, feature2 = -1, threshold = -1, left = -1, right = -1,
            missing_go_left = 1;
    double value = 0.0, gain = 0.0;
    int32_t n_samples = 0, depth = 0;
    uint64_t cat_left[4] = {0, 0, 0, 0};
};

struct FileTree {
    int32_t n_buckets = 1;
    std::vector<FileNode> nodes;
    std::vector<double> values;
    std::vector<int32_t> cover;
    std::vector<double> prefix_values;
    int32_t target = 0, n_targets = 1;
};

struct ModelFile {
    // configuration block
    int32_t n_estimators = 0, max_depth = 0, max_leaves = 0, min_samples_leaf = 0, max_bin = 0, n_time_buckets_cfg = 0;
    double learning_rate = 0, min_child_weight = 0, l2 = 0, min_split_gain = 0, subsample = 1, colsample = 1;
    int64_t horizon = -1, embargo = 0;
    int32_t iid_mode = 0, halo = 0, objective = 0, random_state = 0, n_jobs = 0;
    double quantile = 0.5, huber_delta = 1.0, rank_temperature = 1.0, tree_decay = 1.0;
    // model block
    int32_t n_features = 0, n_bins = 0, n_time_buckets = 1, best_iteration = 0;
    double init_score = 0.0;
    std::vector<double> init_prefix;
    std::vector<std::string> feature_names;
    std::vector<int32_t> is_cat, group_of, monotone;
    int32_t binner_max_bin = 0;
    std::vector<FileFeature> features;
    std::vector<FileTree> trees;
    int32_t objective_kind = 0;   // repeated after the trees
};

// The schema: every field of the file, in order.  A vector is an int32 count and its elements.
template <class Ar>
void walk(Ar& a, ModelFile& m) {
    a(m.n_estimators), a(m.learning_rate), a(m.max_depth), a(m.max_leaves), a(m.min_samples_leaf);
    a(m.min_child_weight), a(m.l2), a(m.min_split_gain), a(m.max_bin), a(m.subsample), a(m.colsample);
    a(m.n_time_buckets_cfg), a(m.horizon), a(m.embargo), a(m.iid_mode), a(m.halo), a(m.objective);
    a(m.quantile), a(m.huber_delta), a(m.rank_temperature), a(m.random_state), a(m.n_jobs), a(m.tree_decay);
    a(m.n_features), a(m.n_bins), a(m.n_time_buckets), a(m.best_iteration), a(m.init_score);
    a.seq(m.init_prefix);
    a.seq(m.feature_names);
    a.seq(m.is_cat);
    a.seq(m.group_of);
    a.seq(m.monotone);
    a(m.binner_max_bin);
    a.seq(m.features, [&](FileFeature& f) {
        a(f.is_cat);
        a.seq(f.edges);
        a.seq(f.cat_values);
        a(f.n_bins);
    });
    a.seq(m.trees, [&](FileTree& t) {
        a(t.n_buckets);
        a.seq(t.nodes, [&](FileNode& n) {
            a(n.is_leaf), a(n.split_type), a(n.feature), a(n.feature2), a(n.threshold), a(n.left), a(n.right);
            a(n.missing_go_left), a(n.value), a(n.gain), a(n.n_samples), a(n.depth);
            for (uint64_t& w : n.cat_left) a(w);
        });
        a.seq(t.values);
        a.seq(t.cover);
        a.seq(t.prefix_values);
        a(t.target), a(t.n_targets);
    });
    a(m.objective_kind);
}

// ------------------------------------------------------------------ archives
// Scalars are stored in their little-endian byte image; strings as an int32 length and bytes.
class Emit {
public:
    std::string out;
    template <class T>
    void operator()(T& v) {
        static_assert(std::is_arithmetic<T>::value, "scalar field");
        char raw[sizeof(T)];
        std::memcpy(raw, &v, sizeof(T));
        out.append(raw, sizeof(T));
    }
    void operator()(std::string& s) {
        int32_t len = (int32_t)s.size();
        (*this)(len);
        out.append(s);
    }
    template <class T>
    void seq(std::vector<T>& v) {
        seq(v, [&](T& x) { (*this)(x); });
    }
    template <class T, class F>
    void seq(std::vector<T>& v, F&& each) {
        int32_t len = (int32_t)v.size();
        (*this)(len);
        for (T& x : v) each(x);
    }
};

class Parse {
public:
    Parse(const std::string& blob, size_t start) : base_(blob.data()), at_(start), size_(blob.size()) {}
    size_t consumed() const { return at_; }
    template <class T>
    void operator()(T& v) {
        static_assert(std::is_arithmetic<T>::value, "scalar field");
        take(&v, sizeof(T));
    }
    void operator()(std::string& s) {
        const size_t len = count(1 << 20);
        need(len);
        s.assign(base_ + at_, len);
        at_ += len;
    }
    template <class T>
    void seq(std::vector<T>& v) {
        seq(v, [&](T& x) { (*this)(x); });
    }
    template <class T, class F>
    void seq(std::vector<T>& v, F&& each) {
        // every element takes at least 4 bytes, so a count larger than the rest of the file is corrupt
        v.assign(count((size_t)(size_ - at_) / 4), T());
        for (T& x : v) each(x);
    }

private:
    const char* base_;
    size_t at_, size_;
    void need(size_t n) const {
        if (size_ - at_ < n)
            throw Error("model file truncated: " + std::to_string(n) + " more bytes needed at offset " +
                        std::to_string(at_) + " of " + std::to_string(size_));
    }
    void take(void* dst, size_t n) {
        need(n);
        std::memcpy(dst, base_ + at_, n);
        at_ += n;
    }
    size_t count(size_t cap) {
        int32_t n = 0;
        (*this)(n);
        if (n < 0 || (size_t)n > cap)
            throw Error("corrupt model file: count " + std::to_string(n) + " at offset " + std::to_string(at_ - 4));
        return (size_t)n;
    }
};

int32_t flag_of(const std::vector<int>& v, int f) { return f < (int)v.size() && v[f] != 0 ? 1 : 0; }

// ------------------------------------------------------------------ Booster <-> record
ModelFile to_file(const Booster& b) {
    ModelFile m;
    const Config& c = b.cfg;
    m.n_estimators = c.n_estimators, m.learning_rate = c.learning_rate, m.max_depth = c.max_depth;
    m.max_leaves = c.max_leaves, m.min_samples_leaf = c.min_samples_leaf, m.min_child_weight = c.min_child_weight;
    m.l2 = c.l2, m.min_split_gain = c.min_split_gain, m.max_bin = c.max_bin, m.subsample = c.subsample;
    m.colsample = c.colsample, m.n_time_buckets_cfg = c.n_time_buckets, m.horizon = c.horizon;
    m.embargo = c.embargo, m.iid_mode = c.iid_mode, m.halo = c.halo, m.objective = c.objective;
    m.quantile = c.quantile, m.huber_delta = c.huber_delta, m.rank_temperature = c.rank_temperature;
    m.random_state = c.random_state, m.n_jobs = c.n_jobs, m.tree_decay = c.tree_decay;
    const int F = b.n_features;
    m.n_features = F, m.n_bins = b.n_bins, m.n_time_buckets = b.n_time_buckets;
    m.best_iteration = b.best_iteration, m.init_score = b.init_score, m.init_prefix = b.init_prefix;
    // the reference names features f0 .. f{F-1} at fit (booster.cpp:244-245); loaded names are kept
    m.feature_names = b.feature_names;
    if ((int)m.feature_names.size() != F) {
        m.feature_names.clear();
        for (int f = 0; f < F; ++f) m.feature_names.push_back("f" + std::to_string(f));
    }
    m.is_cat.resize(F), m.group_of.assign(F, -1), m.monotone.assign(F, 0);
    for (int f = 0; f < F; ++f) {
        m.is_cat[f] = flag_of(b.is_cat, f);
        if ((int)b.group_of.size() == F) m.group_of[f] = b.group_of[f];
        if ((int)c.monotone.size() == F) m.monotone[f] = c.monotone[f];
    }
    m.binner_max_bin = c.max_bin;
    m.features.resize(F);
    for (int f = 0; f < F; ++f) {
        FileFeature& x = m.features[f];
        x.is_cat = m.is_cat[f];
        x.edges = b.edges[f];
        if (x.is_cat) x.cat_values = b.cat_values[f];
        x.n_bins = x.is_cat ? (int32_t)x.cat_values.size() + 1 : (int32_t)x.edges.size() + 2;   // binning.cpp:83,99
    }
    m.trees.resize(b.trees.size());
    for (size_t t = 0; t < b.trees.size(); ++t) {
        const Tree& src = b.trees[t];
        FileTree& dst = m.trees[t];
        dst.n_buckets = src.n_buckets;
        dst.values = src.values, dst.prefix_values = src.prefix_values;
        dst.cover.assign(src.cover.begin(), src.cover.end());
        dst.nodes.resize(src.nodes.size());
        for (size_t k = 0; k < src.nodes.size(); ++k) {
            const Node& n = src.nodes[k];
            FileNode& o = dst.nodes[k];
            o.is_leaf = n.is_leaf ? 1 : 0, o.split_type = n.split_type, o.feature = n.feature;
            o.feature2 = n.feature2, o.threshold = n.threshold, o.left = n.left, o.right = n.right;
            o.missing_go_left = n.missing_go_left, o.value = n.value, o.gain = n.gain;
            o.n_samples = n.n_samples, o.depth = n.depth;
            for (int w = 0; w < 4; ++w) o.cat_left[w] = n.cat_left[w];
        }
    }
    m.objective_kind = c.objective;
    return m;
}

void check(bool ok, const std::string& what) {
    if (!ok) throw Error("invalid model file: " + what);
}

void from_file(const ModelFile& m, Booster& b) {
    Config& c = b.cfg;
    c.n_estimators = m.n_estimators, c.learning_rate = m.learning_rate, c.max_depth = m.max_depth;
    c.max_leaves = m.max_leaves, c.min_samples_leaf = m.min_samples_leaf, c.min_child_weight = m.min_child_weight;
    c.l2 = m.l2, c.min_split_gain = m.min_split_gain, c.max_bin = m.max_bin, c.subsample = m.subsample;
    c.colsample = m.colsample, c.n_time_buckets = m.n_time_buckets_cfg, c.horizon = m.horizon;
    c.embargo = m.embargo, c.iid_mode = m.iid_mode != 0, c.halo = m.halo != 0, c.objective = m.objective_kind;
    c.quantile = m.quantile, c.huber_delta = m.huber_delta, c.rank_temperature = m.rank_temperature;
    c.random_state = m.random_state, c.n_jobs = m.n_jobs, c.tree_decay = m.tree_decay;
    const int F = m.n_features;
    check(F > 0 && F <= 1024, "feature count " + std::to_string(F));
    check((int)m.features.size() == F, "binner describes " + std::to_string(m.features.size()) + " features, model " +
                                           std::to_string(F));
    check(m.is_cat.empty() || (int)m.is_cat.size() == F, "categorical flags for " + std::to_string(m.is_cat.size()) +
                                                               " of " + std::to_string(F) + " features");
    check(m.group_of.empty() || (int)m.group_of.size() == F, "feature groups for " + std::to_string(m.group_of.size()) +
                                                                   " of " + std::to_string(F) + " features");
    check(m.monotone.empty() || (int)m.monotone.size() == F, "monotone constraints for " +
                                                                   std::to_string(m.monotone.size()) + " of " +
                                                                   std::to_string(F) + " features");
    check(m.n_bins >= 2 && m.n_bins <= 256, "bin count " + std::to_string(m.n_bins));
    b.n_features = F, b.n_bins = m.n_bins, b.n_time_buckets = m.n_time_buckets;
    b.best_iteration = m.best_iteration, b.init_score = m.init_score, b.init_prefix = m.init_prefix;
    b.feature_names = m.feature_names;
    b.is_cat.assign(F, 0);
    for (int f = 0; f < (int)m.is_cat.size(); ++f) b.is_cat[f] = m.is_cat[f] != 0 ? 1 : 0;
    b.group_of.clear();
    for (int32_t g : m.group_of)
        if (g >= 0) {
            b.group_of.assign(m.group_of.begin(), m.group_of.end());
            break;
        }
    c.monotone.clear();
    for (int32_t v : m.monotone)
        if (v != 0) {
            c.monotone.assign(m.monotone.begin(), m.monotone.end());
            break;
        }
    b.edges.assign(F, {});
    b.cat_values.assign(F, {});
    for (int f = 0; f < F; ++f) {
        const FileFeature& x = m.features[f];
        check((x.is_cat != 0) == (b.is_cat[f] != 0), "binner and model disagree on categorical feature " +
                                                          std::to_string(f));
        check(x.edges.size() <= 254 && x.cat_values.size() <= 254, "feature " + std::to_string(f) +
                                                                         " has more than 254 edges or categories");
        b.edges[f] = x.edges;
        b.cat_values[f] = x.cat_values;
    }
    if (!b.has_cat()) b.is_cat.clear();
    b.trees.resize(m.trees.size());
    for (size_t t = 0; t < m.trees.size(); ++t) {
        const FileTree& src = m.trees[t];
        Tree& dst = b.trees[t];
        const std::string tn = "tree " + std::to_string(t);
        check(src.target == 0 && src.n_targets == 1, tn + ": multi-head models are not supported on the GPU path");
        const int nn = (int)src.nodes.size();
        check(nn >= 1 && nn <= 256, tn + ": node count " + std::to_string(nn));
        check(src.values.empty() || (int)src.values.size() == nn, tn + ": leaf values for " +
                                                                        std::to_string(src.values.size()) + " nodes");
        dst.n_buckets = src.n_buckets;
        dst.values = src.values, dst.prefix_values = src.prefix_values;
        dst.cover.assign(src.cover.begin(), src.cover.end());
        dst.nodes.resize(nn);
        for (int k = 0; k < nn; ++k) {
            const FileNode& n = src.nodes[k];
            Node& o = dst.nodes[k];
            o.is_leaf = n.is_leaf != 0, o.split_type = n.split_type, o.feature = n.feature, o.feature2 = n.feature2;
            o.threshold = n.threshold, o.left = n.left, o.right = n.right, o.missing_go_left = n.missing_go_left;
            o.value = n.value, o.gain = n.gain, o.n_samples = n.n_samples, o.depth = n.depth;
            for (int w = 0; w < 4; ++w) o.cat_left[w] = n.cat_left[w];
            if (o.is_leaf) continue;
            const std::string at = tn + " node " + std::to_string(k);
            check(o.split_type >= 0 && o.split_type <= 3, at + ": unknown split type " + std::to_string(o.split_type));
            check(o.split_type < 2 || b.has_cat(), at + ": categorical split in a model without categorical features");
            // children come after their parent (the grower appends them, tree.cpp:445-456): a tree
            // walk then always ends
            check(o.left > k && o.left < nn && o.right > k && o.right < nn, at + ": child links " +
                                                                               std::to_string(o.left) + ", " +
                                                                               std::to_string(o.right));
            check(o.feature >= 0 && o.feature < F, at + ": feature " + std::to_string(o.feature));
            check(o.split_type != 1 || (o.feature2 >= 0 && o.feature2 < F), at + ": second feature " +
                                                                                 std::to_string(o.feature2));
        }
    }
    b.fitted = true;
}

}  // namespace

std::string Booster::save_binary() const {
    ModelFile m = to_file(*this);
    Emit e;
    e.out.append(kMagic, 4);
    uint32_t v = kVersion;
    e(v);
    walk(e, m);
    return std::move(e.out);
}

Booster* Booster::load_binary(const std::string& blob) {
    if (blob.size() < 8 || std::memcmp(blob.data(), kMagic, 4) != 0) throw Error("not a TGBM model file (bad magic)");
    uint32_t version = 0;
    std::memcpy(&version, blob.data() + 4, 4);
    if (version != kVersion) throw Error("unsupported TGBM model version " + std::to_string(version));
    ModelFile m;
    Parse p(blob, 8);
    walk(p, m);
    auto b = std::make_unique<Booster>();
    from_file(m, *b);
    return b.release();
}

// ------------------------------------------------------------------ JSON dump
// Same document as the reference's dump (serialize.cpp:266-324): numbers in the stream's default
// format (6 significant digits), one tree per line.
namespace {

class JsonOut {
public:
    std::ostringstream o;
    void key(const char* k, bool first = false) { o << (first ? "" : ",\n") << pad_ << '"' << k << "\": "; }
    template <class R, class F>
    void list(const R& range, F&& each) {
        o << '[';
        bool first = true;
        for (const auto& x : range) {
            if (!first) o << ", ";
            first = false;
            each(x);
        }
        o << ']';
    }
    void str(const std::string& s) {
        o << '"';
        for (char ch : s) (ch == '"' || ch == '\\' ? o << '\\' << ch : o << ch);
        o << '"';
    }
    void indent(const char* p) { pad_ = p; }

private:
    const char* pad_ = "  ";
};

}  // namespace

std::string Booster::save_json() const {
    const ModelFile m = to_file(*this);
    JsonOut j;
    auto num = [&](double x) { j.o << x; };
    j.o << "{\n";
    j.key("format", true), j.o << "\"tidegbm\"";
    j.key("version"), j.o << kVersion;
    j.key("n_features"), j.o << m.n_features;
    j.key("n_bins"), j.o << m.n_bins;
    j.key("best_iteration"), j.o << m.best_iteration;
    j.key("init_score"), num(m.init_score);
    j.key("learning_rate"), num(m.learning_rate);
    j.key("horizon"), j.o << m.horizon;
    j.key("iid_mode"), j.o << (m.iid_mode ? "true" : "false");
    j.key("feature_names"), j.list(m.feature_names, [&](const std::string& s) { j.str(s); });
    j.key("binner"), j.o << "{\n";
    j.indent("    ");
    j.key("max_bin", true), j.o << m.binner_max_bin;
    j.key("edges"), j.list(m.features, [&](const FileFeature& f) { j.list(f.edges, num); });
    j.key("is_cat"), j.list(m.features, [&](const FileFeature& f) { j.o << f.is_cat; });
    j.key("cat_values"), j.list(m.features, [&](const FileFeature& f) { j.list(f.cat_values, num); });
    j.indent("  ");
    j.o << "\n  }";
    j.key("trees"), j.o << "[\n";
    for (size_t t = 0; t < m.trees.size(); ++t) {
        const FileTree& tr = m.trees[t];
        j.o << (t ? ",\n" : "") << "    {\"target\": " << tr.target << ", \"nodes\": ";
        size_t k = 0;
        j.list(tr.nodes, [&](const FileNode& n) {
            // the dump shows the refit leaf value when the tree has one (values[k])
            const double v = tr.values.empty() ? n.value : tr.values[k];
            ++k;
            j.o << "{\"leaf\": " << (n.is_leaf ? "true" : "false") << ", \"f\": " << n.feature
                << ", \"thr\": " << n.threshold << ", \"left\": " << n.left << ", \"right\": " << n.right
                << ", \"miss_left\": " << n.missing_go_left << ", \"v\": ";
            num(v);
            j.o << ", \"gain\": ";
            num(n.gain);
            j.o << '}';
        });
        j.o << '}';
    }
    j.o << "\n  ]\n}\n";
    return j.o.str();
}

}  // namespace tide
