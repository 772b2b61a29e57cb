// TGBM v3 model files, byte-compatible with the reference writer/reader
// (reference cpp/src/serialize.cpp:65-255; magic and version from types.hpp:22-23), and the
// reference's JSON dump (serialize.cpp:266-324).  Fields the GPU path does not use are written
// with the reference defaults (config.hpp:12-113) so the reference can load these files.
#include <cstring>
#include <sstream>

#include "booster.h"

namespace tide {

namespace {

void put_u32(std::string& s, uint32_t v) {
    const char b[4] = {(char)(v & 255), (char)((v >> 8) & 255), (char)((v >> 16) & 255), (char)((v >> 24) & 255)};
    s.append(b, 4);
}
void put_i32(std::string& s, int32_t v) { put_u32(s, (uint32_t)v); }
void put_i64(std::string& s, int64_t v) {
    put_u32(s, (uint32_t)v);
    put_u32(s, (uint32_t)((uint64_t)v >> 32));
}
void put_f64(std::string& s, double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    put_u32(s, (uint32_t)u);
    put_u32(s, (uint32_t)(u >> 32));
}
void put_str(std::string& s, const std::string& x) {
    put_i32(s, (int32_t)x.size());
    s.append(x);
}

struct Reader {
    const char* p;
    const char* e;
    explicit Reader(const std::string& b) : p(b.data()), e(b.data() + b.size()) {}
    uint32_t u32() {
        if (e - p < 4) throw Error("truncated model");
        const uint32_t v = (uint8_t)p[0] | ((uint32_t)(uint8_t)p[1] << 8) | ((uint32_t)(uint8_t)p[2] << 16) |
                           ((uint32_t)(uint8_t)p[3] << 24);
        p += 4;
        return v;
    }
    int32_t i32() { return (int32_t)u32(); }
    int64_t i64() {
        const uint32_t lo = u32(), hi = u32();
        return (int64_t)((uint64_t)lo | ((uint64_t)hi << 32));
    }
    double f64() {
        const uint32_t lo = u32(), hi = u32();
        const uint64_t u = (uint64_t)lo | ((uint64_t)hi << 32);
        double d;
        std::memcpy(&d, &u, 8);
        return d;
    }
    std::string str() {
        const int n = i32();
        if (n < 0 || e - p < n) throw Error("truncated string");
        std::string s(p, p + n);
        p += n;
        return s;
    }
    int count(int max_ok = 1 << 28) {
        const int n = i32();
        if (n < 0 || n > max_ok) throw Error("corrupt model count");
        return n;
    }
};

}  // namespace

std::string Booster::save_binary() const {
    std::string s;
    s.append("TGBM", 4);
    put_u32(s, 3);
    put_i32(s, cfg.n_estimators);
    put_f64(s, cfg.learning_rate);
    put_i32(s, cfg.max_depth);
    put_i32(s, cfg.max_leaves);
    put_i32(s, cfg.min_samples_leaf);
    put_f64(s, cfg.min_child_weight);
    put_f64(s, cfg.l2);
    put_f64(s, cfg.min_split_gain);
    put_i32(s, cfg.max_bin);
    put_f64(s, cfg.subsample);
    put_f64(s, cfg.colsample);
    put_i32(s, cfg.n_time_buckets);
    put_i64(s, cfg.horizon);
    put_i64(s, cfg.embargo);
    put_i32(s, cfg.iid_mode ? 1 : 0);
    put_i32(s, cfg.halo ? 1 : 0);
    put_i32(s, cfg.objective);
    put_f64(s, 0.5);  // quantile
    put_f64(s, 1.0);  // huber_delta
    put_f64(s, 1.0);  // rank_temperature
    put_i32(s, cfg.random_state);
    put_i32(s, cfg.n_jobs);
    put_f64(s, cfg.tree_decay);
    put_i32(s, n_features);
    put_i32(s, n_bins);
    put_i32(s, n_time_buckets);
    put_i32(s, best_iteration);
    put_f64(s, init_score);
    put_i32(s, (int)init_prefix.size());
    for (double v : init_prefix) put_f64(s, v);
    put_i32(s, n_features);  // feature names f0..f{F-1} (booster.cpp:244-245)
    for (int i = 0; i < n_features; ++i) put_str(s, "f" + std::to_string(i));
    auto cat = [&](int f) { return (int)is_cat.size() == n_features && is_cat[f] ? 1 : 0; };
    put_i32(s, n_features);  // is_cat
    for (int i = 0; i < n_features; ++i) put_i32(s, cat(i));
    put_i32(s, n_features);  // group_of
    for (int i = 0; i < n_features; ++i) put_i32(s, -1);
    put_i32(s, n_features);  // monotone
    for (int i = 0; i < n_features; ++i) put_i32(s, 0);
    put_i32(s, cfg.max_bin);  // binner
    put_i32(s, n_features);
    for (int f = 0; f < n_features; ++f) {
        put_i32(s, cat(f));
        put_i32(s, (int)edges[f].size());
        for (double e : edges[f]) put_f64(s, e);
        const std::vector<double> none;
        const std::vector<double>& cv = cat(f) ? cat_values[f] : none;
        put_i32(s, (int)cv.size());
        for (double v : cv) put_f64(s, v);
        put_i32(s, cat(f) ? (int)cv.size() + 1 : (int)edges[f].size() + 2);   // n_bins_col
    }
    put_i32(s, (int)trees.size());
    for (const Tree& t : trees) {
        put_i32(s, t.n_buckets);
        put_i32(s, (int)t.nodes.size());
        for (const Node& n : t.nodes) {
            put_i32(s, n.is_leaf ? 1 : 0);
            put_i32(s, n.split_type);
            put_i32(s, n.feature);
            put_i32(s, n.feature2);
            put_i32(s, n.threshold);
            put_i32(s, n.left);
            put_i32(s, n.right);
            put_i32(s, n.missing_go_left);
            put_f64(s, n.value);
            put_f64(s, n.gain);
            put_i32(s, n.n_samples);
            put_i32(s, n.depth);
            for (int k = 0; k < 4; ++k) put_i64(s, (long long)n.cat_left[k]);
        }
        put_i32(s, (int)t.values.size());
        for (double v : t.values) put_f64(s, v);
        put_i32(s, (int)t.cover.size());
        for (int v : t.cover) put_i32(s, v);
        put_i32(s, (int)t.prefix_values.size());
        for (double v : t.prefix_values) put_f64(s, v);
        put_i32(s, 0);  // target
        put_i32(s, 1);  // n_targets
    }
    put_i32(s, cfg.objective);
    return s;
}

Booster* Booster::load_binary(const std::string& blob) {
    if (blob.size() < 8 || std::memcmp(blob.data(), "TGBM", 4) != 0) throw Error("bad magic");
    Reader r(blob);
    r.p += 4;
    if (r.u32() != 3) throw Error("unsupported model version");
    auto* b = new Booster();
    try {
        Config& c = b->cfg;
        c.n_estimators = r.i32();
        c.learning_rate = r.f64();
        c.max_depth = r.i32();
        c.max_leaves = r.i32();
        c.min_samples_leaf = r.i32();
        c.min_child_weight = r.f64();
        c.l2 = r.f64();
        c.min_split_gain = r.f64();
        c.max_bin = r.i32();
        c.subsample = r.f64();
        c.colsample = r.f64();
        c.n_time_buckets = r.i32();
        c.horizon = r.i64();
        c.embargo = r.i64();
        c.iid_mode = r.i32() != 0;
        c.halo = r.i32() != 0;
        c.objective = r.i32();
        (void)r.f64();
        (void)r.f64();
        (void)r.f64();
        c.random_state = r.i32();
        c.n_jobs = r.i32();
        c.tree_decay = r.f64();
        b->n_features = r.i32();
        b->n_bins = r.i32();
        b->n_time_buckets = r.i32();
        b->best_iteration = r.i32();
        b->init_score = r.f64();
        b->init_prefix.resize(r.count());
        for (double& v : b->init_prefix) v = r.f64();
        const int nn = r.count();
        for (int i = 0; i < nn; ++i) (void)r.str();
        const int nc = r.count();
        b->is_cat.assign(nc, 0);
        for (int i = 0; i < nc; ++i) b->is_cat[i] = r.i32() != 0 ? 1 : 0;
        if (!b->has_cat()) b->is_cat.clear();
        const int ng = r.count();
        for (int i = 0; i < ng; ++i) (void)r.i32();
        const int nm = r.count();
        for (int i = 0; i < nm; ++i)
            if (r.i32() != 0) throw Error("monotone constraints are not supported on the GPU path");
        (void)r.i32();  // binner max_bin
        const int bf = r.count(1 << 20);
        if (bf != b->n_features) throw Error("binner feature count mismatch");
        b->edges.assign(bf, {});
        b->cat_values.assign(bf, {});
        for (int f = 0; f < bf; ++f) {
            const bool cat = r.i32() != 0;
            if (cat != (b->has_cat() && b->is_cat[f] != 0)) throw Error("binner categorical flags mismatch");
            b->edges[f].resize(r.count(1 << 16));
            for (double& e : b->edges[f]) e = r.f64();
            b->cat_values[f].resize(r.count(1 << 16));
            for (double& v : b->cat_values[f]) v = r.f64();
            if (b->cat_values[f].size() > 254 || b->edges[f].size() > 254)
                throw Error("more than 254 bin edges or categories in a feature");
            (void)r.i32();
        }
        b->trees.resize(r.count());
        for (Tree& t : b->trees) {
            t.n_buckets = r.i32();
            t.nodes.resize(r.count(1 << 20));
            for (Node& n : t.nodes) {
                n.is_leaf = r.i32() != 0;
                n.split_type = r.i32();
                n.feature = r.i32();
                n.feature2 = r.i32();
                n.threshold = r.i32();
                n.left = r.i32();
                n.right = r.i32();
                n.missing_go_left = r.i32();
                n.value = r.f64();
                n.gain = r.f64();
                n.n_samples = r.i32();
                n.depth = r.i32();
                for (int k = 0; k < 4; ++k) n.cat_left[k] = (unsigned long long)r.i64();
                if (!n.is_leaf && (n.split_type < 0 || n.split_type > 3)) throw Error("unknown split type");
                if (!n.is_leaf && n.split_type >= 2 && !b->has_cat())
                    throw Error("categorical split in a model without categorical features");
            }
            t.values.resize(r.count());
            for (double& v : t.values) v = r.f64();
            t.cover.resize(r.count());
            for (int& v : t.cover) v = r.i32();
            t.prefix_values.resize(r.count());
            for (double& v : t.prefix_values) v = r.f64();
            if (r.i32() != 0) throw Error("multi-head models are not supported on the GPU path");
            (void)r.i32();
        }
        c.objective = r.i32();
        b->fitted = true;
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}

std::string Booster::save_json() const {  // reference serialize.cpp:266-324
    std::ostringstream o;
    auto cat = [&](int f) { return (int)is_cat.size() == n_features && is_cat[f] ? 1 : 0; };
    o << "{\n  \"format\": \"tidegbm\",\n  \"version\": " << 3 << ",\n";
    o << "  \"n_features\": " << n_features << ",\n";
    o << "  \"n_bins\": " << n_bins << ",\n";
    o << "  \"best_iteration\": " << best_iteration << ",\n";
    o << "  \"init_score\": " << init_score << ",\n";
    o << "  \"learning_rate\": " << cfg.learning_rate << ",\n";
    o << "  \"horizon\": " << cfg.horizon << ",\n";
    o << "  \"iid_mode\": " << (cfg.iid_mode ? "true" : "false") << ",\n";
    o << "  \"feature_names\": [";
    for (int i = 0; i < n_features; ++i) o << (i ? ", " : "") << "\"f" << i << "\"";
    o << "],\n  \"binner\": {\n    \"max_bin\": " << cfg.max_bin << ",\n    \"edges\": [";
    for (int f = 0; f < n_features; ++f) {
        o << (f ? ", " : "") << "[";
        for (size_t i = 0; i < edges[f].size(); ++i) o << (i ? ", " : "") << edges[f][i];
        o << "]";
    }
    o << "],\n    \"is_cat\": [";
    for (int f = 0; f < n_features; ++f) o << (f ? ", " : "") << cat(f);
    o << "],\n    \"cat_values\": [";
    for (int f = 0; f < n_features; ++f) {
        o << (f ? ", " : "") << "[";
        if (cat(f))
            for (size_t i = 0; i < cat_values[f].size(); ++i) o << (i ? ", " : "") << cat_values[f][i];
        o << "]";
    }
    o << "]\n  },\n  \"trees\": [\n";
    for (size_t ti = 0; ti < trees.size(); ++ti) {
        const Tree& t = trees[ti];
        if (ti) o << ",\n";
        o << "    {\"target\": " << 0 << ", \"nodes\": [";
        for (size_t ni = 0; ni < t.nodes.size(); ++ni) {
            const Node& n = t.nodes[ni];
            if (ni) o << ", ";
            const double val = t.values.empty() ? n.value : t.values[ni];
            o << "{\"leaf\": " << (n.is_leaf ? "true" : "false") << ", \"f\": " << n.feature
              << ", \"thr\": " << n.threshold << ", \"left\": " << n.left << ", \"right\": " << n.right
              << ", \"miss_left\": " << n.missing_go_left << ", \"v\": " << val << ", \"gain\": " << n.gain << "}";
        }
        o << "]}";
    }
    o << "\n  ]\n}\n";
    return o.str();
}

}  // namespace tide
