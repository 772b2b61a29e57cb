// Serving loop: resident ensembles over a resident binned panel, mixed predict and refit requests
// admitted by memory budget and coalesced into batched launches.
//
// The reference has no serving layer; its request-level operations are Booster::predict_raw
// (booster.cpp:801-842) and Booster::update (booster.cpp:844-884), each a separate call.  Here:
//
// * Predict requests (panel row ids, or raw feature rows) wait in a FIFO queue.  The predict
//   thread admits every queued request that fits the predictor's staging budget and scores them
//   all with one kernel launch (tgbm_predict_batch), each against the ensemble version current at
//   admission.  Requests arriving meanwhile are admitted at the next launch, as staging frees.
// * Refit requests (Booster::update semantics on a window of panel rows: the window's current
//   predictions, then extra_trees trees on all of its rows, one time bucket) take one of K task
//   slots of a multi-task device session over the same panel.  The refit thread grows one tree
//   for every slot in use per round (one launch per kernel for all of them); requests join at the
//   next round when a slot is free and leave when their trees are done; a publishing thread then
//   builds the new version of their ensemble (device forest and host snapshot) while the next
//   rounds run.  Predict launches run on their own stream and overlap the refit rounds.
// * Unbatched mode (the baseline) serves the same requests one at a time in arrival order with
//   the same kernels: one predict per launch, one refit grown alone to completion.
//
// A result does not depend on what else was in flight: predict scores follow the reference's
// summation order exactly, and a refit's trees depend only on its rows, targets and base version.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "booster.h"
#include "server.h"

namespace tide {

namespace {

using Clock = std::chrono::steady_clock;

double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

std::shared_ptr<tgbm_forest> own(tgbm_forest* f) {
    return std::shared_ptr<tgbm_forest>(f, [](tgbm_forest* p) { tgbm_forest_destroy(p); });
}

}  // namespace

struct Server::Version {
    int version = 0, parent = -1;
    std::shared_ptr<const Snap> snap;
    std::shared_ptr<tgbm_forest> forest;   // the first forest_trees trees of snap
    int forest_trees = 0;
};

struct Server::Req {
    long long id = 0;
    int kind = 0;           // 0 predict rows, 1 predict raw rows, 2 refit
    int model = 0;
    int n = 0;
    std::vector<int> rows;
    std::vector<double> X, y;
    int row0 = 0, extra_trees = 0;
    size_t bytes = 0;
    std::shared_ptr<Version> base;    // version used (predict) or refitted (refit)
    std::vector<double> out;
    int result_version = -1;
    Clock::time_point t_submit, t_done;
    bool done = false;
    std::string error;
};

struct Server::Slot {
    std::shared_ptr<Req> req;
    std::vector<Tree> trees;
};

Server::Server(const std::vector<const Booster*>& models, const double* X, int n, int F, size_t predict_budget,
               int refit_slots, int refit_max_rows, bool batched)
    : batched_(batched), t0_(Clock::now()) {
    if (models.empty()) throw Error("a server needs at least one ensemble");
    if (n <= 0 || F <= 0 || !X) throw Error("a server needs a panel");
    if (refit_slots < 1) throw Error("refit_slots must be >= 1");
    if (refit_max_rows < 1 || refit_max_rows > n) throw Error("refit_max_rows must be in [1, panel rows]");
    const Booster& m0 = *models[0];
    for (const Booster* m : models) {
        if (!m || !m->fitted) throw Error("every ensemble must be fitted");
        if (!m->cfg.interaction_group.empty() || m->cfg.regime_feature >= 0)
            throw Error("serving refits grow on the device: ensembles need no interaction groups or regime feature");
        if (m->n_features != F) throw Error("ensemble and panel feature counts differ");
        if (m->edges != m0.edges || m->n_bins != m0.n_bins || m->is_cat != m0.is_cat ||
            (m->has_cat() && m->cat_values != m0.cat_values))
            throw Error("all ensembles of a server must share bin edges and categories (one binned panel serves "
                        "them all)");
        const Config& a = m->cfg;
        const Config& b = m0.cfg;
        if (a.objective != 0) throw Error("serving refits support the mse objective only");
        if (a.pair_splits || a.force_host_driver) throw Error("serving refits grow on the device: ensembles need pair_splits=0");
        if (a.max_depth != b.max_depth || a.max_leaves != b.max_leaves || a.min_samples_leaf != b.min_samples_leaf ||
            a.min_child_weight != b.min_child_weight || a.l2 != b.l2 || a.l1 != b.l1 ||
            a.max_delta_step != b.max_delta_step || a.min_split_gain != b.min_split_gain ||
            a.learning_rate != b.learning_rate || a.monotone != b.monotone)
            throw Error("ensembles of a server must share their tree-growth settings (refits are grown together)");
    }
    if (tgbm_device_count() < 1) throw Error("GPU device not available: no HIP device was found for device=gpu");
    cfg_ = m0.cfg;
    F_ = F;
    n_rows_ = n;
    cap_rows_ = refit_max_rows;
    K_ = refit_slots;
    // panel bins + refit slots in one session
    std::vector<int> caps(K_, refit_max_rows);
    dev_check(tgbm_session_create(&sess_, n, F, m0.n_bins, K_, caps.data(), cfg_.max_leaves, 1));
    // refit trees grow several leaves per step, like every device-driven fit
    dev_check(tgbm_session_set_spec(sess_, cfg_.spec_width > 1 ? cfg_.spec_width : 0));
    if (!cfg_.monotone.empty()) dev_check(tgbm_session_set_monotone(sess_, cfg_.monotone.data()));
    bin_on_device(sess_, X, m0.edges, m0.is_cat, m0.cat_values);
    has_cat_ = m0.has_cat();
    dev_check(tgbm_predictor_create(&pred_, predict_budget));
    predict_cap_ = std::max<size_t>(predict_budget, 1 << 20);
    for (const Booster* m : models) {
        auto snap = std::make_shared<Snap>();
        auto proto = m->clone_model();
        for (auto& t : proto->trees) snap->trees.push_back(std::make_shared<const Tree>(std::move(t)));
        proto->trees.clear();
        snap->proto = proto;
        auto v = std::make_shared<Version>();
        v->version = 0;
        v->snap = snap;
        v->forest = own(make_forest(*m));
        v->forest_trees = std::min((int)m->trees.size(), m->best_iteration > 0 ? m->best_iteration : m->cfg.n_estimators);
        ModelState ms;
        ms.history.push_back(HistEntry{0, -1, snap});
        ms.current = v;
        ms.shape = v->forest;
        models_.push_back(std::move(ms));
    }
    slots_.resize(K_);
    if (batched_) {
        threads_.emplace_back([this] { predict_loop(); });
        threads_.emplace_back([this] { refit_loop(); });
        threads_.emplace_back([this] { publish_loop(); });
    } else {
        threads_.emplace_back([this] { unbatched_loop(); });
    }
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    cv_pub_.notify_all();
    for (auto& t : threads_) t.join();
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& kv : reqs_) {
        if (!kv.second->done) {
            kv.second->done = true;
            kv.second->error = "server stopped";
        }
    }
    cv_done_.notify_all();
    if (pred_) tgbm_predictor_destroy(pred_);
    if (sess_) tgbm_session_destroy(sess_);
}

long long Server::enqueue(std::shared_ptr<Req> r) {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw Error("server stopped");
    r->id = next_id_++;
    r->t_submit = Clock::now();
    reqs_[r->id] = r;
    if (!batched_) uq_.push_back(r);
    else if (r->kind == 2) rq_.push_back(r);
    else pq_.push_back(r);
    cv_work_.notify_all();
    return r->id;
}

// Staging sizes depend only on the feature count, shared by all versions of a model.
const tgbm_forest* Server::forest_shape(int model) { return models_[model].shape.get(); }

long long Server::submit_predict(int model, const int* rows, int n) {
    if (model < 0 || model >= (int)models_.size()) throw Error("model index out of range");
    if (n < 1) throw Error("a predict request needs at least one row");
    auto r = std::make_shared<Req>();
    r->kind = 0;
    r->model = model;
    r->n = n;
    r->rows.assign(rows, rows + n);
    for (int k = 0; k < n; ++k)
        if (rows[k] < 0 || rows[k] >= n_rows_) throw Error("panel row id out of range");
    r->bytes = tgbm_predict_bytes(forest_shape(model), 0, n);
    if (r->bytes > predict_cap_) throw Error("predict request exceeds the server's predict memory budget");
    return enqueue(r);
}

long long Server::submit_predict_x(int model, const double* X, int n) {
    if (model < 0 || model >= (int)models_.size()) throw Error("model index out of range");
    if (n < 1) throw Error("a predict request needs at least one row");
    auto r = std::make_shared<Req>();
    r->kind = 1;
    r->model = model;
    r->n = n;
    r->X.assign(X, X + (size_t)n * F_);
    r->bytes = tgbm_predict_bytes(forest_shape(model), 1, n);
    if (r->bytes > predict_cap_) throw Error("predict request exceeds the server's predict memory budget");
    return enqueue(r);
}

long long Server::submit_refit(int model, int row0, int n, const double* y, int extra_trees) {
    if (model < 0 || model >= (int)models_.size()) throw Error("model index out of range");
    if (n < 1 || n > cap_rows_) throw Error("refit rows exceed the server's refit slot capacity");
    if (row0 < 0 || (long long)row0 + n > n_rows_) throw Error("refit window outside the panel");
    if (extra_trees < 1) throw Error("extra_trees must be >= 1");
    // categorical splits are decided by the host-driven grower, which the refit slots do not run
    if (has_cat_) throw Error("refits of ensembles with categorical features are not served (fit them with update)");
    auto r = std::make_shared<Req>();
    r->kind = 2;
    r->model = model;
    r->n = n;
    r->row0 = row0;
    r->y.assign(y, y + n);
    r->extra_trees = extra_trees;
    return enqueue(r);
}

int Server::wait(long long id, double* out, int cap, int* version, double* latency_s) {
    std::shared_ptr<Req> r;
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) throw Error("unknown request ticket");
    r = it->second;
    cv_done_.wait(lk, [&] { return r->done; });
    reqs_.erase(id);
    if (!r->error.empty()) throw Error(r->error);
    if (version) *version = r->result_version;
    if (latency_s) *latency_s = secs(r->t_submit, r->t_done);
    if (out && r->kind != 2) std::memcpy(out, r->out.data(), sizeof(double) * std::min(cap, r->n));
    return r->n;
}

std::shared_ptr<Booster> Server::model_version(int model, int version, int* parent) {
    std::shared_ptr<const Snap> snap;
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (model < 0 || model >= (int)models_.size()) throw Error("model index out of range");
        auto& h = models_[model].history;
        if (version < 0) version = models_[model].current->version;
        for (auto& v : h)
            if (v.version == version) {
                if (parent) *parent = v.parent;
                snap = v.snap;
            }
    }
    if (!snap) throw Error("unknown model version (only the most recent versions are kept)");
    auto b = snap->proto->clone_model();
    b->trees.reserve(snap->trees.size());
    for (auto& t : snap->trees) b->trees.push_back(*t);
    if (version > 0) b->best_iteration = (int)b->trees.size();
    return b;
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    return st_;
}

void Server::finish(const std::vector<std::shared_ptr<Req>>& done, const std::string& err) {
    std::lock_guard<std::mutex> lk(mu_);
    const auto now = Clock::now();
    for (auto& r : done) {
        r->t_done = now;
        r->error = err;
        r->done = true;
    }
    cv_done_.notify_all();
}

void Server::run_predicts(tgbm_predictor* p, const std::vector<std::shared_ptr<Req>>& batch) {
    std::vector<tgbm_predict_item> items(batch.size());
    for (size_t i = 0; i < batch.size(); ++i) {
        Req& r = *batch[i];
        r.out.resize(r.n);
        items[i].forest = r.base->forest.get();
        items[i].rows = r.kind == 0 ? r.rows.data() : nullptr;
        items[i].X = r.kind == 1 ? r.X.data() : nullptr;
        items[i].n = r.n;
        items[i].out = r.out.data();
        r.result_version = r.base->version;
    }
    dev_check(tgbm_predict_batch(p, sess_, items.data(), (int)items.size()));
}

void Server::predict_loop() {
    for (;;) {
        std::vector<std::shared_ptr<Req>> batch;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_work_.wait(lk, [&] { return stop_ || !pq_.empty(); });
            if (stop_) return;
            size_t bytes = 0;
            while (!pq_.empty()) {   // FIFO admission while the staging budget lasts
                auto& r = pq_.front();
                if (!batch.empty() && bytes + r->bytes > predict_cap_) break;
                bytes += r->bytes;
                r->base = models_[r->model].current;
                batch.push_back(r);
                pq_.pop_front();
            }
        }
        std::string err;
        const auto t = Clock::now();
        try {
            run_predicts(pred_, batch);
        } catch (const std::exception& e) {
            err = e.what();
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            st_.predict_batches += 1;
            st_.predict_requests += (long long)batch.size();
            for (auto& r : batch) st_.predict_rows += r->n;
            st_.predict_busy_s += secs(t, Clock::now());
        }
        for (auto& r : batch) {   // inputs are no longer needed; results stay until waited for
            std::vector<int>().swap(r->rows);
            std::vector<double>().swap(r->X);
        }
        finish(batch, err);
    }
}

// Bind a free slot to a refit request: the window's rows, its targets and the base version's
// predictions (Booster::update: current model on the new rows, booster.cpp:859).
void Server::start_refit(int slot, const std::shared_ptr<Req>& r) {
    // bind + base predictions + targets, queued without a wait (the next round follows them)
    dev_check(tgbm_task_refit_window(sess_, slot, r->row0, r->n, r->y.data(), r->base->forest.get()));
    slots_[slot].req = r;
    slots_[slot].trees.clear();
    std::vector<double>().swap(r->y);   // inputs are on the device now
}

// Publish base version + new trees as the ensemble's next version (Booster::update appends the
// trees and sets best_iteration to the new count, booster.cpp:880-883).
void Server::publish_refit(Slot& sl) {
    Req& r = *sl.req;
    const Version& base = *r.base;
    auto snap = std::make_shared<Snap>();
    snap->proto = base.snap->proto;
    snap->trees = base.snap->trees;   // shared tree storage
    for (auto& t : sl.trees) snap->trees.push_back(std::make_shared<const Tree>(std::move(t)));
    // device forest: the parent's forest plus every tree after it (the new ones)
    const int keep = base.forest_trees;
    const int add = (int)snap->trees.size() - keep;
    std::vector<int> tn(add), ints;
    std::vector<double> vals, lrs(add, snap->proto->cfg.learning_rate);
    for (int k = 0; k < add; ++k) {
        const Tree& tr = *snap->trees[keep + k];
        tn[k] = (int)tr.nodes.size();
        for (size_t j = 0; j < tr.nodes.size(); ++j) {
            const Node& d = tr.nodes[j];
            const int v[10] = {d.is_leaf, d.split_type, d.feature, d.feature2, d.threshold, d.left, d.right,
                               d.missing_go_left, d.n_samples, d.depth};
            ints.insert(ints.end(), v, v + 10);
            vals.push_back(tr.values.empty() ? d.value : tr.values[j]);
        }
    }
    tgbm_forest* f = nullptr;
    dev_check(tgbm_forest_append(&f, base.forest.get(), keep, add, tn.data(), ints.data(), vals.data(), lrs.data(),
                                 nullptr));
    auto v = std::make_shared<Version>();
    v->snap = snap;
    v->forest = own(f);
    v->forest_trees = (int)snap->trees.size();
    v->parent = base.version;
    std::lock_guard<std::mutex> lk(mu_);
    ModelState& ms = models_[r.model];
    v->version = ++ms.last_version;
    ms.history.push_back(HistEntry{v->version, v->parent, snap});
    while (ms.history.size() > kKeepVersions) ms.history.pop_front();
    ms.current = v;
    r.result_version = v->version;
}

// One boosting round for slots [0, n_active): every active refit grows one more tree.
void Server::refit_round(int n_active) {
    const tgbm_grow_params gp = grow_params(cfg_);
    const double min_cover =
        cfg_.torb_min_cover > 0.0 ? cfg_.torb_min_cover : (double)(2 * std::max(cfg_.min_samples_leaf, 1));
    const int MN = 2 * cfg_.max_leaves - 1;
    nn_.resize(n_active);
    ints_.resize((size_t)n_active * MN * TGBM_NODE_INTS);
    cov_.resize((size_t)n_active * MN);
    dbl_.resize((size_t)n_active * MN * 2);
    vals_.resize((size_t)n_active * MN);
    pref_.resize((size_t)n_active * MN * 256);
    dev_check(tgbm_session_set_active(sess_, 0, n_active));
    dev_check(tgbm_task_grad(sess_, -1, 0));
    dev_check(tgbm_tree_begin(sess_, &gp));
    dev_check(tgbm_tree_grow_device(sess_, &gp));
    dev_check(tgbm_tree_end(sess_, cfg_.learning_rate, min_cover, 1, &gp, nn_.data(), ints_.data(), dbl_.data(),
                            vals_.data(), cov_.data(), pref_.data()));
    for (int i = 0; i < n_active; ++i) {
        const size_t o = (size_t)i * MN;
        Tree t = unpack_tree(nn_[i], &ints_[o * TGBM_NODE_INTS], &dbl_[o * 2], &vals_[o], &cov_[o], &pref_[o * 256], 1);
        for (size_t k = 0; k < t.nodes.size(); ++k) t.cover[k] = t.nodes[k].n_samples;  // update keeps node counts
        t.prefix_values = t.values;
        slots_[i].trees.push_back(std::move(t));
    }
}

void Server::refit_loop() {
    int n_active = 0;
    for (;;) {
        std::vector<std::shared_ptr<Req>> admit;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_work_.wait(lk, [&] { return stop_ || !rq_.empty() || n_active > 0; });
            if (stop_) return;
            while (!rq_.empty() && n_active + (int)admit.size() < K_) {
                auto r = rq_.front();
                rq_.pop_front();
                r->base = models_[r->model].current;
                admit.push_back(r);
            }
        }
        const auto t = Clock::now();
        std::vector<std::shared_ptr<Req>> failed;
        std::string err;
        for (auto& r : admit) {
            try {
                start_refit(n_active, r);
                ++n_active;
            } catch (const std::exception& e) {
                err = e.what();
                failed.push_back(r);
            }
        }
        const auto t_admit = Clock::now();
        if (!failed.empty()) finish(failed, err);
        if (n_active == 0) continue;
        std::vector<std::shared_ptr<Req>> done;
        std::vector<Slot> finished;
        auto t_round = Clock::now();
        try {
            refit_round(n_active);
            t_round = Clock::now();
            for (int i = n_active - 1; i >= 0; --i) {   // retire finished refits, keep slots packed
                if ((int)slots_[i].trees.size() < slots_[i].req->extra_trees) continue;
                finished.push_back(std::move(slots_[i]));
                const int last = n_active - 1;
                if (i != last) {
                    dev_check(tgbm_task_swap(sess_, i, last));
                    std::swap(slots_[i], slots_[last]);
                }
                slots_[last] = Slot{};
                --n_active;
            }
        } catch (const std::exception& e) {
            // a device failure ends every refit in flight
            for (int i = 0; i < n_active; ++i) done.push_back(slots_[i].req);
            for (auto& sl : slots_) sl = Slot{};
            n_active = 0;
            err = e.what();
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            st_.refit_rounds += 1;
            st_.refit_requests += (long long)(done.size() + finished.size());
            st_.refit_busy_s += secs(t, Clock::now());
            st_.refit_admit_s += secs(t, t_admit);
            st_.refit_round_s += secs(t_admit, t_round);
            for (auto& sl : finished) pubq_.push_back(std::move(sl));
        }
        if (!finished.empty()) cv_pub_.notify_one();
        if (!done.empty()) finish(done, err);
    }
}

// Publishing thread: finished refits become new versions in the order their trees were done.
void Server::publish_loop() {
    for (;;) {
        std::vector<Slot> batch;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_pub_.wait(lk, [&] { return stop_ || !pubq_.empty(); });
            if (pubq_.empty()) return;   // stopped, nothing left to publish
            batch.swap(pubq_);
        }
        for (auto& sl : batch) {
            std::string err;
            const auto tp = Clock::now();
            try {
                publish_refit(sl);
            } catch (const std::exception& e) {
                err = e.what();
            }
            {
                std::lock_guard<std::mutex> lk(mu_);
                st_.refit_publish_s += secs(tp, Clock::now());
            }
            finish({sl.req}, err);
        }
    }
}

void Server::unbatched_loop() {
    for (;;) {
        std::shared_ptr<Req> r;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_work_.wait(lk, [&] { return stop_ || !uq_.empty(); });
            if (stop_) return;
            r = uq_.front();
            uq_.pop_front();
            r->base = models_[r->model].current;
        }
        std::string err;
        const auto t = Clock::now();
        try {
            if (r->kind == 2) {
                start_refit(0, r);
                while ((int)slots_[0].trees.size() < r->extra_trees) refit_round(1);
                publish_refit(slots_[0]);
                slots_[0] = Slot{};
            } else {
                run_predicts(pred_, {r});
            }
        } catch (const std::exception& e) {
            err = e.what();
            slots_[0] = Slot{};
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (r->kind == 2) {
                st_.refit_rounds += r->extra_trees;
                st_.refit_requests += 1;
                st_.refit_busy_s += secs(t, Clock::now());
            } else {
                st_.predict_batches += 1;
                st_.predict_requests += 1;
                st_.predict_rows += r->n;
                st_.predict_busy_s += secs(t, Clock::now());
            }
        }
        finish({r}, err);
    }
}

}  // namespace tide
