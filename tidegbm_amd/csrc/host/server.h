// Serving loop over resident ensembles and a resident binned panel (see server.cpp).
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "booster.h"

namespace tide {

struct ServerStats {
    long long predict_batches = 0, predict_requests = 0, predict_rows = 0;
    long long refit_rounds = 0, refit_requests = 0;
    double predict_busy_s = 0.0, refit_busy_s = 0.0;
    // refit thread breakdown: admitting (base predictions, slot binding, row upload), tree
    // rounds, publishing versions (model copy + device forest)
    double refit_admit_s = 0.0, refit_round_s = 0.0, refit_publish_s = 0.0;
};

class Server {
public:
    // models: fitted MSE ensembles sharing bin edges and growth settings, pair_splits=0.
    // X: the panel (n x F raw rows) binned once with the shared edges and kept resident.
    // predict_budget: staging bytes for one predict launch (admission limit of a batch).
    // refit_slots / refit_max_rows: concurrent refits and the largest refit window.
    // batched=false serves requests one at a time in arrival order (baseline).
    Server(const std::vector<const Booster*>& models, const double* X, int n, int F, size_t predict_budget,
           int refit_slots, int refit_max_rows, bool batched);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    long long submit_predict(int model, const int* rows, int n);
    long long submit_predict_x(int model, const double* X, int n);
    long long submit_refit(int model, int row0, int n, const double* y, int extra_trees);
    // Blocks until request `id` is done; copies predict scores to out (cap values), the version
    // used (predict) or published (refit) and the submit-to-done latency.  Returns the row count.
    int wait(long long id, double* out, int cap, int* version, double* latency_s);
    // Snapshot of an ensemble version (-1 = current) and its parent version (-1 for the original).
    std::shared_ptr<Booster> model_version(int model, int version, int* parent);
    ServerStats stats();

private:
    struct Version;
    struct Req;
    struct Slot;
    // Host side of a version: configuration / edges / init scores, and its trees, shared with the
    // versions it was derived from (publishing a refit copies pointers, not trees).
    struct Snap {
        std::shared_ptr<const Booster> proto;   // fitted model fields without trees
        std::vector<std::shared_ptr<const Tree>> trees;
    };
    struct HistEntry {
        int version, parent;
        std::shared_ptr<const Snap> snap;
    };
    // A version's device forest lives while it is current or an admitted request holds it; the
    // host snapshots of the last kKeepVersions versions stay queryable.
    static constexpr size_t kKeepVersions = 64;
    struct ModelState {
        std::deque<HistEntry> history;
        std::shared_ptr<Version> current;
        std::shared_ptr<tgbm_forest> shape;   // version 0's forest: staging sizes of any version
        int last_version = 0;
    };

    long long enqueue(std::shared_ptr<Req> r);
    const tgbm_forest* forest_shape(int model);
    void finish(const std::vector<std::shared_ptr<Req>>& done, const std::string& err);
    void run_predicts(tgbm_predictor* p, const std::vector<std::shared_ptr<Req>>& batch);
    void start_refit(int slot, const std::shared_ptr<Req>& r);
    void publish_refit(Slot& sl);
    void refit_round(int n_active);
    void predict_loop();
    void refit_loop();
    void publish_loop();
    void unbatched_loop();

    bool batched_;
    bool has_cat_ = false;   // the ensembles have categorical features (predicts only)
    std::chrono::steady_clock::time_point t0_;
    Config cfg_;
    int F_ = 0, n_rows_ = 0, cap_rows_ = 0, K_ = 1;
    size_t predict_cap_ = 0;
    tgbm_session* sess_ = nullptr;
    tgbm_predictor* pred_ = nullptr;
    std::vector<ModelState> models_;
    std::vector<Slot> slots_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_, cv_pub_;
    std::vector<Slot> pubq_;   // finished refits for the publishing thread
    std::deque<std::shared_ptr<Req>> pq_, rq_, uq_;
    std::map<long long, std::shared_ptr<Req>> reqs_;
    long long next_id_ = 1;
    bool stop_ = false;
    ServerStats st_;
    std::vector<std::thread> threads_;
    // refit round buffers (refit thread only)
    std::vector<int> nn_, ints_, cov_;
    std::vector<double> dbl_, vals_, pref_;
};

}  // namespace tide
