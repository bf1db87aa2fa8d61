// Continuous-batching solve server.
//
// Requests carry an arrival time on the server clock.  Preparation threads
// take arrived requests in arrival order while a resident slot and device
// memory under the HBM budget are free, builds the mesh, sets up the solve
// on the device (geometry, tiling, initial condition, FP64 audit of the
// initial state) on the solve's own stream, and hands it to the GPU thread.
// The GPU thread admits ready solves into the batch engine between rounds,
// replays `steps_per_round` captured batch steps (every advancing solve
// takes its own CFL step per batch step; solves of at least fifo_min_elems
// elements advance one at a time in arrival order, the smaller ones beside
// them), reads the per-solve done flags, and
// retires finished solves: FP64 audit of the final state, optional state
// copy, device memory returned to the budget.  Preparation of new requests
// overlaps the batch steps of resident ones.
//
// A solve's arithmetic depends only on its own data (per-solve dt, no
// cross-solve reductions), so its result is the same alone and in any mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "dg_batch.h"
#include "dg_server.h"

namespace dgmi {

// device bytes a solve of this size will hold (the solver's allocations)
static int64_t estimate_bytes(int K, int N, int p, int precision) {
    const int64_t Kp = (K + 63) / 64 * 64;
    const int64_t s = precision == DG_FP64 ? 8 : 4;
    // p <= 4 reserves the matrix-core tile layout: modes padded to a multiple
    // of 4 in FP32, to 16 in FP64
    const bool tl = p >= 1 && p <= 4;
    const int64_t np = !tl ? n_modes(p) : precision == DG_FP32 ? 4 * tile_groups(n_modes(p)) : 16;
    const int64_t ntiles = (K + 63) / 64;
    // matrix-core tiles (p >= 3 in a batch): packed geometry, 16 values per element
    const int64_t packed = p >= 3 ? 16 * s : 0;
    return Kp * (4 * N * np * s + 14 * s + 3 * 4 + 6 * 8 + 8 + 8 + 3 * 2 + packed) + ntiles * 64 * 4 + 4096;
}

class ServerImpl : public Server {
   public:
    dg_server_config cfg{};
    BatchEngine* eng = nullptr;
    std::chrono::steady_clock::time_point t0;

    struct Pending {
        uint64_t id;
        dg_request req;
    };
    struct Job {
        uint64_t id = 0;
        dg_request req{};
        Solver* s = nullptr;
        int64_t bytes = 0;
        double admit_s = 0;
        double mass0[4] = {0, 0, 0, 0};
        int status = DG_OK;
    };

    std::mutex mu;
    std::condition_variable cv;       // any state change
    std::deque<Pending> pending;      // submitted, in arrival order
    std::deque<Job> ready;            // prepared, waiting for the GPU thread
    std::vector<Job> active;          // resident in the engine
    std::deque<dg_result> results;
    std::map<uint64_t, std::vector<double>> states;
    bool keep = false;
    int64_t bytes_held = 0;           // budget used by preparing, ready and active solves
    int slots_held = 0;
    uint64_t next_id = 1, n_submitted = 0, n_finished = 0;
    bool stop = false;
    std::string fatal;
    std::vector<std::thread> prep_threads;
    std::thread gpu_thread;

    double now() override {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    ~ServerImpl() override {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        for (std::thread& th : prep_threads)
            if (th.joinable()) th.join();
        if (gpu_thread.joinable()) gpu_thread.join();
        for (Job& j : ready) delete j.s;
        for (Job& j : active) delete j.s;
        delete eng;
    }

    Status start() {
        t0 = std::chrono::steady_clock::now();
        Status st = BatchEngine::create(cfg.precision, cfg.max_slots, 1 << 16, &eng);
        if (!st) return st;
        eng->set_solo_threshold(cfg.solo_min_elems);
        for (int i = 0; i < std::max(1, cfg.prep_threads); i++) prep_threads.emplace_back([this] { prep_loop(); });
        gpu_thread = std::thread([this] { gpu_loop(); });
        return Status::ok();
    }

    Status submit(const dg_request& req, uint64_t* id) override {
        if (req.nx < 1 || req.ny < 1 || req.p < 1 || req.p > 4 || !(req.t_end > 0))
            return Status::err(DG_ERR_ARG, "invalid request");
        Status st = check_problem(req.prob, req.p, cfg.precision);
        if (!st) return st;
        const int64_t need = estimate_bytes(2 * req.nx * req.ny, req.prob.physics, req.p, cfg.precision);
        if (need > cfg.hbm_budget) return Status::err(DG_ERR_BUDGET, "request exceeds the HBM budget on its own");
        std::lock_guard<std::mutex> lk(mu);
        if (!fatal.empty()) return Status::err(DG_ERR_STATE, "server stopped: " + fatal);
        Pending p{next_id++, req};
        if (p.req.arrival_s <= 0) p.req.arrival_s = now();
        // keep arrival order
        auto it = pending.end();
        while (it != pending.begin() && (it - 1)->req.arrival_s > p.req.arrival_s) --it;
        pending.insert(it, p);
        n_submitted++;
        if (id) *id = p.id;
        cv.notify_all();
        return Status::ok();
    }

    Status poll(dg_result* out, int max, int* n) override {
        std::lock_guard<std::mutex> lk(mu);
        int k = 0;
        while (k < max && !results.empty()) {
            if (out) out[k] = results.front();
            results.pop_front();
            k++;
        }
        if (n) *n = k;
        if (!fatal.empty() && k == 0) return Status::err(DG_ERR_STATE, "server stopped: " + fatal);
        return Status::ok();
    }

    Status get_state(uint64_t id, double* c, int64_t n) override {
        std::lock_guard<std::mutex> lk(mu);
        auto it = states.find(id);
        if (it == states.end()) return Status::err(DG_ERR_ARG, "no stored state for this request");
        if ((int64_t)it->second.size() != n) return Status::err(DG_ERR_ARG, "state size mismatch");
        std::copy(it->second.begin(), it->second.end(), c);
        return Status::ok();
    }

    void keep_states(bool k) override {
        std::lock_guard<std::mutex> lk(mu);
        keep = k;
    }

    Status drain() override {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return n_finished == n_submitted || !fatal.empty(); });
        if (!fatal.empty()) return Status::err(DG_ERR_STATE, "server stopped: " + fatal);
        return Status::ok();
    }

    void fail_all(const std::string& why) {
        std::lock_guard<std::mutex> lk(mu);
        fatal = why;
        cv.notify_all();
    }

    void prep_loop() {
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            if (stop) return;
            if (pending.empty() || !fatal.empty()) {
                cv.wait(lk);
                continue;
            }
            const Pending& front = pending.front();
            const double t = now();
            if (front.req.arrival_s > t) {
                cv.wait_for(lk, std::chrono::duration<double>(front.req.arrival_s - t));
                continue;
            }
            const int K = 2 * front.req.nx * front.req.ny;
            const int64_t need = estimate_bytes(K, front.req.prob.physics, front.req.p, cfg.precision);
            if (slots_held >= cfg.max_slots || bytes_held + need > cfg.hbm_budget) {
                cv.wait(lk);
                continue;
            }
            Job j;
            j.id = front.id;
            j.req = front.req;
            j.bytes = need;
            j.admit_s = t;
            pending.pop_front();
            bytes_held += need;
            slots_held += 1;
            lk.unlock();
            prepare(j);
            lk.lock();
            if (j.status != DG_OK) {
                finish_locked(j, nullptr);
            } else {
                ready.push_back(j);
            }
            cv.notify_all();
        }
    }

    void prepare(Job& j) {
        const dg_request& r = j.req;
        Mesh m;
        std::string why = make_periodic_jittered(r.nx, r.ny, r.x0, r.x1, r.y0, r.y1, r.jitter, r.mesh_seed, m);
        if (!why.empty()) {
            j.status = DG_ERR_ARG;
            return;
        }
        Status st = make_batch_solver(m, r.prob, r.p, cfg.precision, &j.s);
        if (st) st = j.s->project_ic();
        dg_audit_out a0{};
        if (st) st = j.s->audit(&a0);
        if (st) st = j.s->begin(r.t_end, true, cfg.cfl);
        if (!st) {
            delete j.s;
            j.s = nullptr;
            j.status = st.code;
            return;
        }
        for (int v = 0; v < 4; v++) j.mass0[v] = a0.mass[v];
    }

    // records the result of a job (mutex held); audit may be null on failure
    void finish_locked(Job& j, const dg_audit_out* a) {
        dg_result res{};
        res.id = j.id;
        res.status = j.status;
        res.arrival_s = j.req.arrival_s;
        res.admit_s = j.admit_s;
        res.finish_s = now();
        for (int v = 0; v < 4; v++) res.mass0[v] = j.mass0[v];
        if (a) {
            res.steps = a->steps;
            res.t = a->t;
            res.n_bad = a->n_bad;
            for (int v = 0; v < 4; v++) res.mass[v] = a->mass[v];
        }
        results.push_back(res);
        bytes_held -= j.bytes;
        slots_held -= 1;
        n_finished++;
    }

    // The solves the engine advances: every resident solve below
    // fifo_min_elems, and of the larger ones (each fills the GPU alone, so
    // running several together gains no throughput and makes the first to
    // arrive wait for the others) only the earliest arrival.  Indices into
    // `active`.
    std::vector<size_t> stepping_set() const {
        std::vector<size_t> order(active.size());
        for (size_t k = 0; k < order.size(); k++) order[k] = k;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t x, size_t y) { return active[x].req.arrival_s < active[y].req.arrival_s; });
        std::vector<size_t> out;
        bool large_taken = false;
        for (size_t k : order) {
            if (cfg.fifo_min_elems > 0 && active[k].s->K >= cfg.fifo_min_elems) {
                if (large_taken) continue;
                large_taken = true;
            }
            out.push_back(k);
        }
        return out;
    }

    void gpu_loop() {
        std::vector<Solver*> resident;
        std::vector<size_t> stepping;
        std::vector<double> h;
        bool changed = false;
        while (true) {
            {
                std::unique_lock<std::mutex> lk(mu);
                while (!ready.empty()) {
                    active.push_back(ready.front());
                    ready.pop_front();
                    changed = true;
                }
                if (active.empty()) {
                    if (stop) return;
                    cv.wait(lk);
                    continue;
                }
            }
            Status st;
            if (changed) {
                stepping = stepping_set();
                resident.clear();
                for (size_t k : stepping) resident.push_back(active[k].s);
                if (!(st = eng->set_slots(resident))) {
                    fail_all(st.msg);
                    return;
                }
                changed = false;
            }
            for (int i = 0; i < cfg.steps_per_round && st; i++) st = eng->step(cfg.use_graph != 0);
            if (st) st = eng->status(h);
            if (!st) {
                fail_all(st.msg);
                return;
            }
            std::vector<bool> done(active.size(), false);
            for (size_t r = 0; r < stepping.size(); r++) {
                const Job& j = active[stepping[r]];
                done[stepping[r]] = h[3 * r] != 0.0 || !(h[3 * r + 2] < j.req.t_end);
            }
            std::vector<Job> still;
            for (size_t k = 0; k < active.size(); k++) {
                Job& j = active[k];
                if (!done[k]) {
                    still.push_back(j);
                    continue;
                }
                dg_audit_out a{};
                Status sa = j.s->refresh();
                if (sa) sa = j.s->audit(&a);
                if (sa && a.n_bad > 0) j.status = DG_ERR_NUMERIC;
                if (!sa) j.status = sa.code;
                std::vector<double> state;
                bool keep_now;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    keep_now = keep;
                }
                if (sa && keep_now) {
                    state.resize((size_t)j.s->N * j.s->NP * j.s->K);
                    if (!j.s->get_state(state.data())) state.clear();
                }
                delete j.s;
                j.s = nullptr;
                std::lock_guard<std::mutex> lk(mu);
                if (!state.empty()) states[j.id] = std::move(state);
                finish_locked(j, sa ? &a : nullptr);
                changed = true;
            }
            if (changed) {
                active.swap(still);
                cv.notify_all();
            }
        }
    }
};

Status Server::create(const dg_server_config& cfg, Server** out) {
    if (cfg.precision != DG_FP32 && cfg.precision != DG_FP64) return Status::err(DG_ERR_ARG, "bad precision");
    if (cfg.max_slots < 1 || cfg.max_slots > 4096) return Status::err(DG_ERR_ARG, "max_slots must be in 1..4096");
    if (cfg.hbm_budget <= 0 || cfg.steps_per_round < 1 || !(cfg.cfl > 0) || cfg.solo_min_elems < 0 ||
        cfg.prep_threads < 0 || cfg.prep_threads > 64 || cfg.fifo_min_elems < 0)
        return Status::err(DG_ERR_ARG, "bad config");
    auto* s = new ServerImpl();
    s->cfg = cfg;
    Status st = s->start();
    if (!st) {
        delete s;
        return st;
    }
    *out = s;
    return Status::ok();
}

}  // namespace dgmi
