// server.cpp — continuous-batching planning server on one GPU.
//
// The reference plans one goal per build+solve, as separate processes (src/Makefile:7-19;
// the terminal cell is fixed per build, src/input_to_build_model.h:33-40). Here a
// long-running server keeps transition models resident in HBM, keyed by environment
// (field seed), and serves planning requests (environment, goal cell, start cell):
//
//   * a request whose model is resident is admitted into a free goal-column slot of
//     that model's live SpMM batch at the next sweep boundary — requests that arrive
//     while others are on the GPU join the running batch (continuous batching);
//   * a column leaves the batch at the first sweep whose own residual is below thresh,
//     which is where its solo solve stops; its slot is reused immediately;
//   * a request whose model is absent triggers a build on one of the builder threads (on
//     its environment's stream), admitted only if the model's estimated footprint fits the HBM budget,
//     after evicting idle models (no live columns) in least-recently-used order;
//   * every model steps on its own stream, at its own pace (a model's next step is
//     launched as soon as its previous one has finished and been harvested), so builds
//     overlap the sweeps on the device; the models whose steps run at once are the
//     max_models (default 1) that hold the oldest running requests, so the oldest request
//     is not slowed by the sweeps of models whose requests came later.
// Column arithmetic does not depend on the other columns or on when a column joined,
// so a request's result is bit-identical alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace e2e {

e2e_status model_build(const e2e_params* p, const e2e_fields* f, const e2e_build_opts* o, Model** out);
struct LiveBatch;
int64_t live_batch_bytes(const Model* m, int kp);
e2e_status live_create(Model* m, int kp, float thresh, hipStream_t s, LiveBatch** out);
void live_free(LiveBatch* L);
e2e_status live_admit(LiveBatch* L, const int32_t* cols, const int32_t* goals, int n);
e2e_status live_step(LiveBatch* L, int nsweeps, int ncols);
e2e_status live_pass(LiveBatch* L, int ncols);
e2e_status live_state(LiveBatch* L, BatchState* st, int32_t* conv, int32_t* start);
e2e_status live_harvest(LiveBatch* L, int col, int64_t start_state, float* v_start, int32_t* a_start,
                        float* vcol, uint32_t* pcol);
int live_kp(const LiveBatch* L);

namespace {

enum ReqStatus { RQ_QUEUED = 0, RQ_RUNNING = 1, RQ_DONE = 2, RQ_REJECTED = 3, RQ_FAILED = 4 };

// Models built at once (each on its environment's stream): a small model's first request
// does not wait for the build of a large one that arrived just before it
constexpr int kBuilders = 2;
enum EnvState { ENV_ABSENT = 0, ENV_BUILDING = 1, ENV_READY = 2, ENV_FAILED = 3 };

struct Req {
    uint64_t id = 0;
    int64_t env = 0;
    int32_t goal = 0, start = 0;
    double t_arrival = 0, t_admit = 0, t_done = 0;
    int status = RQ_QUEUED;
    int sweeps = 0;
    float v_start = 0.f;
    int32_t a_start = 0;
    bool cache_hit = false;
    std::vector<float> vcol;
    std::vector<uint32_t> pcol;
};

struct Env {
    int64_t id = 0;
    e2e_params p{};
    std::vector<float> u, v, ui, vi, yi, s;
    std::vector<int32_t> mask;
    hipStream_t stream = nullptr;
    int state = ENV_ABSENT;
    Model* model = nullptr;
    LiveBatch* batch = nullptr;
    std::vector<uint64_t> slot;   // request id per column slot (0 = free)
    int64_t reserved = 0;         // bytes counted against the budget
    double last_used = 0;
    int live = 0;                 // occupied slots
    std::vector<int32_t> adm_cols, adm_goals;  // admissions for the next boundary
    std::vector<int> steps_since_admit;        // per slot, sweeps launched since admission
    hipEvent_t step_done = nullptr;            // recorded after the launches of a step
    bool inflight = false;                     // a step is on the GPU (not yet harvested)
    bool may_launch = true;                    // among the models that run steps now (max_models)
};

}  // namespace

struct Server {
    e2e_server_opts o{};
    std::mutex mu;
    std::condition_variable wake, done_cv, build_cv;
    std::unordered_map<int64_t, Env*> envs;
    std::unordered_map<uint64_t, Req*> reqs;
    std::deque<uint64_t> pending;
    std::deque<int64_t> build_q;
    std::thread sched;
    std::vector<std::thread> builders;  // kBuilders threads, each building one model at a time
    std::atomic<bool> stop{false};
    uint64_t next_id = 1;
    int64_t budget = 0, resident = 0, peak = 0;
    int64_t n_submitted = 0, n_completed = 0, n_rejected = 0, n_built = 0, n_evicted = 0, n_hits = 0;
    int64_t sweeps_launched = 0;
    double t_first = -1, t_last = 0;
    std::string err;

    int kp() const { return std::max(16, (int)ceil_div(o.max_cols, 16) * 16); }

    int64_t estimate_build_bytes(const Env* e) const {
        const e2e_params& p = e->p;
        const int64_t nc = (int64_t)p.gsize * p.gsize, rows = nc * p.nt * p.num_ac_speeds * p.num_ac_angles;
        const int64_t fields = 4 * nc * p.nt * (4 + 2 * (int64_t)p.nmodes) + 4 * (int64_t)p.nt * p.nrzns * p.nmodes;
        const int64_t model = rows * 12 + rows * 12 * 12;          // row_ptr, R, Rsum + nnz at ~12/row, 8 + 4 B each
        const int64_t temps = rows * (16 * 8 + 4) + std::max<int64_t>(1 << 20, rows * 2) * 8;
        return fields + model + temps;
    }

    // bytes: amount to reserve. Evicts idle resident models (LRU) if needed. Locked.
    bool reserve(int64_t bytes, const Env* requester) {
        while (resident + bytes > budget) {
            Env* victim = nullptr;
            for (auto& kv : envs) {
                Env* e = kv.second;
                if (e == requester || e->state != ENV_READY || e->live > 0) continue;
                if (!victim || e->last_used < victim->last_used) victim = e;
            }
            if (!victim) return false;
            if (victim->batch) live_free(victim->batch);
            victim->batch = nullptr;
            free_model(victim->model);
            victim->model = nullptr;
            resident -= victim->reserved;
            victim->reserved = 0;
            victim->state = ENV_ABSENT;
            n_evicted++;
        }
        resident += bytes;
        peak = std::max(peak, resident);
        return true;
    }

    void complete(Req* r, int status) {
        r->status = status;
        r->t_done = now_ms();
        if (status == RQ_DONE) {
            n_completed++;
            t_last = std::max(t_last, r->t_done);
        }
        done_cv.notify_all();
    }

    void builder_loop() {
        for (;;) {
            Env* e;
            {
                std::unique_lock<std::mutex> lk(mu);
                build_cv.wait(lk, [&] { return stop || !build_q.empty(); });
                if (stop) return;
                e = envs[build_q.front()];
                build_q.pop_front();
            }
            e2e_params p = e->p;
            p.term_subgrid_size = 0;  // goals are per request
            e2e_fields f{e->u.data(), e->v.data(), e->ui.empty() ? nullptr : e->ui.data(),
                         e->vi.empty() ? nullptr : e->vi.data(), e->yi.data(),
                         e->s.empty() ? nullptr : e->s.data(), e->mask.empty() ? nullptr : e->mask.data()};
            e2e_build_opts bo{};
            bo.keep_rsum = 1;
            bo.stream = e->stream;
            Model* m = nullptr;
            e2e_status st = model_build(&p, &f, &bo, &m);
            std::lock_guard<std::mutex> lk(mu);
            if (st == E2E_OK && !m->has_obstacles) {
                e->model = m;
                e->state = ENV_READY;
                n_built++;
                // replace the estimate by what the model actually holds
                const int64_t actual = (int64_t)(m->rows * 12 + m->nnz * 8) +
                                       (m->vi_cz ? (int64_t)(4 * kCompWin + 4) * m->vi_cwin : 0) +
                                       live_batch_bytes_est(e);
                resident += actual - e->reserved;
                e->reserved = actual;
            } else {
                if (m) free_model(m);
                e->state = ENV_FAILED;
                resident -= e->reserved;
                e->reserved = 0;
                err = st == E2E_OK ? "environment has obstacles" : "build failed";
            }
            wake.notify_all();
        }
    }

    void sched_loop() {
        std::vector<Env*> active;
        std::vector<int32_t> conv, start;
        while (!stop) {
            active.clear();
            {
                std::unique_lock<std::mutex> lk(mu);
                // ---- admission (FIFO)
                int running = 0;
                for (auto& kv : envs) running += kv.second->live;
                for (auto it = pending.begin(); it != pending.end();) {
                    Req* r = reqs[*it];
                    Env* e = envs[r->env];
                    bool placed = false;
                    if (e->state == ENV_FAILED) {
                        complete(r, RQ_FAILED);
                        it = pending.erase(it);
                        continue;
                    }
                    if (o.serial && running > 0) break;
                    if (e->state == ENV_ABSENT) {
                        const int64_t need = estimate_build_bytes(e) + live_batch_bytes_est(e);
                        if (need > budget) {
                            complete(r, RQ_REJECTED);
                            n_rejected++;
                            it = pending.erase(it);
                            continue;
                        }
                        if (reserve(need, e)) {
                            e->reserved = need;
                            e->state = ENV_BUILDING;
                            build_q.push_back(e->id);
                            build_cv.notify_one();
                        }
                    } else if (e->state == ENV_READY) {
                        if (!e->batch) {
                            LiveBatch* L = nullptr;
                            if (live_create(e->model, kp(), o.thresh, e->stream, &L) == E2E_OK) {
                                e->batch = L;
                                e->slot.assign(live_kp(L), 0);
                                e->steps_since_admit.assign(live_kp(L), 0);
                            } else {
                                complete(r, RQ_FAILED);
                                it = pending.erase(it);
                                continue;
                            }
                        }
                        if (e->batch) {
                            const int lim = std::min<int>(o.max_cols, (int)e->slot.size());
                            for (int c = 0; c < lim; c++) {
                                if (e->slot[c] == 0) {
                                    e->slot[c] = r->id;
                                    e->live++;
                                    e->adm_cols.push_back(c);
                                    e->adm_goals.push_back(r->goal);
                                    e->steps_since_admit[c] = 0;
                                    r->status = RQ_RUNNING;
                                    r->t_admit = now_ms();
                                    e->last_used = r->t_admit;
                                    placed = true;
                                    running++;
                                    break;
                                }
                            }
                        }
                    }
                    it = placed ? pending.erase(it) : std::next(it);
                }
                for (auto& kv : envs)
                    if (kv.second->batch && kv.second->live > 0) active.push_back(kv.second);
                if (active.empty()) {
                    wake.wait_for(lk, std::chrono::milliseconds(2));
                    continue;
                }
            }
            // ---- each model at its own pace: a model whose last step has finished is
            // harvested and, if it still has running columns, admits and launches its
            // next step at once; a model still on the GPU is left alone. Sweeps per step:
            // the largest running model takes spc, a model with fewer entries
            // proportionally more (up to 16 spc), so that small models are not held to
            // the sweep rate of the largest (extra sweeps of converged columns do nothing)
            const int spc = o.sweeps_per_step > 0 ? o.sweeps_per_step : 4;
            int64_t nnz_max = 1;
            for (Env* e : active) nnz_max = std::max<int64_t>(nnz_max, e->model->nnz);
            // oldest running request first: only the max_models models holding the oldest
            // running requests launch steps (the device has no stream priorities to give
            // the oldest request its share, so the host limits what runs beside it)
            const int max_models = o.max_models == 0 ? 1 : o.max_models;
            if (max_models > 0 && (int)active.size() > max_models) {
                std::vector<std::pair<double, Env*>> age;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    for (Env* e : active) {
                        double oldest = 1e300;
                        for (uint64_t id : e->slot) {
                            const auto it = id ? reqs.find(id) : reqs.end();
                            if (it != reqs.end()) oldest = std::min(oldest, it->second->t_arrival);
                        }
                        age.emplace_back(oldest, e);
                    }
                }
                std::stable_sort(age.begin(), age.end(),
                                 [](const std::pair<double, Env*>& x, const std::pair<double, Env*>& y) {
                                     return x.first < y.first;
                                 });
                for (size_t i = 0; i < age.size(); i++) age[i].second->may_launch = (int)i < max_models;
            } else {
                for (Env* e : active) e->may_launch = true;
            }
            bool progressed = false;
            for (Env* e : active) {
                if (e->inflight) {
                    const hipError_t q = hipEventQuery(e->step_done);
                    if (q == hipErrorNotReady) continue;
                    e->inflight = false;
                    progressed = true;
                    harvest(e, conv, start);
                }
                bool running;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    running = e->live > 0;
                }
                if (!running || !e->batch || !e->may_launch) continue;
                const int64_t ratio = nnz_max / std::max<int64_t>(e->model->nnz, 1);
                const int sw = o.exact_pass ? 1 : (int)std::min<int64_t>((int64_t)spc * 16, (int64_t)spc * std::max<int64_t>(ratio, 1));
                std::vector<int32_t> cols, goals;
                int hi = 0;  // highest occupied slot + 1: the columns a sweep must cover
                {
                    std::lock_guard<std::mutex> lk(mu);
                    cols.swap(e->adm_cols);
                    goals.swap(e->adm_goals);
                    for (int c = 0; c < (int)e->slot.size(); c++) {
                        if (e->slot[c]) {
                            hi = c + 1;
                            e->steps_since_admit[c] += o.exact_pass ? spc : sw;
                        }
                    }
                }
                if (!cols.empty()) live_admit(e->batch, cols.data(), goals.data(), (int)cols.size());
                // sweeps (each column stops at its own first residual below thresh), or one
                // exact backward pass that finishes every running column
                if (o.exact_pass)
                    live_pass(e->batch, hi);
                else
                    live_step(e->batch, sw, hi);
                (void)hipEventRecord(e->step_done, e->stream);
                e->inflight = true;
                progressed = true;
                sweeps_launched += o.exact_pass ? spc : sw;
            }
            // nothing finished or started: poll again (a sleep would add a timer period,
            // tens of microseconds or more, to every short step of a small model)
            if (!progressed) std::this_thread::yield();
        }
    }

    // the converged columns of a model whose step has finished: results to their requests
    void harvest(Env* e, std::vector<int32_t>& conv, std::vector<int32_t>& start) {
        BatchState st{};
        const int k = live_kp(e->batch);
        conv.resize(k);
        start.resize(k);
        if (live_state(e->batch, &st, conv.data(), start.data()) != E2E_OK) return;
        const int max_sweeps = e->p.nt + 4;
        for (int c = 0; c < k; c++) {
            uint64_t rid;
            Req* r;
            int since;
            {
                std::lock_guard<std::mutex> lk(mu);
                rid = e->slot[c];
                if (!rid) continue;
                // a column admitted for the next step has not run yet
                if (std::find(e->adm_cols.begin(), e->adm_cols.end(), c) != e->adm_cols.end()) continue;
                r = reqs[rid];
                since = e->steps_since_admit[c];
            }
            if (conv[c] >= 0) {
                const int64_t s0 = r->start;  // state (t = 0, start cell)
                if (o.keep_full_results) {
                    r->vcol.resize(e->model->ns);
                    r->pcol.resize(e->model->ns);
                }
                live_harvest(e->batch, c, s0, &r->v_start, &r->a_start,
                             o.keep_full_results ? r->vcol.data() : nullptr,
                             o.keep_full_results ? r->pcol.data() : nullptr);
                std::lock_guard<std::mutex> lk(mu);
                r->sweeps = o.exact_pass ? 1 : conv[c] - start[c] + 1;
                e->slot[c] = 0;
                e->live--;
                e->last_used = now_ms();
                complete(r, RQ_DONE);
            } else if (since > max_sweeps + 8) {
                std::lock_guard<std::mutex> lk(mu);
                e->slot[c] = 0;  // the column keeps sweeping until the slot is reused
                e->live--;
                complete(r, RQ_FAILED);
            }
        }
    }

    int64_t live_batch_bytes_est(const Env* e) const {
        const int64_t ns = (int64_t)e->p.gsize * e->p.gsize * e->p.nt;
        return ns * kp() * 4 * 5 + 65536;
    }
};

}  // namespace e2e

using namespace e2e;

extern "C" {

e2e_status e2e_server_create(const e2e_server_opts* opts, void** out) {
    if (!out) {
        set_error("e2e_server_create: out is NULL");
        return E2E_ERR_INVALID;
    }
    *out = nullptr;
    Server* S = new Server();
    if (opts) S->o = *opts;
    // 16 columns per model batch by default: narrower batches keep each sweep short.
    // Measured (six environments of 50^2 to 150^2 cells, Poisson arrivals), p50 / p99 ms
    // at 64, 16 and 8 columns: 20/s 19 / 130, 19 / 126, 19 / 126; 60/s 79 / 208, 62 / 193,
    // 62 / 198; 100/s 224 / 372, 163 / 323, 166 / 258; the same throughput
    if (S->o.max_cols <= 0) S->o.max_cols = 16;
    if (S->o.max_cols > 64) {
        delete S;
        set_error("max_cols must be <= 64");
        return E2E_ERR_INVALID;
    }
    if (S->o.thresh <= 0) S->o.thresh = 1e-6f;
    S->budget = S->o.hbm_budget_bytes;
    if (S->budget <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete S;
            set_error("hipMemGetInfo failed (no GPU?)");
            return E2E_ERR_HIP;
        }
        // idle blocks the library keeps mapped (pool) are free memory for the server too
        S->budget = (int64_t)(0.8 * ((double)fr + (double)pool_cached_bytes()));
    }
    S->sched = std::thread([S] { S->sched_loop(); });
    for (int i = 0; i < kBuilders; i++) S->builders.emplace_back([S] { S->builder_loop(); });
    *out = S;
    return E2E_OK;
}

e2e_status e2e_server_add_env(void* srv, int64_t env_id, const e2e_params* p, const e2e_fields* f) {
    Server* S = static_cast<Server*>(srv);
    if (!S || !p || !f || !f->u || !f->v || !f->yi) {
        set_error("e2e_server_add_env: bad arguments");
        return E2E_ERR_INVALID;
    }
    try {
        Env* e = new Env();
        e->id = env_id;
        e->p = *p;
        const size_t nc = (size_t)p->gsize * p->gsize, f1 = nc * p->nt, fm = f1 * p->nmodes;
        e->u.assign(f->u, f->u + f1);
        e->v.assign(f->v, f->v + f1);
        if (p->nmodes) {
            e->ui.assign(f->ui, f->ui + fm);
            e->vi.assign(f->vi, f->vi + fm);
        }
        e->yi.assign(f->yi, f->yi + (size_t)p->nt * p->nrzns * p->nmodes);
        if (f->s) e->s.assign(f->s, f->s + f1);
        if (f->mask) e->mask.assign(f->mask, f->mask + f1);
        if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
            delete e;
            set_error("hipStreamCreate failed");
            return E2E_ERR_HIP;
        }
        if (hipEventCreateWithFlags(&e->step_done, hipEventDisableTiming) != hipSuccess) {
            (void)hipStreamDestroy(e->stream);
            delete e;
            set_error("hipEventCreate failed");
            return E2E_ERR_HIP;
        }
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->envs.count(env_id)) {
            (void)hipEventDestroy(e->step_done);
            (void)hipStreamDestroy(e->stream);
            delete e;
            set_error("environment %lld already registered", (long long)env_id);
            return E2E_ERR_INVALID;
        }
        S->envs[env_id] = e;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return E2E_ERR_NOMEM;
    }
    return E2E_OK;
}

e2e_status e2e_server_submit(void* srv, int64_t env_id, int32_t goal_cell, int32_t start_cell, uint64_t* req_id) {
    Server* S = static_cast<Server*>(srv);
    if (!S || !req_id) {
        set_error("e2e_server_submit: bad arguments");
        return E2E_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->envs.find(env_id);
    if (it == S->envs.end()) {
        set_error("unknown environment %lld", (long long)env_id);
        return E2E_ERR_INVALID;
    }
    const int nc = it->second->p.gsize * it->second->p.gsize;
    if (goal_cell < 0 || goal_cell >= nc || start_cell < 0 || start_cell >= nc) {
        set_error("goal/start cell out of range");
        return E2E_ERR_INVALID;
    }
    Req* r = new Req();
    r->id = S->next_id++;
    r->env = env_id;
    r->goal = goal_cell;
    r->start = start_cell;
    r->t_arrival = now_ms();
    r->cache_hit = it->second->state == ENV_READY;
    if (r->cache_hit) S->n_hits++;
    if (S->t_first < 0) S->t_first = r->t_arrival;
    S->reqs[r->id] = r;
    S->pending.push_back(r->id);
    S->n_submitted++;
    *req_id = r->id;
    S->wake.notify_all();
    return E2E_OK;
}

static void fill_result(const Req* r, e2e_request_result* out) {
    out->status = r->status;
    out->sweeps = r->sweeps;
    out->value_at_start = r->v_start;
    out->action_at_start = r->a_start;
    out->latency_ms = r->status >= RQ_DONE ? r->t_done - r->t_arrival : 0.0;
    out->queue_ms = r->t_admit > 0 ? r->t_admit - r->t_arrival : 0.0;
    out->cache_hit = r->cache_hit;
}

e2e_status e2e_server_poll(void* srv, uint64_t req_id, e2e_request_result* out) {
    Server* S = static_cast<Server*>(srv);
    if (!S || !out) {
        set_error("e2e_server_poll: bad arguments");
        return E2E_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(req_id);
    if (it == S->reqs.end()) {
        set_error("unknown request");
        return E2E_ERR_INVALID;
    }
    fill_result(it->second, out);
    return E2E_OK;
}

e2e_status e2e_server_wait(void* srv, uint64_t req_id, double timeout_ms, e2e_request_result* out) {
    Server* S = static_cast<Server*>(srv);
    if (!S) {
        set_error("e2e_server_wait: bad arguments");
        return E2E_ERR_INVALID;
    }
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(req_id);
    if (it == S->reqs.end()) {
        set_error("unknown request");
        return E2E_ERR_INVALID;
    }
    Req* r = it->second;
    auto pred = [&] { return r->status >= RQ_DONE; };
    if (timeout_ms > 0) {
        if (!S->done_cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), pred)) {
            if (out) fill_result(r, out);
            set_error("timeout waiting for request");
            return E2E_ERR_TIMEOUT;
        }
    } else {
        S->done_cv.wait(lk, pred);
    }
    if (out) fill_result(r, out);
    return E2E_OK;
}

e2e_status e2e_server_result_columns(void* srv, uint64_t req_id, uint32_t* policy, float* value) {
    Server* S = static_cast<Server*>(srv);
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(req_id);
    if (it == S->reqs.end() || it->second->status != RQ_DONE || it->second->vcol.empty()) {
        set_error("no full result for this request (keep_full_results=0 or not done)");
        return E2E_ERR_INVALID;
    }
    const Req* r = it->second;
    if (value) std::memcpy(value, r->vcol.data(), sizeof(float) * r->vcol.size());
    if (policy) std::memcpy(policy, r->pcol.data(), sizeof(uint32_t) * r->pcol.size());
    return E2E_OK;
}

e2e_status e2e_server_drain(void* srv, double timeout_ms) {
    Server* S = static_cast<Server*>(srv);
    std::unique_lock<std::mutex> lk(S->mu);
    auto pred = [&] {
        for (auto& kv : S->reqs)
            if (kv.second->status < RQ_DONE) return false;
        return true;
    };
    if (timeout_ms > 0) {
        if (!S->done_cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), pred)) {
            set_error("drain timed out");
            return E2E_ERR_TIMEOUT;
        }
    } else {
        S->done_cv.wait(lk, pred);
    }
    return E2E_OK;
}

e2e_status e2e_server_stats_get(void* srv, e2e_server_stats* out) {
    Server* S = static_cast<Server*>(srv);
    if (!S || !out) {
        set_error("e2e_server_stats_get: bad arguments");
        return E2E_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    std::vector<double> lat;
    for (auto& kv : S->reqs)
        if (kv.second->status == RQ_DONE) lat.push_back(kv.second->t_done - kv.second->t_arrival);
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) {
        if (lat.empty()) return 0.0;
        size_t i = (size_t)std::ceil(q * lat.size()) - 1;  // nearest rank
        return lat[std::min(i, lat.size() - 1)];
    };
    std::memset(out, 0, sizeof(*out));
    out->submitted = S->n_submitted;
    out->completed = S->n_completed;
    out->rejected = S->n_rejected;
    out->models_built = S->n_built;
    out->model_evictions = S->n_evicted;
    out->cache_hits = S->n_hits;
    out->peak_reserved_bytes = S->peak;
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    double sum = 0;
    for (double x : lat) sum += x;
    out->mean_ms = lat.empty() ? 0 : sum / lat.size();
    out->max_ms = lat.empty() ? 0 : lat.back();
    out->wall_ms = S->t_first >= 0 ? S->t_last - S->t_first : 0;
    out->solves_per_s = out->wall_ms > 0 ? 1e3 * S->n_completed / out->wall_ms : 0;
    out->sweeps_launched = S->sweeps_launched;
    return E2E_OK;
}

e2e_status e2e_server_destroy(void* srv) {
    Server* S = static_cast<Server*>(srv);
    if (!S) return E2E_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->wake.notify_all();
    S->build_cv.notify_all();
    if (S->sched.joinable()) S->sched.join();
    for (auto& t : S->builders)
        if (t.joinable()) t.join();
    for (auto& kv : S->envs) {
        Env* e = kv.second;
        if (e->batch) live_free(e->batch);
        if (e->model) free_model(e->model);
        if (e->step_done) (void)hipEventDestroy(e->step_done);
        if (e->stream) (void)hipStreamDestroy(e->stream);
        delete e;
    }
    for (auto& kv : S->reqs) delete kv.second;
    delete S;
    return E2E_OK;
}

}  // extern "C"
