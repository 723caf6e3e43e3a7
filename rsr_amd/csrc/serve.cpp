// serve.cpp -- the request loop: admission by budget, step-level continuous batching.
//
// Clients submit requests (any thread).  The server runs lanes: each lane is an internal
// context (own stream, scratch and tick counters) with its own loop thread.  A lane admits
// queued requests in arrival order as in-flight slots and device memory free up -- large
// requests to the large-request lanes, small ones to the others, never both kinds to one
// lane -- builds each admitted
// request's job (jobs.h), and advances its jobs by one tick per iteration: all their
// evaluation steps go into shared launches, then each job's optimizer update runs, then
// one synchronisation of the lane's stream.  Lanes tick independently, so their launches
// overlap on the device and a small request never waits for a large request's step.  A
// job that finishes is completed and its waiters are woken; requests that arrive
// meanwhile join a lane's next tick.
#include <math.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/rsr_amd.h"
#include "jobs.h"
#include "runtime_internal.h"

using namespace rsr;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Pending {
    uint64_t ticket;
    rsr_request req;
    std::vector<uint8_t> mask;
    double t_submit;
};

struct Active {
    uint64_t ticket;
    rsr_programs* programs;
    std::unique_ptr<Job> job;
    size_t bytes;
    double t_submit, t_admit;
};

struct Result {
    int status = RSR_E_NOT_READY;
    double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
    int ticks = 0;
    std::string message;  // the loop thread's error message for a failed request
};

// Give the calling thread the request's error message.
void report(const Result& r) {
    if (r.status != RSR_OK) (void)set_error(r.status, "%s", r.message.c_str());
}

int make_job(rsr_context* ctx, Pending& r, std::unique_ptr<Job>* out) {
    const rsr_request& q = r.req;
    const uint8_t* mask = r.mask.empty() ? nullptr : r.mask.data();
    switch (q.kind) {
        case RSR_REQ_EVAL: return make_eval_job(ctx, q.programs, q.dataset, mask, q.out, out);
        case RSR_REQ_GRAD: return make_grad_job(ctx, q.programs, q.dataset, mask, q.out, q.out2, out);
        case RSR_REQ_LBFGS:
            return make_lbfgs_job(ctx, q.programs, q.dataset, q.iterations, q.history, q.mode, mask, q.out, out);
        case RSR_REQ_CMAES:
            return make_cmaes_job(ctx, q.programs, q.dataset, q.iterations, q.lambda, q.mu, q.seed, nullptr, mask,
                                  q.out, q.iters_out, out);
        case RSR_REQ_NELDER_MEAD:
            return make_nm_job(ctx, q.programs, q.dataset, q.iterations, mask, q.out, q.iters_out, out);
        case RSR_REQ_OPTIMIZE:
            return make_optimize_job(ctx, q.programs, q.dataset, q.iterations, q.seed, 0.05f, mask, q.out, nullptr,
                                     out);
        default: return set_error(RSR_E_INVALID_ARG, "unknown request kind %d", q.kind);
    }
}

}  // namespace

// One lane: an internal context (its own stream, scratch arenas, tick counters and job
// buffers) with its own loop thread and in-flight requests.  Lanes run their ticks
// independently, so their launches overlap on the device and a small request never waits
// for a large request's step.
struct Lane {
    rsr_context* ctx = nullptr;
    bool big = false;              // takes the large requests (and small ones while none waits)
    std::vector<Active> active;    // this lane's thread only
    std::vector<uint8_t> is_big;   // per active request: estimated work above kBigWork
    std::thread th;
    uint64_t seen_version = ~0ull;  // the owner's tuning settings last copied
};

// A request is large when its evaluations are estimated above this many node-evals
// (about a millisecond of device time).
constexpr double kBigWork = 2e9;

double request_work(const rsr_request& q) {
    const rsr_programs* p = q.programs;
    double w = (double)p->code_off[p->n] * (double)q.dataset->num_samples;
    switch (q.kind) {
        case RSR_REQ_EVAL: return w;
        case RSR_REQ_GRAD: return 4.0 * w;
        default: return 16.0 * w;  // fits: many evaluations of (part of) the batch
    }
}

struct rsr_server {
    rsr_context* ctx = nullptr;  // the owner: programs, datasets and settings
    int max_inflight = 64;
    int max_queue = 1024;
    uint64_t budget = 0;

    std::mutex m;                 // guards everything below except each lane's `active`
    std::condition_variable cv;   // wakes the lanes
    std::condition_variable done_cv;
    std::deque<Pending> queue;
    std::unordered_map<uint64_t, Result> results;
    std::unordered_map<rsr_programs*, int> busy;  // batches with a request in flight
    uint64_t next_ticket = 1;
    bool stop = false;
    rsr_server_stats st{};
    int inflight = 0;
    double busy_since = 0.0;
    uint64_t inflight_bytes = 0;
    std::vector<std::unique_ptr<Lane>> lanes;

    void complete(Lane& L, size_t i, int status, double t) {
        Active& a = L.active[i];
        std::lock_guard<std::mutex> g(m);
        Result& r = results[a.ticket];
        r.status = status;
        r.t_submit = a.t_submit;
        r.t_admit = a.t_admit;
        r.t_done = t;
        r.ticks = a.job ? a.job->ticks : 0;
        if (status != RSR_OK) r.message = rsr_last_error_message();
        if (status == RSR_OK)
            st.completed++;
        else
            st.failed++;
        if (a.job) st.node_evals += a.job->node_evals;
        if (--busy[a.programs] == 0) busy.erase(a.programs);
        inflight_bytes -= a.bytes;
        st.inflight_bytes = inflight_bytes;
        if (--inflight == 0) st.busy_seconds += t - busy_since;
        st.inflight = (uint64_t)inflight;
    }

    void fail_pending(Pending& p, int status) {
        std::lock_guard<std::mutex> g(m);
        Result& r = results[p.ticket];
        r.status = status;
        r.t_submit = p.t_submit;
        r.t_admit = r.t_done = now_s();
        r.message = rsr_last_error_message();
        st.failed++;
    }

    // Whether lane L takes request q: never large and small requests in one lane.  A lane
    // advances its requests' ticks together, so a small request in a lane with a large one
    // would take a tick per large step; with the classes apart, a small request shares the
    // device with large ones but never waits for their steps.  Large requests go to the big
    // lanes, to one that holds no small request; small ones to the small lanes, and to a
    // big lane that holds no large request while no large request waits (so the big lanes
    // help with a stream of small requests, and a large one waits at most for the small
    // ones a big lane holds to finish).  A server of one lane takes everything.
    bool takes(const Lane& L, bool big_req, bool large_waiting) const {
        if (lanes.size() == 1) return true;
        bool holds_big = false, holds_small = false;
        for (uint8_t b : L.is_big) (b ? holds_big : holds_small) = true;
        if (big_req) return L.big && !holds_small;
        if (!L.big) return true;
        return !holds_big && !large_waiting;
    }

    // Admit queued requests into lane L in arrival order (skipping ones whose program batch
    // is busy or that belong to another lane).
    void admit(Lane& L) {
        const int lane_cap = std::max(1, max_inflight / (int)lanes.size());
        for (;;) {
            Pending p;
            bool big_req = false;
            {
                std::lock_guard<std::mutex> g(m);
                if (inflight >= max_inflight || (int)L.active.size() >= lane_cap) return;
                bool large_waiting = false;
                for (const Pending& q : queue) large_waiting = large_waiting || request_work(q.req) >= kBigWork;
                auto it = queue.begin();
                std::unordered_set<rsr_programs*> seen;
                for (; it != queue.end(); ++it) {
                    rsr_programs* pr = it->req.programs;
                    const bool blocked = busy.count(pr) || seen.count(pr);
                    seen.insert(pr);  // keep per-batch order: later requests on it wait too
                    if (blocked) continue;
                    const bool b = request_work(it->req) >= kBigWork;
                    if (takes(L, b, large_waiting)) {
                        big_req = b;
                        break;
                    }
                }
                if (it == queue.end()) return;
                p = std::move(*it);
                queue.erase(it);
                st.queued = queue.size();
                busy[p.req.programs]++;  // claimed while the job is built
            }
            std::unique_ptr<Job> job;
            int rc;
            {
                std::lock_guard<std::mutex> g(ctx->mu);  // lowerings are the owner's
                (void)hipSetDevice(ctx->device);
                rc = make_job(L.ctx, p, &job);
            }
            const size_t bytes = job ? job->device_bytes() : 0;
            const uint64_t limit = budget ? budget : ctx->budget;
            bool wait_for_memory = false;
            {
                std::lock_guard<std::mutex> g(m);
                if (rc == RSR_OK && inflight > 0 && inflight_bytes + bytes > limit) rc = RSR_E_BUDGET;
                if ((rc == RSR_E_BUDGET || rc == RSR_E_OOM) && inflight > 0) wait_for_memory = true;
                if (rc != RSR_OK || wait_for_memory)
                    if (--busy[p.req.programs] == 0) busy.erase(p.req.programs);
            }
            if (wait_for_memory) {
                // wait for in-flight requests to free memory; retry from the head of the queue
                job.reset();
                std::lock_guard<std::mutex> g(m);
                queue.push_front(std::move(p));
                st.queued = queue.size();
                return;
            }
            if (rc == RSR_OK && bytes > limit) {
                rc = set_error(RSR_E_BUDGET, "request needs %zu device bytes", bytes);
                std::lock_guard<std::mutex> g(m);
                if (--busy[p.req.programs] == 0) busy.erase(p.req.programs);
            }
            if (rc != RSR_OK) {
                job.reset();
                fail_pending(p, rc);
                done_cv.notify_all();
                continue;
            }
            Active a;
            a.ticket = p.ticket;
            a.programs = p.req.programs;
            a.job = std::move(job);
            a.bytes = bytes;
            a.t_submit = p.t_submit;
            a.t_admit = now_s();
            L.active.push_back(std::move(a));
            L.is_big.push_back(big_req ? 1 : 0);
            std::lock_guard<std::mutex> g(m);
            inflight_bytes += bytes;
            if (inflight++ == 0) busy_since = now_s();
            st.inflight = (uint64_t)inflight;
            st.inflight_bytes = inflight_bytes;
            if ((uint64_t)inflight > st.max_inflight_seen) st.max_inflight_seen = (uint64_t)inflight;
        }
    }

    // The owner's summation and tuning settings, copied into the lane when they changed.
    void sync_settings(Lane& L) {
        if (L.seen_version == ctx->tune_version.load()) return;  // unchanged: no lock taken
        std::lock_guard<std::mutex> g(ctx->mu);
        L.ctx->summation = ctx->summation;
        L.ctx->auto_tune = ctx->auto_tune;
        L.ctx->tuning = ctx->tuning;
        L.ctx->tune_cache = ctx->tune_cache;
        L.ctx->arena_bytes = ctx->arena_bytes;
        L.ctx->launch_mode = ctx->launch_mode;
        L.seen_version = ctx->tune_version.load();
    }

    // The results of the requests that finished this tick, downloaded together into the
    // lane's pinned staging memory with one synchronisation (their finish() then copies
    // them out without one of its own).  Requests whose jobs do not stage finish as before.
    int stage_results(Lane& L) {
        auto round_up = [](size_t b) { return (b + 255) / 256 * 256; };
        size_t need = 0;
        for (auto& a : L.active)
            if (a.job->done) need += round_up(a.job->stage_bytes());
        if (need == 0) return RSR_OK;
        rsr_context* c = L.ctx;
        // nothing is in flight into the old block: the last tick synchronised
        if (ensure_host_stage(c, need) != RSR_OK) return RSR_OK;  // every job finishes in one phase
        char* at = c->host_stage;
        for (auto& a : L.active) {
            if (!a.job->done) continue;
            const size_t b = a.job->stage_bytes();
            if (b == 0) continue;
            const int rc = a.job->stage(at);
            if (rc != RSR_OK) return rc;
            at += round_up(b);
        }
        const hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return set_error(RSR_E_HIP, "result download: %s", hipGetErrorString(e));
        return RSR_OK;
    }

    void loop(Lane& L) {
        (void)hipSetDevice(L.ctx->device);
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return stop || !queue.empty() || !L.active.empty(); });
                if (stop && queue.empty() && L.active.empty()) return;
            }
            admit(L);
            if (L.active.empty()) {
                // nothing this lane may admit (busy batches, another lane's class): sleep briefly
                std::unique_lock<std::mutex> lk(m);
                cv.wait_for(lk, std::chrono::milliseconds(1));
                continue;
            }
            sync_settings(L);
            std::vector<Job*> jobs;
            for (auto& a : L.active)
                if (!a.job->done) jobs.push_back(a.job.get());
            int rc = run_tick(L.ctx, jobs);
            if (rc == RSR_OK) rc = stage_results(L);
            if (rc != RSR_OK) (void)hipStreamSynchronize(L.ctx->stream);
            // complete finished jobs (every job fails with a failed tick)
            bool any_done = false;
            for (size_t i = 0; i < L.active.size();) {
                Active& a = L.active[i];
                if (rc == RSR_OK && !a.job->done) {
                    ++i;
                    continue;
                }
                const int status = rc == RSR_OK ? a.job->finish() : rc;
                complete(L, i, status, now_s());
                a.job.reset();
                L.active.erase(L.active.begin() + (long)i);
                L.is_big.erase(L.is_big.begin() + (long)i);
                any_done = true;
            }
            {
                std::lock_guard<std::mutex> g(m);
                st.ticks++;
            }
            if (any_done) {
                done_cv.notify_all();
                cv.notify_all();  // a freed batch or slot may let another lane admit
            }
        }
    }
};

// Lanes of a server: two for large requests, two for small ones.
constexpr int kLanes = 4;

extern "C" int rsr_server_create(rsr_context* ctx, int max_inflight, uint64_t budget_bytes, int max_queue,
                                 rsr_server** out) {
    if (!ctx || !out) return set_error(RSR_E_INVALID_ARG, "null argument");
    if (max_inflight < 1 || max_inflight > 4096 || max_queue < 1)
        return set_error(RSR_E_INVALID_ARG, "max_inflight must be 1..4096 and max_queue >= 1");
    auto s = std::make_unique<rsr_server>();
    s->ctx = ctx;
    s->max_inflight = max_inflight;
    s->max_queue = max_queue;
    s->budget = budget_bytes;
    const int n_lanes = std::min(kLanes, max_inflight);
    for (int i = 0; i < n_lanes; ++i) {
        auto L = std::make_unique<Lane>();
        int rc = rsr_context_create(ctx->device, ctx->budget, &L->ctx);
        if (rc != RSR_OK) {
            for (auto& l : s->lanes) (void)rsr_context_destroy(l->ctx);
            return rc;
        }
        L->big = n_lanes == 1 || i < n_lanes / 2;
        s->lanes.push_back(std::move(L));
    }
    rsr_server* raw = s.get();
    try {
        for (auto& L : s->lanes) {
            Lane* lp = L.get();
            L->th = std::thread([raw, lp] { raw->loop(*lp); });
        }
    } catch (...) {
        {
            std::lock_guard<std::mutex> g(s->m);
            s->stop = true;
        }
        s->cv.notify_all();
        for (auto& L : s->lanes)
            if (L->th.joinable()) L->th.join();
        for (auto& L : s->lanes) (void)rsr_context_destroy(L->ctx);
        return set_error(RSR_E_UNSUPPORTED, "could not start the server threads");
    }
    ctx->live_servers++;
    *out = s.release();
    return RSR_OK;
}

extern "C" int rsr_server_submit(rsr_server* s, const rsr_request* req, uint64_t* ticket) {
    if (!s || !req || !ticket) return set_error(RSR_E_INVALID_ARG, "null argument");
    if (!req->programs || !req->dataset) return set_error(RSR_E_INVALID_ARG, "request without programs or dataset");
    if (req->kind < RSR_REQ_EVAL || req->kind > RSR_REQ_OPTIMIZE)
        return set_error(RSR_E_INVALID_ARG, "unknown request kind %d", req->kind);
    const int n = req->programs->n;
    if (n > 0 && !req->out) return set_error(RSR_E_INVALID_ARG, "request without output buffer");
    if (req->kind == RSR_REQ_GRAD && n > 0 && !req->out2) return set_error(RSR_E_INVALID_ARG, "gradient output missing");
    if (req->programs->ctx != s->ctx || req->dataset->ctx != s->ctx)
        return set_error(RSR_E_INVALID_ARG, "programs and dataset must belong to the server's context");
    Pending p;
    p.req = *req;
    if (req->mask) p.mask.assign(req->mask, req->mask + n);
    p.req.mask = nullptr;
    p.t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(s->m);
        if (s->stop) return set_error(RSR_E_INVALID_ARG, "server is stopping");
        if ((int)s->queue.size() >= s->max_queue) return set_error(RSR_E_QUEUE_FULL, "request queue full");
        p.ticket = s->next_ticket++;
        *ticket = p.ticket;
        s->results[p.ticket] = Result{};
        s->queue.push_back(std::move(p));
        s->st.submitted++;
        s->st.queued = s->queue.size();
    }
    s->cv.notify_all();
    return RSR_OK;
}

extern "C" int rsr_server_poll(rsr_server* s, uint64_t ticket, int* status) {
    if (!s) return set_error(RSR_E_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> g(s->m);
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return set_error(RSR_E_INVALID_ARG, "unknown ticket");
    if (it->second.status == RSR_E_NOT_READY) return RSR_E_NOT_READY;
    if (status) *status = it->second.status;
    report(it->second);
    return RSR_OK;
}

extern "C" int rsr_server_wait(rsr_server* s, uint64_t ticket, double timeout_seconds, int* status) {
    if (!s) return set_error(RSR_E_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->m);
    auto ready = [&] {
        auto it = s->results.find(ticket);
        return it == s->results.end() || it->second.status != RSR_E_NOT_READY;
    };
    if (timeout_seconds < 0) {
        s->done_cv.wait(lk, ready);
    } else if (!s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_seconds), ready)) {
        return RSR_E_TIMEOUT;
    }
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return set_error(RSR_E_INVALID_ARG, "unknown ticket");
    if (status) *status = it->second.status;
    report(it->second);
    return RSR_OK;
}

extern "C" int rsr_server_timing(rsr_server* s, uint64_t ticket, double* queued_seconds, double* run_seconds,
                                 int32_t* ticks) {
    if (!s) return set_error(RSR_E_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> g(s->m);
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return set_error(RSR_E_INVALID_ARG, "unknown ticket");
    if (it->second.status == RSR_E_NOT_READY) return RSR_E_NOT_READY;
    if (queued_seconds) *queued_seconds = it->second.t_admit - it->second.t_submit;
    if (run_seconds) *run_seconds = it->second.t_done - it->second.t_admit;
    if (ticks) *ticks = it->second.ticks;
    return RSR_OK;
}

extern "C" int rsr_server_release(rsr_server* s, uint64_t ticket) {
    if (!s) return set_error(RSR_E_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> g(s->m);
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return set_error(RSR_E_INVALID_ARG, "unknown ticket");
    if (it->second.status == RSR_E_NOT_READY) return set_error(RSR_E_NOT_READY, "request still running");
    s->results.erase(it);
    return RSR_OK;
}

extern "C" int rsr_server_get_stats(rsr_server* s, rsr_server_stats* out) {
    if (!s || !out) return set_error(RSR_E_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(s->m);
    *out = s->st;
    return RSR_OK;
}

extern "C" int rsr_server_destroy(rsr_server* s) {
    if (!s) return RSR_OK;
    {
        std::lock_guard<std::mutex> g(s->m);
        s->stop = true;
    }
    s->cv.notify_all();
    for (auto& L : s->lanes)
        if (L->th.joinable()) L->th.join();
    for (auto& L : s->lanes) (void)rsr_context_destroy(L->ctx);
    s->ctx->live_servers--;
    delete s;
    return RSR_OK;
}
