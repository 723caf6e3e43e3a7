// serve.cpp -- the request loop: admission by budget, step-level continuous batching.
//
// One device thread owns the loop.  Clients submit requests (any thread); the loop admits
// them in arrival order as in-flight slots and device memory free up, builds each
// admitted request's job (jobs.h), and advances every job by one tick per iteration:
// all jobs' evaluation steps go into shared launches, then each job's optimizer update
// runs, then one synchronisation.  A job that finishes is completed and its waiters are
// woken; requests that arrive meanwhile join the next tick.
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

struct rsr_server {
    rsr_context* ctx = nullptr;
    int max_inflight = 64;
    int max_queue = 1024;
    uint64_t budget = 0;

    std::mutex m;                 // guards everything below except `active` (loop thread only)
    std::condition_variable cv;   // wakes the loop
    std::condition_variable done_cv;
    std::deque<Pending> queue;
    std::unordered_map<uint64_t, Result> results;
    uint64_t next_ticket = 1;
    bool stop = false;
    rsr_server_stats st{};

    std::vector<Active> active;  // loop thread only
    uint64_t inflight_bytes = 0;
    std::thread th;

    void complete(Active& a, int status, double t) {
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
        st.inflight = active.size() > 0 ? active.size() - 1 : 0;
        inflight_bytes -= a.bytes;
        st.inflight_bytes = inflight_bytes;
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

    // Admit queued requests in arrival order (skipping ones whose program batch is busy).
    void admit() {
        for (;;) {
            Pending p;
            {
                std::lock_guard<std::mutex> g(m);
                if ((int)active.size() >= max_inflight) return;
                std::unordered_set<rsr_programs*> busy;
                for (const auto& a : active) busy.insert(a.programs);
                auto it = queue.begin();
                std::unordered_set<rsr_programs*> seen;
                for (; it != queue.end(); ++it) {
                    rsr_programs* pr = it->req.programs;
                    if (!busy.count(pr) && !seen.count(pr)) break;
                    seen.insert(pr);  // keep per-batch order: later requests on it wait too
                }
                if (it == queue.end()) return;
                p = std::move(*it);
                queue.erase(it);
                st.queued = queue.size();
            }
            std::unique_ptr<Job> job;
            int rc;
            {
                std::lock_guard<std::mutex> g(ctx->mu);
                (void)hipSetDevice(ctx->device);
                rc = make_job(ctx, p, &job);
            }
            const size_t bytes = job ? job->device_bytes() : 0;
            const uint64_t limit = budget ? budget : ctx->budget;
            if (rc == RSR_OK && !active.empty() && inflight_bytes + bytes > limit) rc = RSR_E_BUDGET;
            if ((rc == RSR_E_BUDGET || rc == RSR_E_OOM) && !active.empty()) {
                // wait for in-flight requests to free memory; retry from the head of the queue
                {
                    std::lock_guard<std::mutex> g(ctx->mu);
                    job.reset();
                }
                std::lock_guard<std::mutex> g(m);
                queue.push_front(std::move(p));
                st.queued = queue.size();
                return;
            }
            if (rc == RSR_OK && bytes > limit) rc = set_error(RSR_E_BUDGET, "request needs %zu device bytes", bytes);
            if (rc != RSR_OK) {
                {
                    std::lock_guard<std::mutex> g(ctx->mu);
                    job.reset();
                }
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
            inflight_bytes += bytes;
            active.push_back(std::move(a));
            std::lock_guard<std::mutex> g(m);
            st.inflight = active.size();
            st.inflight_bytes = inflight_bytes;
            if (active.size() > st.max_inflight_seen) st.max_inflight_seen = active.size();
        }
    }

    void loop() {
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return stop || !queue.empty() || !active.empty(); });
                if (stop && queue.empty() && active.empty()) return;
            }
            admit();
            if (active.empty()) {
                // nothing admissible (every queued request waits on a busy batch): sleep briefly
                std::unique_lock<std::mutex> lk(m);
                cv.wait_for(lk, std::chrono::milliseconds(1));
                continue;
            }
            const double t0 = now_s();
            int rc = RSR_OK;
            {
                std::lock_guard<std::mutex> g(ctx->mu);
                (void)hipSetDevice(ctx->device);
                std::vector<Job*> jobs;
                for (auto& a : active)
                    if (!a.job->done) jobs.push_back(a.job.get());
                rc = run_tick(ctx, jobs);
                if (rc != RSR_OK) (void)hipStreamSynchronize(ctx->stream);
                // complete finished jobs (every job fails with a failed tick)
                for (size_t i = 0; i < active.size();) {
                    Active& a = active[i];
                    if (rc == RSR_OK && !a.job->done) {
                        ++i;
                        continue;
                    }
                    const int status = rc == RSR_OK ? a.job->finish() : rc;
                    complete(a, status, now_s());
                    a.job.reset();
                    active.erase(active.begin() + (long)i);
                }
            }
            {
                std::lock_guard<std::mutex> g(m);
                st.ticks++;
                st.busy_seconds += now_s() - t0;
                st.inflight = active.size();
            }
            done_cv.notify_all();
        }
    }
};

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
    rsr_server* raw = s.get();
    try {
        s->th = std::thread([raw] { raw->loop(); });
    } catch (...) {
        return set_error(RSR_E_UNSUPPORTED, "could not start the server thread");
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
    s->cv.notify_one();
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
    if (s->th.joinable()) s->th.join();
    s->ctx->live_servers--;
    delete s;
    return RSR_OK;
}
