// jobs.h -- resumable requests (evaluate, gradient, L-BFGS, CMA-ES) and the tick that
// advances many of them together.
//
// A job is a state machine over evaluation steps.  One tick:
//   1. every job enqueues its pre-step kernels and describes its evaluation step(s);
//   2. all MSE steps of all jobs run as merged launches (segment tables), then all
//      gradient steps;
//   3. every job enqueues its post-step kernels (optimizer updates), which add to the
//      job's two device counters;
//   4. the counters of the jobs that decide on them this tick come back in one
//      synchronisation (L-BFGS and CMA-ES follow fixed schedules decided on the device and
//      need none: their ticks are enqueued back to back);
//   5. every job advances its state.
// The direct C-ABI calls run one job until it is done; the server runs the ticks of
// every admitted request together (continuous batching at the step level).
#pragma once
#include <stdint.h>

#include <memory>
#include <vector>

#include "../../include/rsr_amd.h"
#include "kernels.h"
#include "optim.h"
#include "runtime_internal.h"

namespace rsr {

// Device buffers owned by a job (released with it; charged to the context budget).
// Buffers come zero-filled; uploads are asynchronous from host copies the job keeps.
struct JobBuffers {
    rsr_context* ctx;
    std::vector<std::unique_ptr<DevBuf>> bufs;
    std::vector<std::vector<char>> keep;
    explicit JobBuffers(rsr_context* c) : ctx(c) {}
    // A zero-filled buffer (clear = false: its contents are written before any read).
    int raw(size_t bytes, void** out, bool clear = true);
    int upload(void* dst, const void* src, size_t bytes);
    template <typename T>
    int get(size_t count, T** out) {
        return raw(count * sizeof(T), (void**)out);
    }
    // A buffer holding a copy of v (uploaded, so not cleared first).
    template <typename T>
    int put(const std::vector<T>& v, T** out) {
        int rc = raw(v.size() * sizeof(T), (void**)out, v.empty());
        if (rc || v.empty()) return rc;
        return upload(*out, v.data(), v.size() * sizeof(T));
    }
    size_t bytes() const;
};

class Job {
   public:
    virtual ~Job() = default;
    // Enqueue pre-step kernels and append this tick's evaluation steps.
    virtual int plan(std::vector<EvalStep>& mse, std::vector<EvalStep>& grad) = 0;
    // Enqueue post-step kernels (after the merged evaluation launches).
    virtual int post() { return RSR_OK; }
    // Consume this tick's counters (host copies) and move to the next state.
    virtual int advance(const int* counters) = 0;
    // Copy results to the caller's buffers (stream is synchronised afterwards).
    virtual int finish() = 0;
    // Two-phase finish for many jobs at once: stage_bytes() host bytes of results, which
    // stage() enqueues into pinned host memory without synchronising; after the caller has
    // synchronised the stream, finish() copies them out with no synchronisation of its own.
    // A job that returns 0 finishes in one phase.
    virtual size_t stage_bytes() const { return 0; }
    virtual int stage(char*) { return RSR_OK; }
    // Device bytes the job holds (admission accounting).
    virtual size_t device_bytes() const = 0;
    // Whether this tick must read the job's counters back (a host synchronisation) for
    // advance(); when false, advance() gets null and the next tick is enqueued at once.
    virtual bool needs_counters() const { return true; }
    // The jobs a tick advances on behalf of this one (a composite job lists its parts).
    virtual void leaves(std::vector<Job*>& out) {
        if (!done) out.push_back(this);
    }
    // Called after every tick (a composite job moves between its stages here).
    virtual int refresh() { return RSR_OK; }
    bool done = false;
    // Two device ints of the job's own, zero at creation: its kernels add to them, and they
    // only ever grow; a read takes the increase since the previous read (counters_seen).
    int* d_counters = nullptr;
    uint32_t counters_seen[2] = {0u, 0u};
    uint64_t node_evals = 0;    // node-evals launched so far (statistics; fits estimate from active counts)
    int ticks = 0;
};

// compute_mse of every program (or the masked ones); mse_out host array [n].
int make_eval_job(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, const uint8_t* mask, float* mse_out,
                  std::unique_ptr<Job>* out);
// compute_mse_with_gradient; mse_out [n], grad_out [n][32].
int make_grad_job(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, const uint8_t* mask, float* mse_out,
                  float* grad_out, std::unique_ptr<Job>* out);
// run_lbfgs over the masked programs; fitness_out [n] (NaN if not fitted).
int make_lbfgs_job(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int iterations, int history, int mode,
                   const uint8_t* mask, float* fitness_out, std::unique_ptr<Job>* out);
// run_cma_es over the masked programs; fitness_out [n], iters_out [n] optional.
int make_cmaes_job(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int iterations, int lambda, int mu,
                   uint64_t seed, const float* normals, const uint8_t* mask, float* fitness_out, int32_t* iters_out,
                   std::unique_ptr<Job>* out);

// run_nelder_mead over the masked programs with Float constants; fitness_out [n].
int make_nm_job(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int iterations, const uint8_t* mask,
                float* fitness_out, int32_t* iters_out, std::unique_ptr<Job>* out);
// optimize_constants: per-program dispatch to L-BFGS / CMA-ES / Nelder-Mead, constants
// below `threshold` in magnitude set to zero, then the MSE of the result; mse_out [n],
// method_out [n] (RSR_FIT_*, optional).
int make_optimize_job(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int iterations, uint64_t seed,
                      float threshold, const uint8_t* mask, float* mse_out, uint8_t* method_out,
                      std::unique_ptr<Job>* out);

// Advance every job by one tick (caller holds ctx->mu).
int run_tick(rsr_context* ctx, const std::vector<Job*>& jobs);
// Run one job to completion and finish it (caller holds ctx->mu).
int run_job(rsr_context* ctx, Job& job);
// compute_mse through a cached captured graph (caller holds ctx->mu); RSR_E_UNSUPPORTED
// (nothing done) when the evaluation needs more than one launch round or the graph's
// buffers do not fit the budget, and the caller then runs it eagerly.
int eval_mse_graph(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, float* mse_out);
// Time every tuning candidate on this batch and dataset and cache the fastest for their
// workload shape (caller holds ctx->mu).
int autotune(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int* ppl_out, int* group_out);

}  // namespace rsr
