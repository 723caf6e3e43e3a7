// jobs.cpp -- evaluation, gradient, L-BFGS and CMA-ES requests as resumable jobs, and
// the tick that merges the evaluation steps of many jobs into shared launches.
#include "jobs.h"

#include <math.h>
#include <float.h>
#include <string.h>

#include <algorithm>
#include <numeric>

#include "../../include/rsr_amd.h"

#define HIP_TRY(expr, what)                               \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_error(e_, what); \
    } while (0)

namespace rsr {

int JobBuffers::raw(size_t bytes, void** out, bool clear) {
    const size_t n = std::max<size_t>(bytes, 16);
    bufs.emplace_back(new DevBuf());
    int rc = bufs.back()->alloc(ctx, n);
    if (rc) return rc;
    if (clear) HIP_TRY(hipMemsetAsync(bufs.back()->ptr, 0, n, ctx->stream), "job buffer clear");
    *out = bufs.back()->ptr;
    return RSR_OK;
}

int JobBuffers::upload(void* dst, const void* src, size_t bytes) {
    keep.emplace_back((const char*)src, (const char*)src + bytes);
    HIP_TRY(hipMemcpyAsync(dst, keep.back().data(), bytes, hipMemcpyHostToDevice, ctx->stream), "job upload");
    return RSR_OK;
}

size_t JobBuffers::bytes() const {
    size_t s = 0;
    for (const auto& b : bufs) s += b->bytes;
    return s;
}

namespace {

uint64_t prog_nodes(const rsr_programs* p, int i) { return (uint64_t)(p->code_off[i + 1] - p->code_off[i]); }

// Copy the device parameters back into the host constants (unflatten, program.rs:147-192),
// on the stream of the context that ran the fit (a server lane's, or the batch's own).
int download_params(rsr_context* ctx, rsr_programs* p) {
    HIP_TRY(hipMemcpyAsync(p->params.data(), p->dparams.ptr, p->params.size() * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream),
            "params download");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "params download");
    for (int i = 0; i < p->n; ++i) {
        int64_t v = p->value_off[i];
        uint32_t f = p->param_off[i];
        for (int64_t c = p->const_off[i]; c < p->const_off[i + 1]; ++c) {
            const int s = vt_size(p->ctypes[c]);
            for (int k = 0; k < s; ++k) p->values[v + k] = p->params[f + k];
            v += s ? s : 1;
            f += s;
        }
    }
    return RSR_OK;
}

// The lowering's cost-sorted scheduling order restricted to the programs with on[i] != 0,
// uploaded into the job's buffers: steps over a subset launch work for the subset only.
// The big programs stay last; *n_big counts them.
int subset_order(JobBuffers& jb, const Lowering* L, const std::vector<uint8_t>& on, int32_t** d_order, int* count,
                 int* n_big = nullptr) {
    std::vector<int32_t> o;
    int nb = 0;
    for (int32_t i : L->order)
        if (on[i]) {
            o.push_back(i);
            nb += L->big[i];
        }
    *count = (int)o.size();
    if (n_big) *n_big = nb;
    return jb.put(o, d_order);
}

// A step's view of a batch over a subset order.
void set_order(EvalStep& st, const int32_t* order, int count, int n_big) {
    st.P.order = order;
    st.P.n_progs = count;
    st.P.n_big = n_big;
}

EvalStep mse_step(const rsr_programs* p, const Lowering* L, const rsr_dataset* ds, const uint8_t* active,
                  float* d_mse) {
    EvalStep st;
    st.P = dev_view(p, L);
    st.D = dev_view(ds);
    st.active = active;
    st.mse_out = d_mse;
    return st;
}

EvalStep grad_step(const rsr_programs* p, const Lowering* Ld, const rsr_dataset* ds, const uint8_t* active,
                   float* d_mse, float* d_grad) {
    EvalStep st = mse_step(p, Ld, ds, active, d_mse);
    st.grad_out = d_grad;
    st.nparams = Ld->dnparams;
    st.max_params = Ld->max_params;
    return st;
}

// ---------------------------------------------------------------------------------
// compute_mse (evaluator.rs:153-174)
// ---------------------------------------------------------------------------------
class EvalJob final : public Job {
   public:
    explicit EvalJob(rsr_context* c) : ctx(c), jb(c) {}
    rsr_context* ctx;
    rsr_programs* p = nullptr;
    const rsr_dataset* ds = nullptr;
    Lowering* L = nullptr;
    JobBuffers jb;
    float* d_mse = nullptr;
    uint8_t* d_act = nullptr;
    int32_t* d_order = nullptr;  // subset order when not every program is evaluated
    int n_order = 0, n_order_big = 0;
    float* out = nullptr;
    std::vector<uint8_t> on;
    int n_bad = 0;
    uint64_t nodes = 0;
    const float* staged = nullptr;  // results in pinned host memory (two-phase finish)

    int plan(std::vector<EvalStep>& mse, std::vector<EvalStep>&) override {
        EvalStep st = mse_step(p, L, ds, d_act, d_mse);
        if (d_order) set_order(st, d_order, n_order, n_order_big);
        mse.push_back(st);
        node_evals += nodes * (uint64_t)ds->num_samples;
        return RSR_OK;
    }
    int advance(const int*) override {
        done = true;
        return RSR_OK;
    }
    bool needs_counters() const override { return false; }
    size_t stage_bytes() const override { return (size_t)p->n * sizeof(float); }
    int stage(char* host) override {
        HIP_TRY(hipMemcpyAsync(host, d_mse, (size_t)p->n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
                "mse download");
        staged = (const float*)host;
        return RSR_OK;
    }
    int finish() override {
        if (staged) {
            memcpy(out, staged, (size_t)p->n * sizeof(float));
        } else {
            HIP_TRY(hipMemcpyAsync(out, d_mse, (size_t)p->n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
                    "mse download");
            HIP_TRY(hipStreamSynchronize(ctx->stream), "mse download");
        }
        // over a dataset without points no tile runs: the reference's 0 / 0 is not finite,
        // which it reports as f32::MAX (evaluator.rs:173)
        for (int i = 0; i < p->n; ++i)
            out[i] = !on[i] ? NAN : ds->num_samples == 0 ? FLT_MAX : out[i];
        if (n_bad) return set_error(RSR_E_PROGRAM, "%d program(s) rejected by lowering", n_bad);
        return RSR_OK;
    }
    size_t device_bytes() const override { return jb.bytes(); }
};

// ---------------------------------------------------------------------------------
// compute_mse_with_gradient (evaluator.rs:177-224)
// ---------------------------------------------------------------------------------
class GradJob final : public Job {
   public:
    explicit GradJob(rsr_context* c) : ctx(c), jb(c) {}
    rsr_context* ctx;
    rsr_programs* p = nullptr;
    const rsr_dataset* ds = nullptr;
    Lowering *Lp = nullptr, *Ld = nullptr;
    JobBuffers jb;
    float *d_mse = nullptr, *d_grad = nullptr;
    uint8_t *d_act0 = nullptr, *d_act1 = nullptr;
    bool any_zero = false, any_pos = false;
    int32_t *d_order0 = nullptr, *d_order1 = nullptr;
    int n_order0 = 0, n_order1 = 0, n_big0 = 0, n_big1 = 0;
    float *mse_out = nullptr, *grad_out = nullptr;
    std::vector<uint8_t> on;
    int n_bad = 0;
    uint64_t nodes = 0;

    int plan(std::vector<EvalStep>& mse, std::vector<EvalStep>& grad) override {
        // zero-parameter programs are scored by compute_mse on the primal lowering (evaluator.rs:187-189)
        if (any_zero) {
            EvalStep st = mse_step(p, Lp, ds, d_act0, d_mse);
            set_order(st, d_order0, n_order0, n_big0);
            mse.push_back(st);
        }
        if (any_pos) {
            EvalStep st = grad_step(p, Ld, ds, d_act1, d_mse, d_grad);
            set_order(st, d_order1, n_order1, n_big1);
            grad.push_back(st);
        }
        node_evals += nodes * (uint64_t)ds->num_samples;
        return RSR_OK;
    }
    int advance(const int*) override {
        done = true;
        return RSR_OK;
    }
    bool needs_counters() const override { return false; }
    size_t stage_bytes() const override { return (size_t)p->n * 33 * sizeof(float); }
    int stage(char* host) override {  // the MSEs, then the gradients
        HIP_TRY(hipMemcpyAsync(host, d_mse, (size_t)p->n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
                "mse download");
        HIP_TRY(hipMemcpyAsync(host + (size_t)p->n * sizeof(float), d_grad, (size_t)p->n * 32 * sizeof(float),
                               hipMemcpyDeviceToHost, ctx->stream),
                "grad download");
        staged = (const float*)host;
        return RSR_OK;
    }
    const float* staged = nullptr;  // results in pinned host memory (two-phase finish)
    int finish() override {
        if (staged) {
            memcpy(mse_out, staged, (size_t)p->n * sizeof(float));
            memcpy(grad_out, staged + p->n, (size_t)p->n * 32 * sizeof(float));
        } else {
            HIP_TRY(hipMemcpyAsync(mse_out, d_mse, (size_t)p->n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
                    "mse download");
            HIP_TRY(hipMemcpyAsync(grad_out, d_grad, (size_t)p->n * 32 * sizeof(float), hipMemcpyDeviceToHost,
                                   ctx->stream),
                    "grad download");
            HIP_TRY(hipStreamSynchronize(ctx->stream), "grad download");
        }
        for (int i = 0; i < p->n; ++i) {
            if (!on[i]) {
                mse_out[i] = NAN;
            } else if (ds->num_samples == 0) {  // (f32::MAX, 0) for a non-finite MSE (evaluator.rs:214-216)
                mse_out[i] = FLT_MAX;
                for (int k = 0; k < 32; ++k) grad_out[(size_t)i * 32 + k] = 0.0f;
            }
        }
        if (n_bad) return set_error(RSR_E_PROGRAM, "%d program(s) rejected by lowering", n_bad);
        return RSR_OK;
    }
    size_t device_bytes() const override { return jb.bytes(); }
};

// ---------------------------------------------------------------------------------
// run_lbfgs (src/engine/optimize/lbfgs.rs:9-165), one state machine for the batch:
//   INIT   : gradient of every individual; begin; first direction (counter = working)
//   SEARCH : MSE of the searching individuals; Armijo check (counter = still searching);
//            at most 12 rounds per iteration
//   GRAD   : gradient of the moved individuals; curvature update; next direction
// ---------------------------------------------------------------------------------
class LbfgsJob final : public Job {
   public:
    enum Stage { INIT, SEARCH, GRAD };
    explicit LbfgsJob(rsr_context* c) : ctx(c), jb(c) {}
    rsr_context* ctx;
    rsr_programs* p = nullptr;
    const rsr_dataset* ds = nullptr;
    Lowering *Lp = nullptr, *Ld = nullptr;
    JobBuffers jb;
    LbfgsDev st;
    std::vector<int> prog;
    uint8_t *act_fit = nullptr, *act_ls = nullptr, *act_grad = nullptr;
    float *d_mse = nullptr, *d_grad = nullptr, *d_fit = nullptr;
    int32_t *order_p = nullptr, *order_d = nullptr;  // scheduling orders over the fitted programs
    int big_p = 0, big_d = 0;
    DevPrograms CP{};           // one-round line search: the candidate batch (st.K per individual)
    float* d_cmse = nullptr;
    float* fitness_out = nullptr;
    int iters = 0, it = 0, ls_round = 0;
    Stage stage = INIT;
    uint64_t mean_nodes = 0;

    int plan(std::vector<EvalStep>& mse, std::vector<EvalStep>& grad) override {
        if (stage == SEARCH && st.K) {
            HIP_TRY(lbfgs_ls_fill(st, ctx->stream), "lbfgs_ls_fill");
            EvalStep s;
            s.P = CP;
            s.D = dev_view(ds);
            s.active = st.cand_active;
            s.mse_out = d_cmse;
            mse.push_back(s);
        } else if (stage == SEARCH) {
            EvalStep st = mse_step(p, Lp, ds, act_ls, d_mse);
            set_order(st, order_p, (int)prog.size(), big_p);
            mse.push_back(st);
        } else {
            EvalStep st = grad_step(p, Ld, ds, stage == INIT ? act_fit : act_grad, d_mse, d_grad);
            set_order(st, order_d, (int)prog.size(), big_d);
            grad.push_back(st);
        }
        return RSR_OK;
    }
    int post() override {
        float* params = (float*)p->dparams.ptr;
        switch (stage) {
            case INIT:
                HIP_TRY(lbfgs_begin(st, params, d_mse, d_grad, ctx->stream), "lbfgs_begin");
                if (iters > 0)
                    HIP_TRY(lbfgs_direction(st, params, act_ls, d_counters + 1, ctx->stream), "lbfgs_direction");
                break;
            case SEARCH:
                if (st.K)
                    HIP_TRY(lbfgs_ls_pick(st, params, d_cmse, act_ls, act_grad, d_counters, ctx->stream), "ls_pick");
                else
                    HIP_TRY(lbfgs_ls_check(st, params, d_mse, act_ls, act_grad, d_counters, ctx->stream), "ls_check");
                break;
            case GRAD:
                HIP_TRY(lbfgs_update(st, d_grad, act_grad, ctx->stream), "lbfgs_update");
                if (it + 1 < iters)
                    HIP_TRY(lbfgs_direction(st, params, act_ls, d_counters + 1, ctx->stream), "lbfgs_direction");
                break;
        }
        return RSR_OK;
    }
    // The schedule is decided on the device: every iteration runs at most 12 line-search
    // rounds (the reference's Armijo cap, lbfgs.rs:113-125) and then the gradient of the
    // moved individuals.  An individual that accepted earlier, or converged (|g| < 1e-5), is
    // inactive in the remaining rounds and iterations -- its kernels skip it.  With
    // `adaptive`, each tick's counters (individuals still searching; individuals with a new
    // direction) come back to the host, which ends the line search once no individual
    // searches and the fit once none works; without it every round and iteration is enqueued
    // back to back with no host synchronisation.  The results are the same bits either way.
    bool adaptive = false;
    bool needs_counters() const override { return adaptive; }
    int advance(const int* c) override {
        switch (stage) {
            case INIT:
                if (iters == 0 || (c && c[1] == 0)) {
                    done = true;
                } else {
                    stage = SEARCH;
                    ls_round = 0;
                }
                break;
            case SEARCH:
                if (++ls_round == 12 || st.K || (c && c[0] == 0)) stage = GRAD;
                break;
            case GRAD:
                if (++it >= iters || (c && c[1] == 0)) {
                    done = true;
                } else {
                    stage = SEARCH;
                    ls_round = 0;
                }
                break;
        }
        return RSR_OK;
    }
    int finish() override {
        if (prog.empty()) return RSR_OK;
        HIP_TRY(lbfgs_finish(st, (float*)p->dparams.ptr, d_fit, ctx->stream), "lbfgs_finish");
        std::vector<float> fit(prog.size());
        int cnt[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(fit.data(), d_fit, fit.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
                "fitness");
        HIP_TRY(hipMemcpyAsync(cnt, d_counters, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream), "counters");
        int rc = download_params(ctx, p);  // synchronises the stream
        if (rc) return rc;
        for (size_t k = 0; k < prog.size(); ++k) fitness_out[prog[k]] = fit[k];
        // evaluations run: gradients of all at the start and of every working individual per
        // iteration (counter 1), line-search MSEs of the working individuals (counter 1) and of
        // those still searching after a round (counter 0)
        const uint64_t evals = prog.size() + 2 * (uint64_t)(uint32_t)cnt[1] + (uint64_t)(uint32_t)cnt[0];
        node_evals += evals * mean_nodes * (uint64_t)ds->num_samples;
        return RSR_OK;
    }
    size_t device_bytes() const override { return jb.bytes(); }
};

// ---------------------------------------------------------------------------------
// run_cma_es (src/engine/optimize/cmaes.rs:19-192), one state machine for the batch:
//   INIT : MSE of every parent; best = doubled mean
//   RUN  : sample lambda candidates per individual; MSE of every live candidate (one
//          candidate batch: each program repeated lambda times with its own parameter
//          block); selection and update (counter = individuals still working)
// ---------------------------------------------------------------------------------
class CmaesJob final : public Job {
   public:
    enum Stage { INIT, RUN };
    explicit CmaesJob(rsr_context* c) : ctx(c), jb(c) {}
    rsr_context* ctx;
    rsr_programs* p = nullptr;
    const rsr_dataset* ds = nullptr;
    Lowering* Lp = nullptr;
    JobBuffers jb;
    CmaDev st;
    std::vector<int> prog;
    uint8_t* act_fit = nullptr;
    int32_t* order_fit = nullptr;
    int big_fit = 0;
    float *d_mse0 = nullptr, *d_cmse = nullptr, *d_cparams = nullptr, *d_fit = nullptr;
    int* d_iters = nullptr;
    DevPrograms CP{};
    float* fitness_out = nullptr;
    int32_t* iters_out = nullptr;
    int iters = 0, it = 0;
    Stage stage = INIT;
    uint64_t mean_nodes = 0;

    int plan(std::vector<EvalStep>& mse, std::vector<EvalStep>&) override {
        if (stage == INIT) {
            EvalStep st = mse_step(p, Lp, ds, act_fit, d_mse0);
            set_order(st, order_fit, (int)prog.size(), big_fit);
            mse.push_back(st);
            return RSR_OK;
        }
        HIP_TRY(cmaes_sample(st, d_cparams, it, ctx->stream), "cmaes_sample");
        EvalStep s;
        s.P = CP;
        s.D = dev_view(ds);
        s.active = st.cand_active;
        s.mse_out = d_cmse;
        mse.push_back(s);
        return RSR_OK;
    }
    int post() override {
        if (stage == INIT)
            HIP_TRY(cmaes_begin(st, d_mse0, ctx->stream), "cmaes_begin");
        else
            HIP_TRY(cmaes_update(st, d_cmse, d_counters, ctx->stream), "cmaes_update");
        return RSR_OK;
    }
    // A fixed schedule: `iters` generations; an individual that converged (best MSE below
    // 1e-8, cmaes.rs:140-147) has no active candidates and its update is skipped, so the
    // results are those of stopping early, without a host synchronisation per generation.
    bool needs_counters() const override { return false; }
    int advance(const int*) override {
        if (stage == INIT) {
            stage = RUN;
            if (iters == 0) done = true;
            return RSR_OK;
        }
        if (++it >= iters) done = true;
        return RSR_OK;
    }
    int finish() override {
        if (prog.empty()) return RSR_OK;
        HIP_TRY(cmaes_finish(st, (float*)p->dparams.ptr, d_fit, d_iters, ctx->stream), "cmaes_finish");
        std::vector<float> fit(prog.size());
        std::vector<int> its(prog.size());
        int cnt[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(fit.data(), d_fit, fit.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
                "fit");
        HIP_TRY(hipMemcpyAsync(its.data(), d_iters, its.size() * sizeof(int), hipMemcpyDeviceToHost, ctx->stream),
                "iters");
        HIP_TRY(hipMemcpyAsync(cnt, d_counters, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream), "counters");
        int rc = download_params(ctx, p);  // synchronises the stream
        if (rc) return rc;
        for (size_t k = 0; k < prog.size(); ++k) {
            fitness_out[prog[k]] = fit[k];
            if (iters_out) iters_out[prog[k]] = its[k];
        }
        // evaluations run: the parents, then lambda candidates of every individual still
        // working in each generation (counter 0: the working count after each update)
        const uint64_t n = prog.size();
        const uint64_t gens = std::min<uint64_t>(n * (uint64_t)iters, n + (uint64_t)(uint32_t)cnt[0]);
        node_evals += (n + gens * (uint64_t)st.lambda) * mean_nodes * (uint64_t)ds->num_samples;
        return RSR_OK;
    }
    size_t device_bytes() const override { return jb.bytes(); }
};

// ---------------------------------------------------------------------------------
// run_nelder_mead (src/engine/optimize/nelder_mead.rs:9-165), one state machine for the
// batch: every tick evaluates the points each individual requested (its reflected,
// expanded or contracted point, or its shrunk simplex) through 33 candidate slots per
// individual; the step kernel consumes them and requests the next.
// ---------------------------------------------------------------------------------
class NmJob final : public Job {
   public:
    explicit NmJob(rsr_context* c) : ctx(c), jb(c) {}
    rsr_context* ctx;
    rsr_programs* p = nullptr;
    const rsr_dataset* ds = nullptr;
    Lowering* Lp = nullptr;
    JobBuffers jb;
    NmDev st;
    std::vector<int> prog;
    float *d_cmse = nullptr, *d_cparams = nullptr, *d_fit = nullptr;
    int* d_iters = nullptr;
    DevPrograms CP{};
    float* fitness_out = nullptr;
    int32_t* iters_out = nullptr;
    bool started = false;
    uint64_t mean_nodes = 0, mean_dim = 0;

    int plan(std::vector<EvalStep>& mse, std::vector<EvalStep>&) override {
        if (!started) HIP_TRY(nm_begin(st, (const float*)p->dparams.ptr, d_cparams, ctx->stream), "nm_begin");
        EvalStep s;
        s.P = CP;
        s.D = dev_view(ds);
        s.active = st.cand_active;
        s.mse_out = d_cmse;
        mse.push_back(s);
        return RSR_OK;
    }
    int post() override {
        HIP_TRY(nm_step(st, d_cparams, d_cmse, d_counters, ctx->stream), "nm_step");
        return RSR_OK;
    }
    // The counter (individuals that stepped) is read every eighth tick: when none stepped in
    // the ticks since the last read, every individual is done (done is absorbing, and a done
    // individual's step is skipped, so the extra ticks change nothing).
    bool needs_counters() const override { return ticks % 8 == 7; }
    int advance(const int* c) override {
        started = true;
        if (c && c[0] == 0) done = true;
        return RSR_OK;
    }
    int finish() override {
        if (prog.empty()) return RSR_OK;
        HIP_TRY(nm_finish(st, (float*)p->dparams.ptr, d_fit, d_iters, ctx->stream), "nm_finish");
        std::vector<float> fit(prog.size());
        std::vector<int> its(prog.size());
        int cnt[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(fit.data(), d_fit, fit.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
                "fit");
        HIP_TRY(hipMemcpyAsync(its.data(), d_iters, its.size() * sizeof(int), hipMemcpyDeviceToHost, ctx->stream),
                "iters");
        HIP_TRY(hipMemcpyAsync(cnt, d_counters, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream), "counters");
        int rc = download_params(ctx, p);  // synchronises the stream
        if (rc) return rc;
        for (size_t k = 0; k < prog.size(); ++k) {
            fitness_out[prog[k]] = fit[k];
            if (iters_out) iters_out[prog[k]] = its[k];
        }
        // evaluations run: the initial simplex, then one step per working individual per tick
        const uint64_t evals = prog.size() * (mean_dim + 1) + (uint64_t)(uint32_t)cnt[0];
        node_evals += evals * mean_nodes * (uint64_t)ds->num_samples;
        return RSR_OK;
    }
    size_t device_bytes() const override { return jb.bytes(); }
};

// ---------------------------------------------------------------------------------
// optimize_constants (src/engine/search/individual.rs:65-75 with
// src/engine/optimize/mod.rs:10-59): per program, L-BFGS (15 iterations) when it has no
// matrix inverse, else CMA-ES (max(iterations / 16, 2)) when it has a vector or matrix
// constant, else Nelder-Mead (iterations); the three fits run as parts of one job (their
// steps share ticks).  Then every constant component below the threshold in magnitude
// is set to zero and the MSE of the resulting programs is evaluated (the caller's
// calculate_mse after optimize_constants, island.rs:233-236).
// ---------------------------------------------------------------------------------
class OptimizeJob final : public Job {
   public:
    explicit OptimizeJob(rsr_context* c) : ctx(c) {}
    rsr_context* ctx;
    rsr_programs* p = nullptr;
    const rsr_dataset* ds = nullptr;
    std::vector<std::unique_ptr<Job>> fits;
    std::unique_ptr<Job> eval;
    std::vector<float> fit_tmp, mse_tmp;
    std::vector<int32_t> iters_tmp;
    std::vector<uint8_t> chosen;
    float threshold = 0.05f;
    float* mse_out = nullptr;
    int first_rc = RSR_OK;

    int plan(std::vector<EvalStep>&, std::vector<EvalStep>&) override { return RSR_OK; }
    int advance(const int*) override { return RSR_OK; }
    void leaves(std::vector<Job*>& out) override {
        if (eval) {
            eval->leaves(out);
            return;
        }
        for (auto& f : fits) f->leaves(out);
    }
    int refresh() override {
        if (eval) {
            done = eval->done;
            return RSR_OK;
        }
        for (auto& f : fits)
            if (!f->done) return RSR_OK;
        for (auto& f : fits) {
            const int rc = f->finish();
            if (rc && !first_rc) first_rc = rc;
        }
        // apply_threshold (scalar.rs:60-100) on every constant component of the fitted programs
        bool changed = false;
        for (int i = 0; i < p->n; ++i) {
            if (!chosen[i]) continue;
            for (uint32_t f = p->param_off[i]; f < p->param_off[i + 1]; ++f) {
                if (__builtin_fabsf(p->params[f]) < threshold && p->params[f] != 0.0f) {
                    p->params[f] = 0.0f;
                    changed = true;
                }
            }
            int64_t v = p->value_off[i];
            uint32_t f = p->param_off[i];
            for (int64_t c = p->const_off[i]; c < p->const_off[i + 1]; ++c) {
                const int s = vt_size(p->ctypes[c]);
                for (int k = 0; k < s; ++k) p->values[v + k] = p->params[f + k];
                v += s ? s : 1;
                f += s;
            }
        }
        int rc = changed ? upload_params(p, ctx) : RSR_OK;
        if (rc) return rc;
        rc = make_eval_job(ctx, p, ds, chosen.data(), mse_tmp.data(), &eval);
        if (rc) return rc;
        return RSR_OK;
    }
    int finish() override {
        const int rc = eval ? eval->finish() : RSR_OK;
        for (int i = 0; i < p->n; ++i) mse_out[i] = chosen[i] ? mse_tmp[i] : NAN;
        if (first_rc) return first_rc;
        return rc == RSR_E_PROGRAM ? RSR_OK : rc;  // rejected programs are reported per program (NaN)
    }
    size_t device_bytes() const override {
        size_t b = 0;
        for (const auto& f : fits) b += f->device_bytes();
        return b + (eval ? eval->device_bytes() : 0);
    }
};

}  // namespace

// ---------------------------------------------------------------------------------
// job construction
// ---------------------------------------------------------------------------------
int make_eval_job(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, const uint8_t* mask, float* mse_out,
                  std::unique_ptr<Job>* out) {
    auto j = std::make_unique<EvalJob>(ctx);
    j->p = p;
    j->ds = ds;
    j->out = mse_out;
    int rc = get_lowering(p, ds->num_features, false, &j->L);
    if (rc) return rc;
    j->on.assign(p->n, 0);
    bool all = true;
    for (int i = 0; i < p->n; ++i) {
        const bool want = mask == nullptr || mask[i] != 0;
        const bool ok = j->L->status[i] == PROG_OK;
        if (want && !ok) j->n_bad++;
        j->on[i] = want && ok;
        if (!j->on[i]) all = false;
        if (j->on[i]) j->nodes += prog_nodes(p, i);
    }
    // not cleared: the evaluation writes the MSE of every evaluated program, and finish()
    // sets the others
    if ((rc = j->jb.raw((size_t)p->n * sizeof(float), (void**)&j->d_mse, false))) return rc;
    if (!all && ((rc = j->jb.put(j->on, &j->d_act)) || (rc = subset_order(j->jb, j->L, j->on, &j->d_order, &j->n_order, &j->n_order_big))))
        return rc;
    *out = std::move(j);
    return RSR_OK;
}

int make_grad_job(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, const uint8_t* mask, float* mse_out,
                  float* grad_out, std::unique_ptr<Job>* out) {
    auto j = std::make_unique<GradJob>(ctx);
    j->p = p;
    j->ds = ds;
    j->mse_out = mse_out;
    j->grad_out = grad_out;
    int rc = get_lowering(p, ds->num_features, false, &j->Lp);
    if (rc) return rc;
    if ((rc = get_lowering(p, ds->num_features, true, &j->Ld))) return rc;
    const Lowering* Lp = j->Lp;
    const Lowering* Ld = j->Ld;
    std::vector<uint8_t> act0(p->n, 0), act1(p->n, 0);
    j->on.assign(p->n, 0);
    for (int i = 0; i < p->n; ++i) {
        const bool want = mask == nullptr || mask[i] != 0;
        if (!want) continue;
        const bool zero = Ld->nparams[i] == 0;
        const bool ok = zero ? Lp->status[i] == PROG_OK : Ld->status[i] == PROG_OK;
        if (!ok) {
            j->n_bad++;
            continue;
        }
        j->on[i] = 1;
        j->nodes += prog_nodes(p, i);
        if (zero) {
            act0[i] = 1;
            j->any_zero = true;
        } else {
            act1[i] = 1;
            j->any_pos = true;
        }
    }
    if ((rc = j->jb.get((size_t)p->n, &j->d_mse)) || (rc = j->jb.get((size_t)p->n * 32, &j->d_grad)) ||
        (rc = j->jb.put(act0, &j->d_act0)) || (rc = j->jb.put(act1, &j->d_act1)) ||
        (rc = subset_order(j->jb, Lp, act0, &j->d_order0, &j->n_order0, &j->n_big0)) ||
        (rc = subset_order(j->jb, Ld, act1, &j->d_order1, &j->n_order1, &j->n_big1)))
        return rc;
    *out = std::move(j);
    return RSR_OK;
}

int make_lbfgs_job(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int iterations, int history, int mode,
                   const uint8_t* mask, float* fitness_out, std::unique_ptr<Job>* out) {
    if (ds->num_samples == 0) return set_error(RSR_E_INVALID_ARG, "the dataset has no points to fit on");
    if (history < 1 || history > 8 || iterations < 0)
        return set_error(RSR_E_INVALID_ARG, "history must be 1..8, iterations >= 0");
    if (mode != RSR_LBFGS_REFERENCE && mode != RSR_LBFGS_FP64) return set_error(RSR_E_INVALID_ARG, "bad mode");
    auto j = std::make_unique<LbfgsJob>(ctx);
    j->p = p;
    j->ds = ds;
    j->iters = iterations;
    j->fitness_out = fitness_out;
    int rc = get_lowering(p, ds->num_features, false, &j->Lp);
    if (rc) return rc;
    if ((rc = get_lowering(p, ds->num_features, true, &j->Ld))) return rc;
    std::vector<int> nparam;
    std::vector<uint32_t> poff;
    uint64_t nodes = 0;
    for (int i = 0; i < p->n; ++i) {
        fitness_out[i] = NAN;
        if (mask && !mask[i]) continue;
        if (j->Ld->status[i] != PROG_OK || j->Lp->status[i] != PROG_OK) continue;
        if (j->Ld->nparams[i] == 0) continue;  // lbfgs.rs:18-21: nothing to fit, fitness untouched
        j->prog.push_back(i);
        nparam.push_back(j->Ld->nparams[i]);
        poff.push_back(p->param_off[i]);
        nodes += prog_nodes(p, i);
    }
    const int n = (int)j->prog.size();
    j->mean_nodes = n ? nodes / n : 0;
    if (n == 0) {
        j->done = true;
        *out = std::move(j);
        return RSR_OK;
    }
    // Reading the counters back costs a synchronisation per tick, which pays when a round's
    // evaluation is large (the skipped rounds' launches cover every program group: the Mat3
    // workload's 20 iterations over 16,384 x 20,000 took 879 instead of 955 ms) and does not
    // when it is small (with tens of individuals one of them usually searches all 12 rounds:
    // the islands' step took 12.5 instead of 9.8 ms).
    j->adaptive = (double)n * (double)ds->num_samples * (double)j->mean_nodes >= 2e8;
    JobBuffers& sc = j->jb;
    LbfgsDev& st = j->st;
    st.n = n;
    st.M = history;
    st.fp64 = mode == RSR_LBFGS_FP64 ? 1 : 0;
    const size_t tb = st.fp64 ? sizeof(double) : sizeof(float);
    int *d_prog, *d_np;
    uint32_t* d_poff;
    if ((rc = sc.put(j->prog, &d_prog)) || (rc = sc.put(nparam, &d_np)) || (rc = sc.put(poff, &d_poff))) return rc;
    st.prog = d_prog;
    st.nparam = d_np;
    st.param_off = d_poff;
    if ((rc = sc.get((size_t)kMaxP * n, &st.x)) || (rc = sc.get((size_t)kMaxP * n, &st.g)) ||
        (rc = sc.get((size_t)kMaxP * n, &st.p)) || (rc = sc.get((size_t)n, &st.step)) ||
        (rc = sc.get((size_t)n, &st.cur_mse)) || (rc = sc.get((size_t)n, &st.next_mse)))
        return rc;
    if ((rc = sc.raw((size_t)8 * kMaxP * n * tb, &st.S)) || (rc = sc.raw((size_t)8 * kMaxP * n * tb, &st.Y)) ||
        (rc = sc.raw((size_t)8 * n * tb, &st.rho)) || (rc = sc.raw((size_t)8 * n * tb, &st.alpha)) ||
        (rc = sc.raw((size_t)n * tb, &st.ddg)))
        return rc;
    if ((rc = sc.get((size_t)n, &st.hist)) || (rc = sc.get((size_t)n, &st.head)) || (rc = sc.get((size_t)n, &st.ls)) ||
        (rc = sc.get((size_t)n, &st.flags)) || (rc = sc.get(2, &j->d_counters)))
        return rc;
    std::vector<uint8_t> fitmask(p->n, 0);
    for (int i : j->prog) fitmask[i] = 1;
    int cnt;
    if ((rc = sc.put(fitmask, &j->act_fit)) || (rc = sc.get((size_t)p->n, &j->act_ls)) ||
        (rc = sc.get((size_t)p->n, &j->act_grad)) || (rc = sc.get((size_t)p->n, &j->d_mse)) ||
        (rc = sc.get((size_t)p->n * 32, &j->d_grad)) || (rc = sc.get((size_t)n, &j->d_fit)) ||
        (rc = subset_order(sc, j->Lp, fitmask, &j->order_p, &cnt, &j->big_p)) ||
        (rc = subset_order(sc, j->Ld, fitmask, &j->order_d, &cnt, &j->big_d)))
        return rc;
    // Small fits search in one round: the 12 steps of every searching individual as one
    // candidate batch (each candidate a copy of its individual's parameter block, sharing the
    // micro-code, scheduled in the parents' cost order), then the first step that passes.
    // Twelve times the line-search evaluations, in one launch instead of twelve rounds of
    // three: below 1e8 node-evaluations per round the launches cost more than the work.
    constexpr int kRounds = 12;
    if (!j->adaptive && (double)n * (double)ds->num_samples * (double)j->mean_nodes * kRounds <= 1e8) {
        const Lowering* Lp = j->Lp;
        std::vector<int64_t> coff((size_t)n * kRounds);
        std::vector<float> cparams;
        std::vector<ProgDesc> hdesc((size_t)n * kRounds);
        for (int k = 0; k < n; ++k) {
            const int i = j->prog[k];
            const float* pp = &p->params[p->param_off[i]];
            const int pc = (int)(p->param_off[i + 1] - p->param_off[i]);
            for (int r = 0; r < kRounds; ++r) {
                coff[(size_t)k * kRounds + r] = (int64_t)cparams.size();
                cparams.insert(cparams.end(), pp, pp + pc);
                ProgDesc d = Lp->desc[i];
                d.param_off = (uint32_t)coff[(size_t)k * kRounds + r];
                hdesc[(size_t)k * kRounds + r] = d;
            }
        }
        cparams.push_back(0.0f);
        std::vector<int> rank(p->n);
        for (int r = 0; r < p->n; ++r) rank[Lp->order[r]] = r;
        std::vector<int32_t> corder((size_t)n * kRounds);
        std::iota(corder.begin(), corder.end(), 0);
        std::stable_sort(corder.begin(), corder.end(),
                         [&](int a, int b) { return rank[j->prog[a / kRounds]] < rank[j->prog[b / kRounds]]; });
        int64_t* d_coff;
        ProgDesc* d_desc;
        int32_t* d_corder;
        if ((rc = sc.put(coff, &d_coff)) || (rc = sc.put(cparams, &st.cand_params)) ||
            (rc = sc.put(hdesc, &d_desc)) || (rc = sc.put(corder, &d_corder)) ||
            (rc = sc.get((size_t)n * kRounds, &st.cand_active)) || (rc = sc.get((size_t)n * kRounds, &j->d_cmse)))
            return rc;
        st.K = kRounds;
        st.cand_off = d_coff;
        j->CP.code = Lp->dcode;
        j->CP.progs = d_desc;
        j->CP.params = st.cand_params;
        j->CP.order = d_corder;
        j->CP.n_progs = n * kRounds;
        j->CP.spill_slots = Lp->spill_slots;
        j->CP.n_big = j->big_p * kRounds;  // the candidates of big parents sort last, as their parents
        j->CP.gslots = Lp->gslots;
    }
    *out = std::move(j);
    return RSR_OK;
}

int make_cmaes_job(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int iterations, int lambda, int mu,
                   uint64_t seed, const float* normals, const uint8_t* mask, float* fitness_out, int32_t* iters_out,
                   std::unique_ptr<Job>* out) {
    if (ds->num_samples == 0) return set_error(RSR_E_INVALID_ARG, "the dataset has no points to fit on");
    if (lambda < 2 || lambda > kMaxLambda || mu < 1 || mu > lambda || iterations < 0)
        return set_error(RSR_E_INVALID_ARG, "lambda must be 2..%d and 1 <= mu <= lambda", kMaxLambda);
    auto j = std::make_unique<CmaesJob>(ctx);
    j->p = p;
    j->ds = ds;
    j->iters = iterations;
    j->fitness_out = fitness_out;
    j->iters_out = iters_out;
    int rc = get_lowering(p, ds->num_features, false, &j->Lp);
    if (rc) return rc;
    const Lowering* Lp = j->Lp;
    std::vector<int> nparam, ndim;
    std::vector<uint32_t> poff;
    uint64_t nodes = 0;
    for (int i = 0; i < p->n; ++i) {
        fitness_out[i] = NAN;
        if (iters_out) iters_out[i] = 0;
        if (mask && !mask[i]) continue;
        if (Lp->status[i] != PROG_OK) continue;
        const int pc = (int)(p->param_off[i + 1] - p->param_off[i]);
        if (pc == 0) continue;  // cmaes.rs:25-28
        j->prog.push_back(i);
        nparam.push_back(std::min(pc, kMaxP));
        ndim.push_back(std::min(std::min(pc, kMaxP) + pc, kMaxP));
        poff.push_back(p->param_off[i]);
        nodes += prog_nodes(p, i);
    }
    const std::vector<int>& prog = j->prog;
    const int n = (int)prog.size();
    j->mean_nodes = n ? nodes / n : 0;
    if (n == 0) {
        j->done = true;
        *out = std::move(j);
        return RSR_OK;
    }
    // recombination weights and strategy constants in f32 as cmaes.rs:85-102
    std::vector<float> w(mu);
    float sumw = 0.0f;
    const float lmu = logf((float)mu + 0.5f);
    for (int i = 0; i < mu; ++i) {
        const float v = lmu - logf((float)(i + 1));
        w[i] = v > 0.0f ? v : 0.0f;
        sumw += w[i];
    }
    float me = 0.0f;
    for (int i = 0; i < mu; ++i) {
        w[i] = w[i] / sumw;
        me += w[i] * w[i];
    }
    me = 1.0f / me;
    std::vector<float> cc(n), ccov(n), mean((size_t)kMaxP * n, 0.0f), cdiag((size_t)kMaxP * n, 1.0f),
        best((size_t)kMaxP * n, 0.0f), sigma(n, 0.5f);
    std::vector<int64_t> cand_off((size_t)n * lambda);
    std::vector<float> cand_params;
    for (int k = 0; k < n; ++k) {
        const int i = prog[k];
        const float nf = (float)ndim[k];
        cc[k] = (4.0f + me / nf) / (nf + 4.0f + 2.0f * me / nf);
        const float t1 = nf + 1.414f, t2 = nf + 2.0f;
        ccov[k] = (1.0f / me) * (2.0f / (t1 * t1)) + (1.0f - 1.0f / me) * ((2.0f * me - 1.0f) / (t2 * t2 + me));
        const float* pp = &p->params[p->param_off[i]];
        const int pc = (int)(p->param_off[i + 1] - p->param_off[i]);
        for (int jj = 0; jj < nparam[k]; ++jj) mean[(size_t)jj * n + k] = pp[jj];
        for (int jj = nparam[k]; jj < ndim[k]; ++jj) mean[(size_t)jj * n + k] = pp[jj - nparam[k]];
        for (int jj = 0; jj < ndim[k]; ++jj) best[(size_t)jj * n + k] = mean[(size_t)jj * n + k];
        for (int c = 0; c < lambda; ++c) {
            cand_off[(size_t)k * lambda + c] = (int64_t)cand_params.size();
            cand_params.insert(cand_params.end(), pp, pp + pc);
        }
    }
    cand_params.push_back(0.0f);
    JobBuffers& sc = j->jb;
    CmaDev& st = j->st;
    std::vector<uint8_t> fitmask(p->n, 0);
    for (int i : prog) fitmask[i] = 1;
    int cnt;
    if ((rc = sc.put(fitmask, &j->act_fit)) || (rc = sc.get((size_t)p->n, &j->d_mse0)) ||
        (rc = subset_order(sc, Lp, fitmask, &j->order_fit, &cnt, &j->big_fit)))
        return rc;
    st.n = n;
    st.lambda = lambda;
    st.mu = mu;
    st.mu_eff = me;
    st.sigma_decay = expf(-0.01f);  // (-0.01_f32).exp() (cmaes.rs:178)
    st.seed = seed;
    float *d_w, *d_normals = nullptr, *d_cc, *d_ccov;
    int *d_prog, *d_np, *d_nd;
    uint32_t* d_poff;
    int64_t* d_coff;
    if ((rc = sc.put(w, &d_w)) || (rc = sc.put(prog, &d_prog)) || (rc = sc.put(nparam, &d_np)) ||
        (rc = sc.put(ndim, &d_nd)) || (rc = sc.put(poff, &d_poff)) || (rc = sc.put(cand_off, &d_coff)) ||
        (rc = sc.put(cc, &d_cc)) || (rc = sc.put(ccov, &d_ccov)))
        return rc;
    if (normals && iterations > 0) {
        const size_t nn = (size_t)iterations * n * lambda * kMaxP;
        if ((rc = sc.get(nn, &d_normals)) || (rc = sc.upload(d_normals, normals, nn * sizeof(float)))) return rc;
    }
    st.weights = d_w;
    st.normals = d_normals;
    st.prog = d_prog;
    st.nparam = d_np;
    st.ndim = d_nd;
    st.param_off = d_poff;
    st.cand_off = d_coff;
    st.c_c = d_cc;
    st.c_cov = d_ccov;
    if ((rc = sc.put(mean, &st.mean)) || (rc = sc.put(cdiag, &st.cdiag)) || (rc = sc.get((size_t)kMaxP * n, &st.pc)) ||
        (rc = sc.put(best, &st.best)) || (rc = sc.put(sigma, &st.sigma)) || (rc = sc.get((size_t)n, &st.best_mse)) ||
        (rc = sc.get((size_t)n, &st.best_fit)) || (rc = sc.get((size_t)n * lambda * kMaxP, &st.steps)) ||
        (rc = sc.get((size_t)n * lambda * kMaxP, &st.pop)) || (rc = sc.get((size_t)n * lambda, &st.l1)) ||
        (rc = sc.get((size_t)n * lambda, &st.cand_active)) || (rc = sc.get((size_t)n, &st.flags)) ||
        (rc = sc.get((size_t)n, &st.iters)))
        return rc;
    // candidate batch: every individual's program repeated lambda times, own parameter blocks,
    // scheduled in the parent batch's cost order
    if ((rc = sc.put(cand_params, &j->d_cparams))) return rc;
    std::vector<ProgDesc> hdesc((size_t)n * lambda);
    for (int k = 0; k < n; ++k)
        for (int c = 0; c < lambda; ++c) {
            ProgDesc d = Lp->desc[prog[k]];
            d.param_off = (uint32_t)cand_off[(size_t)k * lambda + c];
            hdesc[(size_t)k * lambda + c] = d;
        }
    std::vector<int> rank(p->n);
    for (int r = 0; r < p->n; ++r) rank[Lp->order[r]] = r;
    std::vector<int32_t> corder((size_t)n * lambda);
    std::iota(corder.begin(), corder.end(), 0);
    std::stable_sort(corder.begin(), corder.end(),
                     [&](int a, int b) { return rank[prog[a / lambda]] < rank[prog[b / lambda]]; });
    ProgDesc* d_desc;
    int32_t* d_order;
    if ((rc = sc.put(hdesc, &d_desc)) || (rc = sc.put(corder, &d_order)) ||
        (rc = sc.get((size_t)n * lambda, &j->d_cmse)) || (rc = sc.get((size_t)n, &j->d_fit)) ||
        (rc = sc.get((size_t)n, &j->d_iters)) || (rc = sc.get(2, &j->d_counters)))
        return rc;
    j->CP.code = Lp->dcode;
    j->CP.progs = d_desc;
    j->CP.params = j->d_cparams;
    j->CP.order = d_order;
    j->CP.n_progs = n * lambda;
    j->CP.spill_slots = Lp->spill_slots;
    j->CP.n_big = j->big_fit * lambda;  // the candidates of big parents sort last, as their parents
    j->CP.gslots = Lp->gslots;
    *out = std::move(j);
    return RSR_OK;
}

int make_nm_job(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int iterations, const uint8_t* mask,
                float* fitness_out, int32_t* iters_out, std::unique_ptr<Job>* out) {
    if (ds->num_samples == 0) return set_error(RSR_E_INVALID_ARG, "the dataset has no points to fit on");
    if (iterations < 0) return set_error(RSR_E_INVALID_ARG, "iterations must be >= 0");
    auto j = std::make_unique<NmJob>(ctx);
    j->p = p;
    j->ds = ds;
    j->fitness_out = fitness_out;
    j->iters_out = iters_out;
    int rc = get_lowering(p, ds->num_features, false, &j->Lp);
    if (rc) return rc;
    const Lowering* Lp = j->Lp;
    std::vector<int> ndim;
    std::vector<std::vector<int32_t>> offs;
    std::vector<uint32_t> poff;
    uint64_t nodes = 0, dims = 0;
    for (int i = 0; i < p->n; ++i) {
        fitness_out[i] = NAN;
        if (iters_out) iters_out[i] = 0;
        if (mask && !mask[i]) continue;
        if (Lp->status[i] != PROG_OK) continue;
        // the Float constants, in constant order, at most 32 (nelder_mead.rs:19-30)
        std::vector<int32_t> o;
        uint32_t f = 0;
        for (int64_t c = p->const_off[i]; c < p->const_off[i + 1]; ++c) {
            const int s = vt_size(p->ctypes[c]);
            if (p->ctypes[c] == VtFloat && o.size() < 32) o.push_back((int32_t)f);
            f += s;
        }
        if (o.empty()) continue;  // nelder_mead.rs:32-34
        j->prog.push_back(i);
        ndim.push_back((int)o.size());
        offs.push_back(std::move(o));
        poff.push_back(p->param_off[i]);
        nodes += prog_nodes(p, i);
        dims += ndim.back();
    }
    const int n = (int)j->prog.size();
    if (n == 0) {
        j->done = true;
        *out = std::move(j);
        return RSR_OK;
    }
    j->mean_nodes = nodes / n;
    j->mean_dim = dims / n;
    NmDev& st = j->st;
    JobBuffers& sc = j->jb;
    st.n = n;
    st.max_iter = iterations;
    std::vector<int32_t> opt((size_t)kMaxP * n, 0);
    for (int k = 0; k < n; ++k)
        for (size_t jj = 0; jj < offs[k].size(); ++jj) opt[jj * n + k] = offs[k][jj];
    // candidate slots: full copies of each individual's parameter block
    std::vector<int64_t> cand_off((size_t)n * kNmSlots);
    std::vector<float> cand_params;
    for (int k = 0; k < n; ++k) {
        const int i = j->prog[k];
        const float* pp = &p->params[p->param_off[i]];
        const int pc = (int)(p->param_off[i + 1] - p->param_off[i]);
        for (int c = 0; c < kNmSlots; ++c) {
            cand_off[(size_t)k * kNmSlots + c] = (int64_t)cand_params.size();
            cand_params.insert(cand_params.end(), pp, pp + pc);
        }
    }
    cand_params.push_back(0.0f);
    int *d_nd;
    int32_t* d_opt;
    uint32_t* d_poff;
    int64_t* d_coff;
    if ((rc = sc.put(ndim, &d_nd)) || (rc = sc.put(opt, &d_opt)) || (rc = sc.put(poff, &d_poff)) ||
        (rc = sc.put(cand_off, &d_coff)) || (rc = sc.put(cand_params, &j->d_cparams)))
        return rc;
    st.ndim = d_nd;
    st.opt_off = d_opt;
    st.param_off = d_poff;
    st.cand_off = d_coff;
    if ((rc = sc.get((size_t)kNmSlots * n, &st.fit)) || (rc = sc.get((size_t)kNmSlots * n, &st.mse)) ||
        (rc = sc.get((size_t)kNmSlots * kMaxP * n, &st.pt)) || (rc = sc.get((size_t)kMaxP * n, &st.cen)) ||
        (rc = sc.get((size_t)kMaxP * n, &st.refl)) || (rc = sc.get((size_t)kMaxP * n, &st.trial)) ||
        (rc = sc.get((size_t)n, &st.rf)) || (rc = sc.get((size_t)n, &st.rm)) || (rc = sc.get((size_t)n, &st.lim)) ||
        (rc = sc.get((size_t)kNmSlots * n, &st.idx)) || (rc = sc.get((size_t)n, &st.phase)) ||
        (rc = sc.get((size_t)n, &st.iter)) || (rc = sc.get((size_t)n * kNmSlots, &st.cand_active)) ||
        (rc = sc.get(2, &j->d_counters)))
        return rc;
    std::vector<ProgDesc> hdesc((size_t)n * kNmSlots);
    for (int k = 0; k < n; ++k)
        for (int c = 0; c < kNmSlots; ++c) {
            ProgDesc d = Lp->desc[j->prog[k]];
            d.param_off = (uint32_t)cand_off[(size_t)k * kNmSlots + c];
            hdesc[(size_t)k * kNmSlots + c] = d;
        }
    std::vector<int> rank(p->n);
    for (int r = 0; r < p->n; ++r) rank[Lp->order[r]] = r;
    std::vector<int32_t> corder((size_t)n * kNmSlots);
    std::iota(corder.begin(), corder.end(), 0);
    const std::vector<int>& prog = j->prog;
    std::stable_sort(corder.begin(), corder.end(),
                     [&](int a, int b) { return rank[prog[a / kNmSlots]] < rank[prog[b / kNmSlots]]; });
    ProgDesc* d_desc;
    int32_t* d_order;
    if ((rc = sc.put(hdesc, &d_desc)) || (rc = sc.put(corder, &d_order)) ||
        (rc = sc.get((size_t)n * kNmSlots, &j->d_cmse)) || (rc = sc.get((size_t)n, &j->d_fit)) ||
        (rc = sc.get((size_t)n, &j->d_iters)))
        return rc;
    j->CP.code = Lp->dcode;
    j->CP.progs = d_desc;
    j->CP.params = j->d_cparams;
    j->CP.order = d_order;
    j->CP.n_progs = n * kNmSlots;
    j->CP.spill_slots = Lp->spill_slots;
    int nb = 0;
    for (int i : prog) nb += Lp->big[i];
    j->CP.n_big = nb * kNmSlots;  // the candidates of big parents sort last, as their parents
    j->CP.gslots = Lp->gslots;
    *out = std::move(j);
    return RSR_OK;
}

int make_optimize_job(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int iterations, uint64_t seed,
                      float threshold, const uint8_t* mask, float* mse_out, uint8_t* method_out,
                      std::unique_ptr<Job>* out) {
    if (ds->num_samples == 0) return set_error(RSR_E_INVALID_ARG, "the dataset has no points to fit on");
    if (iterations < 0) return set_error(RSR_E_INVALID_ARG, "iterations must be >= 0");
    if (!(threshold >= 0.0f)) return set_error(RSR_E_INVALID_ARG, "threshold must be >= 0");
    auto j = std::make_unique<OptimizeJob>(ctx);
    j->p = p;
    j->ds = ds;
    j->threshold = threshold;
    j->mse_out = mse_out;
    Lowering* Lp;
    int rc = get_lowering(p, ds->num_features, false, &Lp);
    if (rc) return rc;
    std::vector<uint8_t> ml(p->n, 0), mc(p->n, 0), mn(p->n, 0);
    j->chosen.assign(p->n, 0);
    bool any_l = false, any_c = false, any_n = false;
    for (int i = 0; i < p->n; ++i) {
        if (method_out) method_out[i] = RSR_FIT_NONE;
        if ((mask && !mask[i]) || Lp->status[i] != PROG_OK) continue;
        bool vecmat = false;
        for (int64_t c = p->const_off[i]; c < p->const_off[i + 1]; ++c)
            if (p->ctypes[c] != VtFloat) vecmat = true;
        int m;
        if (Lp->differentiable[i]) {
            ml[i] = 1;
            any_l = true;
            m = RSR_FIT_LBFGS;
        } else if (vecmat) {
            mc[i] = 1;
            any_c = true;
            m = RSR_FIT_CMAES;
        } else {
            mn[i] = 1;
            any_n = true;
            m = RSR_FIT_NELDER_MEAD;
        }
        j->chosen[i] = 1;
        if (method_out) method_out[i] = (uint8_t)m;
    }
    j->fit_tmp.assign(p->n, NAN);
    j->mse_tmp.assign(p->n, NAN);
    j->iters_tmp.assign(p->n, 0);
    std::unique_ptr<Job> f;
    if (any_l) {
        if ((rc = make_lbfgs_job(ctx, p, ds, 15, 6, RSR_LBFGS_REFERENCE, ml.data(), j->fit_tmp.data(), &f))) return rc;
        j->fits.push_back(std::move(f));
    }
    if (any_c) {
        const int ci = std::max(iterations / 16, 2);
        if ((rc = make_cmaes_job(ctx, p, ds, ci, 16, 8, seed, nullptr, mc.data(), j->fit_tmp.data(),
                                 j->iters_tmp.data(), &f)))
            return rc;
        j->fits.push_back(std::move(f));
    }
    if (any_n) {
        if ((rc = make_nm_job(ctx, p, ds, iterations, mn.data(), j->fit_tmp.data(), j->iters_tmp.data(), &f)))
            return rc;
        j->fits.push_back(std::move(f));
    }
    if ((rc = j->refresh())) return rc;  // no fits at all: straight to the final evaluation
    *out = std::move(j);
    return RSR_OK;
}

// ---------------------------------------------------------------------------------
// ticks
// ---------------------------------------------------------------------------------
namespace {

// Run steps of one kind (MSE or gradient) at one points-per-lane.  The per-tile Kahan
// partials (or, in reference order, the per-point terms) live in a shared arena of at
// most ctx->arena_bytes; when the steps need more, every step is split into the same
// number of tile ranges, evaluated round by round with the summation state carried
// between rounds (the sums are the same as unsplit).
int launch_rounds(rsr_context* ctx, std::vector<EvalStep>& steps, bool grad, int ppl, int* queue = nullptr) {
    const bool ref = ctx->summation == RSR_SUM_REFERENCE;
    EvalTuning t = ctx->tuning;
    t.ppl = ppl;
    int rc;
    // device spill areas of the big-program launches (the same for every round)
    BigSpill big;
    big.floats = big_spill_floats(steps.data(), (int)steps.size(), ppl, grad);
    if (big.floats) {
        if ((rc = ensure_scratch(ctx, ctx->scratch_gspill, big.floats * sizeof(float)))) return rc;
        big.ptr = (float*)ctx->scratch_gspill.ptr;
    }
    const bool terms = ref;  // reference order: per-point terms, folded in order
    auto part_floats = [&](const EvalStep& st) {
        return grad ? (terms ? grad_terms_floats(st) : grad_partials_floats(st)) : mse_partials_floats(st, ppl, terms);
    };
    auto state_floats = [&](const EvalStep& st) {
        return grad ? (terms ? grad_terms_state_floats(st) : grad_state_floats(st)) : mse_state_floats(st, terms);
    };
    size_t total = 0;
    std::vector<int64_t> tiles(steps.size());
    for (size_t i = 0; i < steps.size(); ++i) {
        tiles[i] = eval_tiles(steps[i].D, ppl);
        total += part_floats(steps[i]);
    }
    total *= sizeof(float);
    const size_t cap = std::max<size_t>(ctx->arena_bytes, 4096);
    const int64_t rounds = (int64_t)((total + cap - 1) / cap);
    float* state = nullptr;
    if (rounds > 1) {
        size_t sf = 0;
        for (const auto& st : steps) sf += state_floats(st);
        DevBuf& sb = grad ? ctx->scratch_gstate : ctx->scratch_state;
        if ((rc = ensure_scratch(ctx, sb, sf * sizeof(float)))) return rc;
        state = (float*)sb.ptr;
    }
    for (int64_t r = 0; r < rounds; ++r) {
        std::vector<EvalStep> round;
        size_t need = 0;
        float* st_next = state;
        for (size_t i = 0; i < steps.size(); ++i) {
            EvalStep s = steps[i];
            const int64_t chunk = (tiles[i] + rounds - 1) / rounds;
            s.tile_begin = r * chunk;
            s.tile_count = std::min<int64_t>(chunk, tiles[i] - s.tile_begin);
            s.first = r == 0;
            s.last = s.tile_begin + s.tile_count >= tiles[i];
            s.kstate = st_next;
            if (state) st_next += state_floats(s);
            if (s.tile_count <= 0) continue;
            need += part_floats(s);
            round.push_back(s);
        }
        if (round.empty()) continue;
        DevBuf& arena_buf = grad ? ctx->scratch_gpartials : ctx->scratch_partials;
        if ((rc = ensure_scratch(ctx, arena_buf, std::max<size_t>(need, 1) * sizeof(float)))) return rc;
        float* arena = (float*)arena_buf.ptr;
        for (auto& s : round) {
            s.partials = arena;
            arena += part_floats(s);
        }
        HIP_TRY(grad ? launch_grad_steps(round.data(), (int)round.size(), terms, ctx->stream,
                                         ctx->next_timer(), big)
                     : launch_mse_steps(round.data(), (int)round.size(), t, terms, ctx->stream, ctx->next_timer(), big,
                                        queue),
                "evaluation launch");
    }
    return RSR_OK;
}

// MSE steps: each step at the tuning's points per lane, lowered until its own spill area
// fits the workgroup LDS (per step, so a request's result does not depend on the others).
int run_mse(rsr_context* ctx, std::vector<EvalStep>& steps) {
    std::vector<EvalStep> by_ppl[5][2];  // by points per lane (1 to 4) and launch form
    for (const auto& s0 : steps) {
        const EvalTuning t = pick_tuning(ctx, s0);
        EvalStep s = s0;
        s.group = t.group;
        int ppl = t.ppl;
        while (ppl > 1 && eval_mse_lds_bytes(s.P.spill_slots, ppl) > ctx->max_lds) ppl /= 2;
        if (eval_mse_lds_bytes(s.P.spill_slots, ppl) > ctx->max_lds)
            return set_error(RSR_E_UNSUPPORTED, "spill area of %zu bytes exceeds the workgroup LDS limit",
                             eval_mse_lds_bytes(s.P.spill_slots, ppl));
        by_ppl[ppl][ctx->queue_for(t) ? 1 : 0].push_back(s);
    }
    int rc;
    for (int k = 4; k >= 1; --k)
        for (int q = 0; q < 2; ++q)
            if (!by_ppl[k][q].empty() &&
                (rc = launch_rounds(ctx, by_ppl[k][q], false, k, q ? (int*)ctx->queue.ptr : nullptr)))
                return rc;
    return RSR_OK;
}

int run_grad(rsr_context* ctx, std::vector<EvalStep>& steps) {
    for (const auto& s : steps)
        if (eval_grad_lds_bytes(s.P.spill_slots) > ctx->max_lds)
            return set_error(RSR_E_UNSUPPORTED, "dual spill area of %zu bytes exceeds the workgroup LDS limit",
                             eval_grad_lds_bytes(s.P.spill_slots));
    return launch_rounds(ctx, steps, true, 1);
}

}  // namespace

int run_tick(rsr_context* ctx, const std::vector<Job*>& top) {
    std::vector<Job*> jobs;
    for (Job* j : top) j->leaves(jobs);
    const size_t nj = jobs.size();
    if (nj == 0) {
        for (Job* j : top) {
            const int rc = j->refresh();
            if (rc) return rc;
        }
        return RSR_OK;
    }
    int rc;
    std::vector<EvalStep> mse, grad;
    for (size_t k = 0; k < nj; ++k)
        if ((rc = jobs[k]->plan(mse, grad))) return rc;
    if (!mse.empty() && (rc = run_mse(ctx, mse))) return rc;
    if (!grad.empty() && (rc = run_grad(ctx, grad))) return rc;
    for (Job* j : jobs)
        if ((rc = j->post())) return rc;
    // counters of the jobs that decide on them this tick: one copy each, one synchronisation;
    // without any, the next tick is enqueued right away (the next synchronisation is a job's
    // finish() or a later tick's counters)
    std::vector<size_t> want;
    std::vector<uint8_t> reads(nj, 0);
    for (size_t k = 0; k < nj; ++k)
        if (jobs[k]->needs_counters() && jobs[k]->d_counters) {
            want.push_back(k);
            reads[k] = 1;
        }
    std::vector<int> delta(2 * nj, 0);
    if (!want.empty()) {
        if (ctx->host_counters_cap < 2 * want.size()) {
            if (ctx->host_counters) (void)hipHostFree(ctx->host_counters);
            ctx->host_counters = nullptr;
            ctx->host_counters_cap = 0;
            const size_t cap = std::max<size_t>(2 * want.size(), 64);
            HIP_TRY(hipHostMalloc((void**)&ctx->host_counters, cap * sizeof(int), hipHostMallocDefault), "counters");
            ctx->host_counters_cap = cap;
        }
        int* raw = ctx->host_counters;
        for (size_t w = 0; w < want.size(); ++w)
            HIP_TRY(hipMemcpyAsync(raw + 2 * w, jobs[want[w]]->d_counters, 2 * sizeof(int), hipMemcpyDeviceToHost,
                                   ctx->stream),
                    "counter download");
        HIP_TRY(hipStreamSynchronize(ctx->stream), "tick");
        for (size_t w = 0; w < want.size(); ++w) {
            Job* j = jobs[want[w]];
            for (int c = 0; c < 2; ++c) {  // this tick's increase of each counter
                const uint32_t now = (uint32_t)raw[2 * w + c];
                delta[2 * want[w] + c] = (int)(now - j->counters_seen[c]);
                j->counters_seen[c] = now;
            }
        }
    }
    for (size_t k = 0; k < nj; ++k) {
        jobs[k]->ticks++;
        if ((rc = jobs[k]->advance(reads[k] ? delta.data() + 2 * k : nullptr))) return rc;
    }
    for (Job* j : top) {
        if (std::find(jobs.begin(), jobs.end(), j) == jobs.end()) j->ticks++;
        if ((rc = j->refresh())) return rc;
    }
    return RSR_OK;
}

// compute_mse of every program through a captured graph (built on first use for this
// dataset and these settings, replayed afterwards; at most eight per batch, the least
// recently replayed dropped first, and those of destroyed datasets dropped on the next
// call).  Returns RSR_E_UNSUPPORTED without results when the evaluation does not fit one
// launch round or the graph's buffers do not fit the budget; the caller then runs it
// eagerly.
int eval_mse_graph(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, float* out) {
    Lowering* L;
    int rc = get_lowering(p, ds->num_features, false, &L);
    if (rc) return rc;
    if (L->n_big > 0) return RSR_E_UNSUPPORTED;  // big programs run eagerly (context spill area)
    if (ds->num_samples == 0) return RSR_E_UNSUPPORTED;  // nothing to launch: the eager job fills f32::MAX
    const int summ = ctx->summation;
    const bool terms = summ == RSR_SUM_REFERENCE;
    int n_ok = 0;
    for (int i = 0; i < p->n; ++i) n_ok += L->status[i] == PROG_OK;
    EvalTuning t;
    {
        EvalStep probe = mse_step(p, L, ds, nullptr, nullptr);
        probe.P.n_progs = n_ok;
        t = pick_tuning(ctx, probe);
    }
    while (t.ppl > 1 && eval_mse_lds_bytes(L->spill_slots, t.ppl) > ctx->max_lds) t.ppl /= 2;
    if (eval_mse_lds_bytes(L->spill_slots, t.ppl) > ctx->max_lds)
        return set_error(RSR_E_UNSUPPORTED, "spill area exceeds the workgroup LDS limit");
    // drop the graphs whose dataset is gone (their nodes would read freed memory)
    auto& graphs = p->graphs;
    graphs.erase(std::remove_if(graphs.begin(), graphs.end(),
                                [&](const std::unique_ptr<EvalGraph>& e) { return !ctx->live_datasets.count(e->ds_id); }),
                 graphs.end());
    EvalGraph* g = nullptr;
    for (auto& e : p->graphs)
        if (e->ds_id == ds->id && e->summation == summ && e->ppl == t.ppl && e->group == t.group &&
            e->persistent == (ctx->queue_for(t) != nullptr))
            g = e.get();
    if (!g) {
        std::vector<int32_t> order;
        for (int32_t i : L->order)
            if (L->status[i] == PROG_OK) order.push_back(i);
        EvalStep st = mse_step(p, L, ds, nullptr, nullptr);
        st.P.n_progs = (int)order.size();
        st.group = t.group;
        const size_t need = mse_partials_floats(st, t.ppl, terms) * sizeof(float);
        if (need > ctx->arena_bytes) return RSR_E_UNSUPPORTED;
        // at most kMaxGraphs per batch: the least recently replayed one makes room
        constexpr size_t kMaxGraphs = 8;
        if (graphs.size() >= kMaxGraphs) {
            auto lru = std::min_element(graphs.begin(), graphs.end(), [](const auto& a, const auto& b) {
                return a->last_use < b->last_use;
            });
            graphs.erase(lru);
        }
        auto ng = std::make_unique<EvalGraph>();
        ng->ds_id = ds->id;
        ng->summation = summ;
        ng->ppl = t.ppl;
        ng->group = t.group;
        ng->persistent = ctx->queue_for(t) != nullptr;
        // the graph's own buffers; when they do not fit, the batch's other graphs are
        // released and the allocation tried once more, else the call runs eagerly
        auto setup = [&]() {
            int r;
            if ((r = ng->d_mse.alloc(ctx, std::max<size_t>(p->n, 1) * sizeof(float))) ||
                (r = ng->d_order.upload(ctx, order.data(), std::max<size_t>(order.size(), 1) * sizeof(int32_t))) ||
                (r = ng->arena.alloc(ctx, std::max<size_t>(need, 16))))
                return r;
            if (!ng->host &&
                hipHostMalloc((void**)&ng->host, std::max<size_t>(p->n, 1) * sizeof(float), hipHostMallocDefault) !=
                    hipSuccess) {
                ng->host = nullptr;
                return set_error(RSR_E_HIP, "graph output");
            }
            return RSR_OK;
        };
        if (setup() != RSR_OK) {
            graphs.clear();
            if (setup() != RSR_OK) return RSR_E_UNSUPPORTED;
        }
        HIP_TRY(hipMemsetAsync(ng->d_mse.ptr, 0, (size_t)p->n * sizeof(float), ctx->stream), "graph output clear");
        HIP_TRY(hipStreamSynchronize(ctx->stream), "graph set-up");  // uploads done before capture
        st.P.order = (const int32_t*)ng->d_order.ptr;
        st.mse_out = (float*)ng->d_mse.ptr;
        st.partials = (float*)ng->arena.ptr;
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed), "graph capture");
        hipError_t e = launch_mse_steps(&st, 1, t, terms, ctx->stream, nullptr, BigSpill{}, ctx->queue_for(t));
        if (e == hipSuccess)
            e = hipMemcpyAsync(ng->host, ng->d_mse.ptr, (size_t)p->n * sizeof(float), hipMemcpyDeviceToHost,
                               ctx->stream);
        hipGraph_t graph = nullptr;
        const hipError_t e2 = hipStreamEndCapture(ctx->stream, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return hip_error(e != hipSuccess ? e : e2, "graph capture");
        }
        ng->graph = graph;
        HIP_TRY(hipGraphInstantiate(&ng->exec, graph, nullptr, nullptr, 0), "graph instantiate");
        g = ng.get();
        p->graphs.push_back(std::move(ng));
    }
    g->last_use = ++ctx->graph_clock;
    HIP_TRY(hipGraphLaunch(g->exec, ctx->stream), "graph launch");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "graph replay");
    int bad = 0;
    for (int i = 0; i < p->n; ++i) {
        if (L->status[i] != PROG_OK) {
            out[i] = NAN;
            bad++;
        } else {
            out[i] = g->host[i];
        }
    }
    if (bad) return set_error(RSR_E_PROGRAM, "%d program(s) rejected by lowering", bad);
    return RSR_OK;
}

// Points per lane x programs per work item, each candidate timed on the batch itself
// (best of two evaluations after a warm-up, events on the context's stream); in the
// automatic launch mode also the work-queue form at two and four points per lane (at three
// its work loop spills, and at one it has not paid).
int autotune(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int* ppl_out, int* group_out) {
    static const int kPpl[] = {1, 2, 3, 4}, kGroup[] = {1, 2, 4, 8, 16, 32};
    const bool forms = ctx->launch_mode == RSR_LAUNCH_AUTO;
    const bool saved_auto = ctx->auto_tune;
    const EvalTuning saved = ctx->tuning;
    std::vector<float> out(std::max(p->n, 1));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "autotune event");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_error(RSR_E_HIP, "autotune event");
    }
    int rc = RSR_OK;
    float best = -1.0f;
    EvalTuning pick = saved;
    int n_ok = 0;
    ctx->auto_tune = false;
    for (int ppl : kPpl)
        for (int group : kGroup)
        for (int pers = 0; pers < ((forms && (ppl == 2 || ppl == 4)) ? 2 : 1); ++pers) {
            ctx->tuning.ppl = ppl;
            ctx->tuning.group = group;
            ctx->tuning.persistent = pers;
            float ms = -1.0f;
            for (int rep = 0; rep < 3 && rc == RSR_OK; ++rep) {
                std::unique_ptr<Job> job;
                if ((rc = make_eval_job(ctx, p, ds, nullptr, out.data(), &job))) break;
                (void)hipEventRecord(e0, ctx->stream);
                rc = run_job(ctx, *job);
                if (rc == RSR_E_PROGRAM) rc = RSR_OK;  // rejected programs do not change the timing
                (void)hipEventRecord(e1, ctx->stream);
                (void)hipEventSynchronize(e1);
                float t = 0.0f;
                (void)hipEventElapsedTime(&t, e0, e1);
                if (rep > 0 && (ms < 0.0f || t < ms)) ms = t;  // the first run warms the caches
            }
            if (rc) break;
            if (best < 0.0f || ms < best) {
                best = ms;
                pick.ppl = ppl;
                pick.group = group;
                pick.persistent = pers;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->auto_tune = saved_auto;
    ctx->tuning = saved;
    if (rc) return rc;
    Lowering* L;
    if ((rc = get_lowering(p, ds->num_features, false, &L))) return rc;
    for (int i = 0; i < p->n; ++i) n_ok += L->status[i] == PROG_OK;
    ctx->tune_cache[tuning_key(ds->num_features, ds->count8, n_ok, ctx->summation)] = pick;
    ctx->tune_version++;
    if (ppl_out) *ppl_out = pick.ppl;
    if (group_out) *group_out = pick.group;
    return RSR_OK;
}

int run_job(rsr_context* ctx, Job& job) {
    std::vector<Job*> one{&job};
    while (!job.done) {
        int rc = run_tick(ctx, one);
        if (rc) return rc;
    }
    // the results through the context's pinned block: a download into pinned memory is one
    // DMA, where one into pageable memory also costs the runtime a bounce copy (tens of
    // microseconds, most of a small evaluation's time)
    const size_t b = job.stage_bytes();
    if (b > 0 && ensure_host_stage(ctx, b) == RSR_OK) {
        int rc = job.stage(ctx->host_stage);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(ctx->stream), "result download");
    }
    return job.finish();
}

}  // namespace rsr
