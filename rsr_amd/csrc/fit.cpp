// fit.cpp -- C ABI of the batched constant fitters (L-BFGS, CMA-ES).
#include <hip/hip_runtime.h>

#include <memory>

#include "../../include/rsr_amd.h"
#include "jobs.h"

using namespace rsr;

#define HIP_TRY(expr, what)                               \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_error(e_, what); \
    } while (0)

// run_lbfgs (src/engine/optimize/lbfgs.rs:9-165) for every program with mask[i] != 0.
extern "C" int rsr_fit_lbfgs(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int max_iterations,
                             int history, int mode, const uint8_t* mask, float* fitness_out) {
    if (!ctx || !p || !ds || (p->n && !fitness_out)) return set_error(RSR_E_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    std::unique_ptr<Job> job;
    int rc = make_lbfgs_job(ctx, p, ds, max_iterations, history, mode, mask, fitness_out, &job);
    if (rc) return rc;
    return run_job(ctx, *job);
}

// run_cma_es (src/engine/optimize/cmaes.rs:19-192) for every program with mask[i] != 0.
extern "C" int rsr_fit_cmaes(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int max_iterations,
                             int lambda, int mu, uint64_t seed, const float* normals, const uint8_t* mask,
                             float* fitness_out, int32_t* iters_out) {
    if (!ctx || !p || !ds || (p->n && !fitness_out)) return set_error(RSR_E_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    std::unique_ptr<Job> job;
    int rc = make_cmaes_job(ctx, p, ds, max_iterations, lambda, mu, seed, normals, mask, fitness_out, iters_out, &job);
    if (rc) return rc;
    return run_job(ctx, *job);
}

// run_nelder_mead (src/engine/optimize/nelder_mead.rs:9-165) for every program with mask[i] != 0.
extern "C" int rsr_fit_nelder_mead(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int max_iterations,
                                   const uint8_t* mask, float* fitness_out, int32_t* iters_out) {
    if (!ctx || !p || !ds || (p->n && !fitness_out)) return set_error(RSR_E_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    std::unique_ptr<Job> job;
    int rc = make_nm_job(ctx, p, ds, max_iterations, mask, fitness_out, iters_out, &job);
    if (rc) return rc;
    return run_job(ctx, *job);
}

// Individual::optimize_constants (src/engine/search/individual.rs:65-75) + calculate_mse.
extern "C" int rsr_optimize_constants(rsr_context* ctx, rsr_programs* p, const rsr_dataset* ds, int max_iterations,
                                      uint64_t seed, float threshold, const uint8_t* mask, float* mse_out,
                                      uint8_t* method_out) {
    if (!ctx || !p || !ds || (p->n && !mse_out)) return set_error(RSR_E_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    std::unique_ptr<Job> job;
    int rc = make_optimize_job(ctx, p, ds, max_iterations, seed, threshold, mask, mse_out, method_out, &job);
    if (rc) return rc;
    return run_job(ctx, *job);
}

extern "C" float rsr_cmaes_normal(uint64_t seed, uint32_t individual, uint32_t iteration, uint32_t candidate,
                                  uint32_t dim) {
    return cma_normal_host(seed, individual, iteration, candidate, dim);
}
