// optim.h -- device state and launch interface of the batched constant fitters.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rsr {

constexpr int kMaxP = 32;        // flat parameters an optimizer touches (lbfgs.rs:6, cmaes.rs:8)
constexpr int kMaxLambda = 64;   // CMA-ES population per individual

enum : int { LB_DONE = 1, LB_SEARCH = 2, LB_MOVED = 4 };
enum : int { CMA_DONE = 1 };

// L-BFGS state of n individuals, field-major ([j * n + i]); S, Y: [slot][j][i] of T.
struct LbfgsDev {
    int n = 0, M = 6, fp64 = 0;
    const int* prog;        // program index of individual i
    const int* nparam;      // min(param_count, 32)
    const uint32_t* param_off;
    float *x, *g, *p, *step, *cur_mse, *next_mse;
    void *S, *Y, *rho, *alpha, *ddg;  // float or double
    int *hist, *head, *ls, *flags;
    // Line search in one round (K > 0): every step 2^-r, r < K, of a searching individual is
    // evaluated at once as candidate k * K + r, with its own parameter block at
    // cand_params + cand_off[k * K + r] (a copy of the individual's block)
    int K = 0;
    const int64_t* cand_off = nullptr;
    float* cand_params = nullptr;
    uint8_t* cand_active = nullptr;
};

hipError_t lbfgs_begin(const LbfgsDev& st, const float* params, const float* d_mse, const float* d_grad,
                       hipStream_t s);
hipError_t lbfgs_direction(const LbfgsDev& st, float* params, uint8_t* active_ls, int* counter, hipStream_t s);
hipError_t lbfgs_ls_check(const LbfgsDev& st, float* params, const float* d_mse, uint8_t* active_ls,
                          uint8_t* active_grad, int* counter, hipStream_t s);
hipError_t lbfgs_update(const LbfgsDev& st, const float* d_grad, uint8_t* active_grad, hipStream_t s);
// One-round line search: the K candidate points of every searching individual, then the
// first that passes the Armijo test (lbfgs.rs:113-125), the last forced, as K sequential
// rounds would take it.
hipError_t lbfgs_ls_fill(const LbfgsDev& st, hipStream_t s);
hipError_t lbfgs_ls_pick(const LbfgsDev& st, float* params, const float* cand_mse, uint8_t* active_ls,
                         uint8_t* active_grad, int* counter, hipStream_t s);
hipError_t lbfgs_finish(const LbfgsDev& st, float* params, float* fitness, hipStream_t s);

// Diagonal CMA-ES state of n individuals with lambda candidates each.
struct CmaDev {
    int n = 0, lambda = 16, mu = 8;
    float mu_eff = 0.0f, sigma_decay = 1.0f;
    uint64_t seed = 0;
    const float* weights;      // [mu]
    const float* normals;      // optional injected normals [it][n][lambda][32]
    const int* prog;
    const int* nparam;         // min(param_count, 32)
    const int* ndim;           // doubled dimension (cmaes.rs:31-75)
    const uint32_t* param_off; // individual's own parameter block
    const int64_t* cand_off;   // candidate ic's parameter block
    const float *c_c, *c_cov;
    float *mean, *cdiag, *pc, *best, *sigma, *best_mse, *best_fit;
    float *steps, *pop, *l1;   // [ic][32], [ic][32], [ic]
    uint8_t* cand_active;      // [ic]
    int *flags, *iters;
};

hipError_t cmaes_begin(const CmaDev& st, const float* d_mse, hipStream_t s);
hipError_t cmaes_sample(const CmaDev& st, float* cand_params, int it, hipStream_t s);
hipError_t cmaes_update(const CmaDev& st, const float* cand_mse, int* counter, hipStream_t s);
hipError_t cmaes_finish(const CmaDev& st, float* params, float* fitness, int* iters_out, hipStream_t s);
float cma_normal_host(uint64_t seed, uint32_t ind, uint32_t it, uint32_t cand, uint32_t j);

// Nelder-Mead over the Float constants of n individuals (nelder_mead.rs:9-165).
// Simplex records r = 0..ndim hold (fit, mse, point); idx is the records' sorted order.
// Every individual evaluates its requested points through candidate slots
// (slot s of individual i = candidate i * kNmSlots + s, with its own parameter block).
constexpr int kNmSlots = 33;
enum : int { NM_INIT = 0, NM_REFLECT, NM_EXPAND, NM_CONTRACT, NM_SHRINK, NM_DONE };
struct NmDev {
    int n = 0, max_iter = 0;
    const int* ndim;            // Float constants fitted (1..32)
    const int32_t* opt_off;     // [32][n] offset of Float constant j in the parameter block
    const uint32_t* param_off;  // the individual's block in the batch parameters
    const int64_t* cand_off;    // [n][33] candidate parameter blocks
    float *fit, *mse, *pt;      // [33][n], [33][n], [33][32][n]
    float *cen, *refl, *trial;  // [32][n]
    float *rf, *rm, *lim;       // [n]
    int *idx, *phase, *iter;    // [33][n], [n], [n]
    uint8_t* cand_active;       // [n][33]
};
hipError_t nm_begin(const NmDev& st, const float* params, float* cand_params, hipStream_t s);
hipError_t nm_step(const NmDev& st, float* cand_params, const float* cand_mse, int* counter, hipStream_t s);
hipError_t nm_finish(const NmDev& st, float* params, float* fitness, int* iters_out, hipStream_t s);

}  // namespace rsr
