// kernels.h -- host-side launch interface of the evaluation kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "isa.h"

namespace rsr {

// Device-resident dataset, feature-major (SoA): feat[f * npad + i], target[i].
// Points [num_samples, count8) are the reference's zero padding lanes
// (dataset.rs:96-125) and count in the MSE sum; points [count8, npad) are
// tile padding and are masked out.
struct DevDataset {
    const float* feat;
    const float* target;
    int num_features;
    int64_t num_samples;
    int64_t count8;  // num_samples rounded up to a multiple of 8
    int64_t npad;    // allocation length, a multiple of the largest tile
};

// Device-resident program batch (one lowering per feature count).
struct DevPrograms {
    const UInstr* code;
    const ProgDesc* progs;
    const float* params;
    const int32_t* order;  // scheduling position -> program index (cost-sorted)
    int n_progs;
    int spill_slots;       // LDS spill slots per point per channel (max over the batch)
};

struct EvalTuning {
    int ppl = 2;        // points per lane: 1, 2 or 4
    int group = 8;      // programs per work item
};

// Optional events recorded around the main evaluation kernel of a launch (profiling).
struct KernelTimer {
    hipEvent_t begin = nullptr, end = nullptr;
};

// MSE of every program (compute_mse, evaluator.rs:153-174).  `active` may be null.
// `partials` must hold n_tiles(ds, ppl) * n_progs floats.
hipError_t launch_eval_mse(const DevPrograms& P, const DevDataset& D, const uint8_t* active, const EvalTuning& t,
                           float* partials, float* mse_out, hipStream_t s, const KernelTimer* timer = nullptr);
int64_t eval_tiles(const DevDataset& D, int ppl);

// Dynamic LDS bytes a launch needs for the batch's spill slots.
size_t eval_mse_lds_bytes(const DevPrograms& P, int ppl);
size_t eval_grad_lds_bytes(const DevPrograms& P);

// Reference-order summation variants: per-point terms are folded into 8 serial lane
// accumulators in point order and added horizontally like f32x8::reduce_add, so the
// results are bit-identical to the CPU evaluator.  One wave per program.
hipError_t launch_eval_mse_exact(const DevPrograms& P, const DevDataset& D, const uint8_t* active, float* mse_out,
                                 hipStream_t s, const KernelTimer* timer = nullptr);
hipError_t launch_eval_grad_exact(const DevPrograms& P, const DevDataset& D, const uint8_t* nparams_eff,
                                  const uint8_t* active, int max_params, float* mse_out, float* grad_out,
                                  hipStream_t s, const KernelTimer* timer = nullptr);

// Per-point values of programs [first, first+count) (eval_simd over every batch);
// out[(i - first) * count8 + point].
hipError_t launch_eval_points(const DevPrograms& P, const DevDataset& D, int first, int count, float* out,
                              hipStream_t s);

// MSE and gradient (compute_mse_with_gradient, evaluator.rs:177-224) from the dual
// lowering.  nparams_eff[i] = min(param_count, 32); programs with zero parameters are
// flagged inactive here and handled by the primal path.  partials must hold
// n_tiles(ds, 1) * n_progs * 33 floats.
hipError_t launch_eval_grad(const DevPrograms& P, const DevDataset& D, const uint8_t* nparams_eff,
                            const uint8_t* active, int max_params, float* partials, float* mse_out,
                            float* grad_out, hipStream_t s, const KernelTimer* timer = nullptr);

// Per-point value and tangent for one program and one active parameter (eval_simd_dual).
hipError_t launch_eval_points_dual(const DevPrograms& P, const DevDataset& D, int prog, int active_param,
                                   float* val_out, float* grad_out, hipStream_t s);

}  // namespace rsr
