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
    int spill_slots;       // LDS spill slots per point per channel (max over the batch, <= kMaxSpillSlots)
    // Big programs (spill slots beyond the LDS ones, ProgDesc out_reg & kBigProg) sit at the
    // last n_big scheduling positions; they run in bounded launches of their own whose
    // waves each own `gslots` device-memory slots per point per channel.
    int n_big = 0;
    int gslots = 0;
};

// Device memory for the spill slots of big programs: one area per wave of their launches.
struct BigSpill {
    float* ptr = nullptr;
    size_t floats = 0;
};
// Waves of a big-program launch (a multiple of four) whose per-wave area holds
// (kMaxSpillSlots + gslots) * channels * 64 floats: those launches keep every spill slot in
// device memory (channels: points per lane for values, 1 + tangent channels for gradients).
int64_t big_waves(int gslots, int channels);
// Floats of device memory a big-program launch of these per-point entry points needs.
inline size_t big_points_floats(int gslots, int channels) {
    return (size_t)big_waves(gslots, channels) * (size_t)(kMaxSpillSlots + gslots) * channels * 64;
}


struct EvalTuning {
    int ppl = 2;         // points per lane: 1 to 4
    int group = 8;       // programs per work item
    int persistent = 0;  // the work-queue launch form (RSR_LAUNCH_AUTO; chosen by the autotuner)
};

// Optional events recorded around the main evaluation kernel of a launch (profiling).
struct KernelTimer {
    hipEvent_t begin = nullptr, end = nullptr;
};

// One request's evaluation step (host description).  For MSE steps, grad_out /
// nparams / max_params are unused; for gradient steps nparams[i] = min(param_count, 32).
//
// Kahan summation keeps one partial per (tile, program[, slot]); a step may be split into
// tile ranges (tile_begin, tile_count) evaluated by successive launches, with the Kahan
// state (sum, compensation) carried between them in kstate, so the result does not
// depend on the split.
struct EvalStep {
    DevPrograms P;
    DevDataset D;
    const uint8_t* active = nullptr;  // per program, null = all
    const int* live = nullptr;        // skip the whole step when *live == 0 (device flag)
    float* partials = nullptr;        // Kahan summation: step_partials_floats
    float* mse_out = nullptr;         // [n_progs]
    float* grad_out = nullptr;        // [n_progs][32]
    const uint8_t* nparams = nullptr;
    int max_params = 0;
    int64_t tile_begin = 0, tile_count = -1;  // -1: every tile from tile_begin
    float* kstate = nullptr;                  // step_state_floats, needed unless first && last
    bool first = true, last = true;
    int group = 0;                            // programs per work item (0: the launch's tuning)
};

// Floats of device memory that the big-program launches of these steps need (0: none).
// ppl: points per lane of an MSE launch; grad: a gradient launch (one point per lane).
size_t big_spill_floats(const EvalStep* steps, int n, int ppl, bool grad);

// Kernel-argument segment table (kernarg segment, read with scalar loads).
struct SegDesc {
    const UInstr* code;
    const ProgDesc* progs;
    const float* params;
    const int32_t* order;
    const uint8_t* active;
    const int* live;
    const float* feat;
    const float* target;
    float* partials;
    float* mse_out;
    float* grad_out;
    const uint8_t* nparams;
    float* kstate;
    int64_t npad, count8, tile0, n_tiles, first_wave, first_thread;
    int64_t first_big;  // first work item of the segment in the big-program launch
    int64_t first_fin;  // first thread of the segment in the gradient-terms finalize launch
    float* gsum;        // gradient terms: the folded sum of every (position, slot) row
    int n_progs, group, n_groups, F, slots, first, last, terms, tile_pts;
    int n_small;        // scheduling positions [0, n_small) run in the main launch
    int fold_rows;      // rows per scheduling position in the fold (1: MSE; slots: gradients)
    float n_samples;
};
constexpr int kMaxSegs = 16;
struct SegTable {
    int n;
    SegDesc s[kMaxSegs];
};

int64_t eval_tiles(const DevDataset& D, int ppl);
size_t eval_mse_lds_bytes(int spill_slots, int ppl);
size_t eval_grad_lds_bytes(int spill_slots);
// Tiles of the step's range at the given points per lane.
int64_t step_tiles(const EvalStep& st, int ppl);
// Partial slots per (tile, program): 1 for MSE, 1 + max_params for gradients.
int grad_slots(const EvalStep& st);
// MSE partials: one per (tile, program), or with terms one per (point, program).
size_t mse_partials_floats(const EvalStep& st, int ppl, bool terms);
size_t grad_partials_floats(const EvalStep& st);
// Reference-order gradients from per-point terms: one row per (program, slot) over the step's
// tile range, plus the folded sum of every row.
size_t grad_terms_floats(const EvalStep& st);
size_t grad_terms_state_floats(const EvalStep& st);
size_t mse_state_floats(const EvalStep& st, bool terms);
size_t grad_state_floats(const EvalStep& st);

// MSE of every active program of every step (compute_mse, evaluator.rs:153-174), one
// merged launch per kMaxSegs steps.  terms = false: per-tile partials summed with
// Kahan-compensated FP32 by the finalize kernel.  terms = true: per-point terms folded in
// the reference's order (bit-identical to the CPU evaluator).
// queue: kQueueInts zero ints of device memory (the device work queue of the persistent
// evaluation launch; left zero); null: one-shot launches.
constexpr int kQueueInts = 16;
hipError_t launch_mse_steps(const EvalStep* steps, int n, const EvalTuning& tune, bool terms, hipStream_t s,
                            const KernelTimer* timer = nullptr, BigSpill big = {}, int* queue = nullptr);
// MSE + gradient (compute_mse_with_gradient, evaluator.rs:177-224) of every active
// program with parameters, from the dual lowering.  terms = false: per-tile partials,
// Kahan-compensated finalize; terms = true: per-point terms written by the tile-parallel
// kernel and folded in the reference's order (bit-identical to the CPU evaluator).
hipError_t launch_grad_steps(const EvalStep* steps, int n, bool terms, hipStream_t s,
                             const KernelTimer* timer = nullptr, BigSpill big = {});

// Per-point values of programs [first, first+count) (eval_simd over every batch);
// out[(i - first) * count8 + point].
// Big programs among them need big.floats >= big_points_floats(P.gslots, 1).
hipError_t launch_eval_points(const DevPrograms& P, const DevDataset& D, int first, int count, bool any_big,
                              float* out, hipStream_t s, BigSpill big = {});

// Per-point value and tangent for one program and one active parameter (eval_simd_dual).
// A big program needs big.floats >= big_points_floats(P.gslots, 2).
hipError_t launch_eval_points_dual(const DevPrograms& P, const DevDataset& D, int prog, int active_param, bool is_big,
                                   float* val_out, float* grad_out, hipStream_t s, BigSpill big = {});

}  // namespace rsr
