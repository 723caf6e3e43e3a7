/*
 * rsr_amd.h -- C ABI of the MI355X fitness-evaluation engine for RSR.
 *
 * Plain C types only (pointers, sizes, int status codes).  Every function returns
 * RSR_OK (0) or a negative RSR_E_* code; rsr_last_error_message() gives the
 * calling thread's detail string.  Output buffers are caller-owned.  Calls on one
 * context are serialised internally (one HIP stream per context), so island
 * threads may call concurrently, mirroring the reference's rayon callers
 * (src/engine/search/engine.rs:52).
 *
 * Each entry point names the reference interface it replaces (daruandris/rsr).
 */
#ifndef RSR_AMD_H
#define RSR_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSR_API __attribute__((visibility("default")))

/* ---- status codes ---- */
#define RSR_OK 0
#define RSR_E_INVALID_ARG (-1)
#define RSR_E_HIP (-2)          /* HIP runtime error */
#define RSR_E_OOM (-3)          /* device allocation failed */
#define RSR_E_BUDGET (-4)       /* HBM budget of the context exceeded */
#define RSR_E_PROGRAM (-5)      /* one or more programs were rejected by lowering */
#define RSR_E_NO_DEVICE (-6)
#define RSR_E_UNSUPPORTED (-7)
#define RSR_E_QUEUE_FULL (-8)   /* serving: request queue full (back-pressure) */
#define RSR_E_TIMEOUT (-9)      /* serving: wait timed out */
#define RSR_E_NOT_READY (-10)   /* serving: request still running */

/* ---- per-program lowering status (rsr_programs_status) ---- */
#define RSR_PROG_OK 0
#define RSR_PROG_STACK (-1)    /* typed-stack underflow or depth > 32 (UB in the reference) */
#define RSR_PROG_OPCODE (-2)   /* unknown opcode, or a Bool/Int load (no-op in the reference) */
#define RSR_PROG_FEATURE (-3)  /* feature index outside the dataset */
#define RSR_PROG_CONST (-4)    /* constant index out of range or constant type mismatch */
#define RSR_PROG_ROOT (-5)     /* program does not reduce to exactly one Float value */
#define RSR_PROG_REGS (-6)     /* needs more than the 32-register file */

/* ---- wire formats ----
 * Opcodes: the reference `Instruction` enum in declaration order
 *   (src/engine/domain.rs:46-63, src/domains/basic/mod.rs:9-20,
 *    src/domains/linalg/mod.rs:9-51, composed in src/lib.rs:44-48):
 *   0 LoadVarF 1 LoadConstF 2 LoadVarB 3 LoadConstB 4 LoadVarI 5 LoadConstI
 *   6 LoadVarV2 7 LoadConstV2 8 LoadVarV3 9 LoadConstV3 10 LoadVarM2 11 LoadConstM2
 *   12 LoadVarM3 13 LoadConstM3
 *   14 AddF 15 SubF 16 MulF 17 DivF 18 SinF 19 CosF 20 ExpF 21 SqrF 22 SqrtF 23 LnF
 *   24 MakeVec2 25 MakeVec3 26 GetXV2 27 GetYV2 28 GetXV3 29 GetYV3 30 GetZV3
 *   31 AddV2 32 SubV2 33 ScaleV2 34 DotV2 35 NormV2
 *   36 AddV3 37 SubV3 38 ScaleV3 39 DotV3 40 NormV3 41 CrossV3
 *   42 MakeMat2 43 AddM2 44 SubM2 45 ScaleM2 46 MulM2 47 MulM2V2 48 DetM2 49 TraceM2
 *   50 TransposeM2 51 InverseM2
 *   52 MakeMat3 53 AddM3 54 SubM3 55 ScaleM3 56 MulM3 57 MulM3V3 58 DetM3 59 TraceM3
 *   60 TransposeM3 61 InverseM3
 *   62 NegF (extension, not in the reference: exact sign flip)
 * Operand (u16): feature component index for LoadVar*, constant index for LoadConst*.
 * ValueType codes (src/engine/eval/types.rs:1-10):
 *   0 Float 1 Int 2 Bool 3 Vec2 4 Vec3 5 Mat2 6 Mat3.
 * Constant payloads are flat floats per constant: Float 1, Vec2 2, Vec3 3,
 *   Mat2 4, Mat3 9 (column-major, src/engine/eval/scalar.rs:108-112), Int/Bool 1.
 */

typedef struct rsr_context rsr_context;
typedef struct rsr_dataset rsr_dataset;
typedef struct rsr_programs rsr_programs;

RSR_API const char* rsr_error_string(int code);
RSR_API const char* rsr_last_error_message(void);
RSR_API const char* rsr_version(void);

/* Context: one device, one HIP stream, an HBM budget for datasets and programs
 * (0 = 90% of free device memory).  No reference counterpart (the reference is
 * in-process CPU code). */
RSR_API int rsr_context_create(int device_ordinal, uint64_t hbm_budget_bytes, rsr_context** out);
/* Fails with RSR_E_INVALID_ARG (and destroys nothing) while a server, a dataset or a
 * program batch of the context is alive: destroy those first. */
RSR_API int rsr_context_destroy(rsr_context* ctx);
RSR_API int rsr_context_memory(const rsr_context* ctx, uint64_t* used_bytes, uint64_t* budget_bytes);
/* Profiling: when enabled, HIP events on the context's stream bracket every main
 * evaluation kernel; _read returns the accumulated kernel milliseconds and launch
 * count since the last enable (and keeps accumulating). */
RSR_API int rsr_context_profile(rsr_context* ctx, int enable);
RSR_API int rsr_context_profile_read(rsr_context* ctx, double* kernel_ms, int64_t* launches);
/* Wait for all work queued on the context's stream. */
RSR_API int rsr_context_synchronize(rsr_context* ctx);
/* Summation of the per-point squared errors and gradient terms:
 *   RSR_SUM_REFERENCE (default): the reference's order -- 8 lane accumulators fed point
 *     by point (evaluator.rs:159-170) and the f32x8 horizontal add -- bit-identical to the
 *     CPU evaluator.  MSE keeps the per-point terms in a device arena and folds them in
 *     order; gradients use one wave per (program, tangent pass);
 *   RSR_SUM_KAHAN: per-tile pairwise sums combined with Kahan-compensated FP32: close to
 *     the exact sum (about 1e-6 relative to FP64), so it differs from the reference's
 *     plain f32 sums by their own rounding error, which grows with the number of points
 *     (up to 4e-5 relative at 10^5 points), and it is a few percent faster. */
#define RSR_SUM_KAHAN 0
#define RSR_SUM_REFERENCE 1
RSR_API int rsr_context_set_summation(rsr_context* ctx, int mode);

/* The launch form of the MSE evaluation kernel (no reference counterpart: the reference's
 * evaluator is a CPU loop, evaluator.rs:153-174).
 *   RSR_LAUNCH_ONESHOT: one workgroup per four work items (tile x program group, over every
 *     merged request), dealt to the CUs by the hardware dispatcher;
 *   RSR_LAUNCH_PERSISTENT: a launch with more work items than the device holds at once runs
 *     as a grid of the workgroups that fit, draining the items through a device work queue
 *     with one counter per XCD (a smaller launch stays one-shot);
 *   RSR_LAUNCH_AUTO (default): the form of the workload's tuning -- rsr_autotune times both
 *     (the work queue at two and four points per lane), the heuristic takes one-shot.  On
 *     the Lorentz workload in reference order the work queue was 2.6% faster at four points
 *     per lane; on the serving mix one-shot is faster.
 * Every form gives the same bits (every work item writes its own partial sums or terms). */
#define RSR_LAUNCH_ONESHOT 0
#define RSR_LAUNCH_PERSISTENT 1
#define RSR_LAUNCH_AUTO 2
RSR_API int rsr_context_set_launch_mode(rsr_context* ctx, int mode);
/* Graph mode (default off): rsr_eval_mse captures the evaluation of a program batch over a
 * dataset (its launches and the copy of the results) into a hipGraph on first use and
 * replays it on later calls with the same dataset, summation and tuning; new constants
 * (rsr_programs_set_constants, the fitters) are picked up.  Evaluations that need more than
 * one launch round run eagerly. */
RSR_API int rsr_context_set_graphs(rsr_context* ctx, int enable);
/* Upper bound (bytes, >= 4096) of the device arena that holds the per-tile Kahan partials
 * (or, in reference summation, the per-point terms) of one launch; larger evaluations are
 * split into tile ranges with the summation state carried between them, which gives the
 * same results.  Default 4 GiB. */
RSR_API int rsr_context_set_partials_limit(rsr_context* ctx, uint64_t bytes);
/* Evaluation tuning: points per lane (1 to 4: the wave's tile is 64 of them) and programs
 * per work item (the programs that share one tile's features in registers).  By default
 * (or after set_tuning(ctx, 0, 0)) each evaluation takes the tuning that rsr_autotune
 * found for its workload shape -- feature count, point count and program count, each
 * rounded to a power of two, and the summation mode -- or, before any, a heuristic (three points per lane and 8
 * programs per work item for large evaluations; four, or two for small ones, with up to 4
 * programs per item while that keeps 8,000 work items).  A nonzero setting fixes the tuning for
 * every evaluation; get_tuning returns 0, 0 in the per-workload mode.  No reference
 * counterpart (tuning of the device launches). */
RSR_API int rsr_context_set_tuning(rsr_context* ctx, int points_per_lane, int programs_per_item);
RSR_API int rsr_context_get_tuning(const rsr_context* ctx, int* points_per_lane, int* programs_per_item);
/* The tuning an MSE evaluation of this batch over this dataset takes now: points per lane,
 * programs per work item and launch form (1: the work queue) -- the fixed tuning, or the
 * autotuned or heuristic one of the workload's shape.  Lowers the batch for the dataset's
 * feature count first if that has not been done (as an evaluation would).  No reference
 * counterpart. */
RSR_API int rsr_context_query_tuning(rsr_context* ctx, rsr_programs* progs, const rsr_dataset* ds,
                                     int* points_per_lane, int* programs_per_item, int* persistent);
/* The autotuner: times every tuning (points per lane 1 to 4 x programs per work item 1, 2,
 * 4, 8, 16, 32, and in RSR_LAUNCH_AUTO mode the work-queue launch form at two and four points
 * per lane) on one MSE evaluation of this batch over this dataset and keeps the fastest for
 * the workload's shape, where direct calls, the server and the fitters find it; returns
 * the choice of points per lane and programs per item.  Costs about 108 evaluations (72
 * with a fixed launch form).  No reference counterpart. */
RSR_API int rsr_autotune(rsr_context* ctx, rsr_programs* progs, const rsr_dataset* ds, int* points_per_lane,
                         int* programs_per_item);

/* Dataset::new (src/engine/data/dataset.rs:37-169): rows [num_samples][F]
 * row-major where F = sum of the feature types' sizes; zero padding to a multiple
 * of 8 (counted in MSE sums, divisor = num_samples); optional z-score. */
RSR_API int rsr_dataset_create(rsr_context* ctx, const float* rows, const float* y, int64_t num_samples,
                               const uint8_t* feature_types, int n_feature_types, int normalize,
                               rsr_dataset** out);
/* A dataset already in the reference's AoSoA layout (Dataset.feature_flat /
 * target_batches, dataset.rs:17-34): feature_flat[(b*F + f)*8 + lane]. */
RSR_API int rsr_dataset_create_aosoa(rsr_context* ctx, const float* feature_flat, const float* target_batches,
                                     int num_features, int64_t num_batches, int64_t num_samples, rsr_dataset** out);
/* Datasets on disk.  Dataset::from_csv (dataset.rs:298-357): a header row, then records
 * with the header's field count; the target in column target_col (the last when
 * negative); the features in the other columns, left to right, as many as the feature
 * types' flat size; numbers as f32::from_str parses them (no whitespace, correctly
 * rounded).  Dataset::from_json (dataset.rs:265-291): [{"x": [...], "y": v}, ...] with
 * other keys ignored.  rsr_read_* return the rows [n][F] and targets [n] (both pointers
 * null: n only); rsr_dataset_from_file reads and then builds as rsr_dataset_create. */
#define RSR_FILE_CSV 0
#define RSR_FILE_JSON 1
RSR_API int rsr_read_csv(const char* path, const uint8_t* feature_types, int n_feature_types, int target_col,
                         float* rows_out, float* y_out, int64_t capacity, int64_t* n_out);
RSR_API int rsr_read_json(const char* path, const uint8_t* feature_types, int n_feature_types, float* rows_out,
                          float* y_out, int64_t capacity, int64_t* n_out);
RSR_API int rsr_dataset_from_file(rsr_context* ctx, const char* path, int format, const uint8_t* feature_types,
                                  int n_feature_types, int target_col, int normalize, rsr_dataset** out);
/* Dataset::subset (dataset.rs:360-447): strided sample of target_samples points. */
RSR_API int rsr_dataset_subset(rsr_context* ctx, const rsr_dataset* src, int64_t target_samples, rsr_dataset** out);
RSR_API int rsr_dataset_info(const rsr_dataset* ds, int* num_features, int64_t* num_samples, int64_t* num_batches);
/* z-score statistics (feature means[F], std devs[F], target mean, target std). */
RSR_API int rsr_dataset_stats(const rsr_dataset* ds, float* stats_out);
RSR_API int rsr_dataset_destroy(rsr_dataset* ds);

/* A batch of programs (Program::from_nodes, src/engine/expr/program.rs:34-55):
 * program i is ops/args[code_offsets[i] .. code_offsets[i+1]), its constants are
 * const_types[const_offsets[i] .. const_offsets[i+1]) with payloads starting at
 * const_values[value_offsets[i]]. */
RSR_API int rsr_programs_create(rsr_context* ctx, int n_programs, const uint8_t* ops, const uint16_t* args,
                                const int64_t* code_offsets, const uint8_t* const_types,
                                const int64_t* const_offsets, const float* const_values,
                                const int64_t* value_offsets, rsr_programs** out);
RSR_API int rsr_programs_count(const rsr_programs* progs, int* n_programs, int64_t* n_values);
/* Constants in the create-time layout (what the optimizers write back,
 * lbfgs.rs:154-164, cmaes.rs:181-191). */
RSR_API int rsr_programs_get_constants(const rsr_programs* progs, float* const_values_out);
RSR_API int rsr_programs_set_constants(rsr_programs* progs, const float* const_values);
/* Lowering status of every program against a dataset's feature count. */
RSR_API int rsr_programs_status(rsr_context* ctx, rsr_programs* progs, const rsr_dataset* ds, int32_t* status_out);
RSR_API int rsr_programs_destroy(rsr_programs* progs);

/* compute_mse (src/engine/eval/evaluator.rs:153-174) for every program:
 * non-finite -> FLT_MAX.  Rejected programs get NaN and the call returns
 * RSR_E_PROGRAM after evaluating the others. */
RSR_API int rsr_eval_mse(rsr_context* ctx, rsr_programs* progs, const rsr_dataset* ds, float* mse_out);

/* compute_mse with no device allocation (no reference counterpart: the reference is
 * in-process CPU code).  For callers that cannot allocate -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  The same evaluation as
 * rsr_eval_mse (same summation and tuning, hence the same bits), enqueued on the context's
 * stream without waiting: mse_out_device is a device pointer to n floats (NaN for programs
 * the lowering rejected), complete after rsr_context_synchronize.  Its scratch (per-tile
 * partials, or per-point terms in reference summation, and the spill areas of big programs)
 * is one 64 MiB device array per GPU; larger evaluations fail with RSR_E_INVALID_ARG.  The
 * batch must already be lowered for the dataset's feature count (rsr_programs_status or
 * any evaluation), else RSR_E_INVALID_ARG.  Calls that overlap on one GPU race on the
 * array: calls through one context are serialised; contexts sharing a GPU must not
 * overlap these calls. */
RSR_API int rsr_eval_mse_noalloc(rsr_context* ctx, rsr_programs* progs, const rsr_dataset* ds, float* mse_out_device);

/* eval_simd over every 8-point batch (evaluator.rs:20-150): per-point values of
 * programs [first, first+count), out[(i-first) * num_batches*8 + point],
 * padding lanes included. */
RSR_API int rsr_eval_points(rsr_context* ctx, rsr_programs* progs, int first, int count, const rsr_dataset* ds,
                            float* out);

/* compute_mse_with_gradient (evaluator.rs:177-224): mse_out[n], grad_out[n][32]
 * (gradient of the first min(param_count, 32) flat parameters, zeros after). */
RSR_API int rsr_eval_mse_grad(rsr_context* ctx, rsr_programs* progs, const rsr_dataset* ds, float* mse_out,
                              float* grad_out);

/* eval_simd_dual (evaluator.rs:228-388): value and tangent of one program w.r.t.
 * flat parameter `active_param`, per point. */
RSR_API int rsr_eval_points_dual(rsr_context* ctx, rsr_programs* progs, int program, int active_param,
                                 const rsr_dataset* ds, float* val_out, float* grad_out);

/* ---- constant fitting (src/engine/optimize/) ----
 * Fitted constants are written back into the program batch (device and host copies;
 * read them with rsr_programs_get_constants), like the reference's write-back into
 * Individual.constants (lbfgs.rs:154-164, cmaes.rs:181-191).  fitness_out[i] is the
 * optimizer's final MSE, NaN for programs not fitted (mask[i] == 0, rejected, or no
 * parameters).  mask may be NULL (fit every program). */
#define RSR_LBFGS_REFERENCE 0  /* run_lbfgs (lbfgs.rs:9-165): f32, M = history, its quirks */
#define RSR_LBFGS_FP64 1       /* FP64 two-loop dots, curvature pairs and Armijo test; f32 parameters */
RSR_API int rsr_fit_lbfgs(rsr_context* ctx, rsr_programs* progs, const rsr_dataset* ds, int max_iterations,
                          int history, int mode, const uint8_t* mask, float* fitness_out);
/* run_cma_es (cmaes.rs:19-192), diagonal, lambda candidates / mu parents per
 * individual (reference: 16 / 8), doubled dimension as the reference.  Normals come
 * from a Philox4x32-10 stream keyed by `seed` (see rsr_cmaes_normal), or, when
 * `normals` is non-NULL, from normals[((it * n_fit + k) * lambda + c) * 32 + j] with k
 * the fitted individuals in program order.  iters_out (optional): iterations run. */
RSR_API int rsr_fit_cmaes(rsr_context* ctx, rsr_programs* progs, const rsr_dataset* ds, int max_iterations,
                          int lambda, int mu, uint64_t seed, const float* normals, const uint8_t* mask,
                          float* fitness_out, int32_t* iters_out);
/* run_nelder_mead (nelder_mead.rs:9-165) over the Float constants (at most 32) of every
 * program with mask[i] != 0; programs without Float constants are left alone (fitness
 * NaN).  iters_out (optional): iterations run. */
RSR_API int rsr_fit_nelder_mead(rsr_context* ctx, rsr_programs* progs, const rsr_dataset* ds, int max_iterations,
                                const uint8_t* mask, float* fitness_out, int32_t* iters_out);
/* Individual::optimize_constants (individual.rs:65-75) for every program with mask[i] != 0,
 * followed by the caller's calculate_mse (island.rs:233-236): the dispatch of
 * optimize_individual_constants (optimize/mod.rs:10-59) -- L-BFGS (15 iterations, M = 6,
 * reference mode) for programs without a matrix inverse, else CMA-ES (16 / 8,
 * max(max_iterations / 16, 2) iterations, Philox stream `seed`) for programs with a
 * vector or matrix constant, else Nelder-Mead (max_iterations) -- then every constant
 * component with magnitude below `threshold` (reference: 0.05) set to zero, then the
 * MSE of the result into mse_out (NaN for programs not optimized).  The snapped
 * constants are written back to the batch; the reference's simplify_ast of the snapped
 * tree is the separate host call rsr_simplify_programs (a simplified program is a new
 * batch).  method_out (optional): RSR_FIT_*. */
#define RSR_FIT_NONE 0
#define RSR_FIT_LBFGS 1
#define RSR_FIT_CMAES 2
#define RSR_FIT_NELDER_MEAD 3
RSR_API int rsr_optimize_constants(rsr_context* ctx, rsr_programs* progs, const rsr_dataset* ds, int max_iterations,
                                   uint64_t seed, float threshold, const uint8_t* mask, float* mse_out,
                                   uint8_t* method_out);
/* The standard normal the CMA-ES sampler draws for (seed, program, iteration,
 * candidate, dimension): Philox4x32-10 -> two 24-bit uniforms -> Box-Muller as
 * cmaes.rs:13-17. */
RSR_API float rsr_cmaes_normal(uint64_t seed, uint32_t program, uint32_t iteration, uint32_t candidate, uint32_t dim);

/* ---- the search loop's scoring and selection (host code, no device needed) ----
 * Individual::complexity (individual.rs:108-110): the sum of node weights (node.rs:21-33,
 * basic/mod.rs:156-163, linalg/mod.rs:434-478) of each program in the wire format
 * (ops, code_offsets[n+1] as for rsr_programs_create); NegF weighs 1. */
RSR_API int rsr_complexity(int n_programs, const uint8_t* ops, const int64_t* code_offsets, int32_t* complexity_out);
/* Fitness of island.rs:241-255: mse + complexity * parsimony (reference default 5e-6,
 * config.rs:48), FLT_MAX for a non-finite MSE. */
RSR_API int rsr_fitness(int n, const float* mse, const int32_t* complexity, float parsimony, float* fitness_out);
/* Individual::simplify (individual.rs:77-85) of each program, then Program::from_nodes:
 * simplify_ast (src/engine/expr/simplify.rs:13-135) with the try_simplify rules of
 * basic/mod.rs:45-115 and linalg/mod.rs:227-310, 728-890 (constant folding, identity
 * and zero operands, x - x, x / x, ln(exp x), exp(ln x), sqrt(sqr x), sqr(sqrt x)).
 * Input and output use the layout of rsr_programs_create; the outputs are the compiled
 * programs (constants in postfix order, folded-away constants dropped).  Capacities:
 * the simplified code is never longer than the input code, and code_capacity entries
 * of constants with 9 * code_capacity payload floats always suffice.  Host code. */
RSR_API int rsr_simplify_programs(int n_programs, const uint8_t* ops, const uint16_t* args,
                                  const int64_t* code_offsets, const uint8_t* const_types,
                                  const int64_t* const_offsets, const float* const_values,
                                  const int64_t* value_offsets, int64_t code_capacity, int64_t const_capacity,
                                  int64_t value_capacity, uint8_t* ops_out, uint16_t* args_out,
                                  int64_t* code_offsets_out, uint8_t* const_types_out, int64_t* const_offsets_out,
                                  float* const_values_out, int64_t* value_offsets_out);
/* Island::evaluate_buffer (island.rs:215-261) of n_islands islands in one call.  Island g's
 * buffer is programs [island_offsets[g], island_offsets[g+1]) of the batch given in the
 * rsr_programs_create layout; best_fitness[g] is its best individual's fitness and
 * best_mini_mse[g] that individual's MSE on mini_batch (island.rs:217).  The island's
 * uniform draws (its rng.random::<f32>() stream) are draws[draw_offsets[g] ..
 * draw_offsets[g+1]), consumed in the reference's order: one when an individual's
 * mini-batch MSE is not under twice the baseline (5% rescue), one for every promising
 * individual (random fit with probability opt_prob).  Promising individuals are scored on
 * dataset; those under the island's running best fitness, or drawn, with a finite MSE get
 * Individual::optimize_constants (opt_iterations, threshold, CMA-ES stream seed), i.e. the
 * fit, the snap and simplify_ast, and are rescored.  Outputs per individual: the buffer
 * after the step (fitted individuals carry their simplified programs; capacities as for
 * rsr_simplify_programs over the whole batch), fitness (mse + complexity * parsimony, or
 * FLT_MAX), final MSE (the mini-batch MSE for individuals that were not promising, as in
 * the reference), whether it was fitted; per island: the buffer index of the new best
 * (-1: unchanged) and the draws consumed.  The islands' local hall of fame is the caller's:
 * it updates from (complexity, final MSE) in buffer order.  No reference counterpart for
 * the batching: the reference runs one island per rayon thread (engine.rs:52). */
RSR_API int rsr_islands_evaluate(rsr_context* ctx, const rsr_dataset* dataset, const rsr_dataset* mini_batch,
                                 int n_islands, const int64_t* island_offsets, const uint8_t* ops,
                                 const uint16_t* args, const int64_t* code_offsets, const uint8_t* const_types,
                                 const int64_t* const_offsets, const float* const_values,
                                 const int64_t* value_offsets, const float* best_fitness, const float* best_mini_mse,
                                 const float* draws, const int64_t* draw_offsets, float opt_prob, int opt_iterations,
                                 float parsimony, float threshold, uint64_t seed, int64_t code_capacity,
                                 int64_t const_capacity, int64_t value_capacity, uint8_t* ops_out, uint16_t* args_out,
                                 int64_t* code_offsets_out, uint8_t* const_types_out, int64_t* const_offsets_out,
                                 float* const_values_out, int64_t* value_offsets_out, float* fitness_out,
                                 float* mse_out, uint8_t* optimized_out, int32_t* best_index_out,
                                 int64_t* draws_used_out);
/* assign_rank_and_crowding_distance (operators/selection.rs:39-134): non-dominated fronts
 * on (fitness, age), crowding distance over fitness then age. */
RSR_API int rsr_assign_rank_and_crowding(int n, const float* fitness, const uint64_t* age, uint32_t* rank_out,
                                         float* crowding_out);
/* tournament_selection_pareto (selection.rs:13-26) given the k drawn indices (the caller's
 * random stream): the winner under nsga2_compare (lower rank, then larger crowding). */
RSR_API int rsr_tournament(int n, const uint32_t* rank, const float* crowding, int k, const uint32_t* draws,
                           uint32_t* winner_out);

/* ---- device-side variation (operators/{generator,mutation,crossover}.rs) ----
 * One offspring per task, all produced in one launch (one GPU thread each) from parents
 * given in the rsr_programs_create layout.  Each task draws from its own Xoshiro256++
 * stream keyed by `seed`, consuming draws in the order of the reference operator:
 *   GENERATE   generate_random_ast(type, depth) (generator.rs:9-148), depth <= 5
 *   CROSSOVER  crossover(a, b, max_size) (crossover.rs:7-65)
 *   POINT      point_mutation(a) (mutation.rs:11-57)
 *   CONSTANT   constant_perturbation(a) (mutation.rs:59-116)
 *   SUBTREE    subtree_mutation(a, max_size, depth) (mutation.rs:118-160)
 *   MUTATE     one of POINT / CONSTANT / SUBTREE drawn uniformly first (island.rs:156-179)
 * The offspring are compiled programs (Program::from_nodes order of constants), before
 * simplify_ast; at most 384 nodes each.  kind_out (optional): the operator applied. */
#define RSR_VARY_GENERATE 0
#define RSR_VARY_CROSSOVER 1
#define RSR_VARY_POINT 2
#define RSR_VARY_CONSTANT 3
#define RSR_VARY_SUBTREE 4
#define RSR_VARY_MUTATE 5
typedef struct rsr_vary_task {
    uint8_t kind;     /* RSR_VARY_* */
    uint8_t type;     /* GENERATE: the root's ValueType */
    uint8_t depth;    /* GENERATE: max depth; SUBTREE / MUTATE: mutation_max_depth */
    uint8_t reserved;
    int32_t a, b;     /* parent indices (b: CROSSOVER only) */
    int32_t max_size; /* CROSSOVER / SUBTREE / MUTATE: max_tree_size */
    uint64_t seed;    /* the task's stream */
} rsr_vary_task;
/* The variation streams on the host: raw Xoshiro256++ outputs from a state; the state
 * SplitMix64 seeding (rand_xoshiro seed_from_u64) gives a seed; n samples of the stream
 * seeded with `seed`: kind 0 f32 in [0, 1), 1 bool (0 / 1), 2 uniform integer below
 * `bound`, 3 f32 uniform in [a, b). */
RSR_API int rsr_xoshiro256pp(const uint64_t* state, int n, uint64_t* out);
RSR_API int rsr_xoshiro_seed(uint64_t seed, uint64_t* state_out);
RSR_API int rsr_vary_samples(uint64_t seed, int kind, int n, float a, float b, uint32_t bound, float* out);
RSR_API int rsr_vary(rsr_context* ctx, int n_tasks, const rsr_vary_task* tasks, int n_parents, const uint8_t* ops,
                     const uint16_t* args, const int64_t* code_offsets, const uint8_t* const_types,
                     const int64_t* const_offsets, const float* const_values, const int64_t* value_offsets,
                     int n_vars, const uint8_t* var_types, const uint8_t* var_indices, int n_allowed,
                     const uint8_t* allowed_ops, int64_t code_capacity, int64_t const_capacity,
                     int64_t value_capacity, uint8_t* ops_out, uint16_t* args_out, int64_t* code_offsets_out,
                     uint8_t* const_types_out, int64_t* const_offsets_out, float* const_values_out,
                     int64_t* value_offsets_out, uint8_t* kind_out);

/* ---- the search loop (src/engine/search/{island,engine}.rs) as a client of the evaluator ----
 * Engine::new / Engine::run with the islands of Island::new / step_generation /
 * fill_next_generation / evaluate_buffer / nuke, migration and the halls of fame.  One
 * generation of every island runs as merged device calls (variation launches, batched
 * evaluations, rsr_islands_evaluate).  Each island's sequential decisions (crossover or
 * mutation, tournaments, the evaluation step's draws) come from its Xoshiro256++ stream
 * seeded with seed + island (reference: 42 + i); the variation operators draw from
 * per-offspring streams (see rsr_vary).  The dataset must outlive the engine. */
typedef struct rsr_engine rsr_engine;
typedef struct rsr_engine_config {  /* Config (config.rs:11-66); defaults from rsr_engine_default_config */
    int32_t num_islands, island_size, max_generations;
    float crossover_rate;
    int32_t tournament_size, migration_interval;
    float parsimony_penalty, opt_prob;
    int32_t opt_iterations, final_opt_iterations, stagnation_threshold;
    float target_mse, min_improvement, random_injection_rate;
    int32_t min_random_injection, max_tree_size, mutation_max_depth, mutation_cycles, mini_batch_size;
    uint64_t seed;
} rsr_engine_config;
typedef struct rsr_engine_info {
    int64_t generation;    /* generations run */
    int64_t nukes;         /* island nukes so far */
    int32_t best_island;   /* get_global_best's island */
    float best_fitness;    /* its best individual's fitness */
    float best_mse;        /* that individual's MSE on the dataset after the last generation */
    int32_t target_reached, finished, hof_size;
} rsr_engine_info;
RSR_API void rsr_engine_default_config(rsr_engine_config* cfg);
/* variables: (ValueType, feature index) pairs, the registry of the reference; allowed_ops:
 * the operators (RefOp codes) the search may use (Strategy::get_allowed_operators). */
RSR_API int rsr_engine_create(rsr_context* ctx, const rsr_engine_config* cfg, const rsr_dataset* dataset,
                              int n_vars, const uint8_t* var_types, const uint8_t* var_indices, int n_allowed,
                              const uint8_t* allowed_ops, rsr_engine** out);
/* One iteration of Engine::run's loop: every island's step_generation, the halls of fame
 * merged, the global best rescored on the dataset (target check), migration every
 * migration_interval generations. */
RSR_API int rsr_engine_step(rsr_engine* e, int* target_reached);
/* Engine::run: steps until max_generations or the target, then the global best fitted
 * with final_opt_iterations, snapped, simplified and rescored as island 0's best. */
RSR_API int rsr_engine_run(rsr_engine* e, int* generations_out);
RSR_API int rsr_engine_get_info(const rsr_engine* e, rsr_engine_info* info);
/* Programs of the engine in the rsr_programs_create layout, with one value and one count
 * each: POPULATION (island: fitness, age), ISLAND_BEST (island: fitness, age), GLOBAL_BEST
 * (fitness, age), HALL_OF_FAME (global: MSE, complexity), ISLAND_HOF (island: MSE,
 * complexity), PARETO (get_pareto_front, engine.rs:131-155: MSE, complexity). */
#define RSR_ENGINE_POPULATION 0
#define RSR_ENGINE_ISLAND_BEST 1
#define RSR_ENGINE_GLOBAL_BEST 2
#define RSR_ENGINE_HALL_OF_FAME 3
#define RSR_ENGINE_ISLAND_HOF 4
#define RSR_ENGINE_PARETO 5
RSR_API int rsr_engine_programs(const rsr_engine* e, int which, int island, int64_t program_capacity,
                                int64_t code_capacity, int64_t const_capacity, int64_t value_capacity, int* n_out,
                                uint8_t* ops_out, uint16_t* args_out, int64_t* code_offsets_out,
                                uint8_t* const_types_out, int64_t* const_offsets_out, float* const_values_out,
                                int64_t* value_offsets_out, float* value_out, int64_t* count_out);
RSR_API int rsr_engine_destroy(rsr_engine* e);

/* ---------------------------------------------------------------------------------
 * Serving (the "async submit/poll" of the boundary; the reference's callers are rayon
 * workers calling the evaluator in-process, engine.rs:52, island.rs:150-238).
 *
 * A server owns one device thread.  Every admitted request is a state machine over
 * evaluation steps (an evaluation, a gradient, one L-BFGS line-search round, one CMA-ES
 * generation, ...).  Each loop iteration (a tick) runs the next step of every admitted
 * request together in merged launches, then each request's optimizer update, so requests
 * that arrive while others are on the GPU join at the next tick (continuous batching).
 * Requests are admitted in arrival order while fewer than max_inflight are in flight and
 * their device state fits budget_bytes (0: the context's budget); two requests on the
 * same program batch run one after the other.  A request's results are the same alone
 * and in any mix.
 * ------------------------------------------------------------------------------- */
typedef struct rsr_server rsr_server;
#define RSR_REQ_EVAL 0   /* compute_mse of the masked programs: out = mse[n] */
#define RSR_REQ_GRAD 1   /* compute_mse_with_gradient: out = mse[n], out2 = grad[n][32] */
#define RSR_REQ_LBFGS 2  /* rsr_fit_lbfgs: out = fitness[n]; constants updated in place */
#define RSR_REQ_CMAES 3  /* rsr_fit_cmaes: out = fitness[n], iters_out[n] (optional) */
#define RSR_REQ_NELDER_MEAD 4 /* rsr_fit_nelder_mead: out = fitness[n], iters_out[n] (optional) */
#define RSR_REQ_OPTIMIZE 5    /* rsr_optimize_constants (threshold 0.05): out = mse[n] after the fit */
typedef struct rsr_request {
    int kind;                    /* RSR_REQ_* */
    rsr_programs* programs;      /* not used by any other call until the request completes */
    const rsr_dataset* dataset;  /* read only; requests may share it */
    const uint8_t* mask;         /* optional [n]; copied at submit */
    int iterations;              /* L-BFGS, CMA-ES */
    int history;                 /* L-BFGS history (reference: 6) */
    int mode;                    /* L-BFGS: RSR_LBFGS_* */
    int lambda, mu;              /* CMA-ES */
    uint64_t seed;               /* CMA-ES normal stream */
    float* out;                  /* caller-owned outputs, valid until completion */
    float* out2;
    int32_t* iters_out;
} rsr_request;
typedef struct rsr_server_stats {
    uint64_t submitted, completed, failed;
    uint64_t queued, inflight, max_inflight_seen;
    uint64_t ticks;               /* loop iterations that ran at least one request */
    uint64_t node_evals;          /* program nodes x points evaluated by completed requests */
    uint64_t inflight_bytes;      /* device bytes held by in-flight requests */
    double busy_seconds;          /* wall time with at least one request in flight */
} rsr_server_stats;
RSR_API int rsr_server_create(rsr_context* ctx, int max_inflight, uint64_t budget_bytes, int max_queue,
                              rsr_server** out);
/* Queue a request; returns RSR_E_QUEUE_FULL when max_queue requests are waiting. */
RSR_API int rsr_server_submit(rsr_server* s, const rsr_request* req, uint64_t* ticket);
/* RSR_OK with *status = the request's result code once it completed, else RSR_E_NOT_READY. */
RSR_API int rsr_server_poll(rsr_server* s, uint64_t ticket, int* status);
/* Block until the request completes (timeout_seconds < 0: no limit); RSR_E_TIMEOUT otherwise. */
RSR_API int rsr_server_wait(rsr_server* s, uint64_t ticket, double timeout_seconds, int* status);
/* Seconds from submit to admission and from admission to completion, and the ticks it took. */
RSR_API int rsr_server_timing(rsr_server* s, uint64_t ticket, double* queued_seconds, double* run_seconds,
                              int32_t* ticks);
/* Forget a completed request (its timing is no longer available). */
RSR_API int rsr_server_release(rsr_server* s, uint64_t ticket);
RSR_API int rsr_server_get_stats(rsr_server* s, rsr_server_stats* out);
/* Stop after every queued and in-flight request has completed. */
RSR_API int rsr_server_destroy(rsr_server* s);

/* The search loop as a client of the serving loop (no reference counterpart for the
 * sharing: the reference's islands call the evaluator in-process, engine.rs:52): the
 * engine's evaluations and constant fits become requests to `server` (NULL: direct calls
 * again), batched with those of every other client of it, for instance other engines on
 * other threads; its variation stays a direct device launch.  The server's context must be
 * the engine's.  The results are the same bits as direct calls. */
RSR_API int rsr_engine_set_server(rsr_engine* e, rsr_server* server);

/* Lowering self-check (test hook, no reference counterpart, no GPU): lower one
 * program exactly as the evaluation entry points do and run the emitted micro-code
 * on the host over n points (feat[f * n + i]).  Returns the RSR_PROG_* status. */
RSR_API int rsr_lower_check(const uint8_t* ops, const uint16_t* args, int n_code, const uint8_t* const_types,
                            const float* const_values, int n_consts, int num_features, int dual, int active_param,
                            const float* feat, int64_t n, float* vals_out, float* tan_out, int32_t* n_uops,
                            int32_t* max_reg);

/* Lowered micro-code (test hook, no reference counterpart, no GPU): the micro-ops that
 * the evaluation entry points would run for one program, four words each (opcode and
 * operand words with their flags), into words_out[4 * capacity].  *n_uops receives the
 * full count.  Returns the RSR_PROG_* status. */
RSR_API int rsr_lower_code(const uint8_t* ops, const uint16_t* args, int n_code, const uint8_t* const_types,
                           const float* const_values, int n_consts, int num_features, int dual,
                           uint32_t* words_out, int32_t capacity, int32_t* n_uops);

#ifdef __cplusplus
}
#endif
#endif /* RSR_AMD_H */
