// backend.h -- where the search loop's evaluations and fits run: direct calls on a context,
// or requests to a server (rsr_server_*), where they are batched with every other client's.
#pragma once
#include <stdint.h>

#include "../../include/rsr_amd.h"

namespace rsr {

struct Backend {
    rsr_context* ctx = nullptr;
    rsr_server* srv = nullptr;  // null: direct calls on ctx
    // compute_mse of every program (rsr_eval_mse; RSR_E_PROGRAM when some were rejected)
    int eval(rsr_programs* p, const rsr_dataset* ds, float* out) const;
    // rsr_optimize_constants of every program (the server's requests snap at 0.05; another
    // threshold runs as a direct call)
    int optimize(rsr_programs* p, const rsr_dataset* ds, int iterations, uint64_t seed, float threshold,
                 float* mse_out) const;
};

// rsr_islands_evaluate on a backend (the same arguments after the context).
int islands_evaluate(const Backend& B, const rsr_dataset* dataset, const rsr_dataset* mini_batch, int n_islands,
                     const int64_t* island_offsets, const uint8_t* ops, const uint16_t* args,
                     const int64_t* code_offsets, const uint8_t* const_types, const int64_t* const_offsets,
                     const float* const_values, const int64_t* value_offsets, const float* best_fitness,
                     const float* best_mini_mse, const float* draws, const int64_t* draw_offsets, float opt_prob,
                     int opt_iterations, float parsimony, float threshold, uint64_t seed, int64_t code_capacity,
                     int64_t const_capacity, int64_t value_capacity, uint8_t* ops_out, uint16_t* args_out,
                     int64_t* code_offsets_out, uint8_t* const_types_out, int64_t* const_offsets_out,
                     float* const_values_out, int64_t* value_offsets_out, float* fitness_out, float* mse_out,
                     uint8_t* optimized_out, int32_t* best_index_out, int64_t* draws_used_out);

}  // namespace rsr
