// variation_host.h -- host entry of the device variation kernel (variation.hip).
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/rsr_amd.h"
#include "wire.h"

namespace rsr {

// The island's variable registry (type, feature index) and allowed operators.
struct VaryEnv {
    std::vector<uint8_t> var_type, var_idx, ops;
};

// Run every task on the device; out receives one offspring per task, in task order.
// applied (optional): the operator each task applied (RSR_VARY_*; a mutate task reports
// the mutation it drew).  Takes ctx->mu.
int vary(rsr_context* ctx, const std::vector<rsr_vary_task>& tasks, const Wire& parents, const VaryEnv& env,
         Wire& out, std::vector<uint8_t>* applied);

}  // namespace rsr
