// lower.h -- host lowering of reference postfix programs to device micro-code.
#pragma once
#include <stdint.h>
#include <vector>
#include "isa.h"

namespace rsr {

// One program as the C ABI carries it (reference Program, program.rs:17-23):
// the postfix Instruction stream and the constants in postfix order.
struct RefProgram {
    const uint8_t* ops;
    const uint16_t* args;
    int n_code;
    const uint8_t* const_types;  // ValueType per constant
    const float* const_values;   // payloads; Int/Bool constants take one slot
    int n_consts;
};

struct Lowered {
    int32_t status = PROG_OK;
    std::vector<UInstr> code;
    uint32_t out_reg = 0;
    uint32_t n_params = 0;      // flat parameter count (program.rs:85-98), uncapped
    uint32_t max_reg = 0;       // highest register written + 1
    uint32_t spill_slots = 0;   // LDS spill slots used (0 = none)
    bool differentiable = true; // no InverseM2/M3 (linalg/mod.rs:493-495)
    bool has_vecmat_const = false; // any Vec/Mat constant (optimize/mod.rs:40-48)
    uint32_t n_float_consts = 0;   // Float constants (Nelder-Mead dimension, nelder_mead.rs:23-31)
};

// Lower one program for a dataset with `num_features` flat feature
// components.  `dual` selects the forward-AD variant: Inverse ops become the
// identity, matching the reference's no-op dual inverse
// (src/domains/linalg/mod.rs:643, 687).
Lowered lower_program(const RefProgram& p, int num_features, bool dual);

// Flat parameters of a program: constants flattened in postfix order
// (Parameterized::flatten_params, program.rs:100-145), uncapped.
void flatten_params(const RefProgram& p, std::vector<float>& out);

}  // namespace rsr
