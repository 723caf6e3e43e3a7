// isa.h -- instruction sets shared by the host lowering and the HIP kernels.
//
// Two instruction sets live here:
//   1. RefOp: the reference's postfix `Instruction` enum, numbered in
//      declaration order (src/engine/domain.rs:46-63 load variants, then
//      Basic (src/domains/basic/mod.rs:9-20) then Linalg
//      (src/domains/linalg/mod.rs:9-51), composed in src/lib.rs:44-48).
//      This is the wire format of the C ABI (opcode u8 + operand u16).
//   2. UOp: the device micro-ISA.  Every micro-op writes ONE f32 register
//      component and reads up to four register components, so vectors and
//      matrices are carried component by component in a flat per-lane
//      register file.  The host lowering maps the typed stacks onto that
//      file statically (the stack pointer of every typed stack is known at
//      every pc of a postfix program), so the kernel never keeps a stack
//      pointer.
#pragma once
#include <stdint.h>

namespace rsr {

enum RefOp : uint8_t {
    LoadVarF = 0, LoadConstF, LoadVarB, LoadConstB, LoadVarI, LoadConstI,
    LoadVarV2, LoadConstV2, LoadVarV3, LoadConstV3, LoadVarM2, LoadConstM2,
    LoadVarM3, LoadConstM3,
    AddF = 14, SubF, MulF, DivF, SinF, CosF, ExpF, SqrF, SqrtF, LnF,
    MakeVec2 = 24, MakeVec3, GetXV2, GetYV2, GetXV3, GetYV3, GetZV3,
    AddV2, SubV2, ScaleV2, DotV2, NormV2,
    AddV3, SubV3, ScaleV3, DotV3, NormV3, CrossV3,
    MakeMat2, AddM2, SubM2, ScaleM2, MulM2, MulM2V2, DetM2, TraceM2, TransposeM2, InverseM2,
    MakeMat3, AddM3, SubM3, ScaleM3, MulM3, MulM3V3, DetM3, TraceM3, TransposeM3, InverseM3,
    // Not a reference opcode: exact sign flip, used by the scalar workload's `neg`.
    NegF = 62,
    RefOpCount = 63
};

// ValueType (src/engine/eval/types.rs:1-10)
enum ValueType : uint8_t { VtFloat = 0, VtInt = 1, VtBool = 2, VtVec2 = 3, VtVec3 = 4, VtMat2 = 5, VtMat3 = 6 };

static inline int vt_size(int vt) {
    switch (vt) {
        case VtFloat: return 1;
        case VtVec2: return 2;
        case VtVec3: return 3;
        case VtMat2: return 4;
        case VtMat3: return 9;
        default: return 0;
    }
}

// Device micro-ops.  Encoding (four dwords, one scalar load, no field extraction for
// the register operands):
//   w0 = op | x << 8     x = IMM: flat parameter index; SPILL / FILL: LDS slot;
//                            otherwise c | e << 5 | flags << 10
//   w1 = d, w2 = a, w3 = b   register indices (< 32; 0 when unused) plus flags
//                            (kFwd, kVec3, kNoWb below)
// Semantics (every arithmetic step is one IEEE binary32 operation, left to
// right, never contracted into an FMA):
enum UOp : uint8_t {
    U_IMM = 0,    // d = params[x]
    U_ADD,        // d = a + b
    U_SUB,        // d = a - b
    U_MUL,        // d = a * b
    U_DIV,        // d = a / b
    U_SQR,        // d = a * a
    U_SQRT,       // d = sqrt(a)
    U_SIN,        // d = sin(a)
    U_COS,        // d = cos(a)
    U_EXP,        // d = exp(a)
    U_LN,         // d = ln(a)
    U_NEG,        // d = -a
    U_MULADD,     // d = (a * b) + (c * e)
    U_MULSUB,     // d = (a * b) - (c * e)
    U_ADDMUL,     // d = a + (b * c)
    U_SQRTACC,    // d = sqrt(a + (b * b))
    U_ADD3,       // d = (a + b) + c
    U_RCPSAFE,    // d = 1 / ((|a| < 1e-9) ? 1 : a)
    U_INVSEL,     // d = (|c| < 1e-9) ? K : (A * b), A = flags&1 ? -a : a, K = flags&2 ? 1 : 0
    U_SPILL,      // spill slot x <- a (d = a: the register is rewritten with its own value)
    U_FILL,       // d <- spill slot x
    // Vec3 micro-ops on three consecutive registers (operands a..a+2, b..b+2, results
    // d..d+2); every operand component is read before any result is written, so a
    // result block may overlap its operands
    U_ADDV3,      // d+k = a+k + b+k
    U_SUBV3,      // d+k = a+k - b+k
    U_SCALEV3,    // d+k = a+k * b                    (b: one register)
    U_IMMV3,      // d+k = params[x + k]
    U_CROSS3,     // d = (a1*b2) - (a2*b1), d+1 = (a2*b0) - (a0*b2), d+2 = (a0*b1) - (a1*b0)
    U_DOT3,       // d = ((a0*b0) + (a1*b1)) + (a2*b2) (one result register)
    U_NORM3,      // d = sqrt(((a0*a0) + (a1*a1)) + (a2*a2))
    // a Float constant as operand, read from the parameters (p = params[x]); the
    // constant is never materialised in a register
    U_ADDK,       // d = a + p
    U_SUBK,       // d = a - p
    U_KSUB,       // d = p - a
    U_MULK,       // d = a * p
    U_DIVK,       // d = a / p
    U_KDIV,       // d = p / a
    U_SCALEV3K,   // d+k = a+k * p
    // Mat3 (column-major, nine consecutive registers) block ops
    U_MULM3V3,    // d+i = ((a+i * b) + (a+3+i * b+1)) + (a+6+i * b+2), i < 3
    U_SCALEM3,    // d+k = a+k * b, k < 9
    U_SCALEM3K,   // d+k = a+k * p
    // Mat2 (column-major, four consecutive registers) block ops
    U_ADDV4,      // d+k = a+k + b+k, k < 4
    U_SUBV4,      // d+k = a+k - b+k
    U_SCALEV4,    // d+k = a+k * b
    U_SCALEV4K,   // d+k = a+k * p
    U_MULM2,      // d+r+2c = (a+r * b+2c) + (a+r+2 * b+2c+1)
    U_INVM2,      // inverse with the singular -> identity blend (no dual form: never lowered dual)
    // operand-swapped forms, emitted so that a forwarded operand is always a (value code only)
    U_RSUB,       // d = b - a
    U_RDIV,       // d = b / a
    U_RSUBV3,     // d+k = b+k - a+k
    U_RCROSS3,    // cross(b, a): d = (b1*a2) - (b2*a1), d+1 = (b2*a0) - (b0*a2), d+2 = (b0*a1) - (b1*a0)
    U_SCALEV3S,   // d+k = b+k * a                    (a: one register, b: Vec3)
    U_MULM3V3R,   // d+i = ((b+i * a) + (b+3+i * a+1)) + (b+6+i * a+2), i < 3   (a: Vec3, b: Mat3)
    // an operand that holds the same value at every point (it derives from constants
    // only): computed once per lane and copied to the lane's other points (value code only)
    U_SINU,       // d = sin(a)
    U_COSU,       // d = cos(a)
    U_EXPU,       // d = exp(a)
    U_LNU,        // d = ln(a)
    U_SQRTU,      // d = sqrt(a)
    // unary ops of a Float constant p = params[x], computed once per lane (value code only)
    U_SINK,       // d = sin(p)
    U_COSK,       // d = cos(p)
    U_EXPK,       // d = exp(p)
    U_LNK,        // d = ln(p)
    U_SQRTK,      // d = sqrt(p)
    U_NEGK,       // d = -p
    U_SQRK,       // d = p * p
    U_COUNT
};

// Spill slots of one program (per point of a lane, per channel).  Slots [0, kMaxSpillSlots)
// live in the workgroup's LDS; slots [kMaxSpillSlots, kMaxSlots) live in a per-wave area of
// device memory, which only the kernels for such programs ("big" programs, ProgDesc
// out_reg & kBigProg) provide.  kMaxSlots covers the reference's five typed stacks at their
// full depth of 32 (state.rs:12-37: 32 * (1 + 2 + 3 + 4 + 9) = 608 components) plus the
// temporaries of one operator.
constexpr int kMaxSpillSlots = 16;
constexpr int kMaxSlots = 640;
// ProgDesc.out_reg flag: the program spills beyond the LDS slots (see above).
constexpr uint32_t kBigProg = 1u << 8;

// Register file geometry: 32 f32 components per point per lane.  Dataset
// feature components occupy registers [0, F); temporaries use [F, 32).
constexpr int kRegs = 32;

// per-program lowering status (mirrors RSR_PROG_* in include/rsr_amd.h)
enum ProgStatus : int32_t {
    PROG_OK = 0,
    PROG_STACK = -1,
    PROG_OPCODE = -2,
    PROG_FEATURE = -3,
    PROG_CONST = -4,
    PROG_ROOT = -5,
    PROG_REGS = -6,
};

struct alignas(16) UInstr {
    uint32_t w0, w1, w2, w3;
};

static inline UInstr make_uinstr(uint32_t op, uint32_t d, uint32_t a, uint32_t b, uint32_t x) {
    return UInstr{op | (x << 8), d, a, b};
}
// Operand and result flags in the high bits of w1 (d), w2 (a) and w3 (b).  Register
// indices are the low five bits.  Only the value lowering sets them; the dual
// interpreter ignores them.
//   kFwd  on a: operand a is the whole result of the previous micro-op of the same
//         program, still held in the interpreter's accumulator (never reread from the
//         register file).
//   kVec3 on a or b: the operand has three components (a..a+2), staged before the
//         dispatch; on d: the result has three components (d..d+2), written after it.
//   kNone on a or b: the micro-op has no such operand (nothing is staged).
//   kNoWb on d: the result is not written to the register file.  Set only when the
//         next micro-op takes it through kFwd and no later micro-op reads it.
constexpr uint32_t kFwd = 1u << 8, kVec3 = 1u << 9, kNone = 1u << 10, kNoWb = 1u << 11;

// Register groups a micro-op reads and writes (component counts; 0 = unused).
struct UGroups {
    uint8_t na, nb, nc, ne, nd;
};
static inline UGroups ugroups(uint32_t op) {
    switch (op) {
        case U_IMM: case U_FILL: return {0, 0, 0, 0, 1};
        case U_ADD: case U_SUB: case U_MUL: case U_DIV: case U_SQRTACC: return {1, 1, 0, 0, 1};
        case U_SQR: case U_SQRT: case U_SIN: case U_COS: case U_EXP: case U_LN: case U_NEG: case U_RCPSAFE:
        case U_SPILL: case U_ADDK: case U_SUBK: case U_KSUB: case U_MULK: case U_DIVK: case U_KDIV:
            return {1, 0, 0, 0, 1};
        case U_MULADD: case U_MULSUB: return {1, 1, 1, 1, 1};
        case U_ADDMUL: case U_ADD3: case U_INVSEL: return {1, 1, 1, 0, 1};
        case U_ADDV3: case U_SUBV3: case U_CROSS3: return {3, 3, 0, 0, 3};
        case U_SCALEV3: return {3, 1, 0, 0, 3};
        case U_SCALEV3K: return {3, 0, 0, 0, 3};
        case U_IMMV3: return {0, 0, 0, 0, 3};
        case U_DOT3: return {3, 3, 0, 0, 1};
        case U_NORM3: return {3, 0, 0, 0, 1};
        case U_MULM3V3: return {9, 3, 0, 0, 3};
        case U_SCALEM3: return {9, 1, 0, 0, 9};
        case U_SCALEM3K: return {9, 0, 0, 0, 9};
        case U_ADDV4: case U_SUBV4: case U_MULM2: return {4, 4, 0, 0, 4};
        case U_SCALEV4: return {4, 1, 0, 0, 4};
        case U_SCALEV4K: case U_INVM2: return {4, 0, 0, 0, 4};
        case U_RSUB: case U_RDIV: return {1, 1, 0, 0, 1};
        case U_RSUBV3: case U_RCROSS3: return {3, 3, 0, 0, 3};
        case U_SCALEV3S: return {1, 3, 0, 0, 3};
        case U_MULM3V3R: return {3, 9, 0, 0, 3};
        case U_SINU: case U_COSU: case U_EXPU: case U_LNU: case U_SQRTU: return {1, 0, 0, 0, 1};
        case U_SINK: case U_COSK: case U_EXPK: case U_LNK: case U_SQRTK: case U_NEGK: case U_SQRK:
            return {0, 0, 0, 0, 1};
        default: return {0, 0, 0, 0, 0};
    }
}

// operand fields of x
static inline uint32_t ux_c(uint32_t x) { return x & 31u; }
static inline uint32_t ux_e(uint32_t x) { return (x >> 5) & 31u; }
static inline uint32_t ux_flags(uint32_t x) { return (x >> 10) & 3u; }
constexpr uint32_t kInvNeg = 1u, kInvOne = 2u;  // INVSEL flags: negate a; K = 1

// Per-program descriptor consumed by the kernels.
struct ProgDesc {
    uint32_t code_off;   // first micro-op in the code array
    uint32_t code_len;   // number of micro-ops
    uint32_t param_off;  // first flat parameter in the params array
    uint32_t out_reg;    // register holding the Float result (low five bits), | kBigProg
};

}  // namespace rsr
