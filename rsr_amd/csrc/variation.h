// variation.h -- the search loop's variation operators (operators/generator.rs,
// mutation.rs, crossover.rs) over programs in the compiled wire format, for the device
// variation kernel (one offspring per wave).
//
// A program is the postfix code of Program::from_nodes (program.rs:34-55): one opcode and
// one operand per node, constants in postfix order (LoadConst operand = the constant's
// rank among the program's LoadConst nodes), payloads packed in the same order.  The
// reference edits Node trees whose constant indices can point anywhere in
// Individual.constants and compiles them afterwards; editing the compiled form directly
// and renumbering as it is written gives the same compiled program, which is all the
// evaluator and the rest of the search ever read.
//
// Every operator consumes its random draws in the order of the reference function it
// follows (which draws happen, and when, under which conditions).  The draws come from a
// per-offspring Xoshiro256++ stream (the generator the reference's islands use,
// island.rs:24), sampled as documented at VRng below.
#pragma once
#include <stdint.h>

#include "isa.h"

namespace rsr {

constexpr int kVMaxNodes = 384;    // a depth-5 tree of arity-3 operators has 364 nodes
constexpr int kVMaxVals = 9 * kVMaxNodes;
constexpr int kVMaxDepth = 6;
constexpr uint8_t kVNone = 0xFF;

// ---------------------------------------------------------------------------------
// Random stream: Xoshiro256++ (Blackman & Vigna), seeded through SplitMix64 as
// rand_xoshiro's seed_from_u64.  Sampling (rand's published single-sample methods):
//   u32      the upper 32 bits of the 64-bit output
//   f32      (u32 >> 8) * 2^-24                        (rng.random::<f32>())
//   bool     the top bit of u32                        (rng.random::<bool>())
//   [0, n)   Canon's widening multiply on u32 with one correction draw when the low word
//            can carry                                  (rng.random_range(0..n))
//   [a, b)   ((u32 >> 9 | 0x3f800000) as f32 - 1) * (b - a) + a, redrawn if it reaches b
// ---------------------------------------------------------------------------------
struct VRng {
    uint64_t s[4];
};

__host__ __device__ inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline VRng vrng_seed(uint64_t seed) {
    VRng r;
    for (int i = 0; i < 4; ++i) r.s[i] = splitmix64(seed);
    return r;
}
__host__ __device__ inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
__host__ __device__ inline uint64_t vrng_next(VRng& r) {
    const uint64_t out = rotl64(r.s[0] + r.s[3], 23) + r.s[0];
    const uint64_t t = r.s[1] << 17;
    r.s[2] ^= r.s[0];
    r.s[3] ^= r.s[1];
    r.s[1] ^= r.s[2];
    r.s[0] ^= r.s[3];
    r.s[2] ^= t;
    r.s[3] = rotl64(r.s[3], 45);
    return out;
}
__host__ __device__ inline uint32_t vrng_u32(VRng& r) { return (uint32_t)(vrng_next(r) >> 32); }
__host__ __device__ inline float vrng_f32(VRng& r) { return (float)(vrng_u32(r) >> 8) * (1.0f / 16777216.0f); }
__host__ __device__ inline bool vrng_bool(VRng& r) { return (int32_t)vrng_u32(r) < 0; }
__host__ __device__ inline uint32_t vrng_below(VRng& r, uint32_t n) {  // n >= 1
    const uint64_t m = (uint64_t)vrng_u32(r) * n;
    uint32_t hi = (uint32_t)(m >> 32);
    const uint32_t lo = (uint32_t)m;
    if (lo > (uint32_t)(0u - n)) {
        const uint32_t hi2 = (uint32_t)(((uint64_t)vrng_u32(r) * n) >> 32);
        hi += (uint32_t)(lo + hi2 < lo);  // the carry of lo + hi2
    }
    return hi;
}
__host__ __device__ inline float vrng_range(VRng& r, float a, float b) {
    const float scale = b - a;
    for (;;) {
        const float u = __builtin_bit_cast(float, (vrng_u32(r) >> 9) | 0x3f800000u) - 1.0f;
        const float v = u * scale;
        const float res = v + a;
        if (res < b) return res;
    }
}

// ---------------------------------------------------------------------------------
// Opcode metadata (domain.rs:112-198; basic/mod.rs:117-195; linalg/mod.rs:312-491)
// ---------------------------------------------------------------------------------
__host__ __device__ inline bool v_is_load(int op) { return op < AddF; }
__host__ __device__ inline bool v_is_const(int op) { return op < AddF && (op & 1); }
__host__ __device__ inline int v_load_type(int op) {
    switch (op >> 1) {
        case 0: return VtFloat;
        case 1: return VtBool;
        case 2: return VtInt;
        case 3: return VtVec2;
        case 4: return VtVec3;
        case 5: return VtMat2;
        default: return VtMat3;
    }
}
__host__ __device__ inline int v_load_op(int vt, bool konst) {
    int base = 0;
    switch (vt) {
        case VtFloat: base = LoadVarF; break;
        case VtBool: base = LoadVarB; break;
        case VtInt: base = LoadVarI; break;
        case VtVec2: base = LoadVarV2; break;
        case VtVec3: base = LoadVarV3; break;
        case VtMat2: base = LoadVarM2; break;
        default: base = LoadVarM3; break;
    }
    return base + (konst ? 1 : 0);
}
// Constant payload floats (a Bool or Int constant carries one).
__host__ __device__ inline int v_payload(int vt) {
    switch (vt) {
        case VtVec2: return 2;
        case VtVec3: return 3;
        case VtMat2: return 4;
        case VtMat3: return 9;
        default: return 1;
    }
}
__host__ __device__ inline int v_arity(int op) {
    if (v_is_load(op)) return 0;
    switch (op) {
        case AddF: case SubF: case MulF: case DivF:
        case MakeVec2: case AddV2: case SubV2: case ScaleV2: case DotV2:
        case AddV3: case SubV3: case ScaleV3: case DotV3: case CrossV3:
        case MakeMat2: case AddM2: case SubM2: case ScaleM2: case MulM2: case MulM2V2:
        case AddM3: case SubM3: case ScaleM3: case MulM3: case MulM3V3:
            return 2;
        case MakeVec3: case MakeMat3: return 3;
        default: return 1;
    }
}
__host__ __device__ inline int v_ret(int op) {
    if (v_is_load(op)) return v_load_type(op);
    switch (op) {
        case MakeVec2: case AddV2: case SubV2: case ScaleV2: case MulM2V2: return VtVec2;
        case MakeVec3: case AddV3: case SubV3: case ScaleV3: case CrossV3: case MulM3V3: return VtVec3;
        case MakeMat2: case AddM2: case SubM2: case ScaleM2: case MulM2: case TransposeM2: case InverseM2: return VtMat2;
        case MakeMat3: case AddM3: case SubM3: case ScaleM3: case MulM3: case TransposeM3: case InverseM3: return VtMat3;
        default: return VtFloat;
    }
}
// Type of child k (expected_types order: the children's subtrees appear in this order).
__host__ __device__ inline int v_child(int op, int k) {
    switch (op) {
        case GetXV2: case GetYV2: case NormV2: case AddV2: case SubV2: case DotV2: return VtVec2;
        case ScaleV2: return k == 0 ? VtFloat : VtVec2;
        case GetXV3: case GetYV3: case GetZV3: case NormV3: case AddV3: case SubV3: case DotV3: case CrossV3:
            return VtVec3;
        case ScaleV3: return k == 0 ? VtFloat : VtVec3;
        case MakeMat2: return VtVec2;
        case AddM2: case SubM2: case MulM2: case DetM2: case TraceM2: case TransposeM2: case InverseM2: return VtMat2;
        case ScaleM2: return k == 0 ? VtFloat : VtMat2;
        case MulM2V2: return k == 0 ? VtMat2 : VtVec2;
        case MakeMat3: return VtVec3;
        case AddM3: case SubM3: case MulM3: case DetM3: case TraceM3: case TransposeM3: case InverseM3: return VtMat3;
        case ScaleM3: return k == 0 ? VtFloat : VtMat3;
        case MulM3V3: return k == 0 ? VtMat3 : VtVec3;
        default: return VtFloat;  // the Basic domain, MakeVec2 / MakeVec3
    }
}
__host__ __device__ inline bool v_same_children(int a, int b) {
    const int n = v_arity(a);
    if (n != v_arity(b)) return false;
    for (int k = 0; k < n; ++k)
        if (v_child(a, k) != v_child(b, k)) return false;
    return true;
}
// Instruction::is_forbidden_child: the domain rule when parent and child share a domain.
__host__ __device__ inline bool v_forbidden_child(int parent, int child) {
    switch (parent) {
        case SinF: case CosF: return child == SinF || child == CosF || child == ExpF || child == LnF;
        case ExpF: return child == ExpF || child == SinF || child == CosF || child == SqrF || child == LnF;
        case SqrtF:
            return child == SqrtF || child == SqrF || child == SinF || child == CosF || child == LnF || child == ExpF;
        case SqrF: return child == SqrF || child == SqrtF;
        case LnF: return child == LnF || child == ExpF || child == SinF || child == CosF;
        case TransposeM2: return child == TransposeM2;
        case TransposeM3: return child == TransposeM3;
        case InverseM2: return child == InverseM2;
        case InverseM3: return child == InverseM3;
        case GetXV2: case GetYV2: return child == MakeVec2;
        case GetXV3: case GetYV3: case GetZV3: return child == MakeVec3;
        default: return false;
    }
}

// ---------------------------------------------------------------------------------
// Program views
// ---------------------------------------------------------------------------------
struct VProg {  // read-only, compiled form
    const uint8_t* op;
    const uint16_t* arg;
    const uint8_t* ctype;  // per constant
    const float* cval;     // payloads, packed
    int len, nconst;
};
struct VOut {  // fixed-capacity output
    uint8_t* op;
    uint16_t* arg;
    uint8_t* ctype;
    float* cval;
    int len, nconst, nval;
    bool overflow;
};

struct VEnv {  // the island's variable registry and allowed operators
    const uint8_t* var_type;
    const uint8_t* var_idx;
    int n_vars;
    const uint8_t* ops;
    int n_ops;
};

__host__ __device__ inline void vout_node(VOut& o, int op, int arg) {
    if (o.len >= kVMaxNodes) {
        o.overflow = true;
        return;
    }
    o.op[o.len] = (uint8_t)op;
    o.arg[o.len] = (uint16_t)arg;
    o.len++;
}
// A constant node with its payload (the operand is its postfix rank).
__host__ __device__ inline void vout_const(VOut& o, int vt, const float* v) {
    const int n = v_payload(vt);
    if (o.len >= kVMaxNodes || o.nval + n > kVMaxVals) {
        o.overflow = true;
        return;
    }
    o.op[o.len] = (uint8_t)v_load_op(vt, true);
    o.arg[o.len] = (uint16_t)o.nconst;
    o.len++;
    o.ctype[o.nconst++] = (uint8_t)vt;
    for (int i = 0; i < n; ++i) o.cval[o.nval++] = v[i];
}
// Copy nodes [a, b) of p (constants renumbered, payloads copied).  `val` is the payload
// offset of the first constant at or after node a.
__host__ __device__ inline void vout_copy(VOut& o, const VProg& p, int a, int b, int& k, int& val) {
    for (int i = a; i < b; ++i) {
        const int op = p.op[i];
        if (v_is_const(op)) {
            vout_const(o, p.ctype[k], p.cval + val);
            val += v_payload(p.ctype[k]);
            k++;
        } else {
            vout_node(o, op, p.arg[i]);
        }
    }
}
// Constant rank and payload offset of the first constant at or after node i.
__host__ __device__ inline void v_const_pos(const VProg& p, int i, int& k, int& val) {
    k = 0;
    val = 0;
    for (int j = 0; j < i; ++j)
        if (v_is_const(p.op[j])) {
            val += v_payload(p.ctype[k]);
            k++;
        }
}
__host__ __device__ inline int v_node_type(const VProg& p, int i) { return v_ret(p.op[i]); }

// Individual::get_subtree_bounds (individual.rs:91-106)
__host__ __device__ inline void v_subtree(const VProg& p, int root, int& start, int& end) {
    int needed = 1;
    int cur = root;
    end = root;
    for (;;) {
        needed += v_arity(p.op[cur]) - 1;
        if (needed == 0) {
            start = cur;
            return;
        }
        if (cur == 0) break;
        cur--;
    }
    start = 0;
}

// ---------------------------------------------------------------------------------
// generator.rs
// ---------------------------------------------------------------------------------
// random_constant (generator.rs:185-221): every component uniform in [-5, 5); Bool a
// fair bit, Int uniform in [-10, 10).
__host__ __device__ inline void v_random_constant(VRng& r, int vt, float* v) {
    if (vt == VtBool) {
        v[0] = vrng_bool(r) ? 1.0f : 0.0f;
        return;
    }
    if (vt == VtInt) {
        v[0] = (float)((int)vrng_below(r, 20) - 10);
        return;
    }
    const int n = v_payload(vt);
    for (int i = 0; i < n; ++i) v[i] = vrng_range(r, -5.0f, 5.0f);
}
// random_operator (generator.rs:160-183): uniform over the allowed operators returning
// `vt` that the parent does not forbid; -1 when there is none (no draw then).
__host__ __device__ inline int v_random_operator(VRng& r, const VEnv& env, int vt, int parent) {
    int count = 0;
    for (int i = 0; i < env.n_ops; ++i) {
        const int op = env.ops[i];
        if (v_ret(op) == vt && !(parent != kVNone && v_forbidden_child(parent, op))) count++;
    }
    if (count == 0) return -1;
    int pick = (int)vrng_below(r, (uint32_t)count);
    for (int i = 0; i < env.n_ops; ++i) {
        const int op = env.ops[i];
        if (v_ret(op) == vt && !(parent != kVNone && v_forbidden_child(parent, op)))
            if (pick-- == 0) return op;
    }
    return -1;
}
// A leaf of type vt (generator.rs:52-82): the constant is drawn first in every case; with
// a variable of the type available, a fair bit chooses the variable (then uniform over
// them) or the constant.
__host__ __device__ inline void v_leaf(VOut& o, VRng& r, const VEnv& env, int vt) {
    float c[9];
    v_random_constant(r, vt, c);
    int n_valid = 0;
    for (int i = 0; i < env.n_vars; ++i) n_valid += env.var_type[i] == vt;
    if (n_valid > 0 && vrng_bool(r)) {
        int pick = (int)vrng_below(r, (uint32_t)n_valid);
        for (int i = 0; i < env.n_vars; ++i)
            if (env.var_type[i] == vt && pick-- == 0) {
                vout_node(o, v_load_op(vt, false), env.var_idx[i]);
                return;
            }
    }
    vout_const(o, vt, c);
}
// generate_random_ast (generator.rs:9-148) appended to o: build_ast_recursive with an
// explicit stack (frame d is the node at depth d).  A node is a leaf at max_depth, or
// (below the root) with probability 0.2; otherwise an allowed operator whose children
// are built in expected_types order and the operator emitted after them; with no
// allowed operator for the type it falls back to a leaf.
__host__ __device__ inline void v_generate(VOut& o, VRng& r, const VEnv& env, int vt, int max_depth) {
    if (max_depth > kVMaxDepth) max_depth = kVMaxDepth;
    uint8_t f_type[kVMaxDepth + 1], f_op[kVMaxDepth + 1], f_next[kVMaxDepth + 1];
    int sp = 0;
    f_type[0] = (uint8_t)vt;
    f_op[0] = kVNone;
    f_next[0] = 0;
    while (sp >= 0 && !o.overflow) {
        if (f_op[sp] == kVNone) {  // first visit
            const bool term = sp >= max_depth || (sp > 0 && vrng_f32(r) < 0.2f);
            int op = -1;
            if (!term) op = v_random_operator(r, env, f_type[sp], sp > 0 ? f_op[sp - 1] : kVNone);
            if (op < 0) {
                v_leaf(o, r, env, f_type[sp]);
                if (--sp >= 0) f_next[sp]++;
                continue;
            }
            f_op[sp] = (uint8_t)op;
        }
        const int op = f_op[sp];
        if (f_next[sp] < v_arity(op)) {
            const int child = v_child(op, f_next[sp]);
            sp++;
            f_type[sp] = (uint8_t)child;
            f_op[sp] = kVNone;
            f_next[sp] = 0;
            continue;
        }
        vout_node(o, op, 0);
        if (--sp >= 0) f_next[sp]++;
    }
}

// ---------------------------------------------------------------------------------
// mutation.rs
// ---------------------------------------------------------------------------------
// point_mutation (mutation.rs:11-57): one uniform node; a leaf becomes a new leaf of its
// type (constant drawn first, then the variable-or-constant choice); an operator becomes
// a random operator of its return type (no parent restriction) only when that has the
// same children types.
__host__ __device__ inline void v_point_mutation(VOut& o, VRng& r, const VEnv& env, const VProg& p) {
    int k = 0, val = 0;
    if (p.len == 0) return;
    const int idx = (int)vrng_below(r, (uint32_t)p.len);
    const int old = p.op[idx];
    const int vt = v_node_type(p, idx);
    vout_copy(o, p, 0, idx, k, val);
    if (v_arity(old) == 0) {
        float c[9];
        v_random_constant(r, vt, c);
        int n_valid = 0;
        for (int i = 0; i < env.n_vars; ++i) n_valid += env.var_type[i] == vt;
        bool done = false;
        if (n_valid > 0 && vrng_bool(r)) {
            int pick = (int)vrng_below(r, (uint32_t)n_valid);
            for (int i = 0; i < env.n_vars && !done; ++i)
                if (env.var_type[i] == vt && pick-- == 0) {
                    vout_node(o, v_load_op(vt, false), env.var_idx[i]);
                    done = true;
                }
        }
        if (!done) vout_const(o, vt, c);
        if (v_is_const(old)) {  // the replaced constant's payload is dropped
            val += v_payload(p.ctype[k]);
            k++;
        }
    } else {
        const int nop = v_random_operator(r, env, vt, kVNone);
        vout_node(o, (nop >= 0 && v_same_children(nop, old)) ? nop : old, 0);
    }
    vout_copy(o, p, idx + 1, p.len, k, val);
}

// perturb_constant (mutation.rs:82-116)
__host__ __device__ inline void v_perturb(VRng& r, int vt, float* v) {
    const float u = vrng_f32(r);
    if (vt == VtFloat) {
        if (u < 0.8f) {
            v[0] = v[0] * vrng_range(r, 0.9f, 1.1f);
        } else if (u < 0.9f) {
            v[0] = v[0] + vrng_range(r, -0.1f, 0.1f);
        } else {
            v[0] = vrng_range(r, -5.0f, 5.0f);
        }
        return;
    }
    if (vt == VtBool || vt == VtInt) return;
    const int n = v_payload(vt);
    for (int i = 0; i < n; ++i) v[i] = v[i] + vrng_range(r, -0.5f, 0.5f);
}
// constant_perturbation (mutation.rs:59-80): reservoir choice over the constant nodes
// (one draw per constant node), then perturb_constant of the chosen one.
__host__ __device__ inline void v_constant_perturbation(VOut& o, VRng& r, const VProg& p) {
    int target = -1, count = 0, kc = 0;
    for (int i = 0; i < p.len; ++i)
        if (v_is_const(p.op[i])) {
            count++;
            if (vrng_below(r, (uint32_t)count) == 0) target = kc;
            kc++;
        }
    int k = 0, val = 0;
    vout_copy(o, p, 0, p.len, k, val);
    if (target < 0 || o.overflow) return;
    int off = 0;
    for (int j = 0; j < target; ++j) off += v_payload(p.ctype[j]);
    v_perturb(r, p.ctype[target], o.cval + off);
}

// subtree_mutation (mutation.rs:118-160): a uniform node's subtree is replaced by a new
// random tree of its type and depth mutation_max_depth, unless the rest of the program
// alone already reaches max_size (then the program is unchanged and nothing more is drawn).
__host__ __device__ inline void v_subtree_mutation(VOut& o, VRng& r, const VEnv& env, const VProg& p, int max_size,
                                                   int mut_depth) {
    int k = 0, val = 0;
    if (p.len == 0) return;
    const int point = (int)vrng_below(r, (uint32_t)p.len);
    int start, end;
    v_subtree(p, point, start, end);
    const int vt = v_node_type(p, end);
    const int rest = p.len - (end - start + 1);
    if (max_size - rest <= 0) {
        vout_copy(o, p, 0, p.len, k, val);
        return;
    }
    vout_copy(o, p, 0, start, k, val);
    v_generate(o, r, env, vt, mut_depth);
    int k2, val2;
    v_const_pos(p, end + 1, k2, val2);
    vout_copy(o, p, end + 1, p.len, k2, val2);
}

// ---------------------------------------------------------------------------------
// crossover.rs
// ---------------------------------------------------------------------------------
// select_node_index (crossover.rs:67-101): among the first 64 nodes of the required type,
// an internal node with probability 0.9 when there is one, else any of them.
__host__ __device__ inline int v_select_node(VRng& r, const VProg& p, int required) {
    constexpr int kMax = 64;
    int internal[kMax], all[kMax];
    int n_int = 0, n_all = 0;
    for (int i = 0; i < p.len; ++i) {
        if (required >= 0 && v_node_type(p, i) != required) continue;
        if (n_all < kMax) all[n_all++] = i;
        if (v_arity(p.op[i]) > 0 && n_int < kMax) internal[n_int++] = i;
    }
    if (n_all == 0) return -1;
    if (n_int > 0 && vrng_f32(r) < 0.9f) return internal[vrng_below(r, (uint32_t)n_int)];
    return all[vrng_below(r, (uint32_t)n_all)];
}
// crossover (crossover.rs:7-65): a's subtree at a chosen node is replaced by b's subtree
// at a chosen node of the same type; a copy of a when either is empty, no node of the
// type exists in b, or the child would exceed max_size.
__host__ __device__ inline void v_crossover(VOut& o, VRng& r, const VProg& a, const VProg& b, int max_size) {
    int k = 0, val = 0;
    if (a.len == 0 || b.len == 0) {
        vout_copy(o, a, 0, a.len, k, val);
        return;
    }
    const int ra = v_select_node(r, a, -1);
    const int rb = ra < 0 ? -1 : v_select_node(r, b, v_node_type(a, ra));
    if (rb < 0) {
        vout_copy(o, a, 0, a.len, k, val);
        return;
    }
    int sa, ea, sb, eb;
    v_subtree(a, ra, sa, ea);
    v_subtree(b, rb, sb, eb);
    const int new_len = sa + (eb - sb + 1) + (a.len - ea - 1);
    if (new_len > max_size) {
        vout_copy(o, a, 0, a.len, k, val);
        return;
    }
    vout_copy(o, a, 0, sa, k, val);
    int kb, vb;
    v_const_pos(b, sb, kb, vb);
    vout_copy(o, b, sb, eb + 1, kb, vb);
    int ka, va;
    v_const_pos(a, ea + 1, ka, va);
    vout_copy(o, a, ea + 1, a.len, ka, va);
}

}  // namespace rsr
