// simplify.cpp -- the post-fit program simplification of Individual::optimize_constants
// (src/engine/search/individual.rs:65-85), host code.
//
// One postfix pass, as simplify_ast (src/engine/expr/simplify.rs:13-135): every emitted
// subtree is tracked by where it starts in the output and, when it is a single constant,
// by that constant's value.  An operator whose arguments allow it is folded into a new
// constant, replaced by one of its arguments, or (ln∘exp, exp∘ln, sqrt∘sqr, sqr∘sqrt)
// cancelled against its child; the rules are the domains' try_simplify
// (src/domains/basic/mod.rs:45-115, src/domains/linalg/mod.rs:227-310 with the folds of
// :728-890).  The result is then compiled as Program::from_nodes
// (src/engine/expr/program.rs:34-52): constants renumbered in postfix order, and
// constants whose nodes were dropped are not carried.
//
// Folded transcendentals call the C library's sinf / cosf / expf / logf, the routines
// Rust's f32::sin / cos / exp / ln call on Linux; every other fold is a sequence of IEEE
// binary32 operations in the reference's order (this file is built without FP
// contraction).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/rsr_amd.h"
#include "isa.h"
#include "runtime_internal.h"

using namespace rsr;

namespace {

struct Const {
    uint8_t vt = VtFloat;
    float v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

Const cfloat(float x) {
    Const c;
    c.vt = VtFloat;
    c.v[0] = x;
    return c;
}
Const czero(uint8_t vt) {
    Const c;
    c.vt = vt;
    return c;
}
template <int N>
Const cvec(uint8_t vt, const float (&x)[N]) {
    Const c;
    c.vt = vt;
    for (int i = 0; i < N; ++i) c.v[i] = x[i];
    return c;
}

// Scalar::is_zero / is_one / is_identity (src/engine/eval/scalar.rs:16-57)
bool is_zero(const Const& c) {
    const int n = c.vt == VtFloat ? 1 : vt_size(c.vt);
    if (c.vt == VtInt || c.vt == VtBool) return false;
    for (int i = 0; i < n; ++i)
        if (!(fabsf(c.v[i]) < 1e-6f)) return false;
    return true;
}
bool is_one(const Const& c) { return c.vt == VtFloat && fabsf(c.v[0] - 1.0f) < 1e-6f; }
bool is_identity(const Const& c) {
    if (c.vt == VtMat2) {
        const float* m = c.v;
        return fabsf(m[0] - 1.0f) < 1e-6f && fabsf(m[1]) < 1e-6f && fabsf(m[2]) < 1e-6f && fabsf(m[3] - 1.0f) < 1e-6f;
    }
    if (c.vt == VtMat3) {
        const float* m = c.v;
        for (int i = 0; i < 9; ++i) {
            const bool diag = i == 0 || i == 4 || i == 8;
            if (!(fabsf(diag ? m[i] - 1.0f : m[i]) < 1e-6f)) return false;
        }
        return true;
    }
    return false;
}

bool is_load_var(int op) { return op <= LoadConstM3 && (op & 1) == 0; }
bool is_load_const(int op) { return op <= LoadConstM3 && (op & 1) == 1; }
uint8_t load_type(int op) {
    static const uint8_t t[7] = {VtFloat, VtBool, VtInt, VtVec2, VtVec3, VtMat2, VtMat3};
    return t[op >> 1];
}
uint8_t load_const_op(uint8_t vt) {
    switch (vt) {
        case VtFloat: return LoadConstF;
        case VtBool: return LoadConstB;
        case VtInt: return LoadConstI;
        case VtVec2: return LoadConstV2;
        case VtVec3: return LoadConstV3;
        case VtMat2: return LoadConstM2;
        default: return LoadConstM3;
    }
}

// Domain::arity and Domain::return_type (basic/mod.rs:117-133, linalg/mod.rs:312-390);
// NegF is unary and Float.
int arity(int op) {
    switch (op) {
        case AddF: case SubF: case MulF: case DivF: return 2;
        case MakeVec3: case MakeMat3: return 3;
        case MakeVec2: case MakeMat2: case AddV2: case SubV2: case ScaleV2: case DotV2: case AddV3: case SubV3:
        case ScaleV3: case DotV3: case CrossV3: case AddM2: case SubM2: case ScaleM2: case MulM2: case MulM2V2:
        case AddM3: case SubM3: case ScaleM3: case MulM3: case MulM3V3:
            return 2;
        default: return 1;
    }
}
uint8_t return_type(int op) {
    switch (op) {
        case MakeVec2: case AddV2: case SubV2: case ScaleV2: case MulM2V2: return VtVec2;
        case MakeVec3: case AddV3: case SubV3: case ScaleV3: case CrossV3: case MulM3V3: return VtVec3;
        case MakeMat2: case AddM2: case SubM2: case ScaleM2: case MulM2: case TransposeM2: case InverseM2: return VtMat2;
        case MakeMat3: case AddM3: case SubM3: case ScaleM3: case MulM3: case TransposeM3: case InverseM3: return VtMat3;
        default: return VtFloat;
    }
}

enum Act { None_, Replace, Keep };
struct Action {
    Act act = None_;
    Const c;   // Replace
    int keep = 0;  // Keep: index of the argument that replaces the operator
};
Action replace(const Const& c) {
    Action a;
    a.act = Replace;
    a.c = c;
    return a;
}
Action keep(int i) {
    Action a;
    a.act = Keep;
    a.keep = i;
    return a;
}

// Basic domain (basic/mod.rs:45-115).  NegF (not a reference opcode) folds to the exact
// negation and has no other rule.
Action basic_rule(int op, const Const* const* cv, int n, bool args_equal) {
    bool all_const = n > 0;
    for (int i = 0; i < n; ++i) all_const = all_const && cv[i] != nullptr;
    if (all_const && cv[0]->vt == VtFloat) {
        const float a = cv[0]->v[0];
        if (n == 1) {
            float r = NAN;
            switch (op) {
                case SinF: r = sinf(a); break;
                case CosF: r = cosf(a); break;
                case ExpF: r = expf(a); break;
                case SqrF: r = a * a; break;
                case SqrtF: r = sqrtf(a); break;
                case LnF: r = logf(a); break;
                case NegF: r = -a; break;
                default: break;
            }
            if (isfinite(r)) return replace(cfloat(r));
        } else if (n == 2 && cv[1]->vt == VtFloat) {
            const float b = cv[1]->v[0];
            float r = NAN;
            switch (op) {
                case AddF: r = a + b; break;
                case SubF: r = a - b; break;
                case MulF: r = a * b; break;
                case DivF: r = a / b; break;
                default: break;
            }
            if (isfinite(r)) return replace(cfloat(r));
        }
    }
    if (n == 2) {
        const bool a0 = cv[0] && is_zero(*cv[0]), b0 = cv[1] && is_zero(*cv[1]);
        const bool a1 = cv[0] && is_one(*cv[0]), b1 = cv[1] && is_one(*cv[1]);
        switch (op) {
            case AddF:
                if (b0) return keep(0);
                if (a0) return keep(1);
                break;
            case MulF:
                if (b1) return keep(0);
                if (a1) return keep(1);
                if (a0 || b0) return replace(cfloat(0.0f));
                break;
            case SubF:
                if (b0) return keep(0);
                if (args_equal) return replace(cfloat(0.0f));
                break;
            case DivF:
                if (b1) return keep(0);
                if (a0 && !b0) return replace(cfloat(0.0f));
                if (args_equal) return replace(cfloat(1.0f));
                break;
            default: break;
        }
    }
    return Action{};
}

// zero_for_op (linalg/mod.rs:728-741)
Const zero_for(int op) {
    switch (op) {
        case SubV2: case ScaleV2: case MulM2V2: return czero(VtVec2);
        case SubV3: case ScaleV3: case CrossV3: case MulM3V3: return czero(VtVec3);
        case SubM2: case ScaleM2: case MulM2: return czero(VtMat2);
        case SubM3: case ScaleM3: case MulM3: return czero(VtMat3);
        default: return cfloat(0.0f);
    }
}

float det3(const float* m) {
    return ((m[0] * ((m[4] * m[8]) - (m[5] * m[7]))) - (m[3] * ((m[1] * m[8]) - (m[2] * m[7])))) +
           (m[6] * ((m[1] * m[5]) - (m[2] * m[4])));
}

// fold_linalg_constants (linalg/mod.rs:743-890): the operator on constant arguments of
// the types it expects; false when the types do not match or the inverse is singular.
bool linalg_fold(int op, const Const* const* cv, int n, Const& out) {
    auto is = [&](int i, uint8_t vt) { return i < n && cv[i]->vt == vt; };
    const float* a = cv[0]->v;
    const float* b = n > 1 ? cv[1]->v : nullptr;
    switch (op) {
        case MakeVec2:
            if (!(n == 2 && is(0, VtFloat) && is(1, VtFloat))) return false;
            out = cvec<2>(VtVec2, {a[0], b[0]});
            return true;
        case MakeVec3:
            if (!(n == 3 && is(0, VtFloat) && is(1, VtFloat) && is(2, VtFloat))) return false;
            out = cvec<3>(VtVec3, {a[0], b[0], cv[2]->v[0]});
            return true;
        case GetXV2: case GetYV2:
            if (!is(0, VtVec2)) return false;
            out = cfloat(a[op == GetXV2 ? 0 : 1]);
            return true;
        case GetXV3: case GetYV3: case GetZV3:
            if (!is(0, VtVec3)) return false;
            out = cfloat(a[op - GetXV3]);
            return true;
        case AddV2: case SubV2:
            if (!(is(0, VtVec2) && is(1, VtVec2))) return false;
            out = op == AddV2 ? cvec<2>(VtVec2, {a[0] + b[0], a[1] + b[1]}) : cvec<2>(VtVec2, {a[0] - b[0], a[1] - b[1]});
            return true;
        case ScaleV2:
            if (!(is(0, VtFloat) && is(1, VtVec2))) return false;
            out = cvec<2>(VtVec2, {a[0] * b[0], a[0] * b[1]});
            return true;
        case DotV2:
            if (!(is(0, VtVec2) && is(1, VtVec2))) return false;
            out = cfloat((a[0] * b[0]) + (a[1] * b[1]));
            return true;
        case NormV2:
            if (!is(0, VtVec2)) return false;
            out = cfloat(sqrtf((a[0] * a[0]) + (a[1] * a[1])));
            return true;
        case AddV3: case SubV3:
            if (!(is(0, VtVec3) && is(1, VtVec3))) return false;
            out = op == AddV3 ? cvec<3>(VtVec3, {a[0] + b[0], a[1] + b[1], a[2] + b[2]})
                              : cvec<3>(VtVec3, {a[0] - b[0], a[1] - b[1], a[2] - b[2]});
            return true;
        case ScaleV3:
            if (!(is(0, VtFloat) && is(1, VtVec3))) return false;
            out = cvec<3>(VtVec3, {a[0] * b[0], a[0] * b[1], a[0] * b[2]});
            return true;
        case DotV3:
            if (!(is(0, VtVec3) && is(1, VtVec3))) return false;
            out = cfloat(((a[0] * b[0]) + (a[1] * b[1])) + (a[2] * b[2]));
            return true;
        case NormV3:
            if (!is(0, VtVec3)) return false;
            out = cfloat(sqrtf(((a[0] * a[0]) + (a[1] * a[1])) + (a[2] * a[2])));
            return true;
        case CrossV3:
            if (!(is(0, VtVec3) && is(1, VtVec3))) return false;
            out = cvec<3>(VtVec3, {(a[1] * b[2]) - (a[2] * b[1]), (a[2] * b[0]) - (a[0] * b[2]),
                                   (a[0] * b[1]) - (a[1] * b[0])});
            return true;
        case MakeMat2:
            if (!(is(0, VtVec2) && is(1, VtVec2))) return false;
            out = cvec<4>(VtMat2, {a[0], a[1], b[0], b[1]});
            return true;
        case AddM2: case SubM2: {
            if (!(is(0, VtMat2) && is(1, VtMat2))) return false;
            out = czero(VtMat2);
            for (int i = 0; i < 4; ++i) out.v[i] = op == AddM2 ? a[i] + b[i] : a[i] - b[i];
            return true;
        }
        case ScaleM2: {
            if (!(is(0, VtFloat) && is(1, VtMat2))) return false;
            out = czero(VtMat2);
            for (int i = 0; i < 4; ++i) out.v[i] = a[0] * b[i];
            return true;
        }
        case MulM2:
            if (!(is(0, VtMat2) && is(1, VtMat2))) return false;
            out = cvec<4>(VtMat2, {(a[0] * b[0]) + (a[2] * b[1]), (a[1] * b[0]) + (a[3] * b[1]),
                                   (a[0] * b[2]) + (a[2] * b[3]), (a[1] * b[2]) + (a[3] * b[3])});
            return true;
        case MulM2V2:
            if (!(is(0, VtMat2) && is(1, VtVec2))) return false;
            out = cvec<2>(VtVec2, {(a[0] * b[0]) + (a[2] * b[1]), (a[1] * b[0]) + (a[3] * b[1])});
            return true;
        case DetM2:
            if (!is(0, VtMat2)) return false;
            out = cfloat((a[0] * a[3]) - (a[1] * a[2]));
            return true;
        case TraceM2:
            if (!is(0, VtMat2)) return false;
            out = cfloat(a[0] + a[3]);
            return true;
        case TransposeM2:
            if (!is(0, VtMat2)) return false;
            out = cvec<4>(VtMat2, {a[0], a[2], a[1], a[3]});
            return true;
        case InverseM2: {
            if (!is(0, VtMat2)) return false;
            const float det = (a[0] * a[3]) - (a[1] * a[2]);
            if (!(fabsf(det) > 1e-9f)) return false;
            out = cvec<4>(VtMat2, {a[3] / det, (-a[1]) / det, (-a[2]) / det, a[0] / det});
            return true;
        }
        case MakeMat3: {
            if (!(n == 3 && is(0, VtVec3) && is(1, VtVec3) && is(2, VtVec3))) return false;
            const float* c = cv[2]->v;
            out = cvec<9>(VtMat3, {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]});
            return true;
        }
        case AddM3: case SubM3: {
            if (!(is(0, VtMat3) && is(1, VtMat3))) return false;
            out = czero(VtMat3);
            for (int i = 0; i < 9; ++i) out.v[i] = op == AddM3 ? a[i] + b[i] : a[i] - b[i];
            return true;
        }
        case ScaleM3: {
            if (!(is(0, VtFloat) && is(1, VtMat3))) return false;
            out = czero(VtMat3);
            for (int i = 0; i < 9; ++i) out.v[i] = a[0] * b[i];
            return true;
        }
        case MulM3: {
            if (!(is(0, VtMat3) && is(1, VtMat3))) return false;
            out = czero(VtMat3);
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                    out.v[3 * c + r] = ((a[r] * b[3 * c]) + (a[3 + r] * b[3 * c + 1])) + (a[6 + r] * b[3 * c + 2]);
            return true;
        }
        case MulM3V3: {
            if (!(is(0, VtMat3) && is(1, VtVec3))) return false;
            out = czero(VtVec3);
            for (int r = 0; r < 3; ++r) out.v[r] = ((a[r] * b[0]) + (a[3 + r] * b[1])) + (a[6 + r] * b[2]);
            return true;
        }
        case TraceM3:
            if (!is(0, VtMat3)) return false;
            out = cfloat((a[0] + a[4]) + a[8]);
            return true;
        case TransposeM3:
            if (!is(0, VtMat3)) return false;
            out = cvec<9>(VtMat3, {a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]});
            return true;
        case DetM3:
            if (!is(0, VtMat3)) return false;
            out = cfloat(det3(a));
            return true;
        case InverseM3: {
            if (!is(0, VtMat3)) return false;
            const float det = det3(a);
            if (!(fabsf(det) > 1e-9f)) return false;
            const float inv = 1.0f / det;
            const float* m = a;
            out = cvec<9>(VtMat3, {((m[4] * m[8]) - (m[5] * m[7])) * inv, (-((m[1] * m[8]) - (m[2] * m[7]))) * inv,
                                   ((m[1] * m[5]) - (m[2] * m[4])) * inv, (-((m[3] * m[8]) - (m[5] * m[6]))) * inv,
                                   ((m[0] * m[8]) - (m[2] * m[6])) * inv, (-((m[0] * m[5]) - (m[2] * m[3]))) * inv,
                                   ((m[3] * m[7]) - (m[4] * m[6])) * inv, (-((m[0] * m[7]) - (m[1] * m[6]))) * inv,
                                   ((m[0] * m[4]) - (m[1] * m[3])) * inv});
            return true;
        }
        default: return false;
    }
}

// Linalg domain (linalg/mod.rs:227-310)
Action linalg_rule(int op, const Const* const* cv, int n, bool args_equal) {
    bool all_const = n > 0;
    for (int i = 0; i < n; ++i) all_const = all_const && cv[i] != nullptr;
    Const folded;
    if (all_const && linalg_fold(op, cv, n, folded)) return replace(folded);
    if (n == 2) {
        const bool a0 = cv[0] && is_zero(*cv[0]), b0 = cv[1] && is_zero(*cv[1]);
        const bool a1 = cv[0] && is_one(*cv[0]);
        switch (op) {
            case AddV2: case AddV3: case AddM2: case AddM3:
                if (b0) return keep(0);
                if (a0) return keep(1);
                break;
            case SubV2: case SubV3: case SubM2: case SubM3:
                if (b0) return keep(0);
                if (args_equal) return replace(zero_for(op));
                break;
            case ScaleV2: case ScaleV3: case ScaleM2: case ScaleM3:
                if (a0) return replace(zero_for(op));
                if (a1) return keep(1);
                break;
            case DotV2: case DotV3:
                if (a0 || b0) return replace(cfloat(0.0f));
                break;
            case CrossV3:
                if (a0 || b0 || args_equal) return replace(czero(VtVec3));
                break;
            case MulM2: case MulM3: {
                const bool ai = cv[0] && is_identity(*cv[0]), bi = cv[1] && is_identity(*cv[1]);
                if (ai) return keep(1);
                if (bi) return keep(0);
                if (a0 || b0) return replace(zero_for(op));
                break;
            }
            case MulM2V2: case MulM3V3: {
                const bool ai = cv[0] && is_identity(*cv[0]);
                if (ai) return keep(1);
                if (a0 || b0) return replace(zero_for(op));
                break;
            }
            default: break;
        }
    }
    return Action{};
}

struct Node {
    uint8_t op;
    uint16_t arg;  // variable or constant index; 0 for operators
    bool operator==(const Node& o) const { return op == o.op && arg == o.arg; }
};
struct Item {
    size_t start;    // first output node of the subtree
    int cidx = -1;   // the subtree is one constant: its index in the constant list
};

struct Simplified {
    std::vector<Node> code;
    std::vector<Const> consts;
};

Simplified simplify_one(const uint8_t* ops, const uint16_t* args, int64_t n_code, const uint8_t* ctypes,
                        const float* const* cvals, int64_t n_consts) {
    std::vector<Node> out;
    std::vector<Const> consts;  // every constant emitted, dropped ones included
    std::vector<Item> st;
    out.reserve((size_t)n_code);
    for (int64_t pc = 0; pc < n_code; ++pc) {
        const uint8_t op = ops[pc];
        if (is_load_const(op)) {
            Const c;
            const uint16_t k = args[pc];
            if (k < n_consts) {
                c.vt = ctypes[k];
                const int s = vt_size(c.vt);
                for (int i = 0; i < (s ? s : 1); ++i) c.v[i] = cvals[k][i];
            } else {
                c.vt = load_type(op);  // out-of-range index: an unresolved constant, kept as is
            }
            consts.push_back(c);
            st.push_back(Item{out.size(), (int)consts.size() - 1});
            out.push_back(Node{op, (uint16_t)(consts.size() - 1)});
            continue;
        }
        if (is_load_var(op) || op >= RefOpCount) {
            st.push_back(Item{out.size(), -1});
            out.push_back(Node{op, args[pc]});
            continue;
        }
        const int n = arity(op);
        if ((int)st.size() < n) {  // malformed: the operator is kept as an opaque value
            st.push_back(Item{out.size(), -1});
            out.push_back(Node{op, 0});
            continue;
        }
        Item a[3];
        for (int i = n - 1; i >= 0; --i) {
            a[i] = st.back();
            st.pop_back();
        }
        bool args_equal = false;
        if (n == 2 && a[0].start < a[1].start) {
            const size_t la = a[1].start - a[0].start, lb = out.size() - a[1].start;
            args_equal = la == lb && std::equal(out.begin() + a[0].start, out.begin() + a[1].start, out.begin() + a[1].start);
        }
        const Const* cv[3] = {nullptr, nullptr, nullptr};
        for (int i = 0; i < n; ++i) cv[i] = a[i].cidx >= 0 ? &consts[a[i].cidx] : nullptr;
        const Action act = op <= LnF || op == NegF ? basic_rule(op, cv, n, args_equal) : linalg_rule(op, cv, n, args_equal);
        if (act.act == Replace) {
            out.resize(a[0].start);
            consts.push_back(act.c);
            st.push_back(Item{out.size(), (int)consts.size() - 1});
            out.push_back(Node{load_const_op(return_type(op)), (uint16_t)(consts.size() - 1)});
        } else if (act.act == Keep) {
            const int k = act.keep;
            const size_t begin = a[k].start, end = k == n - 1 ? out.size() : a[k + 1].start;
            const size_t dst = a[0].start;
            if (begin > dst) std::copy(out.begin() + begin, out.begin() + end, out.begin() + dst);
            out.resize(dst + (end - begin));
            st.push_back(Item{dst, a[k].cidx});
        } else {
            // an inverse pair: the operator and its child's root cancel
            if (n == 1 && out.size() > a[0].start) {
                const uint8_t child = out.back().op;
                if ((op == LnF && child == ExpF) || (op == ExpF && child == LnF) || (op == SqrtF && child == SqrF) ||
                    (op == SqrF && child == SqrtF)) {
                    out.pop_back();
                    st.push_back(Item{a[0].start, -1});
                    continue;
                }
            }
            out.push_back(Node{op, 0});
            st.push_back(Item{a[0].start, -1});
        }
    }
    // Program::from_nodes: the constants the remaining nodes load, in postfix order
    Simplified r;
    r.code.reserve(out.size());
    for (const Node& nd : out) {
        if (is_load_const(nd.op)) {
            r.consts.push_back(consts[nd.arg]);
            r.code.push_back(Node{nd.op, (uint16_t)(r.consts.size() - 1)});
        } else {
            r.code.push_back(nd);
        }
    }
    return r;
}

}  // namespace

extern "C" int rsr_simplify_programs(int n, const uint8_t* ops, const uint16_t* args, const int64_t* code_offsets,
                                     const uint8_t* const_types, const int64_t* const_offsets,
                                     const float* const_values, const int64_t* value_offsets, int64_t code_capacity,
                                     int64_t const_capacity, int64_t value_capacity, uint8_t* ops_out,
                                     uint16_t* args_out, int64_t* code_offsets_out, uint8_t* const_types_out,
                                     int64_t* const_offsets_out, float* const_values_out,
                                     int64_t* value_offsets_out) {
    if (n < 0 || !code_offsets || !const_offsets || !value_offsets || !code_offsets_out || !const_offsets_out ||
        !value_offsets_out)
        return set_error(RSR_E_INVALID_ARG, "null argument");
    int64_t nc = 0, nk = 0, nv = 0;
    code_offsets_out[0] = const_offsets_out[0] = value_offsets_out[0] = 0;
    for (int i = 0; i < n; ++i) {
        const int64_t c0 = code_offsets[i], c1 = code_offsets[i + 1];
        const int64_t k0 = const_offsets[i], k1 = const_offsets[i + 1];
        if (c1 < c0 || k1 < k0 || (c1 > c0 && (!ops || !args)) || (k1 > k0 && (!const_types || !const_values)))
            return set_error(RSR_E_INVALID_ARG, "program %d: bad offsets or null arrays", i);
        // payload pointer of each constant (Int/Bool constants take one slot)
        std::vector<const float*> cp((size_t)(k1 - k0));
        int64_t v = value_offsets[i];
        for (int64_t k = k0; k < k1; ++k) {
            cp[(size_t)(k - k0)] = const_values + v;
            const int s = vt_size(const_types[k]);
            v += s ? s : 1;
        }
        const Simplified s = simplify_one(ops + c0, args + c0, c1 - c0, const_types + k0, cp.data(), k1 - k0);
        int64_t vals = 0;
        for (const Const& c : s.consts) vals += vt_size(c.vt) ? vt_size(c.vt) : 1;
        if (nc + (int64_t)s.code.size() > code_capacity || nk + (int64_t)s.consts.size() > const_capacity ||
            nv + vals > value_capacity)
            return set_error(RSR_E_INVALID_ARG, "program %d: output capacity exceeded", i);
        if ((!s.code.empty() && (!ops_out || !args_out)) || (!s.consts.empty() && (!const_types_out || !const_values_out)))
            return set_error(RSR_E_INVALID_ARG, "null output arrays");
        for (const Node& nd : s.code) {
            ops_out[nc] = nd.op;
            args_out[nc] = nd.arg;
            ++nc;
        }
        for (const Const& c : s.consts) {
            const_types_out[nk++] = c.vt;
            const int sz = vt_size(c.vt) ? vt_size(c.vt) : 1;
            for (int q = 0; q < sz; ++q) const_values_out[nv++] = c.v[q];
        }
        code_offsets_out[i + 1] = nc;
        const_offsets_out[i + 1] = nk;
        value_offsets_out[i + 1] = nv;
    }
    return RSR_OK;
}
