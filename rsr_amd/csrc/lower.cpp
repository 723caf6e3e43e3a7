// lower.cpp -- postfix Instruction stream -> register-slot micro-code.
//
// The reference interprets a typed-stack machine (src/engine/eval/evaluator.rs:20-150,
// VmState in src/engine/eval/state.rs:12-49).  A postfix program has no branches, so the
// depth of every typed stack is a static function of the pc.  The lowering therefore
// simulates the five typed stacks once on the host, with each stack entry holding the
// register numbers of its components, and emits only the arithmetic:
//   * LoadVar*, MakeVec*, MakeMat*, Get*, Transpose* (and Inverse* in the dual variant)
//     emit nothing: they re-label registers (feature components are resident in
//     registers [0, F) for the whole tile);
//   * every other reference op expands into per-component micro-ops whose arithmetic is
//     the reference formula, operation for operation (src/domains/basic/eval.rs,
//     src/domains/linalg/eval.rs, and for the dual variant src/engine/eval/autodiff.rs);
//   * destinations are assigned by a per-op linear scan that reuses a component register
//     only after its last read inside the op, so multi-output ops (CrossV3, MulM3, ...)
//     never read a clobbered input;
//   * when the register file is full, the oldest value still on a stack (the one consumed
//     last in postfix order) is spilled to a per-lane slot and filled back when its
//     consuming op pops it; when the operands of one op do not fit the file, the op's
//     micro-ops fill and spill its own values on demand.  The first kMaxSpillSlots slots
//     are LDS, the rest (up to kMaxSlots, enough for the reference's 32-deep typed stacks)
//     device memory, so every program within the reference's stack limits lowers.
// Validation replaces the reference's undefined behaviour with status codes: stack
// underflow or depth > 32 (state.rs:14), out-of-range feature or constant indices,
// constant type mismatch, Bool/Int loads (silent no-ops in the reference,
// evaluator.rs:144), and programs that do not reduce to exactly one Float.
#include "lower.h"
#include <algorithm>
#include <array>

namespace rsr {
namespace {

// Component locations: [0, kRegs) a register; kSlotLoc + s spill slot s; kKonstLoc + p a
// Float constant (flat parameter p) not yet in a register.
constexpr int kSlotLoc = 64;
constexpr int kKonstLoc = 1 << 14;
inline bool in_reg(int l) { return l < kRegs; }
inline bool in_slot(int l) { return l >= kSlotLoc && l < kKonstLoc; }

struct Val {
    uint8_t n = 0;
    int16_t loc[9] = {0};    // location of each component (see above)
    uint32_t seq = 0;        // push order (older values are consumed later)
    int32_t konst = -1;      // >= 0: a Float constant (flat parameter index) not yet in a register
};

struct Src {
    bool virt;
    int id;  // physical register or virtual id
};

struct MI {
    UOp op;
    int ns;
    Src s[4];
    uint32_t extra;  // parameter index (IMM and the constant-operand ops), or INVSEL flags
};

constexpr int kMaxDepth = 32;
constexpr int kInf = 1 << 30;

using RegSet = uint32_t;  // bit r: register r
RegSet range_bits(uint32_t r, int n) {
    return n ? (((n >= 32 ? ~0u : ((1u << n) - 1u)) << r)) : 0u;
}

// Operand-swapped form of a micro-op whose forwarded operand is b, so that it becomes
// a: the same op for the commutative ones (IEEE + and * are commutative, and every
// step of these formulas is symmetric in a and b, INVSEL's negation included since
// (-a) * b == (-b) * a), a reversed op otherwise, U_COUNT when there is none.
UOp swapped_op(uint32_t op) {
    switch (op) {
        case U_ADD: case U_MUL: case U_MULADD: case U_MULSUB: case U_ADD3: case U_INVSEL:
        case U_ADDV3: case U_DOT3:
            return (UOp)op;
        case U_SUB: return U_RSUB;
        case U_DIV: return U_RDIV;
        case U_SUBV3: return U_RSUBV3;
        case U_CROSS3: return U_RCROSS3;
        case U_SCALEV3: return U_SCALEV3S;
        case U_MULM3V3: return U_MULM3V3R;
        default: return U_COUNT;
    }
}

// Value-code post-pass.  A postfix program consumes most results in the very next
// operation, so the interpreter keeps the last result's components in an accumulator:
// a consumer whose whole operand a is exactly the previous result takes it from there
// (kFwd; a consumer of it through b is first rewritten to its operand-swapped form),
// and a forwarded result that no later micro-op reads is not written to the register
// file at all (kNoWb).  kVec3 and kNone describe the operands and the result for the
// staging and the write-back.  Liveness is exact: a backward scan over register reads
// and writes, with the program result live at the end.
void annotate_forwarding(std::vector<UInstr>& code, uint32_t out_reg) {
    const int n = (int)code.size();
    // 0. transcendentals of point-invariant values (derived from constants only) are
    //    computed once per lane
    {
        RegSet uni = 0;  // registers holding a point-invariant value
        for (int i = 0; i < n; ++i) {
            UInstr& u = code[i];
            const uint32_t op = u.w0 & 255u, x = u.w0 >> 8;
            const UGroups g = ugroups(op);
            bool inv = op != U_FILL;  // a filled value is not tracked
            if (g.na) inv = inv && (range_bits(u.w2 & 31u, g.na) & ~uni) == 0;
            if (g.nb) inv = inv && (range_bits(u.w3 & 31u, g.nb) & ~uni) == 0;
            if (g.nc) inv = inv && (range_bits(ux_c(x), 1) & ~uni) == 0;
            if (g.ne) inv = inv && (range_bits(ux_e(x), 1) & ~uni) == 0;
            if (inv && g.na == 1) {
                switch (op) {
                    case U_SIN: u.w0 = (u.w0 & ~255u) | U_SINU; break;
                    case U_COS: u.w0 = (u.w0 & ~255u) | U_COSU; break;
                    case U_EXP: u.w0 = (u.w0 & ~255u) | U_EXPU; break;
                    case U_LN: u.w0 = (u.w0 & ~255u) | U_LNU; break;
                    case U_SQRT: u.w0 = (u.w0 & ~255u) | U_SQRTU; break;
                    default: break;
                }
            }
            const RegSet w = range_bits(u.w1 & 31u, g.nd);
            uni = inv ? (uni | w) : (uni & ~w);
        }
    }
    // 1. swap forwarded b operands into a
    for (int i = 1; i < n; ++i) {
        UInstr& u = code[i];
        const UGroups gp = ugroups(code[i - 1].w0 & 255u), g = ugroups(u.w0 & 255u);
        const uint32_t d = code[i - 1].w1 & 31u;
        if (gp.nd != 1 && gp.nd != 3) continue;
        if (g.na == gp.nd && (u.w2 & 31u) == d) continue;
        if (!(g.nb == gp.nd && (u.w3 & 31u) == d)) continue;
        const UOp s = swapped_op(u.w0 & 255u);
        if (s == U_COUNT) continue;
        u.w0 = (u.w0 & ~255u) | s;
        std::swap(u.w2, u.w3);
    }
    // 2. liveness
    std::vector<RegSet> live_after(n);
    auto reads = [&](const UInstr& u, bool skip_a) {
        const UGroups g = ugroups(u.w0 & 255u);
        const uint32_t x = u.w0 >> 8;
        RegSet s = 0;
        if (!skip_a) s |= range_bits(u.w2 & 31u, g.na);
        s |= range_bits(u.w3 & 31u, g.nb);
        if (g.nc) s |= range_bits(ux_c(x), 1);
        if (g.ne) s |= range_bits(ux_e(x), 1);
        return s;
    };
    RegSet live = range_bits(out_reg, 1);
    for (int i = n - 1; i >= 0; --i) {
        live_after[i] = live;
        const UGroups g = ugroups(code[i].w0 & 255u);
        live = (live & ~range_bits(code[i].w1 & 31u, g.nd)) | reads(code[i], false);
    }
    // 3. operand shapes
    for (int i = 0; i < n; ++i) {
        UInstr& u = code[i];
        const UGroups g = ugroups(u.w0 & 255u);
        if (g.na == 3) u.w2 |= kVec3;
        if (g.na == 0) u.w2 |= kNone;
        if (g.nb == 3) u.w3 |= kVec3;
        if (g.nb == 0) u.w3 |= kNone;
        if (g.nd == 3) u.w1 |= kVec3;
    }
    // 4. forwarding and write-back elision
    for (int i = 1; i < n; ++i) {
        UInstr& u = code[i];
        UInstr& p = code[i - 1];
        const UGroups gp = ugroups(p.w0 & 255u), g = ugroups(u.w0 & 255u);
        if (gp.nd != 1 && gp.nd != 3) continue;
        const uint32_t d = p.w1 & 31u;
        if (!(g.na == gp.nd && (u.w2 & 31u) == d)) continue;
        u.w2 |= kFwd;
        // the old value is dead after the consumer where the consumer overwrites it
        const RegSet w = range_bits(d, gp.nd);
        const RegSet survives = w & ~range_bits(u.w1 & 31u, g.nd);
        if (!(w & reads(u, true)) && !(survives & live_after[i])) p.w1 |= kNoWb;
    }
}

class Lowerer {
   public:
    Lowerer(const RefProgram& p, int F, bool dual) : p_(p), F_(F), dual_(dual) {
        for (int r = F; r < kRegs; ++r) free_[r] = true;
        for (int s = 0; s < kMaxSlots; ++s) slot_free_[s] = true;
        // constant payload offsets and flat parameter starts
        int off = 0, flat = 0;
        coff_.resize(p.n_consts);
        cflat_.resize(p.n_consts);
        for (int i = 0; i < p.n_consts; ++i) {
            coff_[i] = off;
            cflat_[i] = flat;
            int s = vt_size(p.const_types[i]);
            off += s ? s : 1;
            flat += s;
        }
        out_.n_params = (uint32_t)flat;
    }

    Lowered run() {
        if (F_ < 0 || F_ > kRegs) return fail(PROG_FEATURE);
        for (int pc = 0; pc < p_.n_code; ++pc) {
            int st = step(p_.ops[pc], p_.args[pc]);
            if (st == PROG_OK && !ok_) st = PROG_REGS;
            if (st != PROG_OK) return fail(st);
        }
        if (stk_[0].size() != 1 || stk_[1].size() || stk_[2].size() || stk_[3].size() || stk_[4].size())
            return fail(PROG_ROOT);
        Val root = pop(VtFloat);  // fills the root if it was spilled
        if (!in_reg(root.loc[0])) {  // nothing else is live: any free temporary will do
            const int r = alloc();
            if (r < 0) return fail(PROG_REGS);
            to_reg(root.loc[0], r);
            root.loc[0] = (int16_t)r;
        }
        if (!ok_) return fail(PROG_REGS);
        out_.out_reg = (uint32_t)root.loc[0];
        out_.status = PROG_OK;
        if (!dual_) annotate_forwarding(out_.code, out_.out_reg);
        return std::move(out_);
    }

   private:
    const RefProgram& p_;
    int F_;
    bool dual_;
    bool free_[kRegs] = {false};
    bool slot_free_[kMaxSlots];
    bool ok_ = true;
    uint32_t seq_ = 0;
    std::vector<Val> stk_[5];  // F, V2, V3, M2, M3
    std::vector<int> coff_, cflat_;
    Lowered out_;

    Lowered fail(int st) {
        Lowered l;
        l.status = st;
        return l;
    }

    static int sidx(int vt) {
        switch (vt) {
            case VtFloat: return 0;
            case VtVec2: return 1;
            case VtVec3: return 2;
            case VtMat2: return 3;
            default: return 4;
        }
    }
    bool can_pop(int vt, int n = 1) { return (int)stk_[sidx(vt)].size() >= n; }

    // Emit the micro-op that moves a component from a spill slot or a pending constant
    // into register r, and free the slot.
    void to_reg(int loc, int r) {
        if (in_slot(loc)) {
            const int slot = loc - kSlotLoc;
            out_.code.push_back(make_uinstr(U_FILL, (uint32_t)r, 0, 0, (uint32_t)slot));
            slot_free_[slot] = true;
        } else {
            out_.code.push_back(make_uinstr(U_IMM, (uint32_t)r, 0, 0, (uint32_t)(loc - kKonstLoc)));
        }
    }
    int new_slot(int limit = kMaxSlots) {
        for (int s = 0; s < limit; ++s)
            if (slot_free_[s]) {
                slot_free_[s] = false;
                if ((uint32_t)s + 1 > out_.spill_slots) out_.spill_slots = s + 1;
                return s;
            }
        return -1;
    }
    // Register r -> a new spill slot; returns the slot location, or -1 when every slot is taken.
    int spill_reg(int r, int limit = kMaxSlots) {
        const int slot = new_slot(limit);
        if (slot < 0) return -1;
        out_.code.push_back(make_uinstr(U_SPILL, (uint32_t)r, (uint32_t)r, 0, (uint32_t)slot));
        free_[r] = true;
        return kSlotLoc + slot;
    }

    // Pop a value; a pending Float constant is materialised into a register.
    Val pop(int vt) {
        Val v = pop_raw(vt);
        if (v.konst >= 0) materialize(v);
        return v;
    }
    // Pop a value as it is (a pending constant stays pending).  Spilled components are
    // filled back into registers while registers can be found; the others stay in their
    // slots, and the consuming op reads them from there (emit fills them on demand).
    Val pop_raw(int vt) {
        Val v = stk_[sidx(vt)].back();
        stk_[sidx(vt)].pop_back();
        for (int k = 0; k < v.n; ++k)
            if (in_slot(v.loc[k])) {
                const int r = alloc();
                if (r < 0) break;
                to_reg(v.loc[k], r);
                v.loc[k] = (int16_t)r;
            }
        return v;
    }
    bool push(int vt, Val v) {
        if ((int)stk_[sidx(vt)].size() >= kMaxDepth) return false;
        v.seq = ++seq_;
        stk_[sidx(vt)].push_back(v);
        return true;
    }
    // Spill one temporary component of the oldest stacked value (the one consumed last)
    // into a slot below `limit`; false if no stacked value holds a temporary register or
    // every such slot is taken.
    bool spill_one(int limit = kMaxSlots) {
        Val* victim = nullptr;
        for (auto& st : stk_)
            for (auto& v : st) {
                if (v.konst >= 0) continue;  // a pending constant holds no register
                bool has_temp = false;
                for (int k = 0; k < v.n; ++k)
                    if (is_temp(v.loc[k])) has_temp = true;
                if (has_temp && (!victim || v.seq < victim->seq)) victim = &v;
            }
        if (!victim) return false;
        for (int k = 0; k < victim->n; ++k) {
            if (!is_temp(victim->loc[k])) continue;
            const int l = spill_reg(victim->loc[k], limit);
            if (l < 0) return false;
            victim->loc[k] = (int16_t)l;
            return true;  // one component per call
        }
        return false;
    }
    bool is_temp(int l) const { return l >= F_ && l < kRegs; }
    // A value component dies: its register or spill slot is free again.
    void release(int l) {
        if (is_temp(l)) free_[l] = true;
        else if (in_slot(l)) slot_free_[l - kSlotLoc] = true;
    }
    int alloc() {
        for (;;) {
            for (int r = F_; r < kRegs; ++r)
                if (free_[r]) {
                    free_[r] = false;
                    if ((uint32_t)r + 1 > out_.max_reg) out_.max_reg = r + 1;
                    return r;
                }
            if (!spill_one()) return -1;
        }
    }

    // A value held in consecutive registers.
    static bool contig(const Val& v) {
        for (int k = 0; k < v.n; ++k)
            if (!in_reg(v.loc[k]) || v.loc[k] != v.loc[0] + k) return false;
        return true;
    }

    // First block of n consecutive temporaries that are free or in `dying`; -1 if none.
    int find_block(int n, const std::vector<int>& dying) const {
        bool f[kRegs];
        for (int r = 0; r < kRegs; ++r) f[r] = free_[r];
        for (int r : dying)
            if (is_temp(r)) f[r] = true;
        for (int r = F_; r + n <= kRegs; ++r) {
            bool ok = true;
            for (int k = 0; k < n && ok; ++k) ok = f[r + k];
            if (ok) return r;
        }
        return -1;
    }

    // One block micro-op: operands a (and b) consumed, result in n_out consecutive
    // registers (n_out = 1: any register).  The operands must be held in registers.  The
    // operand temporaries are released before the result is placed, so the result may
    // reuse them (the device reads every operand component first).  Returns false, with
    // nothing emitted, when no block can be found even after spilling; the caller then
    // lowers the op component by component.
    bool block_op(UOp op, int n_out, const Val& A, const Val* B, uint32_t x, Val& result) {
        std::vector<int> dying;
        for (int k = 0; k < A.n; ++k) dying.push_back(A.loc[k]);
        if (B)
            for (int k = 0; k < B->n; ++k) dying.push_back(B->loc[k]);
        for (int l : dying)
            if (!in_reg(l)) return false;
        // room for the block is made with LDS slots only: where that is not enough, the
        // component-wise form (which may spill to device memory) is the cheaper code
        int d = find_block(n_out, dying);
        while (d < 0) {
            if (!spill_one(kMaxSpillSlots)) return false;
            d = find_block(n_out, dying);
        }
        for (int r : dying) release(r);
        for (int k = 0; k < n_out; ++k) free_[d + k] = false;
        if ((uint32_t)(d + n_out) > out_.max_reg) out_.max_reg = d + n_out;
        out_.code.push_back(make_uinstr(op, (uint32_t)d, (uint32_t)A.loc[0], B ? (uint32_t)B->loc[0] : 0u, x));
        result = Val{};
        result.n = (uint8_t)n_out;
        for (int k = 0; k < n_out; ++k) result.loc[k] = (int16_t)(d + k);
        return true;
    }

    static bool param_operand(UOp op) {
        switch (op) {
            case U_IMM: case U_ADDK: case U_SUBK: case U_KSUB: case U_MULK: case U_DIVK: case U_KDIV:
            case U_SINK: case U_COSK: case U_EXPK: case U_LNK: case U_SQRTK: case U_NEGK: case U_SQRK:
                return true;
            default:
                return false;
        }
    }

    // Emit a micro-op sequence with its own register allocation.  `owned` lists the
    // operand components consumed by the reference op (their locations); `outs` names the
    // result components (virtual ids or operand locations passed through).  Sources are
    // operand locations (P) or earlier results of the sequence (V).  Before each micro-op
    // its sources are brought into registers (a FILL from a spill slot, an IMM for a
    // pending constant); a register is found by taking a free one, else by spilling the
    // oldest stacked value, else by spilling the live value of this op whose next read is
    // furthest away (never a source of the current micro-op).  Sources are released after
    // their last read, results that are still live may end in spill slots.  Returns false
    // only when the register file cannot hold one micro-op's operands.
    bool emit(const std::vector<MI>& seq, const std::vector<int>& owned, const std::vector<Src>& outs,
              Val& result) {
        const int n = (int)seq.size();
        // value table: operand components (by their location at the start) and the virtuals
        struct Ent {
            int loc;   // current location (register, slot or pending constant); -1 dead or unborn
            int last;  // last read (kInf for a result)
        };
        std::vector<Ent> ent;
        std::vector<int> opd_orig;  // entry k < opd_orig.size(): operand with start location opd_orig[k]
        auto opd_entry = [&](int orig) {
            for (size_t k = 0; k < opd_orig.size(); ++k)
                if (opd_orig[k] == orig) return (int)k;
            return -1;
        };
        for (int l : owned)
            if (!in_reg(l) || is_temp(l)) {
                if (opd_entry(l) >= 0) continue;
                opd_orig.push_back(l);
                ent.push_back(Ent{l, -1});
            }
        const int nopd = (int)opd_orig.size();
        for (int i = 0; i < n; ++i) ent.push_back(Ent{-1, -1});
        // entry of a source (-1: a feature register, read in place)
        auto src_entry = [&](const Src& s) -> int {
            if (s.virt) return nopd + s.id;
            if (in_reg(s.id) && !is_temp(s.id)) return -1;
            const int e = opd_entry(s.id);
            return e;  // a source must be an owned operand component (or a feature)
        };
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < seq[i].ns; ++k) {
                const int e = src_entry(seq[i].s[k]);
                if (e >= 0) ent[e].last = i;
            }
        for (const Src& o : outs) {
            const int e = src_entry(o);
            if (e >= 0) ent[e].last = kInf;
        }
        auto kill = [&](int e) {
            release(ent[e].loc);
            ent[e].loc = -1;
        };
        // owned components never read and not passed through die immediately
        for (int e = 0; e < nopd; ++e)
            if (ent[e].last == -1) kill(e);
        auto next_read = [&](int e, int i) -> int {
            if (ent[e].last == kInf) {
                for (int j = i + 1; j < n; ++j)
                    for (int k = 0; k < seq[j].ns; ++k)
                        if (src_entry(seq[j].s[k]) == e) return j;
                return kInf;
            }
            for (int j = i + 1; j <= ent[e].last && j < n; ++j)
                for (int k = 0; k < seq[j].ns; ++k)
                    if (src_entry(seq[j].s[k]) == e) return j;
            return kInf;
        };
        // a register for micro-op i, never one in `pinned`
        auto ealloc = [&](int i, RegSet pinned) -> int {
            for (;;) {
                for (int r = F_; r < kRegs; ++r)
                    if (free_[r]) {
                        free_[r] = false;
                        if ((uint32_t)r + 1 > out_.max_reg) out_.max_reg = r + 1;
                        return r;
                    }
                if (spill_one()) continue;
                int victim = -1, far = -1;
                for (int e = 0; e < (int)ent.size(); ++e) {
                    const int l = ent[e].loc;
                    if (!is_temp(l) || (pinned >> l) & 1u) continue;
                    const int nr = next_read(e, i - 1);
                    if (nr > far) {
                        far = nr;
                        victim = e;
                    }
                }
                if (victim < 0) return -1;
                const int l = spill_reg(ent[victim].loc);
                if (l < 0) return -1;
                ent[victim].loc = l;
            }
        };
        for (int i = 0; i < n; ++i) {
            const MI& m = seq[i];
            int se[4];
            RegSet pinned = 0;
            for (int k = 0; k < m.ns; ++k) {
                se[k] = src_entry(m.s[k]);
                if (se[k] >= 0 && in_reg(ent[se[k]].loc)) pinned |= 1u << ent[se[k]].loc;
            }
            for (int k = 0; k < m.ns; ++k) {
                if (se[k] < 0 || in_reg(ent[se[k]].loc)) continue;
                const int r = ealloc(i, pinned);
                if (r < 0) return false;
                to_reg(ent[se[k]].loc, r);
                ent[se[k]].loc = r;
                pinned |= 1u << r;
            }
            uint32_t rg[4] = {0, 0, 0, 0};
            for (int k = 0; k < m.ns; ++k) rg[k] = (uint32_t)(se[k] >= 0 ? ent[se[k]].loc : m.s[k].id);
            RegSet live_src = 0;
            for (int k = 0; k < m.ns; ++k) {
                if (se[k] < 0) continue;
                if (ent[se[k]].last == i) {
                    if (ent[se[k]].loc >= 0) kill(se[k]);
                } else {
                    live_src |= 1u << ent[se[k]].loc;
                }
            }
            const int d = ealloc(i, live_src);
            if (d < 0) return false;
            ent[nopd + i].loc = d;
            const uint32_t x = param_operand(m.op) ? m.extra : (rg[2] | rg[3] << 5 | m.extra << 10);
            out_.code.push_back(make_uinstr(m.op, (uint32_t)d, rg[0], rg[1], x));
            if (ent[nopd + i].last == -1) kill(nopd + i);  // no reader and no output role
        }
        result.n = (uint8_t)outs.size();
        for (size_t k = 0; k < outs.size(); ++k) {
            const int e = src_entry(outs[k]);
            result.loc[k] = (int16_t)(e >= 0 ? ent[e].loc : outs[k].id);
        }
        return true;
    }

    static Src P(int r) { return Src{false, r}; }
    static Src V(int id) { return Src{true, id}; }
    static MI mi(UOp op, std::initializer_list<Src> s, uint32_t extra = 0) {
        MI m;
        m.op = op;
        m.ns = (int)s.size();
        int k = 0;
        for (const Src& x : s) m.s[k++] = x;
        m.extra = extra;
        return m;
    }

    static std::vector<int> regs_of(std::initializer_list<const Val*> vs) {
        std::vector<int> r;
        for (const Val* v : vs)
            for (int k = 0; k < v->n; ++k) r.push_back(v->loc[k]);
        return r;
    }

    int load_var(int vt, int idx) {
        int n = vt_size(vt);
        if (idx + n > F_) return PROG_FEATURE;
        Val v;
        v.n = (uint8_t)n;
        for (int k = 0; k < n; ++k) v.loc[k] = (int16_t)(idx + k);
        return push(vt, v) ? PROG_OK : PROG_STACK;
    }

    int load_const(int vt, int c) {
        if (c >= p_.n_consts) return PROG_CONST;
        if (p_.const_types[c] != vt) return PROG_CONST;
        int n = vt_size(vt);
        if (vt != VtFloat) out_.has_vecmat_const = true;
        else out_.n_float_consts++;
        if (vt == VtFloat) {  // pending: its consumer reads it from the parameters
            Val r;
            r.n = 1;
            r.konst = cflat_[c];
            return push(vt, r) ? PROG_OK : PROG_STACK;
        }
        if (vt == VtVec3) {
            Val r, none;
            none.n = 0;
            if (block_op(U_IMMV3, 3, none, nullptr, (uint32_t)cflat_[c], r)) return push(vt, r) ? PROG_OK : PROG_STACK;
        }
        std::vector<MI> seq;
        std::vector<Src> outs;
        for (int k = 0; k < n; ++k) {
            seq.push_back(mi(U_IMM, {}, (uint32_t)(cflat_[c] + k)));
            outs.push_back(V(k));
        }
        Val r;
        if (!emit(seq, {}, outs, r)) return PROG_REGS;
        return push(vt, r) ? PROG_OK : PROG_STACK;
    }

    // componentwise binary op over n components: out_k = a_k op b_k
    int cwise(int vt, UOp op) {
        if (!can_pop(vt, 2)) return PROG_STACK;
        Val b = pop(vt), a = pop(vt);
        if ((vt == VtVec3 || vt == VtMat2) && contig(a) && contig(b) && (op == U_ADD || op == U_SUB)) {
            const UOp bop = vt == VtVec3 ? (op == U_ADD ? U_ADDV3 : U_SUBV3) : (op == U_ADD ? U_ADDV4 : U_SUBV4);
            Val r;
            if (block_op(bop, a.n, a, &b, 0, r)) return push(vt, r) ? PROG_OK : PROG_STACK;
        }
        std::vector<MI> seq;
        std::vector<Src> outs;
        for (int k = 0; k < a.n; ++k) {
            seq.push_back(mi(op, {P(a.loc[k]), P(b.loc[k])}));
            outs.push_back(V(k));
        }
        Val r;
        if (!emit(seq, regs_of({&a, &b}), outs, r)) return PROG_REGS;
        return push(vt, r) ? PROG_OK : PROG_STACK;
    }

    // scale: (F s, T v) -> T, out_k = v_k * s (linalg/eval.rs:106-118, 163-175, 245-257, 384-406)
    int scale(int vt) {
        if (!can_pop(VtFloat) || !can_pop(vt)) return PROG_STACK;
        Val s = pop_raw(VtFloat), v = pop(vt);
        if (s.konst >= 0 && (vt == VtVec3 || vt == VtMat2 || vt == VtMat3) && contig(v)) {
            const UOp bop = vt == VtVec3 ? U_SCALEV3K : vt == VtMat2 ? U_SCALEV4K : U_SCALEM3K;
            Val r;
            if (block_op(bop, v.n, v, nullptr, (uint32_t)s.konst, r)) return push(vt, r) ? PROG_OK : PROG_STACK;
        }
        if (s.konst >= 0) materialize(s);
        if ((vt == VtMat2 || vt == VtMat3) && contig(v) && in_reg(s.loc[0])) {
            Val r;
            if (block_op(vt == VtMat2 ? U_SCALEV4 : U_SCALEM3, v.n, v, &s, 0, r))
                return push(vt, r) ? PROG_OK : PROG_STACK;
        }
        if (vt == VtVec3 && contig(v) && in_reg(s.loc[0])) {
            Val r;
            if (block_op(U_SCALEV3, 3, v, &s, 0, r)) return push(vt, r) ? PROG_OK : PROG_STACK;
        }
        std::vector<MI> seq;
        std::vector<Src> outs;
        for (int k = 0; k < v.n; ++k) {
            seq.push_back(mi(U_MUL, {P(v.loc[k]), P(s.loc[0])}));
            outs.push_back(V(k));
        }
        Val r;
        if (!emit(seq, regs_of({&s, &v}), outs, r)) return PROG_REGS;
        return push(vt, r) ? PROG_OK : PROG_STACK;
    }

    int unary_f(UOp op) {
        if (!can_pop(VtFloat)) return PROG_STACK;
        if (!dual_ && stk_[0].back().konst >= 0) {  // a constant operand: read from the parameters
            UOp k = U_COUNT;
            switch (op) {
                case U_SIN: k = U_SINK; break;
                case U_COS: k = U_COSK; break;
                case U_EXP: k = U_EXPK; break;
                case U_LN: k = U_LNK; break;
                case U_SQRT: k = U_SQRTK; break;
                case U_NEG: k = U_NEGK; break;
                case U_SQR: k = U_SQRK; break;
                default: break;
            }
            if (k != U_COUNT) {
                const Val c = pop_raw(VtFloat);
                Val none, r;
                none.n = 0;
                if (!block_op(k, 1, none, nullptr, (uint32_t)c.konst, r) &&
                    !emit({mi(k, {}, (uint32_t)c.konst)}, {}, {V(0)}, r))
                    return PROG_REGS;
                return push(VtFloat, r) ? PROG_OK : PROG_STACK;
            }
        }
        Val a = pop(VtFloat);
        Val r;
        if (!emit({mi(op, {P(a.loc[0])})}, regs_of({&a}), {V(0)}, r)) return PROG_REGS;
        return push(VtFloat, r) ? PROG_OK : PROG_STACK;
    }

    // Float binary op; a pending constant operand becomes a parameter operand (the
    // commutative add and multiply take it on either side: IEEE + and * are commutative).
    int binary_f(UOp op) {
        if (!can_pop(VtFloat, 2)) return PROG_STACK;
        Val b = pop_raw(VtFloat), a = pop_raw(VtFloat);
        if (a.konst >= 0 && b.konst >= 0) materialize(a);
        if (a.konst < 0 && b.konst < 0) {
            push_back_raw(a);
            push_back_raw(b);
            return cwise(VtFloat, op);
        }
        const bool kb = b.konst >= 0;
        const Val& v = kb ? a : b;  // the register operand
        const uint32_t x = (uint32_t)(kb ? b.konst : a.konst);
        UOp k;
        switch (op) {
            case U_ADD: k = U_ADDK; break;
            case U_MUL: k = U_MULK; break;
            case U_SUB: k = kb ? U_SUBK : U_KSUB; break;
            default: k = kb ? U_DIVK : U_KDIV; break;
        }
        Val r;
        if (!block_op(k, 1, v, nullptr, x, r) && !emit({mi(k, {P(v.loc[0])}, x)}, regs_of({&v}), {V(0)}, r))
            return PROG_REGS;
        return push(VtFloat, r) ? PROG_OK : PROG_STACK;
    }
    // Put a popped value back unchanged (same age).
    void push_back_raw(const Val& v) { stk_[sidx(VtFloat)].push_back(v); }
    // A pending Float constant into a register; when none can be found it stays a pending
    // constant location that emit() loads where it is read.
    void materialize(Val& v) {
        if (v.konst < 0) return;
        const int r = alloc();
        if (r >= 0) {
            out_.code.push_back(make_uinstr(U_IMM, (uint32_t)r, 0, 0, (uint32_t)v.konst));
            v.loc[0] = (int16_t)r;
        } else {
            v.loc[0] = (int16_t)(kKonstLoc + v.konst);
        }
        v.konst = -1;
    }

    // generic "consume operands, produce one value" helper
    int produce(int out_vt, const std::vector<MI>& seq, const std::vector<int>& owned,
                const std::vector<Src>& outs) {
        Val r;
        if (!emit(seq, owned, outs, r)) return PROG_REGS;
        return push(out_vt, r) ? PROG_OK : PROG_STACK;
    }

    int step(int op, int arg) {
        switch (op) {
            case LoadVarF: return load_var(VtFloat, arg);
            case LoadVarV2: return load_var(VtVec2, arg);
            case LoadVarV3: return load_var(VtVec3, arg);
            case LoadVarM2: return load_var(VtMat2, arg);
            case LoadVarM3: return load_var(VtMat3, arg);
            case LoadConstF: return load_const(VtFloat, arg);
            case LoadConstV2: return load_const(VtVec2, arg);
            case LoadConstV3: return load_const(VtVec3, arg);
            case LoadConstM2: return load_const(VtMat2, arg);
            case LoadConstM3: return load_const(VtMat3, arg);
            case LoadVarB: case LoadConstB: case LoadVarI: case LoadConstI: return PROG_OPCODE;

            case AddF: return binary_f(U_ADD);
            case SubF: return binary_f(U_SUB);
            case MulF: return binary_f(U_MUL);
            case DivF: return binary_f(U_DIV);
            case SinF: return unary_f(U_SIN);
            case CosF: return unary_f(U_COS);
            case ExpF: return unary_f(U_EXP);
            case SqrF: return unary_f(U_SQR);
            case SqrtF: return unary_f(U_SQRT);
            case LnF: return unary_f(U_LN);
            case NegF: return unary_f(U_NEG);

            case MakeVec2: case MakeVec3: {
                int n = op == MakeVec2 ? 2 : 3;
                if (!can_pop(VtFloat, n)) return PROG_STACK;
                Val r;
                r.n = (uint8_t)n;
                for (int k = n - 1; k >= 0; --k) r.loc[k] = pop(VtFloat).loc[0];
                return push(op == MakeVec2 ? VtVec2 : VtVec3, r) ? PROG_OK : PROG_STACK;
            }
            case GetXV2: case GetYV2: case GetXV3: case GetYV3: case GetZV3: {
                int vt = (op == GetXV2 || op == GetYV2) ? VtVec2 : VtVec3;
                int k = op == GetXV2 ? 0 : op == GetYV2 ? 1 : op - GetXV3;
                if (!can_pop(vt)) return PROG_STACK;
                Val v = pop(vt);
                for (int j = 0; j < v.n; ++j)
                    if (j != k) release(v.loc[j]);
                Val r;
                r.n = 1;
                r.loc[0] = v.loc[k];
                return push(VtFloat, r) ? PROG_OK : PROG_STACK;
            }
            case AddV2: return cwise(VtVec2, U_ADD);
            case SubV2: return cwise(VtVec2, U_SUB);
            case AddV3: return cwise(VtVec3, U_ADD);
            case SubV3: return cwise(VtVec3, U_SUB);
            case AddM2: return cwise(VtMat2, U_ADD);
            case SubM2: return cwise(VtMat2, U_SUB);
            case AddM3: return cwise(VtMat3, U_ADD);
            case SubM3: return cwise(VtMat3, U_SUB);
            case ScaleV2: return scale(VtVec2);
            case ScaleV3: return scale(VtVec3);
            case ScaleM2: return scale(VtMat2);
            case ScaleM3: return scale(VtMat3);

            case DotV2: {  // (a0*b0) + (a1*b1)
                if (!can_pop(VtVec2, 2)) return PROG_STACK;
                Val b = pop(VtVec2), a = pop(VtVec2);
                return produce(VtFloat, {mi(U_MULADD, {P(a.loc[0]), P(b.loc[0]), P(a.loc[1]), P(b.loc[1])})},
                               regs_of({&a, &b}), {V(0)});
            }
            case DotV3: {  // ((a0*b0) + (a1*b1)) + (a2*b2)
                if (!can_pop(VtVec3, 2)) return PROG_STACK;
                Val b = pop(VtVec3), a = pop(VtVec3);
                if (contig(a) && contig(b)) {
                    Val r;
                    if (block_op(U_DOT3, 1, a, &b, 0, r)) return push(VtFloat, r) ? PROG_OK : PROG_STACK;
                }
                return produce(VtFloat,
                               {mi(U_MULADD, {P(a.loc[0]), P(b.loc[0]), P(a.loc[1]), P(b.loc[1])}),
                                mi(U_ADDMUL, {V(0), P(a.loc[2]), P(b.loc[2])})},
                               regs_of({&a, &b}), {V(1)});
            }
            case NormV2: {  // sqrt((v0*v0) + (v1*v1))
                if (!can_pop(VtVec2)) return PROG_STACK;
                Val v = pop(VtVec2);
                return produce(VtFloat,
                               {mi(U_MULADD, {P(v.loc[0]), P(v.loc[0]), P(v.loc[1]), P(v.loc[1])}),
                                mi(U_SQRT, {V(0)})},
                               regs_of({&v}), {V(1)});
            }
            case NormV3: {  // sqrt(((v0*v0) + (v1*v1)) + (v2*v2))
                if (!can_pop(VtVec3)) return PROG_STACK;
                Val v = pop(VtVec3);
                if (contig(v)) {
                    Val r;
                    if (block_op(U_NORM3, 1, v, nullptr, 0, r)) return push(VtFloat, r) ? PROG_OK : PROG_STACK;
                }
                return produce(VtFloat,
                               {mi(U_MULADD, {P(v.loc[0]), P(v.loc[0]), P(v.loc[1]), P(v.loc[1])}),
                                mi(U_SQRTACC, {V(0), P(v.loc[2])})},
                               regs_of({&v}), {V(1)});
            }
            case CrossV3: {
                if (!can_pop(VtVec3, 2)) return PROG_STACK;
                Val b = pop(VtVec3), a = pop(VtVec3);
                if (contig(a) && contig(b)) {
                    Val r;
                    if (block_op(U_CROSS3, 3, a, &b, 0, r)) return push(VtVec3, r) ? PROG_OK : PROG_STACK;
                }
                auto A = [&](int k) { return P(a.loc[k]); };
                auto B = [&](int k) { return P(b.loc[k]); };
                return produce(VtVec3,
                               {mi(U_MULSUB, {A(1), B(2), A(2), B(1)}), mi(U_MULSUB, {A(2), B(0), A(0), B(2)}),
                                mi(U_MULSUB, {A(0), B(1), A(1), B(0)})},
                               regs_of({&a, &b}), {V(0), V(1), V(2)});
            }
            case MakeMat2: {  // columns c0, c1 -> [c0x, c0y, c1x, c1y]
                if (!can_pop(VtVec2, 2)) return PROG_STACK;
                Val c1 = pop(VtVec2), c0 = pop(VtVec2);
                Val r;
                r.n = 4;
                r.loc[0] = c0.loc[0]; r.loc[1] = c0.loc[1]; r.loc[2] = c1.loc[0]; r.loc[3] = c1.loc[1];
                return push(VtMat2, r) ? PROG_OK : PROG_STACK;
            }
            case MakeMat3: {
                if (!can_pop(VtVec3, 3)) return PROG_STACK;
                Val c2 = pop(VtVec3), c1 = pop(VtVec3), c0 = pop(VtVec3);
                Val r;
                r.n = 9;
                for (int k = 0; k < 3; ++k) { r.loc[k] = c0.loc[k]; r.loc[3 + k] = c1.loc[k]; r.loc[6 + k] = c2.loc[k]; }
                return push(VtMat3, r) ? PROG_OK : PROG_STACK;
            }
            case MulM2: {  // out[r+2c] = a[r]*b[2c] + a[r+2]*b[2c+1]
                if (!can_pop(VtMat2, 2)) return PROG_STACK;
                Val b = pop(VtMat2), a = pop(VtMat2);
                if (contig(a) && contig(b)) {
                    Val r;
                    if (block_op(U_MULM2, 4, a, &b, 0, r)) return push(VtMat2, r) ? PROG_OK : PROG_STACK;
                }
                std::vector<MI> seq;
                for (int c = 0; c < 2; ++c)
                    for (int r = 0; r < 2; ++r)
                        seq.push_back(mi(U_MULADD, {P(a.loc[r]), P(b.loc[2 * c]), P(a.loc[r + 2]), P(b.loc[2 * c + 1])}));
                return produce(VtMat2, seq, regs_of({&a, &b}), {V(0), V(1), V(2), V(3)});
            }
            case MulM2V2: {
                if (!can_pop(VtMat2) || !can_pop(VtVec2)) return PROG_STACK;
                Val m = pop(VtMat2), v = pop(VtVec2);
                return produce(VtVec2,
                               {mi(U_MULADD, {P(m.loc[0]), P(v.loc[0]), P(m.loc[2]), P(v.loc[1])}),
                                mi(U_MULADD, {P(m.loc[1]), P(v.loc[0]), P(m.loc[3]), P(v.loc[1])})},
                               regs_of({&m, &v}), {V(0), V(1)});
            }
            case DetM2: {
                if (!can_pop(VtMat2)) return PROG_STACK;
                Val m = pop(VtMat2);
                return produce(VtFloat, {mi(U_MULSUB, {P(m.loc[0]), P(m.loc[3]), P(m.loc[1]), P(m.loc[2])})},
                               regs_of({&m}), {V(0)});
            }
            case TraceM2: {
                if (!can_pop(VtMat2)) return PROG_STACK;
                Val m = pop(VtMat2);
                return produce(VtFloat, {mi(U_ADD, {P(m.loc[0]), P(m.loc[3])})}, regs_of({&m}), {V(0)});
            }
            case TransposeM2: {
                if (!can_pop(VtMat2)) return PROG_STACK;
                Val m = pop(VtMat2);
                std::swap(m.loc[1], m.loc[2]);
                return push(VtMat2, m) ? PROG_OK : PROG_STACK;
            }
            case InverseM2: {
                out_.differentiable = false;
                if (!can_pop(VtMat2)) return PROG_STACK;
                if (dual_) return PROG_OK;  // dual inverse: stack unchanged
                Val m = pop(VtMat2);
                if (contig(m)) {
                    Val r;
                    if (block_op(U_INVM2, 4, m, nullptr, 0, r)) return push(VtMat2, r) ? PROG_OK : PROG_STACK;
                }
                auto M = [&](int k) { return P(m.loc[k]); };
                // det, inv_d, then blended entries (linalg/eval.rs:316-329)
                return produce(VtMat2,
                               {mi(U_MULSUB, {M(0), M(3), M(1), M(2)}),          // v0 det
                                mi(U_RCPSAFE, {V(0)}),                            // v1 inv_d
                                mi(U_INVSEL, {M(3), V(1), V(0)}, kInvOne),       // K=1
                                mi(U_INVSEL, {M(1), V(1), V(0)}, kInvNeg),       // -m1, K=0
                                mi(U_INVSEL, {M(2), V(1), V(0)}, kInvNeg),       // -m2, K=0
                                mi(U_INVSEL, {M(0), V(1), V(0)}, kInvOne)},      // K=1
                               regs_of({&m}), {V(2), V(3), V(4), V(5)});
            }
            case MulM3: {  // out[r+3c] = (a[r]*b[3c] + a[r+3]*b[3c+1]) + a[r+6]*b[3c+2]
                if (!can_pop(VtMat3, 2)) return PROG_STACK;
                Val b = pop(VtMat3), a = pop(VtMat3);
                std::vector<MI> seq;
                std::vector<Src> outs;
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r) {
                        int t = (int)seq.size();
                        seq.push_back(mi(U_MULADD, {P(a.loc[r]), P(b.loc[3 * c]), P(a.loc[r + 3]), P(b.loc[3 * c + 1])}));
                        seq.push_back(mi(U_ADDMUL, {V(t), P(a.loc[r + 6]), P(b.loc[3 * c + 2])}));
                        outs.push_back(V(t + 1));
                    }
                return produce(VtMat3, seq, regs_of({&a, &b}), outs);
            }
            case MulM3V3: {
                if (!can_pop(VtMat3) || !can_pop(VtVec3)) return PROG_STACK;
                Val m = pop(VtMat3), v = pop(VtVec3);
                if (contig(m) && contig(v)) {
                    Val r;
                    if (block_op(U_MULM3V3, 3, m, &v, 0, r)) return push(VtVec3, r) ? PROG_OK : PROG_STACK;
                }
                std::vector<MI> seq;
                std::vector<Src> outs;
                for (int r = 0; r < 3; ++r) {
                    int t = (int)seq.size();
                    seq.push_back(mi(U_MULADD, {P(m.loc[r]), P(v.loc[0]), P(m.loc[r + 3]), P(v.loc[1])}));
                    seq.push_back(mi(U_ADDMUL, {V(t), P(m.loc[r + 6]), P(v.loc[2])}));
                    outs.push_back(V(t + 1));
                }
                return produce(VtVec3, seq, regs_of({&m, &v}), outs);
            }
            case DetM3: {
                if (!can_pop(VtMat3)) return PROG_STACK;
                Val m = pop(VtMat3);
                auto M = [&](int k) { return P(m.loc[k]); };
                return produce(VtFloat,
                               {mi(U_MULSUB, {M(4), M(8), M(5), M(7)}),  // t1
                                mi(U_MULSUB, {M(1), M(8), M(2), M(7)}),  // t2
                                mi(U_MULSUB, {M(1), M(5), M(2), M(4)}),  // t3
                                mi(U_MULSUB, {M(0), V(0), M(3), V(1)}),  // m0*t1 - m3*t2
                                mi(U_ADDMUL, {V(3), M(6), V(2)})},       // + m6*t3
                               regs_of({&m}), {V(4)});
            }
            case TraceM3: {
                if (!can_pop(VtMat3)) return PROG_STACK;
                Val m = pop(VtMat3);
                return produce(VtFloat, {mi(U_ADD3, {P(m.loc[0]), P(m.loc[4]), P(m.loc[8])})}, regs_of({&m}), {V(0)});
            }
            case TransposeM3: {
                if (!can_pop(VtMat3)) return PROG_STACK;
                Val m = pop(VtMat3);
                Val t = m;
                static const int perm[9] = {0, 3, 6, 1, 4, 7, 2, 5, 8};
                for (int k = 0; k < 9; ++k) t.loc[k] = m.loc[perm[k]];
                return push(VtMat3, t) ? PROG_OK : PROG_STACK;
            }
            case InverseM3: {
                out_.differentiable = false;
                if (!can_pop(VtMat3)) return PROG_STACK;
                if (dual_) return PROG_OK;
                Val m = pop(VtMat3);
                auto M = [&](int k) { return P(m.loc[k]); };
                // det (same sequence as DetM3), inv_d, adjugate * inv_d blended (linalg/eval.rs:476-506)
                std::vector<MI> seq = {
                    mi(U_MULSUB, {M(4), M(8), M(5), M(7)}),  // 0 t1 = adj0 numerator
                    mi(U_MULSUB, {M(1), M(8), M(2), M(7)}),  // 1 t2 (adj1 = -t2)
                    mi(U_MULSUB, {M(1), M(5), M(2), M(4)}),  // 2 t3 = adj2 numerator
                    mi(U_MULSUB, {M(0), V(0), M(3), V(1)}),  // 3
                    mi(U_ADDMUL, {V(3), M(6), V(2)}),        // 4 det
                    mi(U_RCPSAFE, {V(4)}),                   // 5 inv_d
                    mi(U_MULSUB, {M(3), M(8), M(5), M(6)}),  // 6 adj3 = -(.)
                    mi(U_MULSUB, {M(0), M(8), M(2), M(6)}),  // 7 adj4
                    mi(U_MULSUB, {M(0), M(5), M(2), M(3)}),  // 8 adj5 = -(.)
                    mi(U_MULSUB, {M(3), M(7), M(4), M(6)}),  // 9 adj6
                    mi(U_MULSUB, {M(0), M(7), M(1), M(6)}),  // 10 adj7 = -(.)
                    mi(U_MULSUB, {M(0), M(4), M(1), M(3)}),  // 11 adj8
                };
                const int num[9] = {0, 1, 2, 6, 7, 8, 9, 10, 11};
                const bool neg[9] = {false, true, false, true, false, true, false, true, false};
                std::vector<Src> outs;
                for (int k = 0; k < 9; ++k) {
                    uint32_t fl = (neg[k] ? 1u : 0u) | ((k == 0 || k == 4 || k == 8) ? 2u : 0u);
                    outs.push_back(V((int)seq.size()));
                    seq.push_back(mi(U_INVSEL, {V(num[k]), V(5), V(4)}, fl));
                }
                return produce(VtMat3, seq, regs_of({&m}), outs);
            }
            default: return PROG_OPCODE;
        }
    }
};

}  // namespace

Lowered lower_program(const RefProgram& p, int num_features, bool dual) {
    Lowerer l(p, num_features, dual);
    return l.run();
}

void flatten_params(const RefProgram& p, std::vector<float>& out) {
    int off = 0;
    for (int i = 0; i < p.n_consts; ++i) {
        int s = vt_size(p.const_types[i]);
        for (int k = 0; k < s; ++k) out.push_back(p.const_values[off + k]);
        off += s ? s : 1;
    }
}

}  // namespace rsr
