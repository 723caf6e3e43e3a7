"""Reference instruction set, value types and program containers (host side).

Mirrors the reference's `Instruction` enum (src/engine/domain.rs:46-63 load variants,
src/domains/basic/mod.rs:9-20, src/domains/linalg/mod.rs:9-51, composed in
src/lib.rs:44-48), `ValueType` (src/engine/eval/types.rs:1-10), `Scalar`
(src/engine/eval/scalar.rs:3-12) and `Program` (src/engine/expr/program.rs:17-23).
Opcode numbers are the declaration order; they are the C ABI wire format.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np


class Op(enum.IntEnum):
    LoadVarF = 0
    LoadConstF = 1
    LoadVarB = 2
    LoadConstB = 3
    LoadVarI = 4
    LoadConstI = 5
    LoadVarV2 = 6
    LoadConstV2 = 7
    LoadVarV3 = 8
    LoadConstV3 = 9
    LoadVarM2 = 10
    LoadConstM2 = 11
    LoadVarM3 = 12
    LoadConstM3 = 13
    # Basic
    AddF = 14
    SubF = 15
    MulF = 16
    DivF = 17
    SinF = 18
    CosF = 19
    ExpF = 20
    SqrF = 21
    SqrtF = 22
    LnF = 23
    # Linalg
    MakeVec2 = 24
    MakeVec3 = 25
    GetXV2 = 26
    GetYV2 = 27
    GetXV3 = 28
    GetYV3 = 29
    GetZV3 = 30
    AddV2 = 31
    SubV2 = 32
    ScaleV2 = 33
    DotV2 = 34
    NormV2 = 35
    AddV3 = 36
    SubV3 = 37
    ScaleV3 = 38
    DotV3 = 39
    NormV3 = 40
    CrossV3 = 41
    MakeMat2 = 42
    AddM2 = 43
    SubM2 = 44
    ScaleM2 = 45
    MulM2 = 46
    MulM2V2 = 47
    DetM2 = 48
    TraceM2 = 49
    TransposeM2 = 50
    InverseM2 = 51
    MakeMat3 = 52
    AddM3 = 53
    SubM3 = 54
    ScaleM3 = 55
    MulM3 = 56
    MulM3V3 = 57
    DetM3 = 58
    TraceM3 = 59
    TransposeM3 = 60
    InverseM3 = 61
    # extension (not a reference opcode): exact sign flip
    NegF = 62


class VT(enum.IntEnum):
    Float = 0
    Int = 1
    Bool = 2
    Vec2 = 3
    Vec3 = 4
    Mat2 = 5
    Mat3 = 6


VT_SIZE = {VT.Float: 1, VT.Int: 0, VT.Bool: 0, VT.Vec2: 2, VT.Vec3: 3, VT.Mat2: 4, VT.Mat3: 9}

LOAD_VAR = {VT.Float: Op.LoadVarF, VT.Vec2: Op.LoadVarV2, VT.Vec3: Op.LoadVarV3, VT.Mat2: Op.LoadVarM2,
            VT.Mat3: Op.LoadVarM3, VT.Bool: Op.LoadVarB, VT.Int: Op.LoadVarI}
LOAD_CONST = {VT.Float: Op.LoadConstF, VT.Vec2: Op.LoadConstV2, VT.Vec3: Op.LoadConstV3, VT.Mat2: Op.LoadConstM2,
              VT.Mat3: Op.LoadConstM3, VT.Bool: Op.LoadConstB, VT.Int: Op.LoadConstI}

F, V2, V3, M2, M3 = VT.Float, VT.Vec2, VT.Vec3, VT.Mat2, VT.Mat3

# (argument types, return type) per operator (basic/mod.rs:135-154, linalg/mod.rs:340-432)
SIG = {
    Op.AddF: ((F, F), F), Op.SubF: ((F, F), F), Op.MulF: ((F, F), F), Op.DivF: ((F, F), F),
    Op.SinF: ((F,), F), Op.CosF: ((F,), F), Op.ExpF: ((F,), F), Op.SqrF: ((F,), F), Op.SqrtF: ((F,), F),
    Op.LnF: ((F,), F), Op.NegF: ((F,), F),
    Op.MakeVec2: ((F, F), V2), Op.MakeVec3: ((F, F, F), V3),
    Op.GetXV2: ((V2,), F), Op.GetYV2: ((V2,), F), Op.GetXV3: ((V3,), F), Op.GetYV3: ((V3,), F),
    Op.GetZV3: ((V3,), F),
    Op.AddV2: ((V2, V2), V2), Op.SubV2: ((V2, V2), V2), Op.ScaleV2: ((F, V2), V2), Op.DotV2: ((V2, V2), F),
    Op.NormV2: ((V2,), F),
    Op.AddV3: ((V3, V3), V3), Op.SubV3: ((V3, V3), V3), Op.ScaleV3: ((F, V3), V3), Op.DotV3: ((V3, V3), F),
    Op.NormV3: ((V3,), F), Op.CrossV3: ((V3, V3), V3),
    Op.MakeMat2: ((V2, V2), M2), Op.AddM2: ((M2, M2), M2), Op.SubM2: ((M2, M2), M2), Op.ScaleM2: ((F, M2), M2),
    Op.MulM2: ((M2, M2), M2), Op.MulM2V2: ((M2, V2), V2), Op.DetM2: ((M2,), F), Op.TraceM2: ((M2,), F),
    Op.TransposeM2: ((M2,), M2), Op.InverseM2: ((M2,), M2),
    Op.MakeMat3: ((V3, V3, V3), M3), Op.AddM3: ((M3, M3), M3), Op.SubM3: ((M3, M3), M3),
    Op.ScaleM3: ((F, M3), M3), Op.MulM3: ((M3, M3), M3), Op.MulM3V3: ((M3, V3), V3), Op.DetM3: ((M3,), F),
    Op.TraceM3: ((M3,), F), Op.TransposeM3: ((M3,), M3), Op.InverseM3: ((M3,), M3),
}

# f32 operations per point per operator (transcendental = 1, loads/relabels = 0)
ALG_OPS = {
    Op.AddF: 1, Op.SubF: 1, Op.MulF: 1, Op.DivF: 1, Op.SinF: 1, Op.CosF: 1, Op.ExpF: 1, Op.SqrF: 1, Op.SqrtF: 1,
    Op.LnF: 1, Op.NegF: 1,
    Op.AddV2: 2, Op.SubV2: 2, Op.ScaleV2: 2, Op.DotV2: 3, Op.NormV2: 4,
    Op.AddV3: 3, Op.SubV3: 3, Op.ScaleV3: 3, Op.DotV3: 5, Op.NormV3: 6, Op.CrossV3: 9,
    Op.AddM2: 4, Op.SubM2: 4, Op.ScaleM2: 4, Op.MulM2: 12, Op.MulM2V2: 6, Op.DetM2: 3, Op.TraceM2: 1,
    Op.InverseM2: 8,
    Op.AddM3: 9, Op.SubM3: 9, Op.ScaleM3: 9, Op.MulM3: 45, Op.MulM3V3: 15, Op.DetM3: 14, Op.TraceM3: 2,
    Op.InverseM3: 51,
}


@dataclass
class Scalar:
    """A constant (scalar.rs:3-12): value type + flat payload (matrices column-major)."""
    vt: VT
    vals: Tuple[float, ...]

    @staticmethod
    def f(x: float) -> "Scalar":
        return Scalar(VT.Float, (float(x),))


@dataclass
class Program:
    """Postfix program: (opcode, operand) pairs + constants in postfix order (program.rs:17-23)."""
    code: List[Tuple[int, int]] = field(default_factory=list)
    constants: List[Scalar] = field(default_factory=list)

    def __len__(self) -> int:
        return len(self.code)

    def param_count(self) -> int:
        return sum(VT_SIZE[c.vt] for c in self.constants)

    def flatten_params(self) -> List[float]:
        out: List[float] = []
        for c in self.constants:
            if VT_SIZE[c.vt]:
                out.extend(c.vals)
        return out


def pack_programs(progs: Sequence[Program]):
    """Pack programs into the C ABI arrays (ops, args, code_offsets, const_types,
    const_offsets, const_values, value_offsets)."""
    n = len(progs)
    code_off = np.zeros(n + 1, np.int64)
    const_off = np.zeros(n + 1, np.int64)
    val_off = np.zeros(n + 1, np.int64)
    ops, args, ctypes, vals = [], [], [], []
    for i, p in enumerate(progs):
        for o, a in p.code:
            ops.append(int(o))
            args.append(int(a))
        for c in p.constants:
            ctypes.append(int(c.vt))
            payload = list(c.vals) if VT_SIZE[c.vt] else [float(c.vals[0]) if c.vals else 0.0]
            vals.extend(payload)
        code_off[i + 1] = len(ops)
        const_off[i + 1] = len(ctypes)
        val_off[i + 1] = len(vals)
    return (np.asarray(ops, np.uint8), np.asarray(args, np.uint16), code_off,
            np.asarray(ctypes, np.uint8), const_off, np.asarray(vals, np.float32), val_off)


def unpack_programs(ops, args, code_off, ctypes, const_off, vals, val_off) -> List[Program]:
    """Inverse of pack_programs."""
    # Python lists once: indexing numpy arrays element by element costs ~10x more
    co, cf, vf = (np.asarray(a[:len(code_off)]).tolist() for a in (code_off, const_off, val_off))
    pairs = list(zip(np.asarray(ops[:co[-1]]).tolist(), np.asarray(args[:co[-1]]).tolist()))
    cts = np.asarray(ctypes[:cf[-1]]).tolist()
    vv = np.asarray(vals[:vf[-1]], np.float32).tolist()
    vts = {t: (VT(t), VT_SIZE[VT(t)] or 1) for t in set(cts)}
    out = []
    for i in range(len(co) - 1):
        consts = []
        v = vf[i]
        for k in range(cf[i], cf[i + 1]):
            vt, n = vts[cts[k]]
            consts.append(Scalar(vt, tuple(vv[v:v + n])))
            v += n
        out.append(Program(pairs[co[i]:co[i + 1]], consts))
    return out


def feature_count(types: Sequence[int]) -> int:
    """Flat feature components of a schema (dataset.rs:46-55; Int/Bool count one)."""
    return sum(VT_SIZE[VT(t)] or 1 for t in types)


def node_evals(progs: Sequence[Program], n_points: int) -> int:
    """Metric numerator: sum of len(code) (every postfix instruction) times points."""
    return sum(len(p.code) for p in progs) * int(n_points)


def alg_flops(progs: Sequence[Program]) -> int:
    """Algorithmic f32 operations per point summed over programs (+3 per program for the MSE)."""
    return sum(sum(ALG_OPS.get(Op(o), 0) for o, _ in p.code) + 3 for p in progs)


# Device micro-op names in opcode order and the operand flags (rsr_amd/csrc/isa.h); for
# tests that inspect the lowered code through api.lower_code.
UOPS = ("IMM ADD SUB MUL DIV SQR SQRT SIN COS EXP LN NEG MULADD MULSUB ADDMUL SQRTACC ADD3 RCPSAFE INVSEL "
        "SPILL FILL ADDV3 SUBV3 SCALEV3 IMMV3 CROSS3 DOT3 NORM3 ADDK SUBK KSUB MULK DIVK KDIV SCALEV3K "
        "MULM3V3 SCALEM3 SCALEM3K ADDV4 SUBV4 SCALEV4 SCALEV4K MULM2 INVM2 RSUB RDIV RSUBV3 RCROSS3 "
        "SCALEV3S MULM3V3R SINU COSU EXPU LNU SQRTU SINK COSK EXPK LNK SQRTK NEGK SQRK").split()
UFLAG_FWD, UFLAG_VEC3, UFLAG_NONE, UFLAG_NOWB = 1 << 8, 1 << 9, 1 << 10, 1 << 11

