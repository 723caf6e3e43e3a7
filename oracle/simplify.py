"""CPU oracle of the post-fit simplification (Individual::simplify, individual.rs:77-85).

TEST INFRASTRUCTURE ONLY (the checker of rsr_simplify_programs; the product never uses
it).

A tree restatement of simplify_ast (src/engine/expr/simplify.rs:13-135): the postfix
program is parsed into a tree and simplified bottom-up, which visits every operator
after its arguments exactly as the reference's single postfix pass does.  Rules follow
try_simplify of src/domains/basic/mod.rs:45-115 and src/domains/linalg/mod.rs:227-310
with the folds of :728-890; the emitted program follows Program::from_nodes
(src/engine/expr/program.rs:34-52).  Facts of the reference pass that a tree form must
state explicitly:
- every constant node carries its own index, so two argument subtrees compare equal
  (for x - x, x / x, cross(x, x)) only when neither contains a constant;
- an inverse pair (ln(exp t), exp(ln t), sqrt(sqr t), sqr(sqrt t)) leaves t marked as
  not constant even when t is a single constant (simplify.rs:118-121);
- a kept argument keeps its constant mark (simplify.rs:108-110).
Folded f32 arithmetic uses numpy float32 scalars (one IEEE rounding per operation);
folded sin/cos/exp/ln call the C library's sinf/cosf/expf/logf, which is what Rust's
f32 methods call on Linux.  Only well-formed programs are handled.  Pinned by 45 known
answers worked out by hand from the rules, and by MSE preservation where the rules are
exact (tests/test_simplify.py).
"""
from __future__ import annotations

import ctypes
import ctypes.util
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from rsr_amd.isa import LOAD_CONST, SIG, VT, VT_SIZE, Op, Program, Scalar

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
for _n in ("sinf", "cosf", "expf", "logf"):
    getattr(_libm, _n).restype = ctypes.c_float
    getattr(_libm, _n).argtypes = [ctypes.c_float]

f32 = np.float32
LOADS_CONST = {int(v) for v in LOAD_CONST.values()}
LOAD_TYPE = {int(Op.LoadVarF): VT.Float, int(Op.LoadConstF): VT.Float, int(Op.LoadVarB): VT.Bool,
             int(Op.LoadConstB): VT.Bool, int(Op.LoadVarI): VT.Int, int(Op.LoadConstI): VT.Int,
             int(Op.LoadVarV2): VT.Vec2, int(Op.LoadConstV2): VT.Vec2, int(Op.LoadVarV3): VT.Vec3,
             int(Op.LoadConstV3): VT.Vec3, int(Op.LoadVarM2): VT.Mat2, int(Op.LoadConstM2): VT.Mat2,
             int(Op.LoadVarM3): VT.Mat3, int(Op.LoadConstM3): VT.Mat3}


@dataclass
class T:
    """A subtree: a leaf (load) or an operator with children."""
    op: int
    arg: int = 0
    kids: List["T"] = field(default_factory=list)
    const: Optional[Tuple[VT, Tuple[np.float32, ...]]] = None  # the leaf's constant value
    const_mark: bool = False  # the reference pass treats this subtree as one constant
    has_const: bool = False   # some node of the subtree is a constant


def _leaf_const(vt: VT, vals) -> T:
    vals = tuple(f32(v) for v in vals)
    return T(int(LOAD_CONST[vt]), 0, [], (vt, vals), True, True)


def _zero(vt: VT) -> T:
    return _leaf_const(vt, [0.0] * (VT_SIZE[vt] or 1))


def _is_zero(c) -> bool:
    vt, v = c
    if vt in (VT.Int, VT.Bool):
        return False
    return all(abs(x) < f32(1e-6) for x in v)


def _is_one(c) -> bool:
    vt, v = c
    return vt == VT.Float and abs(v[0] - f32(1.0)) < f32(1e-6)


def _is_identity(c) -> bool:
    vt, m = c
    if vt == VT.Mat2:
        want = (1, 0, 0, 1)
    elif vt == VT.Mat3:
        want = (1, 0, 0, 0, 1, 0, 0, 0, 1)
    else:
        return False
    return all(abs(x - f32(w)) < f32(1e-6) for x, w in zip(m, want))


def _emit(t: T, out: list):
    for k in t.kids:
        _emit(k, out)
    out.append(t)


def _equal(a: T, b: T) -> bool:
    """Node-sequence equality of the reference's output slices (constants never equal)."""
    if a.has_const or b.has_const:
        return False
    sa, sb = [], []
    _emit(a, sa)
    _emit(b, sb)
    return len(sa) == len(sb) and all(x.op == y.op and x.arg == y.arg for x, y in zip(sa, sb))


def _fold_basic(op: int, cs) -> Optional[T]:
    if not cs or cs[0][0] != VT.Float:
        return None
    a = cs[0][1][0]
    with np.errstate(all="ignore"):
        if len(cs) == 1:
            r = {Op.SinF: lambda: f32(_libm.sinf(float(a))), Op.CosF: lambda: f32(_libm.cosf(float(a))),
                 Op.ExpF: lambda: f32(_libm.expf(float(a))), Op.LnF: lambda: f32(_libm.logf(float(a))),
                 Op.SqrF: lambda: a * a, Op.SqrtF: lambda: np.sqrt(a), Op.NegF: lambda: -a}.get(Op(op))
            r = r() if r else f32(np.nan)
        elif cs[1][0] == VT.Float:
            b = cs[1][1][0]
            r = {Op.AddF: a + b, Op.SubF: a - b, Op.MulF: a * b, Op.DivF: a / b}.get(Op(op), f32(np.nan))
        else:
            return None
    return _leaf_const(VT.Float, [r]) if np.isfinite(r) else None


def _det3(m):
    return (m[0] * (m[4] * m[8] - m[5] * m[7]) - m[3] * (m[1] * m[8] - m[2] * m[7])) + m[6] * (m[1] * m[5] - m[2] * m[4])


def _fold_linalg(op: int, cs) -> Optional[T]:
    types = tuple(c[0] for c in cs)
    v = [c[1] for c in cs]
    o = Op(op)
    F_, V2_, V3_, M2_, M3_ = VT.Float, VT.Vec2, VT.Vec3, VT.Mat2, VT.Mat3
    if types != SIG[o][0]:
        return None
    with np.errstate(all="ignore"):
        if o == Op.MakeVec2:
            return _leaf_const(V2_, [v[0][0], v[1][0]])
        if o == Op.MakeVec3:
            return _leaf_const(V3_, [v[0][0], v[1][0], v[2][0]])
        if o in (Op.GetXV2, Op.GetXV3):
            return _leaf_const(F_, [v[0][0]])
        if o in (Op.GetYV2, Op.GetYV3):
            return _leaf_const(F_, [v[0][1]])
        if o == Op.GetZV3:
            return _leaf_const(F_, [v[0][2]])
        if o in (Op.AddV2, Op.AddV3, Op.AddM2, Op.AddM3):
            return _leaf_const(SIG[o][1], [x + y for x, y in zip(v[0], v[1])])
        if o in (Op.SubV2, Op.SubV3, Op.SubM2, Op.SubM3):
            return _leaf_const(SIG[o][1], [x - y for x, y in zip(v[0], v[1])])
        if o in (Op.ScaleV2, Op.ScaleV3, Op.ScaleM2, Op.ScaleM3):
            return _leaf_const(SIG[o][1], [v[0][0] * y for y in v[1]])
        if o in (Op.DotV2, Op.DotV3):
            s = v[0][0] * v[1][0]
            for x, y in list(zip(v[0], v[1]))[1:]:
                s = s + x * y
            return _leaf_const(F_, [s])
        if o in (Op.NormV2, Op.NormV3):
            s = v[0][0] * v[0][0]
            for x in v[0][1:]:
                s = s + x * x
            return _leaf_const(F_, [np.sqrt(s)])
        if o == Op.CrossV3:
            a, b = v
            return _leaf_const(V3_, [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])
        if o == Op.MakeMat2:
            return _leaf_const(M2_, list(v[0]) + list(v[1]))
        if o == Op.MakeMat3:
            return _leaf_const(M3_, list(v[0]) + list(v[1]) + list(v[2]))
        if o == Op.MulM2:
            a, b = v
            return _leaf_const(M2_, [a[0] * b[0] + a[2] * b[1], a[1] * b[0] + a[3] * b[1],
                                     a[0] * b[2] + a[2] * b[3], a[1] * b[2] + a[3] * b[3]])
        if o == Op.MulM2V2:
            m, x = v
            return _leaf_const(V2_, [m[0] * x[0] + m[2] * x[1], m[1] * x[0] + m[3] * x[1]])
        if o == Op.DetM2:
            m = v[0]
            return _leaf_const(F_, [m[0] * m[3] - m[1] * m[2]])
        if o == Op.TraceM2:
            return _leaf_const(F_, [v[0][0] + v[0][3]])
        if o == Op.TransposeM2:
            m = v[0]
            return _leaf_const(M2_, [m[0], m[2], m[1], m[3]])
        if o == Op.InverseM2:
            m = v[0]
            det = m[0] * m[3] - m[1] * m[2]
            if not abs(det) > f32(1e-9):
                return None
            return _leaf_const(M2_, [m[3] / det, -m[1] / det, -m[2] / det, m[0] / det])
        if o == Op.MulM3:
            a, b = v
            return _leaf_const(M3_, [(a[r] * b[3 * c] + a[3 + r] * b[3 * c + 1]) + a[6 + r] * b[3 * c + 2]
                                     for c in range(3) for r in range(3)])
        if o == Op.MulM3V3:
            m, x = v
            return _leaf_const(V3_, [(m[r] * x[0] + m[3 + r] * x[1]) + m[6 + r] * x[2] for r in range(3)])
        if o == Op.TraceM3:
            m = v[0]
            return _leaf_const(F_, [(m[0] + m[4]) + m[8]])
        if o == Op.TransposeM3:
            m = v[0]
            return _leaf_const(M3_, [m[0], m[3], m[6], m[1], m[4], m[7], m[2], m[5], m[8]])
        if o == Op.DetM3:
            return _leaf_const(F_, [_det3(v[0])])
        if o == Op.InverseM3:
            m = v[0]
            det = _det3(m)
            if not abs(det) > f32(1e-9):
                return None
            inv = f32(1.0) / det
            cof = [m[4] * m[8] - m[5] * m[7], -(m[1] * m[8] - m[2] * m[7]), m[1] * m[5] - m[2] * m[4],
                   -(m[3] * m[8] - m[5] * m[6]), m[0] * m[8] - m[2] * m[6], -(m[0] * m[5] - m[2] * m[3]),
                   m[3] * m[7] - m[4] * m[6], -(m[0] * m[7] - m[1] * m[6]), m[0] * m[4] - m[1] * m[3]]
            return _leaf_const(M3_, [x * inv for x in cof])
    return None


BASIC = {Op.AddF, Op.SubF, Op.MulF, Op.DivF, Op.SinF, Op.CosF, Op.ExpF, Op.SqrF, Op.SqrtF, Op.LnF, Op.NegF}
ZERO_TYPE = {Op.SubV2: VT.Vec2, Op.ScaleV2: VT.Vec2, Op.MulM2V2: VT.Vec2, Op.SubV3: VT.Vec3, Op.ScaleV3: VT.Vec3,
             Op.CrossV3: VT.Vec3, Op.MulM3V3: VT.Vec3, Op.SubM2: VT.Mat2, Op.ScaleM2: VT.Mat2, Op.MulM2: VT.Mat2,
             Op.SubM3: VT.Mat3, Op.ScaleM3: VT.Mat3, Op.MulM3: VT.Mat3}
INVERSE = {(Op.LnF, Op.ExpF), (Op.ExpF, Op.LnF), (Op.SqrtF, Op.SqrF), (Op.SqrF, Op.SqrtF)}


def _rule(o: Op, kids: List[T]):
    """('const', T) | ('keep', i) | None."""
    marks = [k.const if k.const_mark else None for k in kids]
    all_const = all(m is not None for m in marks)
    if o in BASIC:
        if all_const:
            r = _fold_basic(int(o), marks)
            if r is not None:
                return "const", r
        if len(kids) == 2:
            az = marks[0] is not None and _is_zero(marks[0])
            bz = marks[1] is not None and _is_zero(marks[1])
            a1 = marks[0] is not None and _is_one(marks[0])
            b1 = marks[1] is not None and _is_one(marks[1])
            eq = _equal(kids[0], kids[1])
            if o == Op.AddF:
                if bz: return "keep", 0
                if az: return "keep", 1
            elif o == Op.MulF:
                if b1: return "keep", 0
                if a1: return "keep", 1
                if az or bz: return "const", _zero(VT.Float)
            elif o == Op.SubF:
                if bz: return "keep", 0
                if eq: return "const", _zero(VT.Float)
            elif o == Op.DivF:
                if b1: return "keep", 0
                if az and not bz: return "const", _zero(VT.Float)
                if eq: return "const", _leaf_const(VT.Float, [1.0])
        return None
    if all_const:
        r = _fold_linalg(int(o), marks)
        if r is not None:
            return "const", r
    if len(kids) == 2:
        az = marks[0] is not None and _is_zero(marks[0])
        bz = marks[1] is not None and _is_zero(marks[1])
        a1 = marks[0] is not None and _is_one(marks[0])
        ai = marks[0] is not None and _is_identity(marks[0])
        bi = marks[1] is not None and _is_identity(marks[1])
        eq = _equal(kids[0], kids[1])
        if o in (Op.AddV2, Op.AddV3, Op.AddM2, Op.AddM3):
            if bz: return "keep", 0
            if az: return "keep", 1
        elif o in (Op.SubV2, Op.SubV3, Op.SubM2, Op.SubM3):
            if bz: return "keep", 0
            if eq: return "const", _zero(ZERO_TYPE[o])
        elif o in (Op.ScaleV2, Op.ScaleV3, Op.ScaleM2, Op.ScaleM3):
            if az: return "const", _zero(ZERO_TYPE[o])
            if a1: return "keep", 1
        elif o in (Op.DotV2, Op.DotV3):
            if az or bz: return "const", _zero(VT.Float)
        elif o == Op.CrossV3:
            if az or bz or eq: return "const", _zero(VT.Vec3)
        elif o in (Op.MulM2, Op.MulM3):
            if ai: return "keep", 1
            if bi: return "keep", 0
            if az or bz: return "const", _zero(ZERO_TYPE[o])
        elif o in (Op.MulM2V2, Op.MulM3V3):
            if ai: return "keep", 1
            if az or bz: return "const", _zero(ZERO_TYPE[o])
    return None


def _simplify(t: T) -> T:
    if not t.kids:
        return t
    kids = [_simplify(k) for k in t.kids]
    o = Op(t.op)
    act = _rule(o, kids)
    if act is not None and act[0] == "const":
        return act[1]
    if act is not None:
        return kids[act[1]]
    if len(kids) == 1 and kids[0].kids and (o, Op(kids[0].op)) in INVERSE:
        inner = kids[0].kids[0]
        return T(inner.op, inner.arg, inner.kids, inner.const, False, inner.has_const)
    return T(t.op, 0, kids, None, False, any(k.has_const for k in kids))


def simplify(prog: Program) -> Program:
    """Individual::simplify then Program::from_nodes of one well-formed program."""
    stack: List[T] = []
    for op, arg in prog.code:
        op = int(op)
        if op in LOAD_TYPE:
            if op in LOADS_CONST:
                c = prog.constants[arg]
                stack.append(T(op, 0, [], (VT(c.vt), tuple(f32(x) for x in c.vals)), True, True))
            else:
                stack.append(T(op, int(arg)))
            continue
        n = len(SIG[Op(op)][0])
        kids = stack[len(stack) - n:]
        del stack[len(stack) - n:]
        stack.append(T(op, 0, kids))
    nodes: List[T] = []
    for t in stack:
        _emit(_simplify(t), nodes)
    code, consts = [], []
    for t in nodes:
        if t.op in LOADS_CONST:
            vt, vals = t.const
            consts.append(Scalar(vt, tuple(float(x) for x in vals)))
            code.append((t.op, len(consts) - 1))
        else:
            code.append((t.op, t.arg))
    return Program(code, consts)
