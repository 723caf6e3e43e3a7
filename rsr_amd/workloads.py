"""Synthetic workloads: datasets, program populations and known-answer programs.

Known-answer programs restate the target formulas of the reference's own tests in
the same f32 operation order, so the reference evaluator gives MSE exactly 0 on
them: Lorentz force (tests/linalg.rs:90-98, src/main.rs:31-38), transform error
(tests/linalg.rs:171-179), rigid-body energy with a constant Mat3
(tests/linalg.rs:50-62), the square target (tests/basic.rs:41), the linear target
(tests/basic.rs:95) and the hidden-dimension product (tests/basic.rs:141).
"""
from __future__ import annotations

import numpy as np

from .isa import Op, Program, Scalar, VT, SIG

f32 = np.float32


# ----------------------------------------------------------------------------
# known-answer programs
# ----------------------------------------------------------------------------
def square_program() -> Program:
    """((2.5 * x) * x) - 1.2"""
    return Program([(Op.LoadConstF, 0), (Op.LoadVarF, 0), (Op.MulF, 0), (Op.LoadVarF, 0), (Op.MulF, 0),
                    (Op.LoadConstF, 1), (Op.SubF, 0)], [Scalar.f(2.5), Scalar.f(1.2)])


def lorentz_program() -> Program:
    """||E + v x B|| over features E(0..2), v(3..5), B(6..8)."""
    return Program([(Op.LoadVarV3, 0), (Op.LoadVarV3, 3), (Op.LoadVarV3, 6), (Op.CrossV3, 0), (Op.AddV3, 0),
                    (Op.NormV3, 0)], [])


def transform_program() -> Program:
    """||M v - u|| over features M(0..8), v(9..11), u(12..14)."""
    return Program([(Op.LoadVarM3, 0), (Op.LoadVarV3, 9), (Op.MulM3V3, 0), (Op.LoadVarV3, 12), (Op.SubV3, 0),
                    (Op.NormV3, 0)], [])


RIGID_I = (2.0, 0.1, 0.0, 0.1, 1.5, -0.2, 0.0, -0.2, 1.0)


def rigid_body_program() -> Program:
    """w . (I w) with the constant inertia matrix of tests/linalg.rs:50."""
    return Program([(Op.LoadVarV3, 0), (Op.LoadConstM3, 0), (Op.LoadVarV3, 0), (Op.MulM3V3, 0), (Op.DotV3, 0)],
                   [Scalar(VT.Mat3, tuple(float(f32(v)) for v in RIGID_I))])


def linear_program() -> Program:
    """((2*x0) - (3.5*x1)) + (1.2*x2)"""
    return Program([(Op.LoadConstF, 0), (Op.LoadVarF, 0), (Op.MulF, 0), (Op.LoadConstF, 1), (Op.LoadVarF, 1),
                    (Op.MulF, 0), (Op.SubF, 0), (Op.LoadConstF, 2), (Op.LoadVarF, 2), (Op.MulF, 0), (Op.AddF, 0)],
                   [Scalar.f(2.0), Scalar.f(3.5), Scalar.f(1.2)])


def inverse_trace_program() -> Program:
    """tr(A^-1 B) + det(A) over features A(0..3), B(4..7)."""
    return Program([(Op.LoadVarM2, 0), (Op.InverseM2, 0), (Op.LoadVarM2, 4), (Op.MulM2, 0), (Op.TraceM2, 0),
                    (Op.LoadVarM2, 0), (Op.DetM2, 0), (Op.AddF, 0)], [])


# ----------------------------------------------------------------------------
# datasets (rows [N][F], targets [N]) -- targets use the programs' f32 op order
# ----------------------------------------------------------------------------
def data_c0(n: int = 400, seed: int = 1000):
    """scalar target y = ((2.5 x) x) - 1.2, x ~ U(-3, 3)"""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-3.0, 3.0, n).astype(f32)
    y = ((f32(2.5) * x) * x) - f32(1.2)
    return x.reshape(n, 1), y.astype(f32), [VT.Float]


def lorentz_target(E, v, B):
    cx = v[:, 1] * B[:, 2] - v[:, 2] * B[:, 1]
    cy = v[:, 2] * B[:, 0] - v[:, 0] * B[:, 2]
    cz = v[:, 0] * B[:, 1] - v[:, 1] * B[:, 0]
    fx, fy, fz = E[:, 0] + cx, E[:, 1] + cy, E[:, 2] + cz
    return np.sqrt((fx * fx + fy * fy) + fz * fz).astype(f32)


def data_c1(n: int = 100_000, seed: int = 1001):
    """Lorentz y = ||E + v x B||, E, v, B ~ N(0, 1)"""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 9)).astype(f32)
    y = lorentz_target(X[:, 0:3], X[:, 3:6], X[:, 6:9])
    return X, y, [VT.Vec3, VT.Vec3, VT.Vec3]


def transform_target(M, v, u):
    mv0 = (M[:, 0] * v[:, 0] + M[:, 3] * v[:, 1]) + M[:, 6] * v[:, 2]
    mv1 = (M[:, 1] * v[:, 0] + M[:, 4] * v[:, 1]) + M[:, 7] * v[:, 2]
    mv2 = (M[:, 2] * v[:, 0] + M[:, 5] * v[:, 1]) + M[:, 8] * v[:, 2]
    d0, d1, d2 = mv0 - u[:, 0], mv1 - u[:, 1], mv2 - u[:, 2]
    return np.sqrt((d0 * d0 + d1 * d1) + d2 * d2).astype(f32)


def data_c2(n: int = 20_000, seed: int = 1002):
    """y = ||M v - u||, M, v, u ~ N(0, 1)"""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 15)).astype(f32)
    y = transform_target(X[:, 0:9], X[:, 9:12], X[:, 12:15])
    return X, y, [VT.Mat3, VT.Vec3, VT.Vec3]


def inverse_trace_target(A, B):
    det = A[:, 0] * A[:, 3] - A[:, 1] * A[:, 2]
    sing = np.abs(det) < f32(1e-9)
    inv = f32(1.0) / np.where(sing, f32(1.0), det)
    i0 = np.where(sing, f32(1), A[:, 3] * inv)
    i1 = np.where(sing, f32(0), (-A[:, 1]) * inv)
    i2 = np.where(sing, f32(0), (-A[:, 2]) * inv)
    i3 = np.where(sing, f32(1), A[:, 0] * inv)
    ab0 = i0 * B[:, 0] + i2 * B[:, 1]
    ab3 = i1 * B[:, 2] + i3 * B[:, 3]
    return ((ab0 + ab3) + det).astype(f32)


def data_c3(n: int = 50_000, seed: int = 1003):
    """y = tr(A^-1 B) + det(A), A = I + 0.3 N(0,1) with cond(A) < 50, B ~ N(0, 1)"""
    rng = np.random.default_rng(seed)
    out = []
    while sum(len(o) for o in out) < n:
        A = (np.eye(2, dtype=np.float64)[None] + 0.3 * rng.standard_normal((2 * n, 2, 2)))
        c = np.linalg.cond(A)
        out.append(A[c < 50])
    A = np.concatenate(out)[:n]
    # column-major Mat2 payload [a00, a10, a01, a11]
    Af = np.stack([A[:, 0, 0], A[:, 1, 0], A[:, 0, 1], A[:, 1, 1]], 1).astype(f32)
    Bf = rng.standard_normal((n, 4)).astype(f32)
    X = np.concatenate([Af, Bf], 1)
    return X, inverse_trace_target(Af, Bf), [VT.Mat2, VT.Mat2]


# ----------------------------------------------------------------------------
# program populations
# ----------------------------------------------------------------------------
class _Builder:
    def __init__(self):
        self.code = []
        self.consts = []

    def const_f(self, v):
        self.code.append((Op.LoadConstF, len(self.consts)))
        self.consts.append(Scalar.f(float(f32(v))))

    def prog(self):
        return Program(self.code, self.consts)


def gen_c0(n: int = 1000, seed: int = 2000):
    """Ramped half-and-half, depth 2-6, ops {+,-,*,/,sin,cos,exp,sqrt,neg}, leaves x or c ~ U(-5, 5)."""
    rng = np.random.default_rng(seed)
    binary = [Op.AddF, Op.SubF, Op.MulF, Op.DivF]
    unary = [Op.SinF, Op.CosF, Op.ExpF, Op.SqrtF, Op.NegF]
    ops = binary + unary
    progs = []
    for i in range(n):
        depth = 2 + (i // 2) % 5
        full = (i % 2) == 0
        b = _Builder()

        def rec(d):
            if d == 0 or (not full and d < depth and rng.random() < 0.3):
                if rng.random() < 0.5:
                    b.code.append((Op.LoadVarF, 0))
                else:
                    b.const_f(rng.uniform(-5.0, 5.0))
                return
            op = ops[rng.integers(len(ops))]
            for _ in range(2 if op in binary else 1):
                rec(d - 1)
            b.code.append((op, 0))

        rec(depth)
        progs.append(b.prog())
    return progs


def _gen_typed(rng, target_len, ops_by_type, leaves, b, rtype=VT.Float):
    """Random typed tree of about target_len nodes in postfix order."""

    def rec(t, budget):
        cands = [op for op in ops_by_type.get(t, []) if len(SIG[op][0]) < budget]
        if budget <= 1 or not cands:
            leaves[t](b)
            return
        op = cands[rng.integers(len(cands))]
        args = SIG[op][0]
        rest = budget - 1
        # split the remaining budget among the children (each >= 1)
        if len(args) == 1:
            parts = [rest]
        else:
            cuts = np.sort(rng.choice(np.arange(1, rest), size=len(args) - 1, replace=False)) if rest > len(args) \
                else np.arange(1, len(args))
            parts = np.diff(np.concatenate([[0], cuts, [rest]])).tolist()
        for a, p in zip(args, parts):
            rec(a, max(1, int(p)))
        b.code.append((op, 0))

    rec(rtype, target_len)


def gen_c1(n: int = 8192, seed: int = 2001, min_len: int = 5, max_len: int = 128):
    """Uneven lengths 5-128 over {+,-,*,dot,cross,norm,scale,sin,sqrt}; Vec3 leaves E, v, B; constants U(-2, 2)."""
    rng = np.random.default_rng(seed)
    ops = {VT.Float: [Op.AddF, Op.SubF, Op.MulF, Op.SinF, Op.SqrtF, Op.DotV3, Op.NormV3],
           VT.Vec3: [Op.AddV3, Op.SubV3, Op.CrossV3, Op.ScaleV3]}
    leaves = {VT.Float: lambda b: b.const_f(rng.uniform(-2.0, 2.0)),
              VT.Vec3: lambda b: b.code.append((Op.LoadVarV3, 3 * int(rng.integers(3))))}
    progs = []
    while len(progs) < n:
        b = _Builder()
        _gen_typed(rng, int(rng.integers(min_len, max_len + 1)), ops, leaves, b)
        if len(b.code) >= min_len:
            progs.append(b.prog())
    return progs


def gen_c2(n: int = 16384, seed: int = 2002, min_len: int = 5, max_len: int = 48, min_c: int = 1, max_c: int = 8):
    """Programs over M (Mat3), v, u (Vec3) with 1-8 fittable f32 constants ~ U(-3, 3)."""
    rng = np.random.default_rng(seed)
    ops = {VT.Float: [Op.AddF, Op.SubF, Op.MulF, Op.NormV3, Op.DotV3],
           VT.Vec3: [Op.MulM3V3, Op.AddV3, Op.SubV3, Op.ScaleV3],
           VT.Mat3: [Op.ScaleM3]}
    leaves = {VT.Float: lambda b: b.const_f(rng.uniform(-3.0, 3.0)),
              VT.Vec3: lambda b: b.code.append((Op.LoadVarV3, 9 + 3 * int(rng.integers(2)))),
              VT.Mat3: lambda b: b.code.append((Op.LoadVarM3, 0))}
    progs = []
    while len(progs) < n:
        b = _Builder()
        _gen_typed(rng, int(rng.integers(min_len, max_len + 1)), ops, leaves, b)
        if min_c <= len(b.consts) <= max_c:
            progs.append(b.prog())
    return progs


def gen_c3(n: int = 4096, seed: int = 2003, min_len: int = 5, max_len: int = 40, min_c: int = 2, max_c: int = 6):
    """Programs over A, B (Mat2) using {det, inv, tr, matmul, +, -, *} with 2-6 f32 constants."""
    rng = np.random.default_rng(seed)
    ops = {VT.Float: [Op.AddF, Op.SubF, Op.MulF, Op.DetM2, Op.TraceM2],
           VT.Mat2: [Op.InverseM2, Op.MulM2, Op.AddM2, Op.SubM2, Op.ScaleM2]}
    leaves = {VT.Float: lambda b: b.const_f(rng.uniform(-3.0, 3.0)),
              VT.Mat2: lambda b: b.code.append((Op.LoadVarM2, 4 * int(rng.integers(2))))}
    progs = []
    while len(progs) < n:
        b = _Builder()
        _gen_typed(rng, int(rng.integers(min_len, max_len + 1)), ops, leaves, b)
        if min_c <= len(b.consts) <= max_c:
            progs.append(b.prog())
    return progs
