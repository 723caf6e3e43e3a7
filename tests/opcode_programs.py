"""Program populations that reach every operator of the instruction set and every
micro-op the lowering emits (shared by the CPU lowering tests and the GPU parity tests).

Feature layout for these programs: nine flat Float components, reused by every value
type (Float leaves at 0, 1, 8; Vec2 at 2, 7; Vec3 at 4, 6; Mat2 at 1, 5; Mat3 at 0).
"""
import numpy as np

from rsr_amd import workloads as W
from rsr_amd.isa import Op, Program, Scalar, VT, SIG

f32 = np.float32
N_FEATURES = 9
_FEATURE_LEAVES = {VT.Float: (Op.LoadVarF, [0, 1, 8]), VT.Vec2: (Op.LoadVarV2, [2, 7]),
                   VT.Vec3: (Op.LoadVarV3, [4, 6]), VT.Mat2: (Op.LoadVarM2, [1, 5]), VT.Mat3: (Op.LoadVarM3, [0])}
_CONST_OP = {VT.Float: Op.LoadConstF, VT.Vec2: Op.LoadConstV2, VT.Vec3: Op.LoadConstV3, VT.Mat2: Op.LoadConstM2,
             VT.Mat3: Op.LoadConstM3}
_SIZE = {VT.Float: 1, VT.Vec2: 2, VT.Vec3: 3, VT.Mat2: 4, VT.Mat3: 9}


def all_ops_programs(rng, n):
    """Random programs drawing every operator of the instruction set."""
    by_type = {}
    for op, (args, ret) in SIG.items():
        by_type.setdefault(ret, []).append(op)
    progs = []
    for _ in range(n):
        code, consts = [], []

        def rec(t, d):
            if d == 0 or rng.random() < 0.25:
                if rng.random() < 0.3:
                    code.append((_CONST_OP[t], len(consts)))
                    consts.append(Scalar(t, tuple(float(f32(v)) for v in rng.uniform(-2, 2, _SIZE[t]))))
                else:
                    op, idx = _FEATURE_LEAVES[t]
                    code.append((op, int(rng.choice(idx))))
                return
            ops = by_type[t]
            op = ops[rng.integers(len(ops))]
            for a in SIG[op][0]:
                rec(a, d - 1)
            code.append((op, 0))

        rec(VT.Float, int(rng.integers(1, 5)))
        progs.append(Program(code, consts))
    return progs


def targeted_programs():
    """Hand-built programs for operators whose device paths a random draw reaches rarely:
    ln and sqr of point-varying data, Vec2 / Vec3 / Mat2 / Mat3 constants, trace and
    inverse of Mat3, and deep stacks that spill."""
    c = Scalar.f
    v3 = Scalar(VT.Vec3, (0.5, -1.25, 2.0))
    v2 = Scalar(VT.Vec2, (1.5, -0.75))
    m2 = Scalar(VT.Mat2, (1.0, 0.5, -0.25, 2.0))
    m3 = Scalar(VT.Mat3, (1.0, 0.2, -0.3, 0.4, 2.0, 0.1, -0.5, 0.3, 1.5))
    P = Program
    progs = [
        # ln(x^2 + c): LN of point-varying values, SQR of data (value and dual)
        P([(Op.LoadVarF, 0), (Op.SqrF, 0), (Op.LoadConstF, 0), (Op.AddF, 0), (Op.LnF, 0)], [c(0.5)]),
        # ln(c * x): negative arguments give NaN, the MSE FLT_MAX
        P([(Op.LoadConstF, 0), (Op.LoadVarF, 1), (Op.MulF, 0), (Op.LnF, 0)], [c(1.5)]),
        # sqr(c * x - x8)
        P([(Op.LoadConstF, 0), (Op.LoadVarF, 1), (Op.MulF, 0), (Op.LoadVarF, 8), (Op.SubF, 0), (Op.SqrF, 0)],
          [c(0.75)]),
        # |v3 const + x4:6| and dot with a constant
        P([(Op.LoadConstV3, 0), (Op.LoadVarV3, 4), (Op.AddV3, 0), (Op.NormV3, 0)], [v3]),
        P([(Op.LoadVarV3, 6), (Op.LoadConstV3, 0), (Op.CrossV3, 0), (Op.LoadConstV3, 1), (Op.DotV3, 0)],
          [v3, Scalar(VT.Vec3, (-1.0, 0.25, 0.5))]),
        # Vec2 constants: dot, norm, scale
        P([(Op.LoadConstV2, 0), (Op.LoadVarV2, 2), (Op.SubV2, 0), (Op.NormV2, 0)], [v2]),
        P([(Op.LoadConstF, 0), (Op.LoadConstV2, 1), (Op.ScaleV2, 0), (Op.LoadVarV2, 7), (Op.DotV2, 0)], [c(2.0), v2]),
        # Mat2 constant times data, det and trace
        P([(Op.LoadConstM2, 0), (Op.LoadVarM2, 1), (Op.MulM2, 0), (Op.DetM2, 0)], [m2]),
        P([(Op.LoadVarM2, 5), (Op.LoadConstM2, 0), (Op.MulM2, 0), (Op.TraceM2, 0)], [m2]),
        P([(Op.LoadConstM2, 0), (Op.LoadVarV2, 2), (Op.MulM2V2, 0), (Op.GetYV2, 0)], [m2]),
        # Mat3 constant: product with data, trace, inverse, determinant
        P([(Op.LoadConstM3, 0), (Op.LoadVarV3, 4), (Op.MulM3V3, 0), (Op.NormV3, 0)], [m3]),
        P([(Op.LoadConstM3, 0), (Op.LoadVarM3, 0), (Op.MulM3, 0), (Op.TraceM3, 0)], [m3]),
        P([(Op.LoadVarM3, 0), (Op.TraceM3, 0)], []),
        P([(Op.LoadVarM3, 0), (Op.InverseM3, 0), (Op.TraceM3, 0)], []),
        P([(Op.LoadVarM3, 0), (Op.LoadConstM3, 0), (Op.AddM3, 0), (Op.InverseM3, 0), (Op.LoadVarV3, 6),
           (Op.MulM3V3, 0), (Op.GetZV3, 0)], [m3]),
        P([(Op.LoadConstF, 0), (Op.LoadVarM3, 0), (Op.ScaleM3, 0), (Op.InverseM3, 0), (Op.DetM3, 0)], [c(1.25)]),
        P([(Op.LoadVarM2, 1), (Op.InverseM2, 0), (Op.LoadConstM2, 0), (Op.MulM2, 0), (Op.TraceM2, 0)], [m2]),
        # a Float constant reaching dual SQR and LN paths through the parameters
        P([(Op.LoadConstF, 0), (Op.LoadVarF, 0), (Op.AddF, 0), (Op.SqrF, 0), (Op.LoadConstF, 1), (Op.AddF, 0),
           (Op.LnF, 0)], [c(-0.5), c(3.0)]),
    ]
    # deep stacks: 20 live products x_i * c_i summed at the end (spills the register file)
    code, consts = [], []
    for k in range(20):
        code += [(Op.LoadVarF, k % 9), (Op.LoadConstF, len(consts)), (Op.MulF, 0), (Op.SinF, 0)]
        consts.append(c(0.1 * (k + 1)))
    code += [(Op.AddF, 0)] * 19
    progs.append(P(code, consts))
    return progs


def deep_stack_programs():
    """Programs whose live stack values exceed the 32-register file and the 16 LDS spill
    slots: four live Mat3 temporaries with six live Vec3 ones; sixteen live Mat3 products
    combined pairwise; 32 live Mat2 temporaries (the reference's full stack depth)."""
    c = Scalar.f
    P = Program
    progs = []
    code, consts = [], []
    for k in range(4):
        code += [(Op.LoadConstF, len(consts)), (Op.LoadVarM3, 0), (Op.ScaleM3, 0)]
        consts.append(c(0.5 + 0.25 * k))
    for k in range(6):
        code += [(Op.LoadVarV3, 4 if k % 2 else 6), (Op.LoadConstF, len(consts)), (Op.ScaleV3, 0)]
        consts.append(c(-0.3 * (k + 1)))
    code += [(Op.AddV3, 0)] * 5
    code += [(Op.MulM3V3, 0)] * 4 + [(Op.NormV3, 0)]
    progs.append(P(code, consts))
    # sixteen live Mat3 values (c_k * M), added pairwise at the end: 144 live floats
    code, consts = [], []
    for k in range(16):
        code += [(Op.LoadConstF, len(consts)), (Op.LoadVarM3, 0), (Op.ScaleM3, 0)]
        consts.append(c(0.05 * (k + 1) - 0.4))
    code += [(Op.AddM3, 0)] * 15 + [(Op.LoadVarV3, 4), (Op.MulM3V3, 0), (Op.NormV3, 0)]
    progs.append(P(code, consts))
    # two computed Mat3 operands of one MulM3 (18 live components in a single op)
    code = [(Op.LoadConstF, 0), (Op.LoadVarM3, 0), (Op.ScaleM3, 0), (Op.LoadConstF, 1), (Op.LoadVarM3, 0),
            (Op.TransposeM3, 0), (Op.ScaleM3, 0), (Op.MulM3, 0), (Op.LoadConstF, 2), (Op.LoadVarM3, 0),
            (Op.ScaleM3, 0), (Op.SubM3, 0), (Op.DetM3, 0)]
    progs.append(P(code, [c(0.7), c(-1.3), c(0.25)]))
    # 32 live Mat2 temporaries (the reference's stack limit per type), then reduced
    code, consts = [], []
    for k in range(32):
        code += [(Op.LoadConstF, len(consts)), (Op.LoadVarM2, 1 if k % 2 else 5), (Op.ScaleM2, 0)]
        consts.append(c(0.03 * k - 0.5))
    code += [(Op.MulM2, 0)] * 31 + [(Op.TraceM2, 0)]
    progs.append(P(code, consts))
    # 31 live Float temporaries, each sin(x_k * c_k), summed at the end (the stack peaks at
    # 32 with the last product's operands)
    code, consts = [], []
    for k in range(31):
        code += [(Op.LoadVarF, k % 9), (Op.LoadConstF, len(consts)), (Op.MulF, 0), (Op.SinF, 0)]
        consts.append(c(0.1 * (k + 1)))
    code += [(Op.AddF, 0)] * 30
    progs.append(P(code, consts))
    return progs


def mat3_block_programs(limit=24):
    """Programs of gen_c2(16384, seed=2002) whose Mat3 block micro-ops (MULM3V3, MULM3V3R,
    SCALEM3, SCALEM3K) take a base register of 23, the last one a nine-register block can
    start at (components 23..31), in the value or the dual code."""
    from rsr_amd import api
    from rsr_amd.isa import UOPS
    wanted = {UOPS.index(n) for n in ("MULM3V3", "MULM3V3R", "SCALEM3", "SCALEM3K")}
    out = []
    for p in W.gen_c2(16384, seed=2002):
        hit = False
        for dual in (False, True):
            st, words = api.lower_code(p, 15, dual=dual)
            if st != 0:
                break
            for w in words:
                op = int(w[0]) & 255
                if op in wanted and 23 in (int(w[1]) & 31, int(w[2]) & 31, int(w[3]) & 31):
                    hit = True
        if hit:
            out.append(p)
            if len(out) >= limit:
                break
    return out


def feature_rows(rng, n):
    """n rows of the nine flat features: N(0,1), one zero row (sqrt(0), division by zero,
    singular inverses) and one row whose Mat2 at component 1 is singular."""
    X = rng.standard_normal((n, N_FEATURES)).astype(f32)
    X[min(5, n - 1)] = 0.0
    if n > 6:
        X[6, 1:5] = [1, 2, 2, 4]
    return X
