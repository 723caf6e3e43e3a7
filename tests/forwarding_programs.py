"""Programs that exercise each path of the value-code post-pass: forwarded operands,
operand-swapped micro-ops, constant-operand unary ops and point-invariant
transcendentals.  Each entry names the micro-op it must produce.

Features of the Lorentz dataset: E = 0..2, v = 3..5, B = 6..8 (Vec3).  Features of the
Mat3 dataset: M = 0..8 (Mat3), v = 9..11, u = 12..14 (Vec3).
"""
from rsr_amd.isa import Op, Program, Scalar

L = lambda c: (Op.LoadVarV3, c)
K = lambda i: (Op.LoadConstF, i)
N = (Op.NormV3, 0)


def dot(a, b):
    return [L(a), L(b), (Op.DotV3, 0)]


def lorentz_programs():
    """(name, expected micro-op, program) over the Lorentz features."""
    c = [Scalar.f(v) for v in (0.7, -1.3, 0.25, 2.5, 1.75, -0.6, 1.1, 0.4)]
    return [
        ("rsubv3", "RSUBV3", Program([L(3), L(6), L(0), (Op.SubV3, 0), (Op.SubV3, 0), N], [])),
        ("rcross3", "RCROSS3", Program([L(6), L(0), L(3), (Op.CrossV3, 0), (Op.CrossV3, 0), N], [])),
        ("scalev3s", "SCALEV3S", Program([L(0)] + dot(3, 6) + [(Op.ScaleV3, 0), N], [])),
        ("rsub", "RSUB", Program(dot(0, 3) + dot(3, 6) + [(Op.SubF, 0)], [])),
        ("rdiv", "RDIV", Program(dot(0, 3) + dot(3, 6) + [(Op.DivF, 0)], [])),
        ("addv3_fwd", "ADDV3", Program([L(0), L(3), (Op.AddV3, 0), L(6), (Op.AddV3, 0), L(0), (Op.AddV3, 0), N], [])),
        ("sink", "SINK", Program(dot(0, 3) + [K(0), (Op.SinF, 0), (Op.MulF, 0)], c[:1])),
        ("cosk", "COSK", Program(dot(0, 3) + [K(0), (Op.CosF, 0), (Op.AddF, 0)], c[:1])),
        ("expk", "EXPK", Program(dot(0, 3) + [K(0), (Op.ExpF, 0), (Op.SubF, 0)], c[2:3])),
        ("lnk", "LNK", Program(dot(0, 3) + [K(0), (Op.LnF, 0), (Op.MulF, 0)], c[3:4])),
        ("sqrtk", "SQRTK", Program(dot(0, 3) + [K(0), (Op.SqrtF, 0), (Op.DivF, 0)], c[4:5])),
        ("negk", "NEGK", Program(dot(0, 3) + [K(0), (Op.NegF, 0), (Op.MulF, 0)], c[5:6])),
        ("sqrk", "SQRK", Program(dot(0, 3) + [K(0), (Op.SqrF, 0), (Op.AddF, 0)], c[6:7])),
        ("sinu", "SINU", Program(dot(0, 3) + [K(0), K(1), (Op.AddF, 0), (Op.SinF, 0), (Op.MulF, 0)], c[:2])),
        ("sqrtu", "SQRTU", Program(dot(0, 3) + [K(0), K(1), (Op.MulF, 0), (Op.SqrF, 0), (Op.SqrtF, 0), (Op.AddF, 0)],
                                    c[:2])),
        ("expu", "EXPU", Program(dot(0, 3) + [K(0), (Op.NegF, 0), (Op.ExpF, 0), (Op.MulF, 0)], c[2:3])),
        ("lnu", "LNU", Program(dot(0, 3) + [K(0), (Op.SqrF, 0), (Op.LnF, 0), (Op.SubF, 0)], c[3:4])),
        ("cosu", "COSU", Program(dot(0, 3) + [K(0), K(1), (Op.SubF, 0), (Op.CosF, 0), (Op.AddF, 0)], c[:2])),
        # a long mixed chain: forwarded vectors and scalars, write-backs dropped and kept
        ("chain", "CROSS3", Program([L(0), L(3), L(6), (Op.CrossV3, 0), (Op.AddV3, 0), L(3), (Op.SubV3, 0)]
                                    + dot(6, 0) + [(Op.ScaleV3, 0), L(0), (Op.CrossV3, 0), N, (Op.SinF, 0), K(0),
                                                   (Op.MulF, 0)] + dot(3, 3) + [(Op.SqrtF, 0), (Op.AddF, 0)], c[:1])),
    ]


def mat3_programs():
    """(name, expected micro-op, program) over the Mat3 features."""
    return [
        ("mulm3v3r", "MULM3V3R", Program([(Op.LoadVarM3, 0), L(9), L(12), (Op.AddV3, 0), (Op.MulM3V3, 0), N], [])),
        ("mulm3v3", "MULM3V3", Program([(Op.LoadVarM3, 0), L(9), (Op.MulM3V3, 0), L(12), (Op.SubV3, 0), N], [])),
    ]
