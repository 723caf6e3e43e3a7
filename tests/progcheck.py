"""Structural checks of postfix programs in the wire format, from the operator signatures
of rsr_amd/isa.py (basic/mod.rs:135-154, linalg/mod.rs:340-432): typed-stack discipline,
tree depth, subtree bounds (individual.rs:91-106) and the forbidden parent/child pairs
(basic/mod.rs:165-195, linalg/mod.rs:479-491)."""
from rsr_amd.isa import Op, VT, SIG, Program

LOAD_TYPES = [VT.Float, VT.Float, VT.Bool, VT.Bool, VT.Int, VT.Int, VT.Vec2, VT.Vec2, VT.Vec3, VT.Vec3,
              VT.Mat2, VT.Mat2, VT.Mat3, VT.Mat3]

FORBIDDEN = {
    Op.SinF: {Op.SinF, Op.CosF, Op.ExpF, Op.LnF}, Op.CosF: {Op.SinF, Op.CosF, Op.ExpF, Op.LnF},
    Op.ExpF: {Op.ExpF, Op.SinF, Op.CosF, Op.SqrF, Op.LnF},
    Op.SqrtF: {Op.SqrtF, Op.SqrF, Op.SinF, Op.CosF, Op.LnF, Op.ExpF},
    Op.SqrF: {Op.SqrF, Op.SqrtF}, Op.LnF: {Op.LnF, Op.ExpF, Op.SinF, Op.CosF},
    Op.TransposeM2: {Op.TransposeM2}, Op.TransposeM3: {Op.TransposeM3},
    Op.InverseM2: {Op.InverseM2}, Op.InverseM3: {Op.InverseM3},
    Op.GetXV2: {Op.MakeVec2}, Op.GetYV2: {Op.MakeVec2},
    Op.GetXV3: {Op.MakeVec3}, Op.GetYV3: {Op.MakeVec3}, Op.GetZV3: {Op.MakeVec3},
}


def is_load(op):
    return op < Op.AddF


def is_const(op):
    return op < Op.AddF and op % 2 == 1


def arity(op):
    return 0 if is_load(op) else len(SIG[Op(op)][0])


def node_type(op):
    return LOAD_TYPES[op] if is_load(op) else SIG[Op(op)][1]


def typecheck(p: Program, root=VT.Float):
    """The stack discipline holds, constants are numbered in postfix order with matching
    types, and one value of type `root` remains."""
    st, k = [], 0
    for op, arg in p.code:
        if is_load(op):
            if is_const(op):
                if arg != k or k >= len(p.constants) or p.constants[k].vt != LOAD_TYPES[op]:
                    return False
                k += 1
            st.append(LOAD_TYPES[op])
            continue
        args, ret = SIG[Op(op)]
        if len(st) < len(args) or tuple(st[len(st) - len(args):]) != tuple(args):
            return False
        del st[len(st) - len(args):]
        st.append(ret)
    return st == [root] and k == len(p.constants)


def subtree_bounds(p: Program, root: int):
    needed, cur = 1, root
    while True:
        needed += arity(p.code[cur][0]) - 1
        if needed == 0:
            return cur, root
        if cur == 0:
            return 0, root
        cur -= 1


def depth(p: Program) -> int:
    """Depth of the tree (a single leaf: 0)."""
    st = []
    for op, _ in p.code:
        n = arity(op)
        d = 0 if n == 0 else 1 + max(st[len(st) - n:])
        if n:
            del st[len(st) - n:]
        st.append(d)
    return st[-1]


def children(p: Program, i: int):
    """Root indices of node i's children, first to last."""
    out, j = [], i - 1
    for _ in range(arity(p.code[i][0])):
        s, _e = subtree_bounds(p, j)
        out.append(j)
        j = s - 1
    return out[::-1]


def const_index(p: Program, i: int) -> int:
    """Rank of the constant of node i among the program's constants."""
    return sum(1 for op, _ in p.code[:i] if is_const(op))


def splice(a: Program, sa: int, ea: int, b: Program, sb: int, eb: int) -> Program:
    """a[:sa] + b[sb..=eb] + a[ea+1:] with constants renumbered in postfix order."""
    code, consts = [], []

    def copy(p, lo, hi):
        k = const_index(p, lo)
        for op, arg in p.code[lo:hi]:
            if is_const(op):
                code.append((op, len(consts)))
                consts.append(p.constants[k])
                k += 1
            else:
                code.append((op, arg))

    copy(a, 0, sa)
    copy(b, sb, eb + 1)
    copy(a, ea + 1, len(a.code))
    return Program(code, consts)


def same(p: Program, q: Program) -> bool:
    return list(map(tuple, p.code)) == list(map(tuple, q.code)) and \
        [(c.vt, tuple(c.vals)) for c in p.constants] == [(c.vt, tuple(c.vals)) for c in q.constants]
