"""Host logic: the postfix -> micro-code lowering, checked on the CPU.

rsr_lower_check lowers a program exactly as the GPU path does and runs the emitted
micro-code on the host; its per-point values and tangents must equal the oracle's
bit for bit (same IEEE operations in the same order, same transcendental algorithm).
"""
import numpy as np
import pytest

import oracle as O
from rsr_amd import api
from rsr_amd import workloads as W
from rsr_amd.isa import Op, Program, Scalar, VT, SIG
import opcode_programs as OP

f32 = np.float32


def same_bits(a, b):
    a = np.asarray(a, f32)
    b = np.asarray(b, f32)
    both_nan = np.isnan(a) & np.isnan(b)
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | both_nan | ((a == 0) & (b == 0))))


def check_programs(progs, X, y, dual_params=0):
    F = X.shape[1]
    ds = O.ODataset.new(X, y, F)
    f_soa, _ = ds.soa()
    bad = 0
    for p in progs:
        st, v, _, nu, mr = api.lower_check(p, f_soa)
        if st == -6:  # register budget: reported, not silently wrong
            bad += 1
            continue
        assert st == 0, (st, p.code)
        ref = O.eval_points(p, ds)
        assert same_bits(v, ref), p.code
        for k in range(min(dual_params, p.param_count())):
            st, dv, dt, _, _ = api.lower_check(p, f_soa, dual=True, active_param=k)
            rv, rt = O.eval_points_dual(p, ds, k)
            assert st == 0 and same_bits(dv, rv) and same_bits(dt, rt), (k, p.code)
    return bad


def test_c0_population_bit_exact():
    X, y, _ = W.data_c0(64, seed=1)
    assert check_programs(W.gen_c0(400, seed=2), X, y, dual_params=2) == 0


def test_c1_population_bit_exact():
    X, y, _ = W.data_c1(64, seed=3)
    assert check_programs(W.gen_c1(300, seed=4), X, y, dual_params=1) == 0


def test_c2_population_bit_exact_with_tangents():
    X, y, _ = W.data_c2(64, seed=5)
    assert check_programs(W.gen_c2(150, seed=6), X, y, dual_params=8) == 0


def test_c3_population_bit_exact():
    X, y, _ = W.data_c3(64, seed=7)
    assert check_programs(W.gen_c3(150, seed=8), X, y, dual_params=6) == 0


def test_every_operator_bit_exact_incl_tangents():
    rng = np.random.default_rng(9)
    X = rng.standard_normal((64, 9)).astype(f32)
    X[5] = 0.0  # a zero row: exercises sqrt(0), division by 0, singular inverses
    X[6, 1:5] = [1, 2, 2, 4]  # singular Mat2
    y = rng.standard_normal(64).astype(f32)
    progs = OP.all_ops_programs(rng, 600)
    used = {Op(o) for p in progs for o, _ in p.code}
    assert set(SIG) <= used
    assert check_programs(progs, X, y, dual_params=3) == 0


def test_transcendentals_device_math_equals_oracle_bitwise():
    rng = np.random.default_rng(10)
    xs = np.concatenate([rng.uniform(-50, 50, 3000), rng.uniform(-8192, 8192, 1000), np.exp(rng.uniform(-90, 80, 2000)),
                         [0.0, -0.0, 1e-40, np.inf, -np.inf, np.nan, 1e30, -1e30, 88.8, -104.0]]).astype(f32)
    for op, fn in [(Op.SinF, O.sinf), (Op.CosF, O.cosf), (Op.ExpF, O.expf), (Op.LnF, O.logf)]:
        st, v, _, _, _ = api.lower_check(Program([(Op.LoadVarF, 0), (op, 0)], []), xs.reshape(1, -1))
        ref = np.array([fn(float(x)) for x in xs], f32)
        assert st == 0 and same_bits(v, ref), op


@pytest.mark.parametrize("code,consts,F,status", [
    ([(Op.AddF, 0)], [], 1, -1),                                   # underflow
    ([(Op.LoadVarF, 3)], [], 2, -3),                               # feature out of range
    ([(Op.LoadVarV3, 0)], [], 2, -3),                              # vector beyond F
    ([(Op.LoadConstF, 0)], [Scalar(VT.Vec2, (1.0, 2.0))], 1, -4),  # type mismatch
    ([(Op.LoadConstF, 1)], [Scalar.f(1.0)], 1, -4),                # constant index
    ([(Op.LoadVarB, 0)], [], 1, -2),                               # Bool load (reference no-op)
    ([(Op.LoadVarV3, 0)], [], 3, -5),                              # root is not Float
    ([(Op.LoadVarF, 0), (Op.LoadVarF, 0)], [], 1, -5),             # two values left
    ([(99, 0)], [], 1, -2),                                        # unknown opcode
    ([(Op.LoadVarF, 0)] * 33 + [(Op.AddF, 0)] * 32, [], 1, -1),    # depth 33 > 32
])
def test_lowering_rejects_undefined_behaviour(code, consts, F, status):
    st, _, _, _, _ = api.lower_check(Program(code, consts), np.zeros((F, 1), f32))
    assert st == status


def test_deep_stacks_lower_and_match_oracle():
    """Programs beyond the 32-register file and the 16 LDS spill slots lower (slots 16 and up
    live in device memory) and run bit-exact on the host, tangents included: 31 live Vec3
    constants (93 components), and the deep-stack population (16 live Mat3 values, two
    computed Mat3 operands of one MulM3 at 15 features, 32 live Mat2 values, 31 live Floats)."""
    n = 31
    code = [(Op.LoadConstV3, i) for i in range(n)] + [(Op.AddV3, 0)] * (n - 1) + [(Op.NormV3, 0)]
    consts = [Scalar(VT.Vec3, (1.0 + i, 2.0, -3.0 * i)) for i in range(n)]
    rng = np.random.default_rng(21)
    X = OP.feature_rows(rng, 40)
    y = rng.standard_normal(40).astype(f32)
    progs = [Program(code, consts)] + OP.deep_stack_programs()
    assert check_programs(progs, X, y, dual_params=4) == 0
    for F in (9, 15):
        for p in progs:
            st, words = api.lower_code(p, F)
            assert st == 0, (F, p.code)
    # the reference's full depth: 32 live Mat3 values need more than the LDS slots
    code = []
    cs = []
    for k in range(32):
        code += [(Op.LoadConstF, k), (Op.LoadVarM3, 0), (Op.ScaleM3, 0)]
        cs.append(Scalar.f(0.01 * k - 0.2))
    code += [(Op.AddM3, 0)] * 31 + [(Op.TraceM3, 0)]
    # the Float stack peaks at one value; ScaleM3 pops it -- 32 Mat3 stay on the stack
    p32 = Program(code, cs)
    assert check_programs([p32], X, y, dual_params=2) == 0
    lw = api.lower_check(p32, np.zeros((9, 1), f32))
    assert lw[0] == 0 and lw[4] > 32 + 16  # more than 16 spill slots


def test_register_budget_reported():
    # with 32 flat features there is no temporary register at all: RSR_PROG_REGS
    st, _, _, _, _ = api.lower_check(Program([(Op.LoadVarF, 0), (Op.LoadVarF, 1), (Op.AddF, 0)], []),
                                     np.zeros((32, 1), f32))
    assert st == -6
    # a program that only relabels features still lowers there
    st, _, _, _, _ = api.lower_check(Program([(Op.LoadVarF, 31)], []), np.zeros((32, 1), f32))
    assert st == 0


def test_aliasing_ops_emit_no_code():
    # loads, MakeVec3, GetX, Transpose only relabel registers
    p = Program([(Op.LoadVarF, 0), (Op.LoadVarF, 0), (Op.LoadVarF, 0), (Op.MakeVec3, 0), (Op.GetYV3, 0)], [])
    st, v, _, nu, _ = api.lower_check(p, np.array([[3.0, 4.0]], f32))
    assert st == 0 and nu == 0 and list(v) == [3.0, 4.0]


def _uop_names(prog, F):
    from rsr_amd.isa import UOPS
    st, words = api.lower_code(prog, F)
    assert st == 0
    return [UOPS[w[0] & 255] for w in words], words


def test_uop_name_table_matches_isa_header():
    import os
    import re
    from rsr_amd.isa import UOPS
    text = open(os.path.join(os.path.dirname(__file__), "..", "rsr_amd", "csrc", "isa.h")).read()
    body = text[text.index("enum UOp"):text.index("U_COUNT")]
    names = re.findall(r"\bU_(\w+)\s*(?:=\s*0)?\s*,", body)
    assert names == UOPS


@pytest.mark.parametrize("which", ["lorentz", "mat3"])
def test_forwarding_paths_bit_exact_with_poisoned_file(which):
    """Each program produces the micro-op it is built for, and its host execution (where a
    dropped write-back leaves NaN in the register file) equals the oracle bit for bit."""
    import forwarding_programs as FP
    from rsr_amd.isa import UFLAG_FWD, UFLAG_NOWB
    if which == "lorentz":
        progs, (X, y, t) = FP.lorentz_programs(), W.data_c1(300, seed=61)
    else:
        progs, (X, y, t) = FP.mat3_programs(), W.data_c2(300, seed=62)
    F = X.shape[1]
    fwd = nowb = 0
    for name, want, p in progs:
        names, words = _uop_names(p, F)
        assert want in names, (name, names)
        fwd += int(((words[:, 2] & UFLAG_FWD) != 0).sum())
        nowb += int(((words[:, 1] & UFLAG_NOWB) != 0).sum())
        # dual code is never annotated
        st, dwords = api.lower_code(p, F, dual=True)
        assert st == 0 and not ((dwords[:, 1:] >> 8) & 0xF).any(), name
    assert fwd > 0 and nowb > 0
    assert check_programs([p for _, _, p in progs], X, y, dual_params=1) == 0


def test_forwarded_operand_is_always_a():
    """After the swap pass no micro-op takes the previous result through operand b."""
    progs = W.gen_c1(200, seed=63)
    for p in progs:
        st, words = api.lower_code(p, 9)
        if st != 0:
            continue
        for i in range(1, len(words)):
            d_prev = words[i - 1][1] & 31
            if words[i][2] & (1 << 8):
                assert (words[i][2] & 31) == d_prev


def test_targeted_and_mat3_block_programs_bit_exact_incl_tangents():
    """The hand-built operator programs and the Mat3 block programs whose base register is
    23 (components 23..31) lower and run bit-exact on the host, tangents included."""
    rng = np.random.default_rng(11)
    X = OP.feature_rows(rng, 64)
    y = rng.standard_normal(64).astype(f32)
    assert check_programs(OP.targeted_programs(), X, y, dual_params=3) == 0
    blocks = OP.mat3_block_programs(limit=8)
    assert len(blocks) == 8
    X2, y2, _ = W.data_c2(64, seed=12)
    assert check_programs(blocks, X2, y2, dual_params=8) == 0


def test_every_generated_population_program_lowers():
    """No program of the bench populations is refused any more: the three programs of
    gen_c2(16384, seed=2002) that need more than the LDS slots lower and match the oracle
    on the host, tangents included."""
    allp = W.gen_c2(16384, seed=2002)
    progs = [allp[i] for i in (5633, 5663, 14173)]
    X, y, _ = W.data_c2(48, seed=23)
    assert check_programs(progs, X, y, dual_params=8) == 0
    for p in progs:
        st, words = api.lower_code(p, 15)
        assert st == 0
        assert max(int(w[0]) >> 8 for w in words if (int(w[0]) & 255) in (19, 20)) >= 16
