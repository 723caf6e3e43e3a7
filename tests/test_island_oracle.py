"""Pins of the oracle's restatement of Island::evaluate_buffer (island.rs:215-261) and of
its complexity table, with answers worked out by hand.

Data x = y = 1, 2, 3, 4 (padded to 8 with zeros, divisor 4).  Programs and their MSEs:
x -> 0; x + x -> mean(x^2) = 7.5; x - x -> 7.5, simplified to the constant 0 (same MSE,
complexity 1); 100 * x -> 99^2 * 7.5 = 73507.5.  The island's best is x + x (fitness
7.5 + 3 * 5e-6).
"""
import numpy as np

import oracle as O
from oracle import island as OI
from rsr_amd import workloads as W
from rsr_amd.isa import Op, Program, Scalar, VT

f32 = np.float32
x = (Op.LoadVarF, 0)


def test_complexity_table_known_answers():
    progs = [W.square_program(), W.lorentz_program(), W.rigid_body_program(), W.linear_program(),
             W.inverse_trace_program(),
             Program([x, (Op.SinF, 0), x, (Op.ExpF, 0), (Op.DivF, 0), (Op.LoadConstV3, 0), (Op.NormV3, 0),
                      (Op.MulF, 0)], [Scalar(VT.Vec3, (1.0, 2.0, 3.0))])]
    assert [OI.complexity(p) for p in progs] == [7, 11, 16, 11, 18, 18]


def test_evaluate_buffer_known_answer():
    X = np.array([[1.0], [2.0], [3.0], [4.0]], f32)
    y = np.array([1.0, 2.0, 3.0, 4.0], f32)
    ds = O.ODataset.new(X, y, 1)
    A = Program([x], [])
    B = Program([x, x, (Op.AddF, 0)], [])
    C = Program([x, x, (Op.SubF, 0)], [])
    D = Program([(Op.LoadConstF, 0), x, (Op.MulF, 0)], [Scalar.f(100.0)])
    pen = f32(5e-6)
    best_fit = f32(7.5) + f32(3) * pen
    # C: promising (7.5 < 15), draw 0.5 (no random fit), elite (7.5 < best) -> simplified to 0
    # B: promising, draw 0.5, elite (7.5 < C's 7.5 + 5e-6) -> fitted, unchanged, not a new best
    # D: not promising (73507.5 >= 15): rescue draw 0.9 fails, keeps its mini MSE, no fit
    # A: promising, draw 0.005 (< 0.01: random fit; elite anyway) -> MSE 0, the new best
    draws = [0.5, 0.5, 0.9, 0.005, 0.7]
    progs, fit, mse, opt, best_i, used = OI.evaluate_buffer([C, B, D, A], best_fit, B, ds, ds, draws)
    assert used == 4
    assert opt.tolist() == [True, True, False, True]
    assert mse.tolist() == [7.5, 7.5, 73507.5, 0.0]
    assert fit.tolist() == [f32(7.5) + pen, f32(7.5) + f32(3) * pen, f32(73507.5) + f32(3) * pen, pen]
    assert best_i == 3
    assert progs[0].code == [(int(Op.LoadConstF), 0)] and progs[0].constants[0].vals == (0.0,)
    assert progs[1:] == [B, D, A]
