"""CPU oracle of one island's evaluation step (Island::evaluate_buffer, island.rs:215-261).

TEST INFRASTRUCTURE ONLY (the checker of rsr_islands_evaluate; the product never uses
it).

The reference loop as written: each individual in buffer order, the island's best
updated inside the loop, the uniform draws consumed where the reference consumes them.
Individual::optimize_constants (individual.rs:65-75) is the oracle's optimizer dispatch
of optimize/mod.rs:10-59 (L-BFGS 15 iterations M = 6 for differentiable programs,
Nelder-Mead otherwise; CMA-ES programs are not supported here), then the 0.05 snap
(scalar.rs:60-80) and oracle.simplify.  Complexity is Node::weight (node.rs:21-33) with
the operator weights of basic/mod.rs:156-163 and linalg/mod.rs:434-478.  Pinned by
known answers worked out by hand (tests/test_island_oracle.py): the loop's draw
consumption, rescue, random and elite fits and best updates, and the complexity table.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np

import oracle as O
from oracle import simplify as OS
from rsr_amd.isa import Op, Program, Scalar, VT, VT_SIZE

f32 = np.float32
FMAX = np.finfo(f32).max

_W = {}
for _o in (Op.AddF, Op.SubF, Op.MulF, Op.NegF, Op.AddV2, Op.SubV2, Op.AddV3, Op.SubV3, Op.AddM2, Op.SubM2,
           Op.AddM3, Op.SubM3):
    _W[_o] = 1
for _o in (Op.DivF, Op.SqrF, Op.SqrtF, Op.MakeVec2, Op.GetXV2, Op.GetYV2, Op.ScaleV2, Op.DotV2, Op.MakeVec3,
           Op.GetXV3, Op.GetYV3, Op.GetZV3, Op.ScaleV3, Op.DotV3, Op.ScaleM2, Op.ScaleM3):
    _W[_o] = 2
for _o in (Op.SinF, Op.CosF, Op.NormV2, Op.NormV3, Op.MakeMat2, Op.MulM2, Op.MulM2V2, Op.TraceM2, Op.TransposeM2,
           Op.MulM3V3, Op.TraceM3, Op.TransposeM3):
    _W[_o] = 3
for _o in (Op.ExpF, Op.LnF, Op.CrossV3, Op.DetM2, Op.InverseM2, Op.MulM3):
    _W[_o] = 4
_W[Op.MakeMat3] = _W[Op.DetM3] = 5
_W[Op.InverseM3] = 6
_CONST_W = {Op.LoadConstF: 1, Op.LoadConstB: 1, Op.LoadConstI: 1, Op.LoadConstV2: 2, Op.LoadConstV3: 3,
            Op.LoadConstM2: 4, Op.LoadConstM3: 9}


def complexity(p: Program) -> int:
    w = 0
    for op, _ in p.code:
        op = Op(int(op))
        if op <= Op.LoadConstM3:
            w += _CONST_W.get(op, 1)  # a variable weighs 1
        else:
            w += _W[op]
    return w


def _with_params(p: Program, flat) -> Program:
    cs, k = [], 0
    for c in p.constants:
        n = VT_SIZE[c.vt]
        cs.append(Scalar(c.vt, tuple(float(v) for v in flat[k:k + n])) if n else c)
        k += n
    return Program(list(p.code), cs)


def optimize_individual(p: Program, ds: O.ODataset, iterations: int, threshold: float = 0.05) -> Program:
    """Individual::optimize_constants: fit, snap, simplify."""
    n = p.param_count()
    inverse = any(int(op) in (Op.InverseM2, Op.InverseM3) for op, _ in p.code)
    vecmat = any(c.vt in (VT.Vec2, VT.Vec3, VT.Mat2, VT.Mat3) for c in p.constants)
    q = p
    if n > 0:
        if not inverse:
            _, fitted = O.run_lbfgs(p, ds, iters=15, m=6)
        elif vecmat:
            raise NotImplementedError("CMA-ES programs are not covered by this oracle")
        else:
            _, fitted = O.run_nelder_mead(p, ds, iterations)
        q = _with_params(p, np.asarray(fitted, f32)[:n])
    snapped = [Scalar(c.vt, tuple(0.0 if abs(f32(v)) < f32(threshold) else float(v) for v in c.vals))
               if VT_SIZE[c.vt] else c for c in q.constants]
    return OS.simplify(Program(list(q.code), snapped))


def evaluate_buffer(buffer: Sequence[Program], best_fitness: float, best_program: Program, ds: O.ODataset,
                    mini: O.ODataset, draws: Sequence[float], opt_prob: float = 0.01, opt_iterations: int = 100,
                    parsimony: float = 5e-6):
    """Returns (programs, fitness, final_mse, optimized, best_index or -1, draws_used)."""
    d = 0

    def draw():
        nonlocal d
        u = f32(draws[d])
        d += 1
        return u

    best = f32(best_fitness)
    best_i = -1
    baseline = f32(O.compute_mse(best_program, mini))
    progs: List[Program] = []
    fit, mses, opt = [], [], []
    for i, p in enumerate(buffer):
        mini_mse = f32(O.compute_mse(p, mini))
        promising = mini_mse < baseline * f32(2.0)
        if not promising and draw() < f32(0.05):
            promising = True
        final = mini_mse
        fitted = False
        if promising:
            final = f32(O.compute_mse(p, ds))
            elite = final < best
            random_opt = draw() < f32(opt_prob)
            if (elite or random_opt) and np.isfinite(final):
                p = optimize_individual(p, ds, opt_iterations)
                final = f32(O.compute_mse(p, ds))
                fitted = True
        if np.isfinite(final):
            f = final + f32(complexity(p)) * f32(parsimony)
        else:
            f = FMAX
        if f < best:
            best = f
            best_i = i
        progs.append(p)
        fit.append(f)
        mses.append(final)
        opt.append(fitted)
    return progs, np.asarray(fit, f32), np.asarray(mses, f32), np.asarray(opt), best_i, d
