"""Scoring and selection host functions of the library against the oracle and closed forms.

complexity: known answers computed by hand from the node weights of node.rs:21-33 and the
operator weight tables (basic/mod.rs:156-163, linalg/mod.rs:434-478).  fitness, fronts,
crowding distances and tournaments: bit-identical to the oracle's restatement of
selection.rs / island.rs on random populations with ties, which the oracle's own tests pin
against brute force.
"""
import numpy as np
import pytest

import oracle as O
from rsr_amd import api
from rsr_amd import workloads as W
from rsr_amd.isa import Op, Program, Scalar, VT

FMAX = np.finfo(np.float32).max


def test_complexity_known_answers():
    progs = [W.square_program(),          # 1+1+1+1+1+1+1
             W.lorentz_program(),         # 1+1+1 + cross 4 + add 1 + norm 3
             W.rigid_body_program(),      # 1 + Mat3 const 9 + 1 + MulM3V3 3 + DotV3 2
             W.linear_program(),          # 11 nodes of weight 1
             W.inverse_trace_program(),   # 1 + inv 4 + 1 + mul 3 + tr 3 + 1 + det 4 + add 1
             Program([(Op.LoadVarF, 0), (Op.SinF, 0), (Op.LoadVarF, 0), (Op.ExpF, 0), (Op.DivF, 0),
                      (Op.LoadConstV3, 0), (Op.NormV3, 0), (Op.MulF, 0)],
                     [Scalar(VT.Vec3, (1.0, 2.0, 3.0))])]  # 1+3+1+4+2+3+3+1
    assert api.complexity(progs).tolist() == [7, 11, 16, 11, 18, 18]


def test_complexity_empty_batch():
    assert api.complexity([]).shape == (0,)


def test_fitness_matches_oracle():
    rng = np.random.default_rng(5)
    mse = rng.uniform(0, 10, 200).astype(np.float32)
    mse[::17] = np.inf
    mse[::23] = np.nan
    mse[5] = FMAX
    comp = rng.integers(1, 200, 200).astype(np.int32)
    got = api.fitness(mse, comp, 5e-6)
    ref = np.array([O.fitness(float(m), int(c), 5e-6) for m, c in zip(mse, comp)], np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert (got[::17] == FMAX).all() and (got[::23] == FMAX).all()


@pytest.mark.parametrize("seed", range(8))
def test_rank_and_crowding_match_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 40))
    fit = rng.choice(rng.uniform(0, 5, 8).astype(np.float32), n)  # many ties
    age = rng.integers(0, 6, n).astype(np.uint64)
    r, c = api.assign_rank_and_crowding(fit, age)
    ro, co = O.assign_rank_and_crowding(fit, age)
    assert np.array_equal(r, ro)
    assert np.array_equal(c.view(np.uint32), co.view(np.uint32))


def test_tournament_matches_oracle():
    rng = np.random.default_rng(9)
    for _ in range(200):
        n = int(rng.integers(2, 30))
        fit = rng.uniform(0, 5, n).astype(np.float32)
        age = rng.integers(0, 10, n).astype(np.uint64)
        r, c = api.assign_rank_and_crowding(fit, age)
        k = int(rng.integers(1, 5))
        draws = rng.integers(0, n, k)
        assert api.tournament(draws, r, c) == O.tournament(list(draws), r, c)


def test_tournament_rejects_out_of_range_draw():
    with pytest.raises(api.RsrError):
        api.tournament([0, 5], np.zeros(3, np.uint32), np.zeros(3, np.float32))
