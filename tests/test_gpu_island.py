"""The islands' evaluation step (rsr_islands_evaluate) against the oracle's sequential
restatement of Island::evaluate_buffer (island.rs:215-261).

Bar (reference-order summation): for every island, the buffer after the step (programs
and constant bits), every fitness and final MSE, which individuals were fitted, the new
best's index and the number of draws consumed equal the oracle's.  The islands' bests
are chosen so that some individuals beat them (elite fits) and opt_prob is raised so
that random fits happen too; the batched call speculates on the fits and must keep
exactly the ones the sequential loop makes.
"""
import numpy as np
import pytest

import oracle as O
from oracle import island as OI
from rsr_amd import api
from rsr_amd import workloads as W

pytestmark = pytest.mark.gpu
f32 = np.float32


def _bits(a):
    return np.asarray(a, f32).view(np.uint32)


@pytest.mark.parametrize("opt_prob", [0.01, 0.3])
def test_islands_evaluate_equals_sequential_loop(gpu_ctx, opt_prob):
    X, y, t = W.data_c0(403, seed=95)
    ods = O.ODataset.new(X, y, 1)
    omini = ods.subset(64)
    pool = W.gen_c0(6 * 25 + 6, seed=96)
    rng = np.random.default_rng(97)
    G = 6
    buffers = [pool[g * 25:(g + 1) * 25] for g in range(G)]
    bests = pool[G * 25:]
    # best fitness: the best program's full-dataset fitness, so later individuals can beat it
    best_fit = np.array([f32(O.compute_mse(b, ods)) + f32(OI.complexity(b)) * f32(5e-6) for b in bests], f32)
    best_fit[0] = np.finfo(f32).max  # an island whose best is still unscored: every finite MSE is elite
    best_mini = np.array([O.compute_mse(b, omini) for b in bests], f32)
    draws = [rng.random(200).astype(f32) for _ in range(G)]
    gpu_ctx.set_summation(True)
    try:
        ds = api.Dataset.new(gpu_ctx, X, y, t)
        mini = ds.subset(64)
        res = api.islands_evaluate(gpu_ctx, ds, mini, buffers, best_fit, best_mini, draws, opt_prob=opt_prob,
                                   opt_iterations=40, parsimony=5e-6)
    finally:
        gpu_ctx.set_summation(False)
    fitted = 0
    for g in range(G):
        progs, fit, mse, opt, bi, used = OI.evaluate_buffer(buffers[g], best_fit[g], bests[g], ods, omini, draws[g],
                                                            opt_prob=opt_prob, opt_iterations=40, parsimony=5e-6)
        assert res["draws_used"][g] == used, g
        assert res["best_index"][g] == bi, g
        assert np.array_equal(res["optimized"][g], opt), g
        assert np.array_equal(_bits(res["mse"][g]), _bits(mse)), (g, res["mse"][g], mse)
        assert np.array_equal(_bits(res["fitness"][g]), _bits(fit)), g
        for a, b in zip(res["buffers"][g], progs):
            assert [(int(o), int(x)) for o, x in a.code] == [(int(o), int(x)) for o, x in b.code], g
            assert [_bits(c.vals).tolist() for c in a.constants] == [_bits(c.vals).tolist() for c in b.constants], g
        fitted += int(opt.sum())
    assert fitted > 0


def test_islands_evaluate_edge_cases(gpu_ctx):
    """An empty island, a one-individual island, and a draw stream that runs out (an
    error naming the island, nothing partial returned)."""
    X, y, t = W.data_c0(200, seed=98)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    mini = ds.subset(64)
    progs = W.gen_c0(3, seed=99)
    sq = W.square_program()  # the data's own formula: MSE 0 on both datasets
    res = api.islands_evaluate(gpu_ctx, ds, mini, [[], [sq]], [1e30, 1e30], [1e30, 1e30],
                               [np.zeros(0, f32), np.full(4, 0.5, f32)])
    assert [len(b) for b in res["buffers"]] == [0, 1]
    assert res["draws_used"].tolist() == [0, 1]  # promising by its mini MSE: one draw (no random fit)
    assert res["best_index"].tolist() == [-1, 0]
    assert res["optimized"][1].tolist() == [True]  # elite against the best's 1e30
    assert res["mse"][1].tolist() == [0.0]
    with pytest.raises(api.RsrError):
        api.islands_evaluate(gpu_ctx, ds, mini, [progs], [1e30], [0.0], [np.full(2, 0.9, f32)])
    mini.close()
    ds.close()
