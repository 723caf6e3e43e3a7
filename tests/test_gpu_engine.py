"""The search loop on the library (rsr_engine_*): Island::step_generation /
fill_next_generation / evaluate_buffer / nuke and Engine::run with migration and the halls
of fame (src/engine/search/{island,engine}.rs), every island's generation batched on the GPU.

The variation draws are parity unpinned (test_gpu_variation.py), so a run is checked for
what the reference's loop guarantees whatever the draws: typed populations of the island
size, each island's best fitness never rising, the global best's reported MSE equal to the
evaluator's, halls of fame keyed by the programs' complexity (individual.rs:108-110) with
the global entry the least MSE over the islands, a Pareto front of strictly falling MSE, the
ring migration (engine.rs:97-114), the nuke (island.rs:187-213), the stop at the target,
the final fit's fitness (engine.rs:78-90), and runs that repeat bit for bit."""
import numpy as np
import pytest

from rsr_amd import api
from rsr_amd import workloads as W
from rsr_amd.isa import Op, VT

import progcheck as PC

pytestmark = pytest.mark.gpu

BASIC = [int(o) for o in (Op.AddF, Op.SubF, Op.MulF, Op.DivF, Op.SinF, Op.CosF, Op.ExpF, Op.SqrF, Op.LnF,
                          Op.SqrtF)]
VARS = [(VT.Float, 0)]


@pytest.fixture(scope="module")
def square(gpu_ctx):
    X, y, t = W.data_c0(400, seed=71)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    yield ds
    ds.close()


def small(**kw):
    cfg = dict(num_islands=4, island_size=12, max_generations=8, migration_interval=2, opt_iterations=30,
               final_opt_iterations=200, seed=42)
    cfg.update(kw)
    return cfg


def snapshot(E):
    G = E.config.num_islands
    pops = [E.programs(api.ENGINE_POPULATION, g) for g in range(G)]
    bests = [E.programs(api.ENGINE_ISLAND_BEST, g) for g in range(G)]
    return pops, bests


def test_generations_keep_the_loop_invariants(gpu_ctx, square):
    E = api.Engine(gpu_ctx, square, VARS, BASIC, **small())
    G, size = E.config.num_islands, E.config.island_size
    prev_best = [np.float32(np.finfo(np.float32).max)] * G
    try:
        for gen in range(6):
            E.step()
            pops, bests = snapshot(E)
            for g in range(G):
                progs, fit, age = pops[g]
                assert len(progs) == size
                assert all(PC.typecheck(p) for p in progs)
                assert np.all(age >= 0)
                bp, bf, _ = bests[g]
                assert len(bp) == 1 and PC.typecheck(bp[0])
                assert bf[0] <= prev_best[g]  # an island's best only improves
                prev_best[g] = bf[0]
            info = E.info()
            assert info.generation == gen + 1
            gp, gf, _ = E.programs(api.ENGINE_GLOBAL_BEST)
            assert gf[0] == min(b[1][0] for b in bests)  # get_global_best
            P = api.Programs(gpu_ctx, gp)
            assert api.compute_mse(gpu_ctx, P, square)[0] == np.float32(info.best_mse)
            P.close()
            # halls of fame: keyed by complexity, the global entry the least over the islands
            hp, hm, hc = E.programs(api.ENGINE_HALL_OF_FAME)
            assert np.array_equal(api.complexity(hp), hc)
            for g in range(G):
                ip, im, ic = E.programs(api.ENGINE_ISLAND_HOF, g)
                assert np.array_equal(api.complexity(ip), ic)
                for c, m in zip(ic, im):
                    assert c in set(hc.tolist()) and hm[list(hc).index(c)] <= m
            pp, pm, pc = E.programs(api.ENGINE_PARETO)
            assert np.all(np.diff(pc) > 0) and np.all(np.diff(pm) < 0)
            assert set(pc.tolist()) <= set(hc.tolist())
            # migration after the step of every generation index divisible by the interval
            if gen > 0 and gen % E.config.migration_interval == 0 and not info.target_reached:
                for g in range(G):
                    last = pops[g][0][-1]
                    src = bests[(g - 1) % G][0][0]
                    assert PC.same(last, src), (gen, g)
    finally:
        E.close()


def test_runs_repeat_bit_for_bit(gpu_ctx, square):
    outs = []
    for _ in range(2):
        E = api.Engine(gpu_ctx, square, VARS, BASIC, **small(max_generations=4))
        for _ in range(4):
            E.step()
        pops, bests = snapshot(E)
        outs.append((pops, bests))
        E.close()
    for g in range(4):
        (p1, f1, a1), (p2, f2, a2) = outs[0][0][g], outs[1][0][g]
        assert all(PC.same(x, y) for x, y in zip(p1, p2))
        assert np.array_equal(f1, f2) and np.array_equal(a1, a2)
        assert PC.same(outs[0][1][g][0][0], outs[1][1][g][0][0])


def test_seed_changes_the_run(gpu_ctx, square):
    E1 = api.Engine(gpu_ctx, square, VARS, BASIC, **small(seed=1))
    E2 = api.Engine(gpu_ctx, square, VARS, BASIC, **small(seed=2))
    p1 = E1.programs(api.ENGINE_POPULATION, 0)[0]
    p2 = E2.programs(api.ENGINE_POPULATION, 0)[0]
    assert not all(PC.same(x, y) for x, y in zip(p1, p2))
    E1.close()
    E2.close()


def test_nuke_keeps_the_best_and_regenerates(gpu_ctx, square):
    """stagnation_threshold 0: every island is nuked after every generation; the population
    is then the best plus island_size - 1 new trees scored on the dataset."""
    E = api.Engine(gpu_ctx, square, VARS, BASIC, **small(stagnation_threshold=0, num_islands=3, migration_interval=100))
    try:
        for gen in range(3):
            E.step()
            assert E.info().nukes == 3 * (gen + 1)
            for g in range(3):
                progs, fit, age = E.programs(api.ENGINE_POPULATION, g)
                bp, bf, ba = E.programs(api.ENGINE_ISLAND_BEST, g)
                assert PC.same(progs[0], bp[0]) and fit[0] == bf[0]
                assert np.all(age[1:] == 0)
                P = api.Programs(gpu_ctx, progs[1:])
                mse = api.compute_mse(gpu_ctx, P, square, allow_rejected=True)
                P.close()
                want = api.fitness(mse, api.complexity(progs[1:]), E.config.parsimony_penalty)
                assert np.array_equal(fit[1:], want)
    finally:
        E.close()


def test_run_stops_at_the_target_and_fits_the_best(gpu_ctx, square):
    E = api.Engine(gpu_ctx, square, VARS, BASIC, **small(target_mse=1e30))
    try:
        assert E.run() == 1  # the first generation's best is under any MSE
        info = E.info()
        assert info.target_reached and info.finished
        bp, bf, _ = E.programs(api.ENGINE_ISLAND_BEST, 0)
        P = api.Programs(gpu_ctx, bp)
        mse = api.compute_mse(gpu_ctx, P, square, allow_rejected=True)
        P.close()
        assert bf[0] == api.fitness(mse, api.complexity(bp), E.config.parsimony_penalty)[0]
        with pytest.raises(api.RsrError):
            E.step()
    finally:
        E.close()


def test_run_improves_the_square_target(gpu_ctx, square):
    """y = 2.5 x^2 - 1.2: after a short run the best MSE is far below the first generation's."""
    E = api.Engine(gpu_ctx, square, VARS, BASIC, **small(max_generations=12, num_islands=6, island_size=20))
    try:
        E.step()
        first = E.info().best_mse
        E.run()
        info = E.info()
        assert info.generation == 12 or info.target_reached
        bp, bf, _ = E.programs(api.ENGINE_ISLAND_BEST, 0)
        P = api.Programs(gpu_ctx, bp)
        last = api.compute_mse(gpu_ctx, P, square)[0]
        P.close()
        assert last <= first and last < 0.1 * np.var(W.data_c0(400, seed=71)[1])
    finally:
        E.close()


def test_engines_served_together_equal_alone(gpu_ctx, square):
    """The search loop as a client of the serving loop: two engines (different seeds) stepping
    at once on two threads, their evaluations and fits batched by one server, give the same
    populations, fitnesses and bests bit for bit as each engine alone with direct calls."""
    import threading

    gpu_ctx.set_summation(True)
    try:
        alone = []
        for seed in (5, 6):
            E = api.Engine(gpu_ctx, square, VARS, BASIC, **small(seed=seed, max_generations=5))
            for _ in range(5):
                E.step()
            alone.append(snapshot(E))
            E.close()
        srv = api.Server(gpu_ctx, max_inflight=16)
        engines = [api.Engine(gpu_ctx, square, VARS, BASIC, **small(seed=seed, max_generations=5)) for seed in (5, 6)]
        errors = []

        def run(E):
            try:
                E.set_server(srv)
                for _ in range(5):
                    E.step()
            except Exception as e:  # reported below
                errors.append(repr(e))

        threads = [threading.Thread(target=run, args=(E,)) for E in engines]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        assert not errors, errors
        assert srv.stats()["completed"] > 0
        for E, (pops, bests) in zip(engines, alone):
            p2, b2 = snapshot(E)
            for g in range(E.config.num_islands):
                assert all(PC.same(x, y) for x, y in zip(pops[g][0], p2[g][0]))
                assert np.array_equal(pops[g][1], p2[g][1]) and np.array_equal(pops[g][2], p2[g][2])
                assert PC.same(bests[g][0][0], b2[g][0][0]) and bests[g][1][0] == b2[g][1][0]
            E.close()
        srv.close()
    finally:
        gpu_ctx.set_summation(False)
