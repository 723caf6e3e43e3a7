"""Selection parity on device fitnesses (operators/selection.rs:13-134, island.rs:241-255).

Populations of the scalar and the Lorentz workloads are scored on the GPU with
Kahan-compensated summation (the mode whose MSEs differ from the CPU evaluator's in their
last bits; reference order is bit-identical) and on the oracle in the reference's f32 order; both
go through fitness = mse + complexity * parsimony, NSGA-II ranks and crowding distances on
islands of 25 with random ages, and the same tournament draws (k = 2, config.rs:46).
Wherever the fitness gaps inside an island exceed 1e-5 relative (so both sides see the same
fitness order) and the tournament is not decided by a crowding near-tie, the two sides pick
the same individual."""
import numpy as np
import pytest

import oracle as O
from rsr_amd import api
from rsr_amd import workloads as W

pytestmark = pytest.mark.gpu
f32 = np.float32
FMAX = np.finfo(f32).max
TOL = 1e-5


def _separated(fit):
    f = np.sort(np.asarray(fit, np.float64))
    f = f[f < FMAX]
    if len(f) < 2:
        return True
    return bool(np.all(np.diff(f) > TOL * np.maximum(np.abs(f[1:]), 1e-30)))


@pytest.mark.parametrize("which", ["scalar", "lorentz"])
def test_tournament_picks_match_on_gpu_fitnesses(gpu_ctx, which):
    if which == "scalar":
        X, y, t = W.data_c0(400, seed=1000)
        progs = W.gen_c0(1000, seed=2000)
    else:
        X, y, t = W.data_c1(20000, seed=1001)
        progs = W.gen_c1(1000, seed=2001)
    print("selection parity:", which, flush=True)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    ods = O.ODataset.new(X, y, X.shape[1])
    P = api.Programs(gpu_ctx, progs)
    mse_gpu = api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)
    P.close()
    ds.close()
    mse_ref = np.array([O.compute_mse(p, ods) for p in progs], f32)
    cx = api.complexity(progs)
    fit_gpu = api.fitness(mse_gpu, cx)
    fit_ref = np.array([O.fitness(float(m), int(c), 5e-6) for m, c in zip(mse_ref, cx)], f32)
    rng = np.random.default_rng(3)
    clear = total = 0
    islands = 0
    for g in range(len(progs) // 25):
        idx = np.arange(25 * g, 25 * g + 25)
        if not (_separated(fit_ref[idx]) and _separated(fit_gpu[idx])):
            continue
        islands += 1
        age = rng.integers(0, 8, 25).astype(np.uint64)
        rg, cg = api.assign_rank_and_crowding(fit_gpu[idx], age)
        rr, cr = O.assign_rank_and_crowding(fit_ref[idx], age)
        assert np.array_equal(rg, rr), g  # same fitness order -> same fronts
        for _ in range(40):
            d = rng.integers(0, 25, 2).astype(np.uint32)
            total += 1
            a, b = int(d[0]), int(d[1])
            if rr[a] == rr[b] and a != b:
                # decided by crowding: skip near-ties of the two crowding distances
                ca, cb = float(cr[a]), float(cr[b])
                if abs(ca - cb) <= 1e-4 * max(abs(ca), abs(cb), 1e-30):
                    continue
            clear += 1
            assert api.tournament(d, rg, cg) == O.tournament(list(d), rr, cr), (g, d)
    assert islands >= 10 and clear >= 0.8 * total, (islands, clear, total)
