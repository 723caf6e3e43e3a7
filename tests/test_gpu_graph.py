"""Graph mode: a captured evaluation replays to exactly the eager results.

Every replay is checked against an eager evaluation run in between (its own launches and
synchronisations), and the graph picks up constants changed after capture.
"""
import numpy as np
import pytest

from rsr_amd import api
from rsr_amd import workloads as W

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("reference", [False, True])
@pytest.mark.parametrize("gen,data,npts", [(W.gen_c1, W.data_c1, 3001), (W.gen_c2, W.data_c2, 1207)])
def test_graph_replays_match_eager(gpu_ctx, reference, gen, data, npts):
    X, y, t = data(npts, seed=91)
    progs = gen(300, seed=92)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    gpu_ctx.set_summation(reference)
    try:
        eager = api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)
        for rep in range(3):
            gpu_ctx.set_graphs(True)
            try:
                g = api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)
            finally:
                gpu_ctx.set_graphs(False)
            again = api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)  # eager work between replays
            assert np.array_equal(_bits(g), _bits(eager)), rep
            assert np.array_equal(_bits(again), _bits(eager)), rep
        # new constants after capture: the replay reads them
        vals = P.get_constants()
        P.set_constants((vals * np.float32(1.25)).astype(np.float32))
        eager2 = api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)
        gpu_ctx.set_graphs(True)
        try:
            g2 = api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)
        finally:
            gpu_ctx.set_graphs(False)
        assert np.array_equal(_bits(g2), _bits(eager2))
    finally:
        gpu_ctx.set_summation(False)


def test_graph_cache_bound_and_destroyed_datasets(gpu_ctx):
    """A batch keeps at most eight graphs (the least recently replayed dropped first) and
    drops the graphs of destroyed datasets: replays over eleven datasets, some destroyed
    between calls and new ones made in their place, all equal the eager results."""
    progs = W.gen_c0(200, seed=95)
    P = api.Programs(gpu_ctx, progs)
    sets = {}
    for k in range(11):
        X, y, t = W.data_c0(300 + 37 * k, seed=100 + k)
        sets[k] = api.Dataset.new(gpu_ctx, X, y, t)
    try:
        for rnd in range(3):
            for k in sorted(sets):
                eager = api.compute_mse(gpu_ctx, P, sets[k], allow_rejected=True)
                gpu_ctx.set_graphs(True)
                try:
                    g = api.compute_mse(gpu_ctx, P, sets[k], allow_rejected=True)
                finally:
                    gpu_ctx.set_graphs(False)
                assert np.array_equal(_bits(g), _bits(eager)), (rnd, k)
            for k in list(sets)[:3]:  # destroyed; new datasets take their place
                sets.pop(k).close()
                X, y, t = W.data_c0(500 + k + 50 * rnd, seed=200 + k + 10 * rnd)
                sets[100 * (rnd + 1) + k] = api.Dataset.new(gpu_ctx, X, y, t)
    finally:
        for ds in sets.values():
            ds.close()
        P.close()
