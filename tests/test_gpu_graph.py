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
