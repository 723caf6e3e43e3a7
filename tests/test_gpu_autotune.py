"""The library's autotuner (rsr_autotune) and its per-workload tuning mode.

A fresh context is in the per-workload mode (get_tuning 0, 0).  rsr_autotune returns a
valid choice and caches it for the workload shape and summation mode: later evaluations in
the per-workload mode -- direct calls and the no-allocation entry point -- run with that
tuning, which is visible in Kahan summation (per-tile partials depend on the tile): their
bits equal an evaluation at that tuning fixed by hand.  Reference-order results equal the oracle under
any tuning."""
import numpy as np
import pytest

import oracle as O
from rsr_amd import api
from rsr_amd import workloads as W

pytestmark = pytest.mark.gpu
f32 = np.float32


def test_autotune_choice_is_used_by_per_workload_mode():
    ctx = api.Context(0)
    try:
        assert ctx.tuning() == (0, 0)
        X, y, t = W.data_c1(20000, seed=61)
        progs = W.gen_c1(512, seed=62)
        ds = api.Dataset.new(ctx, X, y, t)
        P = api.Programs(ctx, progs)
        print("autotune: 512 programs x 20000 points", flush=True)
        ctx.set_summation(False)  # the choice is cached per summation mode
        ppl, group = ctx.autotune(P, ds)
        assert ppl in (1, 2, 3, 4) and group in (1, 2, 4, 8, 16, 32)
        assert ctx.tuning() == (0, 0)
        assert ctx.query_tuning(P, ds)[:2] == (ppl, group)
        auto = api.compute_mse(ctx, P, ds)
        ctx.set_tuning(ppl, group)
        fixed = api.compute_mse(ctx, P, ds)
        assert np.array_equal(auto.view(np.uint32), fixed.view(np.uint32))
        # a different fixed tuning sums other tiles: it must not be what auto mode used
        other = (1, 4) if ppl != 1 else (4, 8)
        ctx.set_tuning(*other)
        diff = api.compute_mse(ctx, P, ds)
        assert not np.array_equal(diff.view(np.uint32), auto.view(np.uint32))
        ctx.set_tuning(0, 0)
        ctx.set_summation(True)
        ref = api.compute_mse(ctx, P, ds)
        ods = O.ODataset.new(X, y, 9)
        want = np.array([O.compute_mse(p, ods) for p in progs[:64]], f32)
        assert np.array_equal(ref[:64].view(np.uint32), want.view(np.uint32))
        P.close()
        ds.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("reference", [True, False])
def test_autotune_launch_form_automatic_mode(reference):
    """In the automatic launch mode (the default) the autotuner also times the work-queue
    form; query_tuning reports the choice, and the evaluation's bits equal the same tuning
    fixed by hand under either forced form (the forms differ only in how work items are
    dealt out)."""
    ctx = api.Context(0)
    try:
        X, y, t = W.data_c1(30011, seed=63)
        progs = W.gen_c1(2048, seed=64)
        ds = api.Dataset.new(ctx, X, y, t)
        P = api.Programs(ctx, progs)
        ctx.set_summation(reference)
        print("autotune, automatic launch form: 2048 programs x 30011 points", flush=True)
        ppl, group = ctx.autotune(P, ds)
        q = ctx.query_tuning(P, ds)
        assert q[:2] == (ppl, group)
        assert not q[2] or ppl in (2, 4)
        auto = api.compute_mse(ctx, P, ds)
        ctx.set_tuning(ppl, group)
        for mode in (api.LAUNCH_ONESHOT, api.LAUNCH_PERSISTENT):
            ctx.set_launch_mode(mode)
            assert ctx.query_tuning(P, ds) == (ppl, group, mode == api.LAUNCH_PERSISTENT)
            fixed = api.compute_mse(ctx, P, ds)
            assert np.array_equal(auto.view(np.uint32), fixed.view(np.uint32)), mode
        ctx.set_launch_mode(api.LAUNCH_AUTO)
        assert ctx.query_tuning(P, ds) == (ppl, group, False)  # a fixed tuning is one-shot here
        P.close()
        ds.close()
    finally:
        ctx.close()


def test_heuristic_tuning_rule():
    """Without an autotuned entry the tuning follows the heuristic's rule: three points per
    lane and 8 programs per item, one-shot, for large evaluations (at least 200,000 pairs of a
    192-point tile and a program); four points per lane and up to 4 programs per item while
    8,000 work items remain for mid-size ones; two points per lane below 8,000 pairs of a
    256-point tile and a program."""
    ctx = api.Context(0)
    try:
        cases = [(W.data_c1, W.gen_c1, 100_000, 2048, (3, 8)),   # 524 tiles x 2048 programs
                 (W.data_c1, W.gen_c1, 10_000, 2000, (4, 4)),    # 40 tiles: 40 x 500 items
                 (W.data_c0, W.gen_c0, 4_000, 1200, (4, 2)),     # 16 tiles: 16 x 600 items
                 (W.data_c0, W.gen_c0, 400, 1000, (2, 1))]       # 2 tiles of 256 x 1000 < 8,000
        for data, gen, n, p, want in cases:
            X, y, t = data(n, seed=65)
            ds = api.Dataset.new(ctx, X, y, t)
            P = api.Programs(ctx, gen(p, seed=66))
            try:
                assert ctx.query_tuning(P, ds) == (want[0], want[1], False), (n, p)
            finally:
                P.close()
                ds.close()
    finally:
        ctx.close()
