"""Serving: requests admitted while others run, results the same alone and in the mix.

Bar: every request served in a mix (evaluation, gradient, L-BFGS in both modes, CMA-ES,
different datasets, feature types and sizes, arriving while others are on the GPU) gives
bit-identical outputs and fitted constants to the same request run alone through the
direct entry points.
"""
import time

import numpy as np
import pytest

from rsr_amd import api
from rsr_amd import workloads as W

pytestmark = pytest.mark.gpu


def valid_programs(ctx, progs, ds):
    P = api.Programs(ctx, progs)
    st = P.status(ds)
    P.close()
    return [p for p, s in zip(progs, st) if s == 0]


def setups(ctx):
    """(kind, dataset, programs, kwargs) of six requests of different shapes."""
    out = []
    X, y, t = W.data_c1(3001, seed=61)
    ds = api.Dataset.new(ctx, X, y, t)
    out.append(("eval", ds, valid_programs(ctx, W.gen_c1(300, seed=62), ds), {}))
    X, y, t = W.data_c2(1503, seed=63)
    ds = api.Dataset.new(ctx, X, y, t)
    out.append(("grad", ds, valid_programs(ctx, W.gen_c2(200, seed=64), ds), {}))
    out.append(("lbfgs", ds, valid_programs(ctx, W.gen_c2(120, seed=65), ds),
                {"iterations": 10, "history": 5, "mode": api.LBFGS_FP64}))
    X, y, t = W.data_c3(999, seed=66)
    ds = api.Dataset.new(ctx, X, y, t)
    out.append(("cmaes", ds, valid_programs(ctx, W.gen_c3(64, seed=67), ds),
                {"iterations": 8, "lam": 16, "mu": 8, "seed": 7}))
    X, y, t = W.data_c0(400, seed=68)
    ds = api.Dataset.new(ctx, X, y, t)
    out.append(("eval", ds, valid_programs(ctx, W.gen_c0(1000, seed=69), ds), {}))
    out.append(("lbfgs", ds, valid_programs(ctx, W.gen_c0(200, seed=70), ds),
                {"iterations": 15, "history": 6, "mode": api.LBFGS_REFERENCE}))
    return out


def run_alone(ctx, kind, ds, progs, kw):
    P = api.Programs(ctx, progs)
    if kind == "eval":
        res = (api.compute_mse(ctx, P, ds),)
    elif kind == "grad":
        res = api.compute_mse_with_gradient(ctx, P, ds)
    elif kind == "lbfgs":
        res = (api.fit_lbfgs(ctx, P, ds, **kw),)
    else:
        res = api.fit_cmaes(ctx, P, ds, **kw)
    return [np.array(r) for r in res], P.get_constants()


def submit(srv, kind, ds, P, kw):
    if kind == "eval":
        return srv.submit_eval(P, ds)
    if kind == "grad":
        return srv.submit_grad(P, ds)
    if kind == "lbfgs":
        return srv.submit_lbfgs(P, ds, **kw)
    return srv.submit_cmaes(P, ds, **kw)


def same(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                                 b.view(np.uint32) if b.dtype == np.float32 else b)


@pytest.mark.parametrize("reference", [False, True])
def test_mix_equals_alone(gpu_ctx, reference):
    gpu_ctx.set_summation(reference)
    try:
        reqs = setups(gpu_ctx)
        alone = [run_alone(gpu_ctx, *r) for r in reqs]
        srv = api.Server(gpu_ctx, max_inflight=8)
        try:
            Ps = [api.Programs(gpu_ctx, r[2]) for r in reqs]
            tickets = [submit(srv, r[0], r[1], P, r[3]) for r, P in zip(reqs[:3], Ps[:3])]
            time.sleep(0.002)  # the rest arrive while the first three are on the GPU
            tickets += [submit(srv, r[0], r[1], P, r[3]) for r, P in zip(reqs[3:], Ps[3:])]
            results = [tk.wait(60.0) for tk in tickets]
            st = srv.stats()
        finally:
            srv.close()  # joins the loop thread even when a request failed
        for k, (got, (res, consts)) in enumerate(zip(results, alone)):
            got = got if isinstance(got, tuple) else (got,)
            assert len(got) == len(res)
            for g, r in zip(got, res):
                assert same(g, r), (k, reqs[k][0])
            assert same(Ps[k].get_constants(), consts), (k, reqs[k][0])
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["max_inflight_seen"] >= 2  # requests overlapped on the device
    finally:
        gpu_ctx.set_summation(False)


def test_admission_slots_and_order(gpu_ctx):
    X, y, t = W.data_c0(400, seed=71)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    progs = valid_programs(gpu_ctx, W.gen_c0(300, seed=72), ds)
    srv = api.Server(gpu_ctx, max_inflight=2)
    try:
        P = api.Programs(gpu_ctx, progs)
        before = api.compute_mse(gpu_ctx, P, ds)
        # same program batch: eval, fit, eval run in order; the last sees the fitted constants
        t1 = srv.submit_eval(P, ds)
        t2 = srv.submit_lbfgs(P, ds, iterations=5)
        t3 = srv.submit_eval(P, ds)
        others = [srv.submit_eval(api.Programs(gpu_ctx, progs), ds) for _ in range(5)]
        r1 = t1.wait(60.0)
        fit = t2.wait(60.0)
        after = t3.wait(60.0)
        rest = [o.wait(60.0) for o in others]
        st = srv.stats()
        q, r, ticks = t3.timing()
    finally:
        srv.close()  # joins the loop thread even when a request failed
    assert same(r1, before)
    assert same(after, api.compute_mse(gpu_ctx, P, ds))
    assert np.isnan(fit).sum() < len(fit)  # some programs had constants to fit
    for o in rest:
        assert same(o, before)
    assert st["max_inflight_seen"] <= 2
    assert st["completed"] == 8
    assert q >= 0.0 and r >= 0.0 and ticks >= 1


def test_budget_rejects_oversized_request(gpu_ctx):
    X, y, t = W.data_c2(800, seed=73)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    progs = valid_programs(gpu_ctx, W.gen_c2(2048, seed=74), ds)
    srv = api.Server(gpu_ctx, max_inflight=4, budget_bytes=64 * 1024)
    try:
        tk = srv.submit_lbfgs(api.Programs(gpu_ctx, progs), ds, iterations=3)
        with pytest.raises(api.RsrError) as e:
            tk.wait(60.0)
        assert e.value.code == -4  # RSR_E_BUDGET
        small = srv.submit_eval(api.Programs(gpu_ctx, progs[:16]), ds)
        assert small.wait(60.0).shape == (16,)
    finally:
        srv.close()


def test_context_is_not_destroyed_under_a_server():
    """A context with a live server refuses destruction (the server's loop thread still
    uses its stream and buffers); after the server is destroyed it goes normally."""
    ctx = api.Context(0)
    srv = api.Server(ctx, max_inflight=2)
    try:
        with pytest.raises(api.RsrError) as e:
            ctx.close()
        assert e.value.code == -1  # RSR_E_INVALID_ARG
        assert ctx.h is not None
    finally:
        srv.close()
    ctx.close()
    assert ctx.h is None


def test_context_is_not_destroyed_under_its_datasets_and_programs():
    """Datasets and program batches keep their context: destroying it first is refused,
    and works once they are gone."""
    ctx = api.Context(0)
    X, y, t = W.data_c0(100, seed=3)
    ds = api.Dataset.new(ctx, X, y, t)
    P = api.Programs(ctx, W.gen_c0(4, seed=4))
    with pytest.raises(api.RsrError):
        ctx.close()
    ds.close()
    with pytest.raises(api.RsrError):
        ctx.close()
    assert ctx.h is not None
    P.close()
    ctx.close()
    assert ctx.h is None


def test_served_work_queue_launches_equal_direct(gpu_ctx):
    """With the work-queue launch form forced on the context, the server's lanes (which copy
    its settings) evaluate through their own queues: requests that overlap on the device,
    each with more work items than the device holds at once, give the bits of a direct
    one-shot evaluation (reference order)."""
    gpu_ctx.set_summation(True)
    made = []
    try:
        want = []
        for k, (n, p) in enumerate([(20_011, 2000), (30_007, 1500), (12_289, 3000)]):
            X, y, t = W.data_c1(n, seed=91 + k)
            ds = api.Dataset.new(gpu_ctx, X, y, t)
            P = api.Programs(gpu_ctx, W.gen_c1(p, seed=95 + k))
            made.append((ds, P))
            gpu_ctx.set_launch_mode(api.LAUNCH_ONESHOT)
            want.append(api.compute_mse(gpu_ctx, P, ds, allow_rejected=True))
        gpu_ctx.set_launch_mode(api.LAUNCH_PERSISTENT)
        srv = api.Server(gpu_ctx, max_inflight=8)
        try:
            print("served work-queue evaluations: 3 requests", flush=True)
            tickets = [srv.submit_eval(P, ds) for ds, P in made]
            got = [np.asarray(tk.wait(60.0)) for tk in tickets]
        finally:
            srv.close()
        for k, (g, w) in enumerate(zip(got, want)):
            assert np.array_equal(g.view(np.uint32), w.view(np.uint32)), k
    finally:
        gpu_ctx.set_launch_mode(api.LAUNCH_AUTO)
        gpu_ctx.set_summation(False)
        for ds, P in made:
            P.close()
            ds.close()
