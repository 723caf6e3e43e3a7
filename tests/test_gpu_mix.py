"""A reduced serving job mix (the config's feature types and targets, a 10^6-point dataset
among 400 to 50,000-point ones, 1,000 to 2,000 programs each) through the server with
per-generation evaluation and constant-fit (optimize_constants) requests submitted as the
jobs arrive.  The context's partials limit is small, so the 10^6-point evaluations split
into tile ranges inside the mix.  Every served result equals the same request run alone
through the direct entry points, bit for bit, and the evaluations equal the oracle's f32
evaluator on sampled programs (reference-order summation)."""
import threading

import numpy as np
import pytest

import oracle as O
from rsr_amd import api
from rsr_amd import mix
from rsr_amd import workloads as W

pytestmark = pytest.mark.gpu
f32 = np.float32


def _bits(a):
    return np.asarray(a, f32).view(np.uint32)


def test_reduced_mix_served_equals_direct_and_oracle(gpu_ctx):
    rng = np.random.default_rng(91)
    spec = [(1, 1_000_000, 1000), (0, 400, 1500), (2, 20_000, 1200), (3, 50_000, 1000), (1, 3_001, 2000),
            (0, 9_000, 1000), (3, 700, 1800), (2, 1_003, 1000)]
    pools = [g(4096, seed=4100 + k) for k, g in enumerate(mix.GEN)]
    jobs = []
    for j, (kind, n, pop) in enumerate(spec):
        X, y, t = mix.DATA[kind](n, seed=5000 + j)
        ds = api.Dataset.new(gpu_ctx, X, y, t)
        progs = [pools[kind][i] for i in rng.integers(0, 4096, pop)]
        P = api.Programs(gpu_ctx, progs)
        ok = P.status(ds) == 0
        P.close()
        progs = [p for p, g in zip(progs, ok) if g]
        jobs.append((kind, X, y, ds, progs))
    gpu_ctx.set_summation(True)
    gpu_ctx.set_partials_limit(256 << 20)  # the 10^6-point steps split into tile ranges
    served = [None] * len(jobs)
    errors = []
    srv = api.Server(gpu_ctx, max_inflight=16, budget_bytes=mix.HBM_BUDGET)
    try:
        def client(j):
            try:
                kind, X, y, ds, progs = jobs[j]
                P = api.Programs(gpu_ctx, progs)
                P.status(ds)
                mse = np.array(srv.submit_eval(P, ds).wait(300.0))
                m = mix._fit_mask(mse, 0.01, 8)
                after = np.array(srv.submit_optimize(P, ds, iterations=100, seed=7 + j, mask=m).wait(300.0))
                served[j] = (mse, m, after, P.get_constants())
                P.close()
            except Exception as e:  # reported below
                errors.append(repr(e))

        print("reduced mix: %d jobs through the server" % len(jobs), flush=True)
        threads = [threading.Thread(target=client, args=(j,)) for j in range(len(jobs))]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        st = srv.stats()
    finally:
        srv.close()
    try:
        assert not errors, errors
        assert st["completed"] == 2 * len(jobs) and st["failed"] == 0
        assert st["max_inflight_seen"] >= 2
        print("reduced mix: direct calls and oracle", flush=True)
        for j, (kind, X, y, ds, progs) in enumerate(jobs):
            mse, m, after, consts = served[j]
            P = api.Programs(gpu_ctx, progs)
            d_mse = api.compute_mse(gpu_ctx, P, ds)
            assert np.array_equal(_bits(mse), _bits(d_mse)), j
            d_after, _ = api.optimize_constants(gpu_ctx, P, ds, iterations=100, seed=7 + j, mask=m)
            assert np.array_equal(_bits(after[m == 1]), _bits(d_after[m == 1])), j
            assert np.array_equal(_bits(consts), _bits(P.get_constants())), j
            P.close()
            ods = O.ODataset.new(X, y, X.shape[1])
            for i in np.random.default_rng(j).choice(len(progs), 4 if len(y) > 100_000 else 12, replace=False):
                assert _bits(mse[i]) == _bits(f32(O.compute_mse(progs[i], ods))), (j, i)
    finally:
        gpu_ctx.set_partials_limit(4 << 30)
        gpu_ctx.set_summation(False)
        for job in jobs:
            job[3].close()


def test_small_requests_served_together_equal_direct(gpu_ctx):
    """Many small evaluation requests (every workload type, 400 to 4,000 points) submitted
    at once: the server merges their steps into shared launches, and every request's MSEs
    equal its direct evaluation bit for bit."""
    pools = [g(512, seed=4300 + k) for k, g in enumerate(mix.GEN)]
    gpu_ctx.set_summation(True)
    try:
        r = mix.small_requests(gpu_ctx, pools, n_req=48, pop=200, reps=1)
    finally:
        gpu_ctx.set_summation(False)
    print("small requests:", r, flush=True)
    assert r["served_equals_direct"]
