"""The serving job mix: independent regression jobs arriving at random, each running
per-generation evaluation + constant-fit requests against its device-resident dataset
and population (the island loop of src/engine/search/island.rs:215-261 reduced to its
evaluator traffic: score the population, then fit the constants of the promising
individuals).

Jobs: Poisson arrivals (mean `rate` jobs/s); dataset sizes log-uniform in
[400, 1,000,000] points; feature types and targets of the four workloads (scalar x;
E, v, B Vec3; M Mat3 + v, u Vec3; A, B Mat2); populations uniform in [1,000, 16,384]
programs drawn from per-type program pools.  Device-resident data is admitted under a
200 GB HBM budget.  A generation is one MSE request over the whole population followed by
one constant-fit request (Individual::optimize_constants, individual.rs:65-75, with the
reference's default opt_iterations = 100, search/config.rs:50: L-BFGS 15 iterations for
differentiable programs, CMA-ES max(100 / 16, 2) = 6 iterations for programs with vector or
matrix constants, Nelder-Mead 100 iterations otherwise; then the 0.05 snap) over the best
`fit_frac` of the population -- the individuals an island fits: those under its best
fitness plus a 1% random pass (opt_prob = 0.01, config.rs:49).
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import api
from . import workloads as W

DATA = [W.data_c0, W.data_c1, W.data_c2, W.data_c3]
GEN = [W.gen_c0, W.gen_c1, W.gen_c2, W.gen_c3]
TYPE_NAMES = ["scalar", "3xVec3", "Mat3+2xVec3", "2xMat2"]


@dataclass
class MixJob:
    index: int
    kind: int
    arrival: float
    n_points: int
    programs: list
    ds: object = None
    latencies: list = field(default_factory=list)
    node_evals: int = 0
    error: str = ""


def make_pools(size: int, seed: int = 4100, c1_pool=None):
    pools = []
    for t, g in enumerate(GEN):
        if t == 1 and c1_pool is not None:
            pools.append(list(c1_pool))
        else:
            pools.append(g(size, seed=seed + t))
    return pools


def make_jobs(ctx, pools, n_jobs=64, rate=4.0, seed=4000, pop_range=(1000, 16384), n_range=(400, 1_000_000)):
    """Jobs with device-resident datasets and populations (lowered ahead of the run)."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    arrivals -= arrivals[0]
    kinds = rng.integers(0, 4, n_jobs)
    npts = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), n_jobs)).astype(np.int64)
    pops = rng.integers(pop_range[0], pop_range[1] + 1, n_jobs)
    jobs = []
    for j in range(n_jobs):
        k = int(kinds[j])
        X, y, t = DATA[k](int(npts[j]), seed=5000 + j)
        ds = api.Dataset.new(ctx, X, y, t)
        pool = pools[k]
        idx = rng.integers(0, len(pool), int(pops[j]))
        progs = [pool[i] for i in idx]
        P = api.Programs(ctx, progs)
        ok = P.status(ds) == 0  # an island never holds programs the evaluator rejects
        P.close()
        progs = [p for p, good in zip(progs, ok) if good]
        jobs.append(MixJob(j, k, float(arrivals[j]), int(npts[j]), progs, ds))
    return jobs


def _fit_mask(mse: np.ndarray, frac: float, min_fit: int):
    k = max(min_fit, int(len(mse) * frac))
    k = min(k, len(mse))
    order = np.argsort(np.where(np.isfinite(mse), mse, np.inf), kind="stable")
    m = np.zeros(len(mse), np.uint8)
    m[order[:k]] = 1
    return m


def _nodes(progs, n):
    return int(sum(len(p.code) for p in progs)) * int(n)


HBM_BUDGET = 200 * 10**9  # the job mix's device admission budget


def prepare(ctx, jobs):
    """Upload and lower every job's population before a run's clock starts (as an island
    holds its population on the device), at its original constants, so that every runner
    starts from the same state and no run pays another's set-up."""
    for jb in jobs:
        if getattr(jb, "P", None) is None:
            jb.P = api.Programs(ctx, jb.programs)
            jb.P.status(jb.ds)
            jb.consts0 = jb.P.get_constants()
        else:
            jb.P.set_constants(jb.consts0)


def release(jobs):
    for jb in jobs:
        if getattr(jb, "P", None) is not None:
            jb.P.close()
            jb.P = None


def run_served(ctx, jobs, generations=3, fit_frac=0.01, min_fit=8, max_inflight=64, opt_iters=100, arrivals=True):
    """Every job is a client thread submitting to one server as it arrives (arrivals=False:
    all at once, the server at saturation)."""
    prepare(ctx, jobs)
    srv = api.Server(ctx, max_inflight=max_inflight, budget_bytes=HBM_BUDGET)
    for jb in jobs:
        jb.latencies, jb.error = [], ""
    t_start = time.perf_counter()

    def client(jb: MixJob):
        try:
            P = jb.P
            delay = jb.arrival - (time.perf_counter() - t_start) if arrivals else 0.0
            if delay > 0:
                time.sleep(delay)
            for g in range(generations):
                t0 = time.perf_counter()
                mse = np.array(srv.submit_eval(P, jb.ds).wait())
                m = _fit_mask(mse, fit_frac, min_fit)
                srv.submit_optimize(P, jb.ds, iterations=opt_iters, seed=1000 * jb.index + g, mask=m).wait()
                jb.latencies.append(time.perf_counter() - t0)
        except Exception as e:  # reported in the summary
            jb.error = repr(e)

    threads = [threading.Thread(target=client, args=(jb,)) for jb in jobs]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    wall = time.perf_counter() - t_start
    st = srv.stats()
    srv.close()
    return summarize(jobs, wall, st, generations)


def run_sequential(ctx, jobs, generations=3, fit_frac=0.01, min_fit=8, opt_iters=100):
    """The same requests one at a time through the direct entry points (no server): the
    plain evaluator, one population per launch."""
    prepare(ctx, jobs)
    for jb in jobs:
        jb.latencies, jb.error = [], ""
    t_start = time.perf_counter()
    for jb in jobs:
        P = jb.P
        for g in range(generations):
            t0 = time.perf_counter()
            mse = api.compute_mse(ctx, P, jb.ds, allow_rejected=True)
            m = _fit_mask(mse, fit_frac, min_fit)
            api.optimize_constants(ctx, P, jb.ds, iterations=opt_iters, seed=1000 * jb.index + g, mask=m)
            jb.latencies.append(time.perf_counter() - t0)
    wall = time.perf_counter() - t_start
    return summarize(jobs, wall, None, generations)


def run_fifo(ctx, jobs, generations=3, fit_frac=0.01, min_fit=8, opt_iters=100):
    """The plain evaluator serving the same arrivals: every job is a client thread as in
    run_served, but one worker runs the generations' requests one at a time, in arrival
    order, through the direct entry points (one population per launch, no batching); the
    latency counts the wait in that queue, as the served latency counts the server's."""
    import queue as _q
    prepare(ctx, jobs)
    work = _q.Queue()
    for jb in jobs:
        jb.latencies, jb.error = [], ""
    t_start = time.perf_counter()

    def worker():
        while True:
            item = work.get()
            if item is None:
                return
            fn, done = item
            try:
                done.append(fn())
            except Exception as e:  # handed to the waiting client
                done.append(e)

    def call(fn):
        done = []
        work.put((fn, done))
        while not done:
            time.sleep(0.0002)
        if isinstance(done[0], Exception):
            raise done[0]
        return done[0]

    def client(jb: MixJob):
        try:
            P = jb.P
            delay = jb.arrival - (time.perf_counter() - t_start)
            if delay > 0:
                time.sleep(delay)
            for g in range(generations):
                t0 = time.perf_counter()
                mse = call(lambda: api.compute_mse(ctx, P, jb.ds, allow_rejected=True))
                m = _fit_mask(mse, fit_frac, min_fit)
                call(lambda: api.optimize_constants(ctx, P, jb.ds, iterations=opt_iters, seed=1000 * jb.index + g,
                                                    mask=m))
                jb.latencies.append(time.perf_counter() - t0)
        except Exception as e:  # reported in the summary
            jb.error = repr(e)

    wt = threading.Thread(target=worker)
    wt.start()
    threads = [threading.Thread(target=client, args=(jb,)) for jb in jobs]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    work.put(None)
    wt.join()
    wall = time.perf_counter() - t_start
    return summarize(jobs, wall, None, generations)


def small_requests(ctx, pools, n_req=256, seed=4200, n_range=(400, 4000), pop=1000, reps=3):
    """Many small evaluation requests, the traffic of islands and small regression jobs:
    n_req populations of `pop` programs, each over its own dataset of log-uniform size in
    n_range, every workload type.  plain: one request at a time through rsr_eval_mse (one
    population per launch); served: all submitted to one server at once, their steps merged
    into shared launches.  Same work, the library's default summation; returns the best of
    `reps` timings of each and whether every served MSE equals its direct one."""
    rng = np.random.default_rng(seed)
    reqs = []
    for j in range(n_req):
        k = int(rng.integers(0, 4))
        n = int(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        X, y, t = DATA[k](n, seed=6000 + j)
        ds = api.Dataset.new(ctx, X, y, t)
        progs = [pools[k][i] for i in rng.integers(0, len(pools[k]), pop)]
        P = api.Programs(ctx, progs)
        ok = P.status(ds) == 0
        P.close()
        progs = [q for q, good in zip(progs, ok) if good]
        reqs.append((ds, api.Programs(ctx, progs), _nodes(progs, n)))
    nodes = sum(r[2] for r in reqs)
    direct = [api.compute_mse(ctx, P, ds) for ds, P, _ in reqs]  # warm-up: lowerings, tunings
    best_plain = best_served = None
    same = True
    srv = api.Server(ctx, max_inflight=n_req, budget_bytes=HBM_BUDGET)
    try:
        for _ in range(reps):
            ctx.synchronize()
            t0 = time.perf_counter()
            for ds, P, _ in reqs:
                api.compute_mse(ctx, P, ds)
            dt = time.perf_counter() - t0
            best_plain = dt if best_plain is None else min(best_plain, dt)
            t0 = time.perf_counter()
            tickets = [srv.submit_eval(P, ds) for ds, P, _ in reqs]
            outs = [np.array(tk.wait()) for tk in tickets]
            dt = time.perf_counter() - t0
            best_served = dt if best_served is None else min(best_served, dt)
            same = same and all(np.array_equal(o.view(np.uint32), d.view(np.uint32)) for o, d in zip(outs, direct))
    finally:
        srv.close()
        for ds, P, _ in reqs:
            P.close()
            ds.close()
    return {"requests": n_req, "programs_per_request": pop, "points": list(n_range), "node_evals": nodes,
            "plain_s": best_plain, "served_s": best_served,
            "plain_node_evals_per_s": nodes / best_plain, "served_node_evals_per_s": nodes / best_served,
            "served_over_plain": best_plain / best_served, "served_equals_direct": bool(same)}


def summarize(jobs, wall, stats, generations):
    lat = np.array([x for jb in jobs for x in jb.latencies], np.float64)
    eval_nodes = sum(_nodes(jb.programs, jb.n_points) for jb in jobs if jb.latencies) * generations
    out = {
        "jobs": len(jobs),
        "generations": generations,
        "completed_generations": int(len(lat)),
        "errors": [jb.error for jb in jobs if jb.error][:3],
        "wall_s": wall,
        "p50_gen_latency_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
        "p99_gen_latency_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
        "eval_node_evals": int(eval_nodes),
        "eval_node_evals_per_s": eval_nodes / wall if wall > 0 else None,
    }
    if stats is not None:
        out["server"] = {k: stats[k] for k in ("ticks", "max_inflight_seen", "busy_seconds", "node_evals",
                                                "completed", "failed")}
        out["node_evals_per_s"] = stats["node_evals"] / wall if wall > 0 else None
        out["node_evals_per_busy_s"] = stats["node_evals"] / stats["busy_seconds"] if stats["busy_seconds"] else None
    return out
