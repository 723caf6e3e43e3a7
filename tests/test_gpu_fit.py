"""GPU constant fitting against the oracle's restatements of lbfgs.rs / cmaes.rs.

With reference-order summation the GPU's MSE and gradients are bit-identical to the
oracle's, so L-BFGS (both modes) and CMA-ES (same injected normals) follow the same
trajectories: fitness and fitted constants must match exactly.  With the default
Kahan summation the fitted MSE must agree within 1e-4 relative (the trajectories may
part where an Armijo or sorting decision falls inside the last-bit differences).
"""
import numpy as np
import pytest

import oracle as O
from rsr_amd import api
from rsr_amd import workloads as W
from rsr_amd.isa import Op, Program, Scalar, VT

pytestmark = pytest.mark.gpu
f32 = np.float32
FMAX = np.finfo(f32).max


def _with_consts(p, flat):
    consts, k = [], 0
    for c in p.constants:
        consts.append(Scalar(c.vt, tuple(float(v) for v in flat[k:k + len(c.vals)])))
        k += len(c.vals)
    return Program(p.code, consts)


@pytest.mark.parametrize("fp64", [False, True])
def test_lbfgs_reference_order_bit_exact(gpu_ctx, fp64):
    X, y, t = W.data_c2(400, seed=31)
    progs = W.gen_c2(64, seed=32)
    gpu_ctx.set_summation(True)
    try:
        ds = api.Dataset.new(gpu_ctx, X, y, t)
        ods = O.ODataset.new(X, y, 15)
        P = api.Programs(gpu_ctx, progs)
        fit = api.fit_lbfgs(gpu_ctx, P, ds, iterations=15, history=6 if not fp64 else 5,
                            mode=api.LBFGS_FP64 if fp64 else api.LBFGS_REFERENCE)
        vals = P.get_constants()
    finally:
        gpu_ctx.set_summation(False)
    off = 0
    for i, p in enumerate(progs):
        nv = sum(len(c.vals) for c in p.constants)
        rf, rc = O.run_lbfgs(p, ods, iters=15, m=6 if not fp64 else 5, fp64=fp64)
        assert (fit[i] == rf) or (np.isnan(fit[i]) and np.isnan(rf)), (i, fit[i], rf)
        assert np.array_equal(vals[off:off + nv].view(np.uint32), rc[:nv].view(np.uint32)) or \
            np.array_equal(vals[off:off + nv], rc[:nv], equal_nan=True), i
        off += nv


@pytest.mark.parametrize("n_progs,n_points,schedule", [(64, 403, "one-round line search"),
                                                       (512, 2000, "fixed schedule")])
def test_lbfgs_schedules_bit_identical(gpu_ctx, n_progs, n_points, schedule):
    """The L-BFGS schedules by fit size: small fits (below 1e8 node-evaluations for the 12
    line-search steps of every individual) evaluate all 12 steps at once and take the first
    that passes; mid-size fits run the 12 rounds back to back.  Sampled programs' fitness and
    constants equal the oracle's bit for bit (reference order, FP64 mode)."""
    X, y, t = W.data_c2(n_points, seed=41)
    progs = W.gen_c2(n_progs, seed=42)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    ods = O.ODataset.new(X, y, 15)
    P = api.Programs(gpu_ctx, progs)
    ok = P.status(ds) == 0
    work = sum(len(p.code) for p, g in zip(progs, ok) if g) * n_points
    assert (work * 12 <= 1e8) == (schedule == "one-round line search") and work < 2e8
    gpu_ctx.set_summation(True)
    try:
        fit = api.fit_lbfgs(gpu_ctx, P, ds, iterations=20, history=5, mode=api.LBFGS_FP64, mask=ok.astype(np.uint8))
        vals = P.get_constants()
    finally:
        gpu_ctx.set_summation(False)
    offs = np.cumsum([0] + [sum(len(c.vals) for c in p.constants) for p in progs])
    sample = np.random.default_rng(43).choice(np.flatnonzero(ok), 12, replace=False)
    for i in sample:
        p = progs[i]
        nv = offs[i + 1] - offs[i]
        rf, rc = O.run_lbfgs(p, ods, iters=20, m=5, fp64=True)
        assert (fit[i] == rf) or (np.isnan(fit[i]) and (rf is None or np.isnan(rf))), (i, fit[i], rf)
        assert np.array_equal(vals[offs[i]:offs[i] + nv].view(np.uint32), rc[:nv].view(np.uint32)) or \
            np.array_equal(vals[offs[i]:offs[i] + nv], rc[:nv], equal_nan=True), i
    P.close()
    ds.close()


def test_lbfgs_large_fit_counter_driven_schedule_bit_identical(gpu_ctx):
    """A fit large enough (1,024 Mat3 programs x 20,000 points: over 2e8 node-evaluations per
    round) that the host reads the counters back each tick and ends line searches and the
    fit early: sampled programs' fitness and constants equal the oracle's bit for bit
    (every program's fit is independent of the others in the batch)."""
    X, y, t = W.data_c2(20_000, seed=37)
    progs = W.gen_c2(1024, seed=38)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    ods = O.ODataset.new(X, y, 15)
    P = api.Programs(gpu_ctx, progs)
    ok = P.status(ds) == 0
    assert sum(len(p.code) for p, g in zip(progs, ok) if g) * 20_000 >= 2e8
    gpu_ctx.set_summation(True)
    try:
        fit = api.fit_lbfgs(gpu_ctx, P, ds, iterations=20, history=5, mode=api.LBFGS_FP64, mask=ok.astype(np.uint8))
        vals = P.get_constants()
    finally:
        gpu_ctx.set_summation(False)
    offs = np.cumsum([0] + [sum(len(c.vals) for c in p.constants) for p in progs])
    sample = np.random.default_rng(39).choice(np.flatnonzero(ok), 12, replace=False)
    for i in sample:
        p = progs[i]
        nv = offs[i + 1] - offs[i]
        rf, rc = O.run_lbfgs(p, ods, iters=20, m=5, fp64=True)
        assert (fit[i] == rf) or (np.isnan(fit[i]) and (rf is None or np.isnan(rf))), (i, fit[i], rf)
        assert np.array_equal(vals[offs[i]:offs[i] + nv].view(np.uint32), rc[:nv].view(np.uint32)) or \
            np.array_equal(vals[offs[i]:offs[i] + nv], rc[:nv], equal_nan=True), i
    P.close()
    ds.close()


def test_lbfgs_default_summation_bit_identical_and_kahan_properties(gpu_ctx):
    """The library's default (reference-order) summation: FP64-mode L-BFGS, 20 iterations,
    m = 5, on 128 Mat3 programs gives the oracle's fitness and constants bit for bit, every
    program.  Kahan summation changes the MSEs in their last bits, so its line searches may
    take other steps; what it guarantees is checked instead: the reported fitness is the MSE
    of the written-back constants, and the fits improve the population."""
    X, y, t = W.data_c2(2000, seed=33)
    progs = W.gen_c2(128, seed=34)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    ods = O.ODataset.new(X, y, 15)
    P = api.Programs(gpu_ctx, progs)
    c0 = P.get_constants()
    gpu_ctx.set_summation(True)
    try:
        fit = api.fit_lbfgs(gpu_ctx, P, ds, iterations=20, history=5, mode=api.LBFGS_FP64)
        vals = P.get_constants()
    finally:
        gpu_ctx.set_summation(False)
    off = 0
    for i, p in enumerate(progs):
        nv = sum(len(c.vals) for c in p.constants)
        rf, rc = O.run_lbfgs(p, ods, iters=20, m=5, fp64=True)
        assert (fit[i] == rf) or (np.isnan(fit[i]) and (rf is None or np.isnan(rf))), (i, fit[i], rf)
        assert np.array_equal(vals[off:off + nv].view(np.uint32), rc[:nv].view(np.uint32)) or \
            np.array_equal(vals[off:off + nv], rc[:nv], equal_nan=True), i
        off += nv
    P.set_constants(c0)
    before = api.compute_mse(gpu_ctx, P, ds)
    fit = api.fit_lbfgs(gpu_ctx, P, ds, iterations=20, history=5, mode=api.LBFGS_FP64)
    after = api.compute_mse(gpu_ctx, P, ds)
    # the Armijo loop force-accepts after 11 halvings (lbfgs.rs:113-125), so a fit may end
    # worse than it started; but the reported fitness is the MSE of the written-back constants
    fin = fit < FMAX
    ok = np.abs(after[fin] - fit[fin]) <= 1e-5 * np.maximum(np.abs(fit[fin]), 1e-20)
    assert ok.all()
    assert np.median(fit[fin] / before[fin]) < 0.5


def test_cmaes_injected_normals_bit_exact(gpu_ctx):
    X, y, t = W.data_c3(256, seed=35)
    progs = W.gen_c3(24, seed=36)
    rng = np.random.default_rng(37)
    iters, lam, mu = 12, 16, 8
    normals = rng.standard_normal((iters, len(progs), lam, 32)).astype(f32)
    gpu_ctx.set_summation(True)
    try:
        ds = api.Dataset.new(gpu_ctx, X, y, t)
        ods = O.ODataset.new(X, y, 8)
        P = api.Programs(gpu_ctx, progs)
        fit, its = api.fit_cmaes(gpu_ctx, P, ds, iters, lam, mu, normals=normals)
        vals = P.get_constants()
    finally:
        gpu_ctx.set_summation(False)
    off = 0
    for i, p in enumerate(progs):
        nv = sum(len(c.vals) for c in p.constants)
        rf, rc, rits = O.run_cmaes(p, ods, iters, lam, mu, np.ascontiguousarray(normals[:, i]))
        assert fit[i] == rf and its[i] == rits, (i, fit[i], rf, its[i], rits)
        assert np.array_equal(vals[off:off + nv], rc[:nv]), i
        off += nv


def test_cmaes_philox_improves_and_is_reproducible(gpu_ctx):
    X, y, t = W.data_c3(5000, seed=38)
    progs = W.gen_c3(256, seed=39)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    before = api.compute_mse(gpu_ctx, P, ds)
    c0 = P.get_constants()
    fit1, _ = api.fit_cmaes(gpu_ctx, P, ds, 30, 32, 16, seed=7)
    c1 = P.get_constants()
    P.set_constants(c0)
    fit2, _ = api.fit_cmaes(gpu_ctx, P, ds, 30, 32, 16, seed=7)
    assert np.array_equal(fit1, fit2) and np.array_equal(c1, P.get_constants())
    # best-so-far never worse than the start (cmaes.rs:111-148): fitness includes 0.01*L1
    fin = before < FMAX
    assert (fit1[fin] <= before[fin] + 0.01 * 32 * 10).all()
    assert np.median(fit1[fin] / before[fin]) < 1.0


def test_cmaes_config_lambda32_mu16_30_iterations_bit_exact(gpu_ctx):
    """The Mat2 config's CMA-ES setting -- lambda = 32, mu = 16, 30 iterations -- on its
    programs, with injected normals: fitness, iterations run and fitted constants equal the
    oracle's run_cma_es (cmaes.rs:19-192) bit for bit, in reference-order summation."""
    X, y, t = W.data_c3(5003, seed=1003)
    progs = W.gen_c3(16, seed=2003)
    rng = np.random.default_rng(38)
    iters, lam, mu = 30, 32, 16
    normals = rng.standard_normal((iters, len(progs), lam, 32)).astype(f32)
    gpu_ctx.set_summation(True)
    try:
        ds = api.Dataset.new(gpu_ctx, X, y, t)
        ods = O.ODataset.new(X, y, 8)
        P = api.Programs(gpu_ctx, progs)
        assert (P.status(ds) == 0).all()
        fit, its = api.fit_cmaes(gpu_ctx, P, ds, iters, lam, mu, normals=normals)
        vals = P.get_constants()
    finally:
        gpu_ctx.set_summation(False)
    off = 0
    for i, p in enumerate(progs):
        nv = sum(len(c.vals) for c in p.constants)
        rf, rc, rits = O.run_cmaes(p, ods, iters, lam, mu, np.ascontiguousarray(normals[:, i]))
        assert fit[i] == rf and its[i] == rits, (i, fit[i], rf, its[i], rits)
        assert np.array_equal(vals[off:off + nv], rc[:nv]), i
        off += nv


def test_cmaes_normal_stream_matches_box_muller():
    # host restatement: Philox4x32-10 -> 24-bit uniforms -> Box-Muller with the device ln / cos
    import struct
    z = np.array([api.cmaes_normal(123, 5, it, c, j) for it in range(3) for c in range(8) for j in range(16)])
    assert abs(z.mean()) < 0.15 and 0.85 < z.std() < 1.15
    assert api.cmaes_normal(123, 5, 0, 0, 0) == api.cmaes_normal(123, 5, 0, 0, 0)


@pytest.mark.parametrize("gen,data,nf", [(W.gen_c0, W.data_c0, 1), (W.gen_c3, W.data_c3, 8)])
def test_nelder_mead_reference_order_bit_exact(gpu_ctx, gen, data, nf):
    X, y, t = data(403, seed=81)
    progs = gen(48, seed=82)
    gpu_ctx.set_summation(True)
    try:
        ds = api.Dataset.new(gpu_ctx, X, y, t)
        P = api.Programs(gpu_ctx, progs)
        st = P.status(ds)
        fit, its = api.fit_nelder_mead(gpu_ctx, P, ds, iterations=60, mask=(st == 0).astype(np.uint8))
        vals = P.get_constants()
    finally:
        gpu_ctx.set_summation(False)
    ods = O.ODataset.new(X, y, nf)
    off = 0
    for i, p in enumerate(progs):
        nv = sum(len(c.vals) for c in p.constants)
        if st[i] == 0:
            rf, rc = O.run_nelder_mead(p, ods, 60)
            if rf is None:
                assert np.isnan(fit[i]), i
            else:
                assert fit[i] == rf, (i, fit[i], rf)
                assert np.array_equal(vals[off:off + nv].view(np.uint32), rc[:nv].view(np.uint32)), i
        off += nv


def _mat2_const_program():
    """tr(A^-1 (B + C)) + c with a Mat2 constant C and a Float constant c."""
    return Program([(Op.LoadVarM2, 0), (Op.InverseM2, 0), (Op.LoadVarM2, 4), (Op.LoadConstM2, 0), (Op.AddM2, 0),
                    (Op.MulM2, 0), (Op.TraceM2, 0), (Op.LoadConstF, 1), (Op.AddF, 0)],
                   [Scalar(VT.Mat2, (0.3, -0.2, 0.1, 0.04)), Scalar(VT.Float, (0.7,))])


def test_optimize_constants_dispatch_snap_and_mse(gpu_ctx):
    X, y, t = W.data_c3(600, seed=83)
    progs = W.gen_c3(40, seed=84) + [_mat2_const_program(), _mat2_const_program()]
    gpu_ctx.set_summation(True)
    try:
        ds = api.Dataset.new(gpu_ctx, X, y, t)
        P = api.Programs(gpu_ctx, progs)
        st = P.status(ds)
        ok = (st == 0).astype(np.uint8)
        before = P.get_constants().copy()
        mse, meth = api.optimize_constants(gpu_ctx, P, ds, iterations=40, seed=5, mask=ok)
        after = P.get_constants()
        again = api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)
    finally:
        gpu_ctx.set_summation(False)
    ods = O.ODataset.new(X, y, 8)
    off = 0
    for i, p in enumerate(progs):
        nv = sum(len(c.vals) for c in p.constants)
        if not ok[i]:
            assert meth[i] == api.FIT_NONE and np.isnan(mse[i])
            off += nv
            continue
        has_inv = any(op == Op.InverseM2 for op, _ in p.code)
        vecmat = any(c.vt != VT.Float for c in p.constants)
        want = api.FIT_LBFGS if not has_inv else (api.FIT_CMAES if vecmat else api.FIT_NELDER_MEAD)
        assert meth[i] == want, i
        v = after[off:off + nv]
        assert not np.any((np.abs(v) < 0.05) & (v != 0)), i  # apply_threshold
        assert mse[i] == again[i], i  # calculate_mse of the snapped program
        # deterministic fits follow the oracle exactly, then snap
        if want in (api.FIT_LBFGS, api.FIT_NELDER_MEAD):
            q = _with_consts(p, before[off:off + nv])
            if want == api.FIT_LBFGS:
                _, rc = O.run_lbfgs(q, ods, iters=15, m=6)
            else:
                _, rc = O.run_nelder_mead(q, ods, 40)
            rc = np.where(np.abs(rc[:nv]) < 0.05, f32(0), rc[:nv]).astype(f32)
            assert np.array_equal(v.view(np.uint32), rc.view(np.uint32)) or np.array_equal(v, rc), i
        off += nv


@pytest.mark.parametrize("which", ["scalar", "mat2"])
def test_optimize_individuals_simplify_and_rescore(gpu_ctx, which):
    """Individual::optimize_constants as the island uses it: fit and snap on the GPU, then
    simplify_ast, recompile and rescore.  The simplified program equals the oracle's
    simplification of the snapped program, and its MSE (reference order), complexity and
    fitness equal the oracle's for that program."""
    from oracle import simplify as OS
    if which == "scalar":
        X, y, t = W.data_c0(403, seed=91)
        progs = W.gen_c0(80, seed=92)
    else:
        X, y, t = W.data_c3(500, seed=93)
        progs = W.gen_c3(40, seed=94)
    gpu_ctx.set_summation(True)
    try:
        ds = api.Dataset.new(gpu_ctx, X, y, t)
        P = api.Programs(gpu_ctx, progs)
        ok = (P.status(ds) == 0).astype(np.uint8)
        P.close()
        out, mse, comp, fit, meth = api.optimize_individuals(gpu_ctx, progs, ds, iterations=20, seed=3, mask=ok)
        # the snapped constants of the same (deterministic) fit, for the oracle's simplification
        P2 = api.Programs(gpu_ctx, progs)
        _, meth2 = api.optimize_constants(gpu_ctx, P2, ds, iterations=20, seed=3, mask=ok)
        snapped = P2.get_constants()
    finally:
        gpu_ctx.set_summation(False)
    assert np.array_equal(meth, meth2)
    ods = O.ODataset.new(X, y, X.shape[1])
    off, shrunk = 0, 0
    for i, p in enumerate(progs):
        nv = sum(len(c.vals) for c in p.constants)
        if ok[i]:
            want = OS.simplify(_with_consts(p, snapped[off:off + nv]))
            assert [(int(o), int(a)) for o, a in out[i].code] == [(int(o), int(a)) for o, a in want.code], i
            assert [np.float32(c.vals).tobytes() for c in out[i].constants] == \
                [np.float32(c.vals).tobytes() for c in want.constants], i
            shrunk += len(want.code) < len(p.code)
        else:
            assert out[i] is progs[i]
        m = O.compute_mse(out[i], ods)
        assert np.float32(mse[i]).view(np.uint32) == np.float32(m).view(np.uint32), (i, mse[i], m)
        assert fit[i] == np.float32(O.fitness(float(mse[i]), int(comp[i]), 5e-6)), i
        off += nv
    assert np.array_equal(comp, api.complexity(out))
    assert shrunk > 0  # snapping to zero made some programs simplify
