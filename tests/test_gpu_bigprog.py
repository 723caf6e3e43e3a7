"""GPU parity of programs beyond the register file and the LDS spill slots ("big" programs:
spill slots in device memory, bounded launches of their own), through the C ABI.

The reference evaluates any program within its 32-deep typed stacks
(src/engine/eval/state.rs:12-37); these programs need more than the 32 registers plus 16
LDS slots: the deep-stack population (16 live Mat3 values, two computed Mat3 operands of
one MulM3 at 15 features, 32 live Mat2 values, 31 live Floats, ...) and the three programs
of gen_c2(16384, seed=2002) that the lowering used to refuse (5633, 5663, 14173).  They run
mixed with ordinary programs in one batch, at one, two and four points per lane:
reference-order MSEs, per-point values, gradients and tangents are bit-identical to the
oracle; Kahan MSEs within 1e-5; and the ordinary programs' results do not change when big
ones share their batch."""
import numpy as np
import pytest

import opcode_programs as OP
import oracle as O
from rsr_amd import api
from rsr_amd import workloads as W
from rsr_amd.isa import VT

pytestmark = pytest.mark.gpu
f32 = np.float32
C2_REFUSED = (5633, 5663, 14173)


def _same(a, b):
    a = np.asarray(a, f32)
    b = np.asarray(b, f32)
    ok = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b)) | ((a == 0) & (b == 0))
    return bool(ok.all())


def _is_big(p, F, dual=False):
    st, words = api.lower_code(p, F, dual=dual)
    assert st == 0
    slots = [int(w[0]) >> 8 for w in words if (int(w[0]) & 255) in (19, 20)]  # SPILL, FILL
    return bool(slots) and max(slots) >= 16


@pytest.fixture(scope="module")
def c2_case():
    allp = W.gen_c2(16384, seed=2002)
    big = [allp[i] for i in C2_REFUSED]
    small = allp[:40]
    progs = small[:20] + big + small[20:]  # big programs inside the batch, not at its end
    return progs, [20, 21, 22]


@pytest.mark.parametrize("ppl", [1, 2, 3, 4])
def test_refused_c2_programs_now_evaluate_bit_identical(gpu_ctx, c2_case, ppl):
    progs, big_idx = c2_case
    assert all(_is_big(progs[i], 15) for i in big_idx)
    X, y, t = W.data_c2(3001, seed=1002)
    print("big c2: ppl", ppl, flush=True)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    ods = O.ODataset.new(X, y, 15)
    P = api.Programs(gpu_ctx, progs)
    assert (P.status(ds) == 0).all()
    gpu_ctx.set_tuning(ppl, 4)
    try:
        gpu_ctx.set_summation(True)
        mse_ref = api.compute_mse(gpu_ctx, P, ds)
        gm_ref, gg_ref = api.compute_mse_with_gradient(gpu_ctx, P, ds)
        gpu_ctx.set_summation(False)
        mse_k = api.compute_mse(gpu_ctx, P, ds)
        gm_k, gg_k = api.compute_mse_with_gradient(gpu_ctx, P, ds)
        pts = api.eval_points(gpu_ctx, P, ds)
    finally:
        gpu_ctx.set_summation(False)
        gpu_ctx.set_tuning(2, 8)
    for i, p in enumerate(progs):
        rm = O.compute_mse(p, ods)
        assert _same(mse_ref[i], rm), (i, mse_ref[i], rm)
        assert abs(float(mse_k[i]) - rm) <= 1e-5 * max(abs(rm), 1e-30) or mse_k[i] == rm, (i, mse_k[i], rm)
        ref = O.eval_points(p, ods)
        assert _same(pts[i][: len(ref)], ref), i
        om, og = O.compute_mse_with_gradient(p, ods)
        assert _same(gm_ref[i], om) and _same(gg_ref[i], og), i
        scale = max(abs(om), 1e-30)
        assert abs(float(gm_k[i]) - om) <= 1e-5 * scale or gm_k[i] == om, i
    # the ordinary programs alone give the same bits as beside the big ones
    Ps = api.Programs(gpu_ctx, [p for i, p in enumerate(progs) if i not in big_idx])
    gpu_ctx.set_tuning(ppl, 4)
    try:
        alone = api.compute_mse(gpu_ctx, Ps, ds)
    finally:
        gpu_ctx.set_tuning(2, 8)
    mixed = np.array([m for i, m in enumerate(mse_k) if i not in big_idx], f32)
    assert _same(alone, mixed)
    for i in big_idx:
        for k in range(min(progs[i].param_count(), 3)):
            v, tt = api.eval_points_dual(gpu_ctx, P, ds, i, k)
            rv, rt = O.eval_points_dual(progs[i], ods, k)
            assert _same(v[: len(rv)], rv) and _same(tt[: len(rt)], rt), (i, k)
    Ps.close()
    P.close()
    ds.close()


@pytest.mark.parametrize("ppl", [1, 4])
def test_deep_stack_programs_bit_identical(gpu_ctx, ppl):
    progs = OP.deep_stack_programs() + OP.targeted_programs()[:6]
    rng = np.random.default_rng(77)
    n = 2003
    X = OP.feature_rows(rng, n)
    y = rng.standard_normal(n).astype(f32)
    print("deep stacks: ppl", ppl, flush=True)
    ds = api.Dataset.new(gpu_ctx, X, y, [VT.Float] * OP.N_FEATURES)
    ods = O.ODataset.new(X, y, OP.N_FEATURES)
    P = api.Programs(gpu_ctx, progs)
    assert (P.status(ds) == 0).all()
    gpu_ctx.set_tuning(ppl, 4)
    gpu_ctx.set_summation(True)
    try:
        mse = api.compute_mse(gpu_ctx, P, ds)
        gm, gg = api.compute_mse_with_gradient(gpu_ctx, P, ds)
        pts = api.eval_points(gpu_ctx, P, ds)
    finally:
        gpu_ctx.set_summation(False)
        gpu_ctx.set_tuning(2, 8)
    for i, p in enumerate(progs):
        assert _same(mse[i], O.compute_mse(p, ods)), (i, mse[i], O.compute_mse(p, ods))
        ref = O.eval_points(p, ods)
        assert _same(pts[i][: len(ref)], ref), i
        om, og = O.compute_mse_with_gradient(p, ods)
        assert _same(gm[i], om) and _same(gg[i], og), i
        if ppl == 1:
            for k in range(min(p.param_count(), 2)):
                v, tt = api.eval_points_dual(gpu_ctx, P, ds, i, k)
                rv, rt = O.eval_points_dual(p, ods, k)
                assert _same(v[: len(rv)], rv) and _same(tt[: len(rt)], rt), (i, k)
    P.close()
    ds.close()


def test_big_programs_in_fits(gpu_ctx, c2_case):
    """L-BFGS (reference mode) over a batch holding the big programs: their fitted
    constants and fitness equal the oracle's restatement of run_lbfgs bit for bit."""
    progs, big_idx = c2_case
    X, y, t = W.data_c2(1000, seed=1003)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    ods = O.ODataset.new(X, y, 15)
    P = api.Programs(gpu_ctx, progs)
    gpu_ctx.set_summation(True)
    try:
        fit = api.fit_lbfgs(gpu_ctx, P, ds, iterations=15, history=6, mode=api.LBFGS_REFERENCE)
        vals = P.get_constants()
    finally:
        gpu_ctx.set_summation(False)
    off = 0
    for i, p in enumerate(progs):
        nv = sum(len(c.vals) for c in p.constants)
        if i in big_idx or i < 4:
            rf, rc = O.run_lbfgs(p, ods, iters=15, m=6)
            assert (fit[i] == rf) or (np.isnan(fit[i]) and np.isnan(rf)), (i, fit[i], rf)
            assert _same(vals[off:off + nv], rc[:nv]), i
        off += nv
    P.close()
    ds.close()
