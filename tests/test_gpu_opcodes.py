"""GPU parity of every operator and every micro-op, through the C ABI.

The population is the all-operator random draw plus hand-built programs for the rarely
drawn paths (ln and sqr of data, Vec2 / Vec3 / Mat2 / Mat3 constants, trace and inverse of
Mat3, deep stacks) and the programs whose Mat3 block micro-ops sit at the top of the
register file.  Lowered on the CPU first, the population covers every micro-op of the
value code and of the dual code; on the device, at one, two and four points per lane:
per-point values, reference-order MSEs, reference-order MSE gradients and per-point
tangents are bit-identical to the oracle (same IEEE operations in the same order)."""
import numpy as np
import pytest

import opcode_programs as OP
import oracle as O
from rsr_amd import api
from rsr_amd import workloads as W
from rsr_amd.isa import UOPS, VT

pytestmark = pytest.mark.gpu
f32 = np.float32


def _same(a, b):
    a = np.asarray(a, f32)
    b = np.asarray(b, f32)
    ok = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b)) | ((a == 0) & (b == 0))
    return bool(ok.all())


# value-only micro-ops: the dual lowering never emits them
_VALUE_ONLY = {"RCPSAFE", "INVSEL", "INVM2", "RSUB", "RDIV", "RSUBV3", "RCROSS3", "SCALEV3S", "MULM3V3R", "SINU",
               "COSU", "EXPU", "LNU", "SQRTU", "SINK", "COSK", "EXPK", "LNK", "SQRTK", "NEGK", "SQRK"}


def _population():
    rng = np.random.default_rng(31)
    return OP.all_ops_programs(rng, 400) + OP.targeted_programs()


def _coverage(progs, F, dual):
    seen = set()
    for p in progs:
        st, words = api.lower_code(p, F, dual=dual)
        assert st == 0, p.code
        seen.update(UOPS[int(w[0]) & 255] for w in words)
    return seen


@pytest.fixture(scope="module")
def population():
    progs = _population()
    # every micro-op is reached by the value code, every dual-capable one by the dual code
    assert _coverage(progs, OP.N_FEATURES, False) == set(UOPS)
    assert _coverage(progs, OP.N_FEATURES, True) == set(UOPS) - _VALUE_ONLY
    return progs


@pytest.mark.parametrize("n", [1003, 4096])
@pytest.mark.parametrize("ppl", [1, 2, 4])
def test_every_opcode_values_and_mse_bit_identical(gpu_ctx, population, ppl, n):
    rng = np.random.default_rng(100 + n)
    X = OP.feature_rows(rng, n)
    y = rng.standard_normal(n).astype(f32)
    print("opcodes values+mse: ppl", ppl, "points", n, "programs", len(population), flush=True)
    ds = api.Dataset.new(gpu_ctx, X, y, [VT.Float] * OP.N_FEATURES)
    ods = O.ODataset.new(X, y, OP.N_FEATURES)
    P = api.Programs(gpu_ctx, population)
    gpu_ctx.set_tuning(ppl, 4)
    gpu_ctx.set_summation(True)
    try:
        mse = api.compute_mse(gpu_ctx, P, ds)
        pts = api.eval_points(gpu_ctx, P, ds)
    finally:
        gpu_ctx.set_summation(False)
        gpu_ctx.set_tuning(2, 8)
    for i, p in enumerate(population):
        ref = O.eval_points(p, ods)
        assert _same(pts[i][: len(ref)], ref), (i, p.code)
        assert _same(mse[i], O.compute_mse(p, ods)), (i, mse[i], O.compute_mse(p, ods), p.code)
    P.close()
    ds.close()


def test_every_opcode_gradients_and_tangents_bit_identical(gpu_ctx, population):
    rng = np.random.default_rng(7)
    n = 1003
    X = OP.feature_rows(rng, n)
    y = rng.standard_normal(n).astype(f32)
    print("opcodes gradients: programs", len(population), flush=True)
    ds = api.Dataset.new(gpu_ctx, X, y, [VT.Float] * OP.N_FEATURES)
    ods = O.ODataset.new(X, y, OP.N_FEATURES)
    P = api.Programs(gpu_ctx, population)
    gpu_ctx.set_summation(True)
    try:
        mse, grad = api.compute_mse_with_gradient(gpu_ctx, P, ds)
    finally:
        gpu_ctx.set_summation(False)
    checked = 0
    for i, p in enumerate(population):
        rm, rg = O.compute_mse_with_gradient(p, ods)
        assert _same(mse[i], rm), (i, mse[i], rm, p.code)
        assert _same(grad[i], rg), (i, p.code)
        for k in range(min(p.param_count(), 3)):
            v, t = api.eval_points_dual(gpu_ctx, P, ds, i, k)
            rv, rt = O.eval_points_dual(p, ods, k)
            assert _same(v[: len(rv)], rv) and _same(t[: len(rt)], rt), (i, k, p.code)
            checked += 1
    assert checked > 300
    P.close()
    ds.close()


@pytest.mark.parametrize("ppl", [1, 2, 4])
def test_mat3_blocks_at_register_23_bit_identical(gpu_ctx, ppl):
    """Programs whose nine-register Mat3 blocks start at register 23 (components 23..31),
    in the value and the dual code, against the oracle."""
    progs = OP.mat3_block_programs(limit=24)
    assert len(progs) == 24
    X, y, t = W.data_c2(2011, seed=45)
    print("mat3 blocks at register 23: ppl", ppl, flush=True)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    ods = O.ODataset.new(X, y, X.shape[1])
    P = api.Programs(gpu_ctx, progs)
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
        ref = O.eval_points(p, ods)
        assert _same(pts[i][: len(ref)], ref), i
        assert _same(mse[i], O.compute_mse(p, ods)), i
        rm, rg = O.compute_mse_with_gradient(p, ods)
        assert _same(gm[i], rm) and _same(gg[i], rg), i
        if ppl == 1:
            for k in range(p.param_count()):
                v, tt = api.eval_points_dual(gpu_ctx, P, ds, i, k)
                rv, rt = O.eval_points_dual(p, ods, k)
                assert _same(v[: len(rv)], rv) and _same(tt[: len(rt)], rt), (i, k)
    P.close()
    ds.close()
