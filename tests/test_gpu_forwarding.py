"""GPU parity of the value-code post-pass paths: forwarded operands, dropped write-backs,
operand-swapped micro-ops, constant-operand unary ops and point-invariant
transcendentals, at every points-per-lane setting (one point per lane runs with the
accumulator off, two and four with it on).  Bar: bit-identical per-point values and
reference-order MSEs to the oracle."""
import numpy as np
import pytest

import forwarding_programs as FP
import oracle as O
from rsr_amd import api
from rsr_amd import isa
from rsr_amd import workloads as W

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    ok = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b)) | ((a == 0) & (b == 0))
    return bool(ok.all())


@pytest.mark.parametrize("ppl", [1, 2, 4])
@pytest.mark.parametrize("which", ["lorentz", "mat3"])
def test_forwarding_paths_bit_identical(gpu_ctx, which, ppl):
    if which == "lorentz":
        entries, (X, y, t) = FP.lorentz_programs(), W.data_c1(1003, seed=71)
    else:
        entries, (X, y, t) = FP.mat3_programs(), W.data_c2(1003, seed=72)
    progs = [p for _, _, p in entries]
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    ods = O.ODataset.new(X, y, X.shape[1])
    P = api.Programs(gpu_ctx, progs)
    gpu_ctx.set_summation(True)
    gpu_ctx.set_tuning(ppl, 4)
    try:
        mse = api.compute_mse(gpu_ctx, P, ds)
    finally:
        gpu_ctx.set_summation(False)
        gpu_ctx.set_tuning(2, 8)
    pts = api.eval_points(gpu_ctx, P, ds)
    for i, (name, _, p) in enumerate(entries):
        assert mse[i] == O.compute_mse(p, ods), name
        ref = O.eval_points(p, ods)
        assert _bits_equal(pts[i][: len(ref)], ref), name


def _last_register_programs(gen, X, count, seed, extra):
    """Programs whose dual code writes a one-component result to register 30 or 31 (the
    file's last two registers), and the first `extra` programs of the pool."""
    F = X.shape[1]
    pool = [p for p in gen(count, seed=seed) if api.lower_code(p, F, dual=False)[0] == 0]
    hit = []
    for p in pool:
        st, words = api.lower_code(p, F, dual=True)
        if st == 0 and any((int(w[1]) & 31) >= 30 and not (int(w[1]) & (isa.UFLAG_VEC3 | isa.UFLAG_NOWB))
                           for w in words):
            hit.append(p)
    return hit, pool[:extra]


@pytest.mark.parametrize("which", ["mat3", "mat2"])
def test_dual_last_registers_bit_identical(gpu_ctx, which):
    """Dual code whose one-component results land in registers 30 and 31 next to Vec3
    results: per-point values and tangents of every parameter, and reference-order
    gradients, bit-identical to the oracle (regression: a write-back of components 1
    and 2 speculated past the end of the file overwrote a pending tangent)."""
    if which == "mat3":
        X, y, t = W.data_c2(1503, seed=63)
        hit, rest = _last_register_programs(W.gen_c2, X, 3000, 64, 40)
        assert len(hit) >= 10  # the pool reaches the last registers
    else:  # Mat2 programs stay below register 30; their block ops take scalars from 29..31
        X, y, t = W.data_c3(777, seed=73)
        hit, rest = _last_register_programs(W.gen_c3, X, 3000, 74, 80)
    progs = hit + rest
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    ods = O.ODataset.new(X, y, X.shape[1])
    P = api.Programs(gpu_ctx, progs)
    for i, p in enumerate(progs):
        for k in range(min(p.param_count(), 32)):
            v, g = api.eval_points_dual(gpu_ctx, P, ds, i, k)
            rv, rg = O.eval_points_dual(p, ods, k)
            assert _bits_equal(v[: len(rv)], rv), (i, k)
            assert _bits_equal(g[: len(rg)], rg), (i, k)
    gpu_ctx.set_summation(True)
    try:
        mse, grad = api.compute_mse_with_gradient(gpu_ctx, P, ds)
    finally:
        gpu_ctx.set_summation(False)
    for i, p in enumerate(progs):
        rm, rg = O.compute_mse_with_gradient(p, ods)
        assert _bits_equal(mse[i], rm), (i, mse[i], rm)
        assert _bits_equal(grad[i], rg), i
