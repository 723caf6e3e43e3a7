"""Edge cases of the evaluation entry points on the GPU, against the oracle: an empty program
batch, a dataset without points (the reference's 0 / 0 MSE is not finite and reported as
f32::MAX, evaluator.rs:173; its gradient is then zero, evaluator.rs:214-216), a batch whose
every program the lowering rejects, and wide feature sets (27, 32 and 33 components)."""
import numpy as np
import pytest

import oracle as O
from rsr_amd import api
from rsr_amd import workloads as W
from rsr_amd.isa import Op, Program, Scalar, VT

pytestmark = pytest.mark.gpu
f32 = np.float32
FMAX = np.finfo(f32).max


def test_empty_batch(gpu_ctx):
    X, y, t = W.data_c0(100, seed=1)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, [])
    try:
        assert api.compute_mse(gpu_ctx, P, ds).size == 0
        m, g = api.compute_mse_with_gradient(gpu_ctx, P, ds)
        assert m.size == 0
    finally:
        P.close()
        ds.close()


@pytest.mark.parametrize("reference", [True, False])
def test_dataset_without_points(gpu_ctx, reference):
    X, y, t = W.data_c2(10, seed=3)
    progs = W.gen_c2(24, seed=4)
    ds = api.Dataset.new(gpu_ctx, X[:0], y[:0], t)
    ods = O.ODataset.new(X[:0], y[:0], 15)
    P = api.Programs(gpu_ctx, progs)
    gpu_ctx.set_summation(reference)
    try:
        mse = api.compute_mse(gpu_ctx, P, ds)
        m, g = api.compute_mse_with_gradient(gpu_ctx, P, ds)
        for i, p in enumerate(progs):
            assert mse[i] == O.compute_mse(p, ods) == FMAX
            rm, rg = O.compute_mse_with_gradient(p, ods)
            assert m[i] == rm == FMAX and not np.asarray(g[i]).any() and not np.asarray(rg).any()
        with pytest.raises(api.RsrError):
            api.fit_lbfgs(gpu_ctx, P, ds, iterations=2)
    finally:
        gpu_ctx.set_summation(False)
        P.close()
        ds.close()


def test_every_program_rejected(gpu_ctx):
    X, y, t = W.data_c0(64, seed=5)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    bad = [Program([(Op.LoadVarF, 3), (Op.SinF, 0)], []),   # feature out of range
           Program([(Op.AddF, 0)], []),                     # stack underflow
           Program([(Op.LoadVarF, 0), (Op.LoadVarF, 0)], [])]  # two values left
    P = api.Programs(gpu_ctx, bad)
    try:
        assert (P.status(ds) != 0).all()
        with pytest.raises(api.RsrError):
            api.compute_mse(gpu_ctx, P, ds)
        assert np.isnan(api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)).all()
    finally:
        P.close()
        ds.close()


def _wide_dataset(n, seed, types):
    rng = np.random.default_rng(seed)
    F = sum({VT.Float: 1, VT.Vec3: 3, VT.Mat3: 9}[t] for t in types)
    X = rng.normal(size=(n, F)).astype(f32)
    y = rng.normal(size=n).astype(f32)
    return X, y


def test_wide_feature_sets(gpu_ctx):
    """27 feature components (two Mat3, two Vec3, three Float) leave five registers plus the
    spill slots for the program's values: Mat3 products and sums still lower and run
    bit-identical to the oracle (operands are flat component offsets, the last one 26).  At
    32 components no register is left for a result and the lowering rejects the program
    (RSR_PROG_REGS, no undefined behaviour); a 33rd component is refused at upload."""
    types = [VT.Mat3, VT.Mat3, VT.Vec3, VT.Vec3, VT.Float, VT.Float, VT.Float]
    X, y = _wide_dataset(1001, 7, types)
    progs = [
        Program([(Op.LoadVarF, 26), (Op.LoadVarF, 25), (Op.MulF, 0)], []),
        Program([(Op.LoadVarM3, 9), (Op.LoadVarV3, 21), (Op.MulM3V3, 0), (Op.NormV3, 0), (Op.LoadVarF, 26),
                 (Op.AddF, 0)], []),
        Program([(Op.LoadVarM3, 0), (Op.LoadVarM3, 9), (Op.MulM3, 0), (Op.LoadConstM3, 0), (Op.AddM3, 0),
                 (Op.DetM3, 0), (Op.LoadConstF, 1), (Op.MulF, 0)],
                [Scalar(VT.Mat3, [0.5, -1.0, 2.0, 0.25, 1.5, -0.75, 1.0, 0.0, -2.0]), Scalar(VT.Float, [1.5])]),
    ]
    ds = api.Dataset.new(gpu_ctx, X, y, types)
    ods = O.ODataset.new(X, y, 27)
    P = api.Programs(gpu_ctx, progs)
    gpu_ctx.set_summation(True)
    try:
        assert (P.status(ds) == 0).all()
        mse = api.compute_mse(gpu_ctx, P, ds)
        pts = api.eval_points(gpu_ctx, P, ds)
        for i, p in enumerate(progs):
            assert mse[i].view(np.uint32) == f32(O.compute_mse(p, ods)).view(np.uint32), i
            ref = O.eval_points(p, ods)
            assert np.array_equal(pts[i].view(np.uint32), ref.view(np.uint32)), i
    finally:
        gpu_ctx.set_summation(False)
        P.close()
        ds.close()
    full = [VT.Mat3, VT.Mat3, VT.Mat3, VT.Vec3, VT.Float, VT.Float]  # 32 components
    X32, y32 = _wide_dataset(64, 8, full)
    ds = api.Dataset.new(gpu_ctx, X32, y32, full)
    P = api.Programs(gpu_ctx, [Program([(Op.LoadVarF, 31), (Op.LoadVarF, 30), (Op.MulF, 0)], [])])
    try:
        assert P.status(ds)[0] == -6  # RSR_PROG_REGS
    finally:
        P.close()
        ds.close()
    X33, y33 = _wide_dataset(64, 9, full + [VT.Float])
    with pytest.raises(api.RsrError):
        api.Dataset.new(gpu_ctx, X33, y33, full + [VT.Float])
