"""GPU parity of sin, cos, exp, ln and the square roots over their whole input ranges.

The device computes sin / cos / exp with fused multiply-adds and ldexp, and takes a short
square-root sequence whenever no value of the wave is below 2^-96 in magnitude.  Each
function runs here on datasets built to reach every branch of those sequences: arguments
near multiples of pi/2 and up to 2^24 and beyond, exp arguments whose results are
subnormal or overflow, ln of subnormals, square roots of values below 2^-96 in some waves
and not in others, and inf, NaN, +-0 and negatives.  Per-point values (one point per lane)
are bit-identical to the oracle; and with the oracle's values as targets, the
reference-order MSE is exactly 0 at every points-per-lane setting, which every point of
every wave must meet."""
import numpy as np
import pytest

import oracle as O
from rsr_amd import api
from rsr_amd.isa import Op, Program, VT

pytestmark = pytest.mark.gpu
f32 = np.float32


def _same(a, b):
    a = np.asarray(a, f32)
    b = np.asarray(b, f32)
    return bool(((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))).all())


def _args(fn: str, finite: bool) -> np.ndarray:
    rng = np.random.default_rng(77)
    if fn in ("sin", "cos"):
        k = rng.integers(-40000, 40000, 6000).astype(np.float64)
        near = (k * (np.pi / 2) + rng.uniform(-1e-3, 1e-3, k.size)).astype(f32)  # near multiples of pi/2
        parts = [rng.uniform(-8192, 8192, 8000), near, rng.uniform(-10, 10, 4000),
                 rng.uniform(-16777215, 16777215, 2000), np.float32(2.0) ** rng.uniform(-140, 24, 2000),
                 [0.0, -0.0, 16777215.0, -16777215.0]]
        if not finite:
            parts.append([16777216.0, 3e30, -3e30, np.inf, -np.inf, np.nan])
    elif fn == "exp":
        parts = [rng.uniform(-104, 88.7228, 8000), rng.uniform(-103.99, -87.3, 4000),
                 rng.uniform(88.5, 88.7228, 2000), rng.uniform(-1, 1, 2000), [0.0, -0.0, 88.72283, -103.97, -104.0]]
        if not finite:  # results past FLT_MAX
            parts.append([88.72284, 88.8, 89.0, -1e30, 1e30, np.inf, -np.inf, np.nan])
    elif fn == "ln":
        parts = [np.float32(2.0) ** rng.uniform(-148.5, 127.9, 8000), rng.uniform(0.5, 2.0, 4000),
                 [1.0, 1e-45, 1e-40, 1.17549435e-38, 3.4028235e38]]
        if not finite:
            parts.append([0.0, -0.0, -1.0, np.inf, -np.inf, np.nan])
    else:  # sqrt: some waves hold a value below 2^-96 (the long sequence), most do not
        x = (np.float32(2.0) ** rng.uniform(-95, 127, 16384)).astype(f32)
        tiny = (np.float32(2.0) ** rng.uniform(-148.5, -96.5, 16)).astype(f32)
        x[np.arange(16) * 1024 + rng.integers(0, 64, 16)] = tiny
        parts = [x, [0.0, 2.0 ** -96, 1e-45]]
        if not finite:
            parts.append([-0.0, -1.0, -1e-40, np.inf, -np.inf, np.nan])
    # the hand-picked values first, and a whole number of 8-point batches: no padding lanes
    # (which would add f(0) - 0 to the MSE sums)
    x = np.concatenate([np.asarray(p, np.float64) for p in parts[::-1]]).astype(f32)
    return x[: x.size // 8 * 8]


_PROG = {"sin": Op.SinF, "cos": Op.CosF, "exp": Op.ExpF, "ln": Op.LnF, "sqrt": Op.SqrtF}


@pytest.mark.parametrize("fn", ["sin", "cos", "exp", "ln", "sqrt"])
def test_values_bit_identical_everywhere(gpu_ctx, fn):
    x = _args(fn, finite=False)
    print("transcendental values:", fn, x.size, flush=True)
    prog = Program([(Op.LoadVarF, 0), (_PROG[fn], 0)], [])
    y = np.zeros_like(x)
    ds = api.Dataset.new(gpu_ctx, x.reshape(-1, 1), y, [VT.Float])
    P = api.Programs(gpu_ctx, [prog])
    try:
        got = api.eval_points(gpu_ctx, P, ds)[0]
        ref = O.eval_points(prog, O.ODataset.new(x.reshape(-1, 1), y, 1))
        assert _same(got, ref), np.flatnonzero(~((got.view(np.uint32) == ref.view(np.uint32)) |
                                                 (np.isnan(got) & np.isnan(ref))))[:10]
    finally:
        P.close()
        ds.close()


@pytest.mark.parametrize("fn", ["sin", "cos", "exp", "ln", "sqrt"])
def test_oracle_targets_give_mse_exactly_zero(gpu_ctx, fn):
    x = _args(fn, finite=True)
    prog = Program([(Op.LoadVarF, 0), (_PROG[fn], 0)], [])
    y = O.eval_points(prog, O.ODataset.new(x.reshape(-1, 1), np.zeros_like(x), 1))[: x.size].copy()
    assert np.isfinite(y).all()
    ds = api.Dataset.new(gpu_ctx, x.reshape(-1, 1), y, [VT.Float])
    P = api.Programs(gpu_ctx, [prog])
    gpu_ctx.set_summation(True)
    try:
        for ppl in (1, 2, 3, 4):
            print("transcendental mse:", fn, "ppl", ppl, flush=True)
            gpu_ctx.set_tuning(ppl, 4)
            assert api.compute_mse(gpu_ctx, P, ds)[0] == 0.0, (fn, ppl)
    finally:
        gpu_ctx.set_summation(False)
        gpu_ctx.set_tuning(0, 0)
        P.close()
        ds.close()


@pytest.mark.parametrize("tiny_waves", [False, True])
def test_norm_sqrt_long_and_short_sequences(gpu_ctx, tiny_waves):
    """NormV3 of vectors whose squared length is below 2^-96 in some waves (the long
    square-root sequence for those waves) or in none: per-point values bit-identical, and
    the MSE against the oracle's values exactly 0 at every points-per-lane setting."""
    rng = np.random.default_rng(5)
    v = rng.normal(size=(8192, 3)).astype(f32)
    if tiny_waves:
        idx = np.arange(0, 8192, 512) + 7
        v[idx] = (rng.normal(size=(idx.size, 3)) * 1e-16).astype(f32)
        v[idx[0]] = 0.0
    prog = Program([(Op.LoadVarV3, 0), (Op.NormV3, 0)], [])
    y = O.eval_points(prog, O.ODataset.new(v, np.zeros(8192, f32), 3))[:8192].copy()
    ds = api.Dataset.new(gpu_ctx, v, y, [VT.Vec3])
    P = api.Programs(gpu_ctx, [prog])
    gpu_ctx.set_summation(True)
    try:
        assert _same(api.eval_points(gpu_ctx, P, ds)[0][:8192], y)
        for ppl in (1, 2, 3, 4):
            gpu_ctx.set_tuning(ppl, 4)
            assert api.compute_mse(gpu_ctx, P, ds)[0] == 0.0, ppl
    finally:
        gpu_ctx.set_summation(False)
        gpu_ctx.set_tuning(0, 0)
        P.close()
        ds.close()
