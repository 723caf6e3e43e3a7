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
