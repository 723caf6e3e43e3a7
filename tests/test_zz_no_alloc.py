"""rsr_eval_mse_noalloc: the allocation-free variant of rsr_eval_mse, through the C ABI.

Two calls in a row (the second on a smaller batch and dataset) each equal rsr_eval_mse bit
for bit in both summation modes; the context's device memory is unchanged by them;
rejected programs read NaN; big programs (device spill areas from the same scratch) are
covered; a request beyond the scratch is refused with RSR_E_INVALID_ARG; an unlowered
batch is refused.  The test's own output buffers come from the HIP runtime the library
itself loaded (device-memory plumbing of the test only)."""
import ctypes as C

import numpy as np
import pytest

from rsr_amd import api
from rsr_amd import workloads as W
from rsr_amd.isa import Op, Program

pytestmark = pytest.mark.gpu
f32 = np.float32


class DeviceBuffer:
    """n floats of device memory from the HIP runtime the library loaded."""

    def __init__(self, n):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.n = max(n, 1)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(4 * self.n)) == 0
        host = np.full(self.n, -1.0, f32)
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), C.c_size_t(4 * self.n), 1) == 0

    def read(self, n):
        host = np.zeros(self.n, f32)
        assert self.hip.hipMemcpy(host.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(4 * self.n), 2) == 0
        return host[:n]

    def close(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _run_noalloc(ctx, P, ds, n):
    out = DeviceBuffer(n)
    used0 = ctx.memory()
    api.eval_mse_noalloc(ctx, P, ds, out.ptr.value, allow_rejected=True)
    ctx.synchronize()
    used1 = ctx.memory()
    assert used0 == used1, (used0, used1)  # no device allocation by the call
    got = out.read(n)
    out.close()
    return got


def _bits(a):
    return np.asarray(a, f32).view(np.uint32)


@pytest.mark.parametrize("reference", [False, True])
def test_noalloc_equals_eval_mse_on_two_calls(gpu_ctx, reference):
    allp = W.gen_c2(16384, seed=2002)
    bad = Program([(Op.LoadVarF, 40)], [])  # feature index out of range: rejected
    progs1 = allp[:300] + [allp[5633], bad, allp[14173]] + allp[300:500]
    progs2 = allp[600:700] + [allp[5663]]
    X1, y1, t1 = W.data_c2(20000 if not reference else 3001, seed=1002)
    X2, y2, t2 = W.data_c2(1003, seed=1004)
    gpu_ctx.set_summation(reference)
    try:
        for progs, (X, y, t) in ((progs1, (X1, y1, t1)), (progs2, (X2, y2, t2))):
            print("noalloc: reference", reference, "programs", len(progs), "points", len(y), flush=True)
            ds = api.Dataset.new(gpu_ctx, X, y, t)
            P = api.Programs(gpu_ctx, progs)
            want = api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)  # also lowers the batch
            got = _run_noalloc(gpu_ctx, P, ds, len(progs))
            rej = P.status(ds) != 0
            assert np.isnan(got[rej]).all() and np.isnan(want[rej]).all()
            assert np.array_equal(_bits(got[~rej]), _bits(want[~rej]))
            P.close()
            ds.close()
    finally:
        gpu_ctx.set_summation(False)


def test_noalloc_refuses_what_does_not_fit(gpu_ctx):
    X, y, t = W.data_c1(100000, seed=5)
    ds = api.Dataset.new(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, W.gen_c1(512, seed=6))
    out = DeviceBuffer(512)
    with pytest.raises(api.RsrError) as e:  # not lowered for this dataset yet
        api.eval_mse_noalloc(gpu_ctx, P, ds, out.ptr.value)
    assert e.value.code == api.RSR_E_INVALID_ARG
    P.status(ds)
    gpu_ctx.set_summation(True)  # 512 x 100,000 per-point terms: 205 MB > 64 MiB
    try:
        with pytest.raises(api.RsrError) as e:
            api.eval_mse_noalloc(gpu_ctx, P, ds, out.ptr.value)
        assert e.value.code == api.RSR_E_INVALID_ARG
    finally:
        gpu_ctx.set_summation(False)
    api.eval_mse_noalloc(gpu_ctx, P, ds, out.ptr.value)  # Kahan partials fit
    gpu_ctx.synchronize()
    want = api.compute_mse(gpu_ctx, P, ds)
    assert np.array_equal(_bits(out.read(512)), _bits(want))
    out.close()
    P.close()
    ds.close()


def test_noalloc_dataset_without_points(gpu_ctx):
    """Over a dataset without points every evaluated program reads f32::MAX (the reference's
    0 / 0 is not finite), as from rsr_eval_mse; rejected programs read NaN."""
    X, y, t = W.data_c0(16, seed=81)
    progs = W.gen_c0(40, seed=82) + [Program([(Op.LoadVarF, 3), (Op.SinF, 0)], [])]
    ds = api.Dataset.new(gpu_ctx, X[:0], y[:0], t)
    P = api.Programs(gpu_ctx, progs)
    try:
        want = api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)
        got = _run_noalloc(gpu_ctx, P, ds, len(progs))
        assert np.array_equal(_bits(got), _bits(want))
        assert (want[:-1] == np.finfo(f32).max).all() and np.isnan(want[-1])
    finally:
        P.close()
        ds.close()
