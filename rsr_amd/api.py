"""ctypes binding of librsr_amd.so and the host-side evaluator / fitness interface.

The names follow the reference's evaluator and fitness surface
(src/engine/eval/evaluator.rs: compute_mse, compute_mse_with_gradient, eval_simd,
eval_simd_dual; src/engine/data/dataset.rs: Dataset::new / subset;
src/engine/optimize/mod.rs: optimize_individual_constants), batched over many
programs.  Everything here calls the HIP library; there is no CPU fallback: if the
library is missing or no GPU is present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import time
from typing import List, Optional, Sequence

import numpy as np

from .isa import Program, Scalar, VT, VT_SIZE, pack_programs, unpack_programs, feature_count

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librsr_amd.so")

RSR_OK = 0
RSR_E_INVALID_ARG = -1
LAUNCH_ONESHOT, LAUNCH_PERSISTENT, LAUNCH_AUTO = 0, 1, 2
RSR_E_PROGRAM = -5

_lib = None

_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)
_u16p = C.POINTER(C.c_uint16)

# name -> (restype, argtypes); the exported C ABI (include/rsr_amd.h)
SIGNATURES = {
    "rsr_error_string": (C.c_char_p, [C.c_int]),
    "rsr_last_error_message": (C.c_char_p, []),
    "rsr_version": (C.c_char_p, []),
    "rsr_context_create": (C.c_int, [C.c_int, C.c_uint64, _PP]),
    "rsr_context_destroy": (C.c_int, [_P]),
    "rsr_context_memory": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "rsr_context_set_tuning": (C.c_int, [_P, C.c_int, C.c_int]),
    "rsr_context_get_tuning": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "rsr_context_query_tuning": (C.c_int, [_P, _P, _P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "rsr_context_profile": (C.c_int, [_P, C.c_int]),
    "rsr_context_profile_read": (C.c_int, [_P, C.POINTER(C.c_double), _i64p]),
    "rsr_context_synchronize": (C.c_int, [_P]),
    "rsr_context_set_summation": (C.c_int, [_P, C.c_int]),
    "rsr_context_set_partials_limit": (C.c_int, [_P, C.c_uint64]),
    "rsr_context_set_graphs": (C.c_int, [_P, C.c_int]),
    "rsr_dataset_create": (C.c_int, [_P, _f32p, _f32p, C.c_int64, _u8p, C.c_int, C.c_int, _PP]),
    "rsr_dataset_create_aosoa": (C.c_int, [_P, _f32p, _f32p, C.c_int, C.c_int64, C.c_int64, _PP]),
    "rsr_dataset_subset": (C.c_int, [_P, _P, C.c_int64, _PP]),
    "rsr_dataset_info": (C.c_int, [_P, C.POINTER(C.c_int), _i64p, _i64p]),
    "rsr_dataset_stats": (C.c_int, [_P, _f32p]),
    "rsr_dataset_destroy": (C.c_int, [_P]),
    "rsr_programs_create": (C.c_int, [_P, C.c_int, _u8p, _u16p, _i64p, _u8p, _i64p, _f32p, _i64p, _PP]),
    "rsr_programs_count": (C.c_int, [_P, C.POINTER(C.c_int), _i64p]),
    "rsr_programs_get_constants": (C.c_int, [_P, _f32p]),
    "rsr_programs_set_constants": (C.c_int, [_P, _f32p]),
    "rsr_programs_status": (C.c_int, [_P, _P, _P, _i32p]),
    "rsr_programs_destroy": (C.c_int, [_P]),
    "rsr_eval_mse": (C.c_int, [_P, _P, _P, _f32p]),
    "rsr_eval_points": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _f32p]),
    "rsr_eval_mse_grad": (C.c_int, [_P, _P, _P, _f32p, _f32p]),
    "rsr_eval_points_dual": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _f32p, _f32p]),
}


class RsrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rsr error {code}: {msg}")
        self.code = code


def lib():
    """Load librsr_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run `python -m rsr_amd.build` (hipcc, gfx950)")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def _check(rc: int, allow=()):
    if rc != RSR_OK and rc not in allow:
        msg = lib().rsr_last_error_message().decode(errors="replace")
        raise RsrError(rc, msg)
    return rc


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class Context:
    """One GPU, one HIP stream, an HBM budget (0 = 90% of free memory)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        h = C.c_void_p()
        _check(lib().rsr_context_create(device, hbm_budget_bytes, C.byref(h)))
        self.h = h

    def close(self):
        """Destroy the context.  Raises, and keeps it, while one of its servers, datasets or
        program batches is alive (after collecting garbage that may still hold some)."""
        if self.h:
            rc = lib().rsr_context_destroy(self.h)
            if rc != RSR_OK:
                import gc
                gc.collect()
                rc = lib().rsr_context_destroy(self.h)
            _check(rc)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def memory(self):
        u, b = C.c_uint64(), C.c_uint64()
        _check(lib().rsr_context_memory(self.h, C.byref(u), C.byref(b)))
        return u.value, b.value

    def set_tuning(self, points_per_lane: int, programs_per_item: int):
        """Fix the evaluation tuning; (0, 0) returns to the per-workload (autotuned) tuning."""
        _check(lib().rsr_context_set_tuning(self.h, points_per_lane, programs_per_item))

    def autotune(self, progs: "Programs", ds: "Dataset"):
        """rsr_autotune: time every tuning on this batch and dataset, keep the fastest for the
        workload's shape; returns (points_per_lane, programs_per_item)."""
        fn = lib().rsr_autotune
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        a, b = C.c_int(), C.c_int()
        _check(fn(self.h, progs.h, ds.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def tuning(self):
        a, b = C.c_int(), C.c_int()
        _check(lib().rsr_context_get_tuning(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def query_tuning(self, progs: "Programs", ds: "Dataset"):
        """(points per lane, programs per item, persistent) an evaluation of progs over ds takes."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib().rsr_context_query_tuning(self.h, progs.h, ds.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, bool(c.value)

    def profile(self, enable: bool):
        _check(lib().rsr_context_profile(self.h, int(enable)))

    def profile_read(self):
        ms, n = C.c_double(), C.c_int64()
        _check(lib().rsr_context_profile_read(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_summation(self, reference_order: bool):
        """True (default): the reference's 8-lane order (bit-identical); False: pairwise + Kahan."""
        _check(lib().rsr_context_set_summation(self.h, 1 if reference_order else 0))

    def set_launch_mode(self, mode: int):
        """LAUNCH_ONESHOT: one workgroup per four work items; LAUNCH_PERSISTENT: launches with
        more work than the device holds at once run as a resident grid draining a per-XCD
        device work queue (smaller ones one-shot); LAUNCH_AUTO (default): the form the
        workload's autotuned tuning chose (one-shot without one).  The same bits either way."""
        _check(lib().rsr_context_set_launch_mode(self.h, int(mode)))

    def set_graphs(self, enable: bool):
        """compute_mse replays a captured hipGraph per (programs, dataset, settings)."""
        _check(lib().rsr_context_set_graphs(self.h, 1 if enable else 0))

    def set_partials_limit(self, nbytes: int):
        """Bound the Kahan-partials arena of one launch (larger evaluations split into tile ranges)."""
        _check(lib().rsr_context_set_partials_limit(self.h, int(nbytes)))

    def synchronize(self):
        _check(lib().rsr_context_synchronize(self.h))


class Dataset:
    """Device-resident dataset (Dataset::new, dataset.rs:37-169)."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.h = handle

    @staticmethod
    def new(ctx: Context, rows: np.ndarray, y: np.ndarray, feature_types: Sequence[int],
            normalize: bool = False) -> "Dataset":
        rows = np.ascontiguousarray(rows, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        n = y.shape[0]
        ft = np.asarray([int(t) for t in feature_types], np.uint8)
        F = feature_count(ft)
        if rows.size != n * F:
            raise ValueError(f"rows must be [{n}][{F}]")
        h = C.c_void_p()
        _check(lib().rsr_dataset_create(ctx.h, _ptr(rows, _f32p), _ptr(y, _f32p), n, _ptr(ft, _u8p), len(ft),
                                        int(normalize), C.byref(h)))
        return Dataset(ctx, h)

    @staticmethod
    def from_aosoa(ctx: Context, feature_flat: np.ndarray, target_batches: np.ndarray, num_features: int,
                   num_samples: int) -> "Dataset":
        ff = np.ascontiguousarray(feature_flat, np.float32)
        tb = np.ascontiguousarray(target_batches, np.float32)
        nb = tb.size // 8
        h = C.c_void_p()
        _check(lib().rsr_dataset_create_aosoa(ctx.h, _ptr(ff, _f32p), _ptr(tb, _f32p), num_features, nb,
                                              num_samples, C.byref(h)))
        return Dataset(ctx, h)

    @staticmethod
    def from_file(ctx: Context, path: str, feature_types: Sequence[int], fmt: str = "csv",
                  target_col: Optional[int] = None, normalize: bool = False) -> "Dataset":
        """Dataset::from_csv / from_json (dataset.rs:265-357) read straight into a device dataset."""
        ft = np.asarray([int(t) for t in feature_types] or [0], np.uint8)
        h = C.c_void_p()
        l = lib()
        l.rsr_dataset_from_file.restype = C.c_int
        l.rsr_dataset_from_file.argtypes = [_P, C.c_char_p, C.c_int, _u8p, C.c_int, C.c_int, C.c_int, _PP]
        _check(l.rsr_dataset_from_file(ctx.h, os.fsencode(path), 1 if fmt == "json" else 0, _ptr(ft, _u8p),
                                       len(feature_types), -1 if target_col is None else int(target_col),
                                       1 if normalize else 0, C.byref(h)))
        return Dataset(ctx, h)

    def subset(self, target_samples: int) -> "Dataset":
        h = C.c_void_p()
        _check(lib().rsr_dataset_subset(self.ctx.h, self.h, target_samples, C.byref(h)))
        return Dataset(self.ctx, h)

    def info(self):
        f, s, b = C.c_int(), C.c_int64(), C.c_int64()
        _check(lib().rsr_dataset_info(self.h, C.byref(f), C.byref(s), C.byref(b)))
        return f.value, s.value, b.value

    @property
    def num_features(self):
        return self.info()[0]

    @property
    def num_samples(self):
        return self.info()[1]

    @property
    def num_batches(self):
        return self.info()[2]

    def stats(self):
        F = self.num_features
        out = np.zeros(2 * F + 2, np.float32)
        _check(lib().rsr_dataset_stats(self.h, _ptr(out, _f32p)))
        return out

    def close(self):
        if self.h:
            lib().rsr_dataset_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Programs:
    """A device-resident batch of programs."""

    def __init__(self, ctx: Context, progs: Sequence[Program], packed=None):
        """packed: the pack_programs arrays of progs, when the caller already has them
        (progs is then only counted)."""
        self.ctx = ctx
        self.n = len(progs)
        ops, args, co, ct, cof, cv, vo = packed if packed is not None else pack_programs(progs)
        self._keep = (ops, args, co, ct, cof, cv, vo)
        h = C.c_void_p()
        _check(lib().rsr_programs_create(ctx.h, self.n, _ptr(ops, _u8p), _ptr(args, _u16p), _ptr(co, _i64p),
                                         _ptr(ct, _u8p), _ptr(cof, _i64p), _ptr(cv, _f32p), _ptr(vo, _i64p),
                                         C.byref(h)))
        self.h = h
        self.n_values = int(vo[-1])

    def get_constants(self) -> np.ndarray:
        out = np.zeros(max(self.n_values, 1), np.float32)
        _check(lib().rsr_programs_get_constants(self.h, _ptr(out, _f32p)))
        return out[: self.n_values]

    def set_constants(self, values: np.ndarray):
        v = np.ascontiguousarray(values, np.float32)
        if v.size < self.n_values:
            raise ValueError("too few constant values")
        _check(lib().rsr_programs_set_constants(self.h, _ptr(v, _f32p)))

    def status(self, ds: Dataset) -> np.ndarray:
        out = np.zeros(max(self.n, 1), np.int32)
        _check(lib().rsr_programs_status(self.ctx.h, self.h, ds.h, _ptr(out, _i32p)))
        return out[: self.n]

    def close(self):
        if self.h:
            lib().rsr_programs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compute_mse(ctx: Context, progs: Programs, ds: Dataset, allow_rejected: bool = False) -> np.ndarray:
    """Batched compute_mse (evaluator.rs:153-174)."""
    out = np.zeros(max(progs.n, 1), np.float32)
    _check(lib().rsr_eval_mse(ctx.h, progs.h, ds.h, _ptr(out, _f32p)), (RSR_E_PROGRAM,) if allow_rejected else ())
    return out[: progs.n]


def eval_mse_noalloc(ctx: Context, progs: Programs, ds: Dataset, mse_out_device: int, allow_rejected: bool = False):
    """rsr_eval_mse_noalloc: compute_mse into a device buffer (address mse_out_device, n floats)
    with no device allocation; asynchronous on the context's stream (ctx.synchronize())."""
    fn = lib().rsr_eval_mse_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    _check(fn(ctx.h, progs.h, ds.h, C.c_void_p(mse_out_device)), (RSR_E_PROGRAM,) if allow_rejected else ())


def eval_points(ctx: Context, progs: Programs, ds: Dataset, first: int = 0, count: Optional[int] = None) -> np.ndarray:
    """Per-point values (eval_simd over all batches): [count][num_batches*8]."""
    count = progs.n - first if count is None else count
    nb = ds.num_batches
    out = np.zeros((max(count, 1), nb * 8), np.float32)
    _check(lib().rsr_eval_points(ctx.h, progs.h, first, count, ds.h, _ptr(out, _f32p)))
    return out[:count]


def compute_mse_with_gradient(ctx: Context, progs: Programs, ds: Dataset, allow_rejected: bool = False):
    """Batched compute_mse_with_gradient (evaluator.rs:177-224): (mse[n], grad[n][32])."""
    mse = np.zeros(max(progs.n, 1), np.float32)
    grad = np.zeros((max(progs.n, 1), 32), np.float32)
    _check(lib().rsr_eval_mse_grad(ctx.h, progs.h, ds.h, _ptr(mse, _f32p), _ptr(grad, _f32p)),
           (RSR_E_PROGRAM,) if allow_rejected else ())
    return mse[: progs.n], grad[: progs.n]


def eval_points_dual(ctx: Context, progs: Programs, ds: Dataset, program: int, active_param: int):
    """eval_simd_dual over all batches: (values, tangents) per point."""
    n = ds.num_batches * 8
    v = np.zeros(n, np.float32)
    g = np.zeros(n, np.float32)
    _check(lib().rsr_eval_points_dual(ctx.h, progs.h, program, active_param, ds.h, _ptr(v, _f32p), _ptr(g, _f32p)))
    return v, g


def lower_code(prog: Program, num_features: int, dual: bool = False):
    """Lowered micro-code of one program (test hook, no GPU): (status, words[n][4] uint32)."""
    l = lib()
    fn = l.rsr_lower_code
    fn.restype = C.c_int
    fn.argtypes = [_u8p, _u16p, C.c_int, _u8p, _f32p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint32), C.c_int32,
                   _i32p]
    ops, args, co, ct, cof, cv, vo = pack_programs([prog])
    ops = ops if ops.size else np.zeros(1, np.uint8)
    args = args if args.size else np.zeros(1, np.uint16)
    ct = ct if ct.size else np.zeros(1, np.uint8)
    cv = cv if cv.size else np.zeros(1, np.float32)
    cap = 4 * max(len(prog.code), 1) + 64
    words = np.zeros((cap, 4), np.uint32)
    n = C.c_int32()
    st = fn(_ptr(ops, _u8p), _ptr(args, _u16p), len(prog.code), _ptr(ct, _u8p), _ptr(cv, _f32p),
            len(prog.constants), num_features, int(dual), words.ctypes.data_as(C.POINTER(C.c_uint32)), cap,
            C.byref(n))
    return st, words[: min(n.value, cap)]


def lower_check(prog: Program, rows_soa: np.ndarray, dual: bool = False, active_param: int = 0):
    """Lower one program and run its micro-code on the host (lowering self-check, no GPU).

    rows_soa: [F][n] feature-major.  Returns (status, values[n], tangents[n], n_uops, max_reg).
    """
    l = lib()
    fn = l.rsr_lower_check
    fn.restype = C.c_int
    fn.argtypes = [_u8p, _u16p, C.c_int, _u8p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int, _f32p, C.c_int64,
                   _f32p, _f32p, _i32p, _i32p]
    ops, args, co, ct, cof, cv, vo = pack_programs([prog])
    ops = ops if ops.size else np.zeros(1, np.uint8)
    args = args if args.size else np.zeros(1, np.uint16)
    ct = ct if ct.size else np.zeros(1, np.uint8)
    cv = cv if cv.size else np.zeros(1, np.float32)
    feat = np.ascontiguousarray(rows_soa, np.float32)
    F, n = (feat.shape[0], feat.shape[1]) if feat.ndim == 2 else (0, 0)
    if feat.size == 0:
        feat = np.zeros(1, np.float32)
    vals = np.zeros(max(n, 1), np.float32)
    tans = np.zeros(max(n, 1), np.float32)
    nu, mr = C.c_int32(), C.c_int32()
    st = fn(_ptr(ops, _u8p), _ptr(args, _u16p), len(prog.code), _ptr(ct, _u8p), _ptr(cv, _f32p),
            len(prog.constants), F, int(dual), active_param, _ptr(feat, _f32p), n, _ptr(vals, _f32p),
            _ptr(tans, _f32p), C.byref(nu), C.byref(mr))
    return st, vals[:n], tans[:n], nu.value, mr.value


# ---------------------------------------------------------------------------------
# constant fitting (src/engine/optimize/)
# ---------------------------------------------------------------------------------
LBFGS_REFERENCE = 0
LBFGS_FP64 = 1


def _bind_fit():
    l = lib()
    l.rsr_fit_lbfgs.restype = C.c_int
    l.rsr_fit_lbfgs.argtypes = [_P, _P, _P, C.c_int, C.c_int, C.c_int, _u8p, _f32p]
    l.rsr_fit_cmaes.restype = C.c_int
    l.rsr_fit_cmaes.argtypes = [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_uint64, _f32p, _u8p, _f32p, _i32p]
    l.rsr_cmaes_normal.restype = C.c_float
    l.rsr_cmaes_normal.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
    return l


def _mask(progs: Programs, mask):
    if mask is None:
        return None
    m = np.ascontiguousarray(np.asarray(mask, np.uint8))
    return m


def fit_lbfgs(ctx: Context, progs: Programs, ds: Dataset, iterations: int = 15, history: int = 6,
              mode: int = LBFGS_REFERENCE, mask=None) -> np.ndarray:
    """Batched run_lbfgs (lbfgs.rs:9-165); constants are written back into `progs`."""
    l = _bind_fit()
    fit = np.zeros(max(progs.n, 1), np.float32)
    m = _mask(progs, mask)
    _check(l.rsr_fit_lbfgs(ctx.h, progs.h, ds.h, iterations, history, mode,
                           _ptr(m, _u8p) if m is not None else None, _ptr(fit, _f32p)))
    return fit[: progs.n]


def fit_cmaes(ctx: Context, progs: Programs, ds: Dataset, iterations: int, lam: int = 16, mu: int = 8,
              seed: int = 0, normals: Optional[np.ndarray] = None, mask=None):
    """Batched run_cma_es (cmaes.rs:19-192): returns (fitness[n], iterations_run[n])."""
    l = _bind_fit()
    fit = np.zeros(max(progs.n, 1), np.float32)
    its = np.zeros(max(progs.n, 1), np.int32)
    m = _mask(progs, mask)
    nv = None if normals is None else np.ascontiguousarray(normals, np.float32)
    _check(l.rsr_fit_cmaes(ctx.h, progs.h, ds.h, iterations, lam, mu, seed,
                           _ptr(nv, _f32p) if nv is not None else None,
                           _ptr(m, _u8p) if m is not None else None, _ptr(fit, _f32p), _ptr(its, _i32p)))
    return fit[: progs.n], its[: progs.n]


def fit_nelder_mead(ctx: Context, progs: Programs, ds: Dataset, iterations: int = 100, mask=None):
    """Batched run_nelder_mead (nelder_mead.rs:9-165): returns (fitness[n], iterations_run[n])."""
    l = lib()
    l.rsr_fit_nelder_mead.restype = C.c_int
    l.rsr_fit_nelder_mead.argtypes = [_P, _P, _P, C.c_int, _u8p, _f32p, _i32p]
    fit = np.zeros(max(progs.n, 1), np.float32)
    its = np.zeros(max(progs.n, 1), np.int32)
    m = _mask(progs, mask)
    _check(l.rsr_fit_nelder_mead(ctx.h, progs.h, ds.h, iterations, _ptr(m, _u8p) if m is not None else None,
                                 _ptr(fit, _f32p), _ptr(its, _i32p)))
    return fit[: progs.n], its[: progs.n]


FIT_NONE, FIT_LBFGS, FIT_CMAES, FIT_NELDER_MEAD = 0, 1, 2, 3


def optimize_constants(ctx: Context, progs: Programs, ds: Dataset, iterations: int = 100, seed: int = 0,
                       threshold: float = 0.05, mask=None):
    """Individual::optimize_constants (individual.rs:65-75) + calculate_mse, batched:
    returns (mse_after[n], method[n]) with method one of FIT_*."""
    l = lib()
    l.rsr_optimize_constants.restype = C.c_int
    l.rsr_optimize_constants.argtypes = [_P, _P, _P, C.c_int, C.c_uint64, C.c_float, _u8p, _f32p, _u8p]
    mse = np.zeros(max(progs.n, 1), np.float32)
    meth = np.zeros(max(progs.n, 1), np.uint8)
    m = _mask(progs, mask)
    _check(l.rsr_optimize_constants(ctx.h, progs.h, ds.h, iterations, seed, threshold,
                                    _ptr(m, _u8p) if m is not None else None, _ptr(mse, _f32p), _ptr(meth, _u8p)))
    return mse[: progs.n], meth[: progs.n]


def optimize_individuals(ctx: Context, programs: Sequence[Program], ds: Dataset, iterations: int = 100,
                         seed: int = 0, threshold: float = 0.05, mask=None, parsimony: float = 5e-6):
    """Individual::optimize_constants (individual.rs:65-75) as its callers use it
    (island.rs:230-255, engine.rs:95-98): fit and snap the constants on the GPU, simplify
    and recompile every optimized program (simplify_ast), then calculate_mse, complexity
    and fitness of the result.  Programs with mask[i] == 0 come back unchanged with
    their plain MSE.  Returns (programs[n], mse[n], complexity[n], fitness[n], method[n])."""
    n = len(programs)
    P = Programs(ctx, programs)
    try:
        _, meth = optimize_constants(ctx, P, ds, iterations, seed, threshold, mask)
        vals = P.get_constants()
    finally:
        P.close()
    todo = np.ones(n, bool) if mask is None else np.asarray(mask, np.uint8)[:n] != 0
    fitted, off = [], 0
    for p in programs:
        cs = []
        for c in p.constants:
            k = VT_SIZE[c.vt] or 1
            cs.append(Scalar(c.vt, tuple(float(v) for v in vals[off:off + k])))
            off += k
        fitted.append(Program(list(p.code), cs))
    idx = [i for i in range(n) if todo[i]]
    simplified = simplify_programs([fitted[i] for i in idx])
    out = list(programs)
    for i, q in zip(idx, simplified):
        out[i] = q
    Q = Programs(ctx, out)
    try:
        mse = compute_mse(ctx, Q, ds, allow_rejected=True)
    finally:
        Q.close()
    comp = complexity(out)
    return out, mse, comp, fitness(mse, comp, parsimony), meth


# ---------------------------------------------------------------------------------
# scoring and selection (host functions; selection.rs, individual.rs:108-110, island.rs:241-255)
# ---------------------------------------------------------------------------------
def complexity(progs: Sequence[Program]) -> np.ndarray:
    """Individual::complexity of every program (sum of node weights)."""
    l = lib()
    l.rsr_complexity.restype = C.c_int
    l.rsr_complexity.argtypes = [C.c_int, _u8p, _i64p, _i32p]
    ops, _, co, *_ = pack_programs(progs)
    out = np.zeros(max(len(progs), 1), np.int32)
    _check(l.rsr_complexity(len(progs), _ptr(ops, _u8p), _ptr(co, _i64p), _ptr(out, _i32p)))
    return out[: len(progs)]


def fitness(mse, complexity_, parsimony: float = 5e-6) -> np.ndarray:
    l = lib()
    l.rsr_fitness.restype = C.c_int
    l.rsr_fitness.argtypes = [C.c_int, _f32p, _i32p, C.c_float, _f32p]
    m = np.ascontiguousarray(mse, np.float32)
    c = np.ascontiguousarray(complexity_, np.int32)
    out = np.zeros(max(len(m), 1), np.float32)
    _check(l.rsr_fitness(len(m), _ptr(m, _f32p), _ptr(c, _i32p), parsimony, _ptr(out, _f32p)))
    return out[: len(m)]


def read_dataset_file(path: str, feature_types: Sequence[int], fmt: str = "csv", target_col: Optional[int] = None):
    """The rows [n][F] and targets [n] that Dataset::from_csv / from_json read (host, no GPU)."""
    l = lib()
    ft = np.asarray([int(t) for t in feature_types] or [0], np.uint8)
    F = feature_count(feature_types)
    n = C.c_int64()
    if fmt == "json":
        fn = l.rsr_read_json
        fn.restype = C.c_int
        fn.argtypes = [C.c_char_p, _u8p, C.c_int, _f32p, _f32p, C.c_int64, _i64p]
        call = lambda r, y, cap: fn(os.fsencode(path), _ptr(ft, _u8p), len(feature_types), r, y, cap, C.byref(n))
    else:
        fn = l.rsr_read_csv
        fn.restype = C.c_int
        fn.argtypes = [C.c_char_p, _u8p, C.c_int, C.c_int, _f32p, _f32p, C.c_int64, _i64p]
        tc = -1 if target_col is None else int(target_col)
        call = lambda r, y, cap: fn(os.fsencode(path), _ptr(ft, _u8p), len(feature_types), tc, r, y, cap, C.byref(n))
    _check(call(None, None, 0))
    rows = np.zeros(max(n.value * F, 1), np.float32)
    y = np.zeros(max(n.value, 1), np.float32)
    _check(call(_ptr(rows, _f32p), _ptr(y, _f32p), n.value))
    return rows[: n.value * F].reshape(n.value, F), y[: n.value]


def simplify_programs(progs: Sequence[Program]) -> List[Program]:
    """Individual::simplify + Program::from_nodes of every program (host, no GPU)."""
    l = lib()
    fn = l.rsr_simplify_programs
    fn.restype = C.c_int
    fn.argtypes = [C.c_int, _u8p, _u16p, _i64p, _u8p, _i64p, _f32p, _i64p, C.c_int64, C.c_int64, C.c_int64,
                   _u8p, _u16p, _i64p, _u8p, _i64p, _f32p, _i64p]
    n = len(progs)
    ops, args, co, ct, cof, cv, vo = pack_programs(progs)
    cap = max(int(co[-1]), 1)
    ops_o = np.zeros(cap, np.uint8)
    args_o = np.zeros(cap, np.uint16)
    ct_o = np.zeros(cap, np.uint8)
    cv_o = np.zeros(9 * cap, np.float32)
    co_o = np.zeros(n + 1, np.int64)
    cof_o = np.zeros(n + 1, np.int64)
    vo_o = np.zeros(n + 1, np.int64)
    pad = lambda a, t: a if a.size else np.zeros(1, t)
    _check(fn(n, _ptr(pad(ops, np.uint8), _u8p), _ptr(pad(args, np.uint16), _u16p), _ptr(co, _i64p),
              _ptr(pad(ct, np.uint8), _u8p), _ptr(cof, _i64p), _ptr(pad(cv, np.float32), _f32p), _ptr(vo, _i64p),
              cap, cap, 9 * cap, _ptr(ops_o, _u8p), _ptr(args_o, _u16p), _ptr(co_o, _i64p), _ptr(ct_o, _u8p),
              _ptr(cof_o, _i64p), _ptr(cv_o, _f32p), _ptr(vo_o, _i64p)))
    return unpack_programs(ops_o, args_o, co_o, ct_o, cof_o, cv_o, vo_o)


def islands_evaluate(ctx: Context, dataset: Dataset, mini_batch: Dataset, buffers: Sequence[Sequence[Program]],
                     best_fitness, best_mini_mse, draws: Sequence[np.ndarray], opt_prob: float = 0.01,
                     opt_iterations: int = 100, parsimony: float = 5e-6, threshold: float = 0.05, seed: int = 0,
                     packed=None, unpack: bool = True):
    """Island::evaluate_buffer of every island (rsr_islands_evaluate; defaults of config.rs:40-66).

    buffers[g]: island g's next-generation buffer; best_fitness[g] / best_mini_mse[g]: its
    best individual's fitness and mini-batch MSE; draws[g]: its uniform f32 stream.
    Returns dict(buffers, fitness, mse, optimized (per island lists / arrays),
    best_index[g] (-1: unchanged), draws_used[g]).  packed: pack_programs of the
    concatenated buffers, when the caller has it already; unpack=False leaves the output
    programs in the wire format (buffers: None, packed_out: the arrays), which is what a
    compiled caller gets -- the Python objects cost more to build than the step itself."""
    l = lib()
    fn = l.rsr_islands_evaluate
    fn.restype = C.c_int
    fn.argtypes = ([_P, _P, _P, C.c_int, _i64p, _u8p, _u16p, _i64p, _u8p, _i64p, _f32p, _i64p, _f32p, _f32p, _f32p,
                    _i64p, C.c_float, C.c_int, C.c_float, C.c_float, C.c_uint64, C.c_int64, C.c_int64, C.c_int64,
                    _u8p, _u16p, _i64p, _u8p, _i64p, _f32p, _i64p, _f32p, _f32p, _u8p, _i32p, _i64p])
    G = len(buffers)
    flat = [p for b in buffers for p in b]
    n = len(flat)
    isl = np.zeros(G + 1, np.int64)
    isl[1:] = np.cumsum([len(b) for b in buffers])
    ops, args, co, ct, cof, cv, vo = packed if packed is not None else pack_programs(flat)
    dr = [np.ascontiguousarray(d, np.float32) for d in draws]
    doff = np.zeros(G + 1, np.int64)
    doff[1:] = np.cumsum([d.size for d in dr])
    dall = np.concatenate(dr) if dr and doff[-1] else np.zeros(1, np.float32)
    bf = np.ascontiguousarray(best_fitness, np.float32)
    bm = np.ascontiguousarray(best_mini_mse, np.float32)
    cap = max(int(co[-1]), 1)
    ops_o, args_o = np.zeros(cap, np.uint8), np.zeros(cap, np.uint16)
    ct_o, cv_o = np.zeros(cap, np.uint8), np.zeros(9 * cap, np.float32)
    co_o, cof_o, vo_o = np.zeros(n + 1, np.int64), np.zeros(n + 1, np.int64), np.zeros(n + 1, np.int64)
    fit, mse = np.zeros(max(n, 1), np.float32), np.zeros(max(n, 1), np.float32)
    opt = np.zeros(max(n, 1), np.uint8)
    besti, used = np.zeros(max(G, 1), np.int32), np.zeros(max(G, 1), np.int64)
    pad = lambda a, t: a if a.size else np.zeros(1, t)
    _check(fn(ctx.h, dataset.h, mini_batch.h, G, _ptr(isl, _i64p), _ptr(pad(ops, np.uint8), _u8p),
              _ptr(pad(args, np.uint16), _u16p), _ptr(co, _i64p), _ptr(pad(ct, np.uint8), _u8p), _ptr(cof, _i64p),
              _ptr(pad(cv, np.float32), _f32p), _ptr(vo, _i64p), _ptr(pad(bf, np.float32), _f32p),
              _ptr(pad(bm, np.float32), _f32p), _ptr(dall, _f32p), _ptr(doff, _i64p), opt_prob, opt_iterations,
              parsimony, threshold, seed, cap, cap, 9 * cap, _ptr(ops_o, _u8p), _ptr(args_o, _u16p),
              _ptr(co_o, _i64p), _ptr(ct_o, _u8p), _ptr(cof_o, _i64p), _ptr(cv_o, _f32p), _ptr(vo_o, _i64p),
              _ptr(fit, _f32p), _ptr(mse, _f32p), _ptr(opt, _u8p), _ptr(besti, _i32p), _ptr(used, _i64p)))
    split = lambda a: [a[isl[g]:isl[g + 1]] for g in range(G)]
    out = {"fitness": split(fit[:n]), "mse": split(mse[:n]), "optimized": split(opt[:n].astype(bool)),
           "best_index": besti[:G], "draws_used": used[:G]}
    if unpack:
        out["buffers"] = split(unpack_programs(ops_o, args_o, co_o, ct_o, cof_o, cv_o, vo_o))
    else:
        out["buffers"] = None
        out["packed_out"] = (ops_o, args_o, co_o, ct_o, cof_o, cv_o, vo_o)
    return out


def assign_rank_and_crowding(fitness_, age):
    l = lib()
    l.rsr_assign_rank_and_crowding.restype = C.c_int
    l.rsr_assign_rank_and_crowding.argtypes = [C.c_int, _f32p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), _f32p]
    f = np.ascontiguousarray(fitness_, np.float32)
    a = np.ascontiguousarray(age, np.uint64)
    n = len(f)
    r = np.zeros(max(n, 1), np.uint32)
    c = np.zeros(max(n, 1), np.float32)
    _check(l.rsr_assign_rank_and_crowding(n, _ptr(f, _f32p), a.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          r.ctypes.data_as(C.POINTER(C.c_uint32)), _ptr(c, _f32p)))
    return r[:n], c[:n]


def tournament(draws, rank, crowding) -> int:
    l = lib()
    l.rsr_tournament.restype = C.c_int
    l.rsr_tournament.argtypes = [C.c_int, C.POINTER(C.c_uint32), _f32p, C.c_int, C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32)]
    r = np.ascontiguousarray(rank, np.uint32)
    c = np.ascontiguousarray(crowding, np.float32)
    d = np.ascontiguousarray(draws, np.uint32)
    w = C.c_uint32()
    _check(l.rsr_tournament(len(r), r.ctypes.data_as(C.POINTER(C.c_uint32)), _ptr(c, _f32p), len(d),
                            d.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(w)))
    return int(w.value)


def cmaes_normal(seed: int, program: int, iteration: int, candidate: int, dim: int) -> float:
    """The normal the CMA-ES sampler draws (host restatement of the device function)."""
    return _bind_fit().rsr_cmaes_normal(seed, program, iteration, candidate, dim)


# ---------------------------------------------------------------------------------
# serving: submit / poll / wait (rsr_server_*)
# ---------------------------------------------------------------------------------
REQ_EVAL, REQ_GRAD, REQ_LBFGS, REQ_CMAES, REQ_NELDER_MEAD, REQ_OPTIMIZE = 0, 1, 2, 3, 4, 5
RSR_E_NOT_READY = -10
RSR_E_TIMEOUT = -9


class _Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("programs", C.c_void_p), ("dataset", C.c_void_p), ("mask", C.c_void_p),
                ("iterations", C.c_int), ("history", C.c_int), ("mode", C.c_int), ("lam", C.c_int),
                ("mu", C.c_int), ("seed", C.c_uint64), ("out", C.c_void_p), ("out2", C.c_void_p),
                ("iters_out", C.c_void_p)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("queued", C.c_uint64), ("inflight", C.c_uint64), ("max_inflight_seen", C.c_uint64),
                ("ticks", C.c_uint64), ("node_evals", C.c_uint64), ("inflight_bytes", C.c_uint64),
                ("busy_seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _bind_server():
    l = lib()
    if getattr(l, "_server_bound", False):
        return l
    l.rsr_server_create.restype = C.c_int
    l.rsr_server_create.argtypes = [_P, C.c_int, C.c_uint64, C.c_int, _PP]
    l.rsr_server_submit.restype = C.c_int
    l.rsr_server_submit.argtypes = [_P, C.POINTER(_Request), C.POINTER(C.c_uint64)]
    l.rsr_server_poll.restype = C.c_int
    l.rsr_server_poll.argtypes = [_P, C.c_uint64, C.POINTER(C.c_int)]
    l.rsr_server_wait.restype = C.c_int
    l.rsr_server_wait.argtypes = [_P, C.c_uint64, C.c_double, C.POINTER(C.c_int)]
    l.rsr_server_timing.restype = C.c_int
    l.rsr_server_timing.argtypes = [_P, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double), _i32p]
    l.rsr_server_release.restype = C.c_int
    l.rsr_server_release.argtypes = [_P, C.c_uint64]
    l.rsr_server_get_stats.restype = C.c_int
    l.rsr_server_get_stats.argtypes = [_P, C.POINTER(ServerStats)]
    l.rsr_server_destroy.restype = C.c_int
    l.rsr_server_destroy.argtypes = [_P]
    l._server_bound = True
    return l


class Ticket:
    """A submitted request: its id, its output arrays and the objects it keeps alive."""

    def __init__(self, server: "Server", tid: int, kind: int, n: int, keep, out, out2=None, iters=None):
        self.server, self.id, self.kind, self.n = server, tid, kind, n
        self._keep = keep
        self.out, self.out2, self.iters = out, out2, iters

    def done(self) -> bool:
        return self.server.poll(self)

    def wait(self, timeout: float = -1.0):
        """Block until complete; returns the outputs (raises RsrError if the request failed)."""
        self.server.wait(self, timeout)
        if self.kind == REQ_EVAL:
            return self.out[: self.n]
        if self.kind == REQ_GRAD:
            return self.out[: self.n], self.out2[: self.n]
        if self.kind in (REQ_CMAES, REQ_NELDER_MEAD):
            return self.out[: self.n], self.iters[: self.n]
        return self.out[: self.n]

    def timing(self):
        """(seconds queued, seconds running, ticks)."""
        q, r, t = C.c_double(), C.c_double(), C.c_int32()
        _check(_bind_server().rsr_server_timing(self.server.h, self.id, C.byref(q), C.byref(r), C.byref(t)))
        return q.value, r.value, t.value


class Server:
    """The request loop (rsr_server_*): continuous batching of evaluation and constant-fit
    requests over device-resident datasets and populations."""

    def __init__(self, ctx: Context, max_inflight: int = 64, budget_bytes: int = 0, max_queue: int = 4096):
        l = _bind_server()
        h = C.c_void_p()
        _check(l.rsr_server_create(ctx.h, max_inflight, budget_bytes, max_queue, C.byref(h)))
        self.h = h
        self.ctx = ctx

    def _submit(self, kind, progs: Programs, ds: Dataset, mask=None, iterations=0, history=6,
                mode=0, lam=16, mu=8, seed=0):
        n = progs.n
        out = np.zeros(max(n, 1), np.float32)
        out2 = np.zeros((max(n, 1), 32), np.float32) if kind == REQ_GRAD else None
        iters = np.zeros(max(n, 1), np.int32) if kind in (REQ_CMAES, REQ_NELDER_MEAD) else None
        m = _mask(progs, mask)
        r = _Request(kind, progs.h.value if isinstance(progs.h, C.c_void_p) else progs.h,
                     ds.h.value if isinstance(ds.h, C.c_void_p) else ds.h,
                     m.ctypes.data if m is not None else None, iterations, history, mode, lam, mu, seed,
                     out.ctypes.data, out2.ctypes.data if out2 is not None else None,
                     iters.ctypes.data if iters is not None else None)
        tid = C.c_uint64()
        _check(_bind_server().rsr_server_submit(self.h, C.byref(r), C.byref(tid)))
        return Ticket(self, tid.value, kind, n, (progs, ds, m), out, out2, iters)

    def submit_eval(self, progs: Programs, ds: Dataset, mask=None) -> Ticket:
        return self._submit(REQ_EVAL, progs, ds, mask)

    def submit_grad(self, progs: Programs, ds: Dataset, mask=None) -> Ticket:
        return self._submit(REQ_GRAD, progs, ds, mask)

    def submit_lbfgs(self, progs: Programs, ds: Dataset, iterations: int = 15, history: int = 6,
                     mode: int = LBFGS_REFERENCE, mask=None) -> Ticket:
        return self._submit(REQ_LBFGS, progs, ds, mask, iterations=iterations, history=history, mode=mode)

    def submit_cmaes(self, progs: Programs, ds: Dataset, iterations: int, lam: int = 16, mu: int = 8,
                     seed: int = 0, mask=None) -> Ticket:
        return self._submit(REQ_CMAES, progs, ds, mask, iterations=iterations, lam=lam, mu=mu, seed=seed)

    def submit_nelder_mead(self, progs: Programs, ds: Dataset, iterations: int = 100, mask=None) -> Ticket:
        return self._submit(REQ_NELDER_MEAD, progs, ds, mask, iterations=iterations)

    def submit_optimize(self, progs: Programs, ds: Dataset, iterations: int = 100, seed: int = 0,
                        mask=None) -> Ticket:
        """optimize_constants + calculate_mse (threshold 0.05): the ticket yields mse[n]."""
        return self._submit(REQ_OPTIMIZE, progs, ds, mask, iterations=iterations, seed=seed)

    def poll(self, t: Ticket) -> bool:
        st = C.c_int()
        rc = _bind_server().rsr_server_poll(self.h, t.id, C.byref(st))
        if rc == RSR_E_NOT_READY:
            return False
        _check(rc)
        _check(st.value)
        return True

    def wait(self, t: Ticket, timeout: float = -1.0):
        st = C.c_int()
        rc = _bind_server().rsr_server_wait(self.h, t.id, timeout, C.byref(st))
        _check(rc)
        _check(st.value)

    def stats(self) -> dict:
        s = ServerStats()
        _check(_bind_server().rsr_server_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def close(self):
        if getattr(self, "h", None):
            _bind_server().rsr_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- device-side variation and the search loop ----------------------------------
VARY_GENERATE, VARY_CROSSOVER, VARY_POINT, VARY_CONSTANT, VARY_SUBTREE, VARY_MUTATE = range(6)


class VaryTask(C.Structure):
    _fields_ = [("kind", C.c_uint8), ("type", C.c_uint8), ("depth", C.c_uint8), ("reserved", C.c_uint8),
                ("a", C.c_int32), ("b", C.c_int32), ("max_size", C.c_int32), ("seed", C.c_uint64)]


def vary(ctx: Context, tasks: Sequence[tuple], parents: Sequence[Program], variables: Sequence[tuple],
         allowed_ops: Sequence[int], packed=None, unpack: bool = True, timing: Optional[dict] = None):
    """rsr_vary: one offspring per task on the device.

    tasks: (kind, seed, a=-1, b=-1, type=0, depth=0, max_size=32) tuples; variables: (ValueType,
    feature index) pairs; allowed_ops: RefOp codes.  Returns (offspring, applied kinds).
    packed: pack_programs(parents), when the caller has it; unpack=False returns the
    offspring in the wire format (the seven arrays) instead of Program objects; timing, a
    dict, receives "call_ms", the time of the rsr_vary call alone."""
    l = lib()
    fn = l.rsr_vary
    fn.restype = C.c_int
    fn.argtypes = [_P, C.c_int, C.POINTER(VaryTask), C.c_int, _u8p, _u16p, _i64p, _u8p, _i64p, _f32p, _i64p,
                   C.c_int, _u8p, _u8p, C.c_int, _u8p, C.c_int64, C.c_int64, C.c_int64, _u8p, _u16p, _i64p, _u8p,
                   _i64p, _f32p, _i64p, _u8p]
    n = len(tasks)
    arr = (VaryTask * max(n, 1))()
    for i, t in enumerate(tasks):
        kind, seed = int(t[0]), int(t[1])
        a = int(t[2]) if len(t) > 2 else -1
        b = int(t[3]) if len(t) > 3 else -1
        vt = int(t[4]) if len(t) > 4 else 0
        depth = int(t[5]) if len(t) > 5 else 0
        ms = int(t[6]) if len(t) > 6 else 32
        arr[i] = VaryTask(kind, vt, depth, 0, a, b, ms, seed)
    ops, args, co, ct, cof, cv, vo = packed if packed is not None else pack_programs(list(parents))
    vt_ = np.array([v[0] for v in variables] or [0], np.uint8)
    vi_ = np.array([v[1] for v in variables] or [0], np.uint8)
    al = np.array(list(allowed_ops) or [0], np.uint8)
    cap = 384 * max(n, 1)
    # output capacities (written only up to what the offspring use)
    ops_o, args_o = np.empty(cap, np.uint8), np.empty(cap, np.uint16)
    ct_o, cv_o = np.empty(cap, np.uint8), np.empty(9 * cap, np.float32)
    co_o, cof_o, vo_o = np.zeros(n + 1, np.int64), np.zeros(n + 1, np.int64), np.zeros(n + 1, np.int64)
    kinds = np.zeros(max(n, 1), np.uint8)
    pad = lambda x, t: x if x.size else np.zeros(1, t)
    t0 = time.perf_counter()
    _check(fn(ctx.h, n, arr, len(parents), _ptr(pad(ops, np.uint8), _u8p), _ptr(pad(args, np.uint16), _u16p),
              _ptr(co, _i64p), _ptr(pad(ct, np.uint8), _u8p), _ptr(cof, _i64p), _ptr(pad(cv, np.float32), _f32p),
              _ptr(vo, _i64p), len(variables), _ptr(vt_, _u8p), _ptr(vi_, _u8p), len(allowed_ops), _ptr(al, _u8p),
              cap, cap, 9 * cap, _ptr(ops_o, _u8p), _ptr(args_o, _u16p), _ptr(co_o, _i64p), _ptr(ct_o, _u8p),
              _ptr(cof_o, _i64p), _ptr(cv_o, _f32p), _ptr(vo_o, _i64p), _ptr(kinds, _u8p)))
    if timing is not None:
        timing["call_ms"] = (time.perf_counter() - t0) * 1e3
    wire = (ops_o[:co_o[n]], args_o[:co_o[n]], co_o, ct_o[:cof_o[n]], cof_o, cv_o[:vo_o[n]], vo_o)
    return (unpack_programs(*wire) if unpack else wire), kinds[:n]


class EngineConfig(C.Structure):
    _fields_ = [("num_islands", C.c_int32), ("island_size", C.c_int32), ("max_generations", C.c_int32),
                ("crossover_rate", C.c_float), ("tournament_size", C.c_int32), ("migration_interval", C.c_int32),
                ("parsimony_penalty", C.c_float), ("opt_prob", C.c_float), ("opt_iterations", C.c_int32),
                ("final_opt_iterations", C.c_int32), ("stagnation_threshold", C.c_int32),
                ("target_mse", C.c_float), ("min_improvement", C.c_float),
                ("random_injection_rate", C.c_float), ("min_random_injection", C.c_int32),
                ("max_tree_size", C.c_int32), ("mutation_max_depth", C.c_int32), ("mutation_cycles", C.c_int32),
                ("mini_batch_size", C.c_int32), ("seed", C.c_uint64)]


class EngineInfo(C.Structure):
    _fields_ = [("generation", C.c_int64), ("nukes", C.c_int64), ("best_island", C.c_int32),
                ("best_fitness", C.c_float), ("best_mse", C.c_float), ("target_reached", C.c_int32),
                ("finished", C.c_int32), ("hof_size", C.c_int32)]


ENGINE_POPULATION, ENGINE_ISLAND_BEST, ENGINE_GLOBAL_BEST, ENGINE_HALL_OF_FAME, ENGINE_ISLAND_HOF, ENGINE_PARETO = \
    range(6)


def engine_default_config() -> EngineConfig:
    l = lib()
    l.rsr_engine_default_config.restype = None
    l.rsr_engine_default_config.argtypes = [C.POINTER(EngineConfig)]
    c = EngineConfig()
    l.rsr_engine_default_config(C.byref(c))
    return c


class Engine:
    """The search loop (rsr_engine_*): Engine::new / run over the given dataset."""

    def __init__(self, ctx: Context, dataset: Dataset, variables: Sequence[tuple], allowed_ops: Sequence[int],
                 config: Optional[EngineConfig] = None, **overrides):
        l = lib()
        self._l = l
        cfg = config if config is not None else engine_default_config()
        for k, v in overrides.items():
            setattr(cfg, k, v)
        self.config = cfg
        l.rsr_engine_create.restype = C.c_int
        l.rsr_engine_create.argtypes = [_P, C.POINTER(EngineConfig), _P, C.c_int, _u8p, _u8p, C.c_int, _u8p,
                                        C.POINTER(_P)]
        for name, args in (("rsr_engine_step", [_P, C.POINTER(C.c_int)]),
                           ("rsr_engine_run", [_P, C.POINTER(C.c_int)]),
                           ("rsr_engine_get_info", [_P, C.POINTER(EngineInfo)]),
                           ("rsr_engine_destroy", [_P]),
                           ("rsr_engine_programs", [_P, C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                                    C.POINTER(C.c_int), _u8p, _u16p, _i64p, _u8p, _i64p, _f32p,
                                                    _i64p, _f32p, _i64p])):
            f = getattr(l, name)
            f.restype = C.c_int
            f.argtypes = args
        vt_ = np.array([v[0] for v in variables] or [0], np.uint8)
        vi_ = np.array([v[1] for v in variables] or [0], np.uint8)
        al = np.array(list(allowed_ops) or [0], np.uint8)
        h = _P()
        _check(l.rsr_engine_create(ctx.h, C.byref(cfg), dataset.h, len(variables), _ptr(vt_, _u8p), _ptr(vi_, _u8p),
                                   len(allowed_ops), _ptr(al, _u8p), C.byref(h)))
        self.h = h
        self.dataset = dataset  # kept alive with the engine

    def set_server(self, server=None):
        """Route the evaluations and fits through `server` (None: direct calls)."""
        self._l.rsr_engine_set_server.restype = C.c_int
        self._l.rsr_engine_set_server.argtypes = [_P, _P]
        _check(self._l.rsr_engine_set_server(self.h, server.h if server is not None else None))
        self.server = server  # kept alive with the engine

    def step(self) -> bool:
        t = C.c_int(0)
        _check(self._l.rsr_engine_step(self.h, C.byref(t)))
        return bool(t.value)

    def run(self) -> int:
        g = C.c_int(0)
        _check(self._l.rsr_engine_run(self.h, C.byref(g)))
        return g.value

    def info(self) -> EngineInfo:
        i = EngineInfo()
        _check(self._l.rsr_engine_get_info(self.h, C.byref(i)))
        return i

    def programs(self, which: int, island: int = 0):
        """(programs, values, counts) of rsr_engine_programs."""
        n = C.c_int(0)
        fn = self._l.rsr_engine_programs
        nul = lambda t: C.cast(None, t)
        rc = fn(self.h, which, island, 0, 0, 0, 0, C.byref(n), nul(_u8p), nul(_u16p), nul(_i64p), nul(_u8p),
                nul(_i64p), nul(_f32p), nul(_i64p), nul(_f32p), nul(_i64p))
        if rc not in (0, RSR_E_INVALID_ARG):
            _check(rc)
        cnt = n.value
        cap = 384 * max(cnt, 1)
        ops_o, args_o = np.zeros(cap, np.uint8), np.zeros(cap, np.uint16)
        ct_o, cv_o = np.zeros(cap, np.uint8), np.zeros(9 * cap, np.float32)
        co_o, cof_o, vo_o = np.zeros(cnt + 1, np.int64), np.zeros(cnt + 1, np.int64), np.zeros(cnt + 1, np.int64)
        val, num = np.zeros(max(cnt, 1), np.float32), np.zeros(max(cnt, 1), np.int64)
        _check(fn(self.h, which, island, cnt, cap, cap, 9 * cap, C.byref(n), _ptr(ops_o, _u8p), _ptr(args_o, _u16p),
                  _ptr(co_o, _i64p), _ptr(ct_o, _u8p), _ptr(cof_o, _i64p), _ptr(cv_o, _f32p), _ptr(vo_o, _i64p),
                  _ptr(val, _f32p), _ptr(num, _i64p)))
        return unpack_programs(ops_o, args_o, co_o, ct_o, cof_o, cv_o, vo_o), val[:cnt], num[:cnt]

    def close(self):
        if self.h:
            self._l.rsr_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
