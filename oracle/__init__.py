"""CPU oracle: ctypes binding of oracle/liboracle.so (a plain-C restatement of the
reference evaluator, dual AD, optimizers and selection).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product (rsr_amd/) never uses it.
See rsr_oracle.c for the reference file:line each function follows.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_u16p = C.POINTER(C.c_uint16)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_i32p = C.POINTER(C.c_int32)

_PROG = [_u8p, _u16p, C.c_int, _u8p, _f32p, C.c_int]
SIGS = {
    "orc_sinf": (C.c_float, [C.c_float]),
    "orc_cosf": (C.c_float, [C.c_float]),
    "orc_expf": (C.c_float, [C.c_float]),
    "orc_logf": (C.c_float, [C.c_float]),
    "orc_eval_points": (C.c_int, _PROG + [_f32p, C.c_int, C.c_int64, _f32p]),
    "orc_compute_mse": (C.c_int, _PROG + [_f32p, _f32p, C.c_int, C.c_int64, C.c_int64, _f32p]),
    "orc_compute_mse_with_gradient": (C.c_int, _PROG + [_f32p, _f32p, C.c_int, C.c_int64, C.c_int64, _f32p, _f32p]),
    "orc_eval_points_dual": (C.c_int, _PROG + [_f32p, C.c_int, C.c_int64, C.c_int, _f32p, _f32p]),
    "orc_eval_points_f64": (C.c_int, _PROG + [_f32p, C.c_int, C.c_int64, _f64p]),
    "orc_compute_mse_f64": (C.c_int, _PROG + [_f32p, _f32p, C.c_int, C.c_int64, C.c_int64, _f64p]),
    "orc_dataset_num_batches": (C.c_int64, [C.c_int64]),
    "orc_dataset_new": (C.c_int, [_f32p, _f32p, C.c_int64, C.c_int, C.c_int, _f32p, _f32p, _f32p]),
    "orc_dataset_subset": (C.c_int, [_f32p, _f32p, C.c_int, C.c_int64, C.c_int64, _f32p, _f32p, _i64p]),
    "orc_run_lbfgs": (C.c_int, _PROG + [_f32p, _f32p, C.c_int, C.c_int64, C.c_int64, C.c_int, C.c_int, _f32p]),
    "orc_run_lbfgs_f64": (C.c_int, _PROG + [_f32p, _f32p, C.c_int, C.c_int64, C.c_int64, C.c_int, C.c_int, _f32p]),
    "orc_run_cmaes": (C.c_int, _PROG + [_f32p, _f32p, C.c_int, C.c_int64, C.c_int64, C.c_int, C.c_int, _f32p,
                                        C.c_int, C.c_float, C.c_float, _f32p, _f32p, _i32p]),
    "orc_run_nelder_mead": (C.c_int, _PROG + [_f32p, _f32p, C.c_int, C.c_int64, C.c_int64, C.c_int, _f32p]),
    "orc_assign_rank_and_crowding": (C.c_int, [C.c_int, _f32p, _u64p, _u32p, _f32p]),
    "orc_tournament": (C.c_int, [C.c_int, _i32p, _u32p, _f32p]),
    "orc_fitness": (C.c_float, [C.c_float, C.c_int, C.c_float]),
    "orc_cmaes_constants": (None, [C.c_int, _f32p, _f32p, _f32p]),
    "orc_ulp_sweep": (C.c_int, [C.c_int, C.c_float, C.c_float, C.c_uint32, C.POINTER(C.c_double)]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build as _b
            _b.build()
        l = C.CDLL(LIB_PATH)
        for k, (r, a) in SIGS.items():
            f = getattr(l, k)
            f.restype = r
            f.argtypes = a
        _lib = l
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class OProgram:
    """A program packed for the oracle: ops, args, const types, const payloads."""

    def __init__(self, prog):
        from rsr_amd.isa import VT_SIZE
        self.ops = np.asarray([o for o, _ in prog.code] or [0], np.uint8)
        self.args = np.asarray([a for _, a in prog.code] or [0], np.uint16)
        self.n_code = len(prog.code)
        self.ctypes = np.asarray([int(c.vt) for c in prog.constants] or [0], np.uint8)
        vals = []
        for c in prog.constants:
            vals.extend(list(c.vals) if VT_SIZE[c.vt] else [float(c.vals[0]) if c.vals else 0.0])
        self.cvals = np.asarray(vals or [0.0], np.float32)
        self.n_consts = len(prog.constants)

    def args_(self):
        return [_p(self.ops, _u8p), _p(self.args, _u16p), self.n_code, _p(self.ctypes, _u8p),
                _p(self.cvals, _f32p), self.n_consts]


class ODataset:
    """Reference AoSoA dataset: feature_flat[(b*F+f)*8+lane], target_batches[b*8+lane]."""

    def __init__(self, feature_flat: np.ndarray, target_batches: np.ndarray, num_features: int, num_samples: int,
                 stats=None):
        self.ff = np.ascontiguousarray(feature_flat, np.float32)
        self.tb = np.ascontiguousarray(target_batches, np.float32)
        self.F = int(num_features)
        self.nb = self.tb.size // 8
        self.ns = int(num_samples)
        self.stats = stats

    @staticmethod
    def new(rows: np.ndarray, y: np.ndarray, num_features: int, normalize: bool = False) -> "ODataset":
        rows = np.ascontiguousarray(rows, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        n = y.size
        nb = lib().orc_dataset_num_batches(n)
        ff = np.zeros(max(nb * num_features * 8, 1), np.float32)
        tb = np.zeros(max(nb * 8, 1), np.float32)
        st = np.zeros(2 * num_features + 2, np.float32)
        lib().orc_dataset_new(_p(rows, _f32p), _p(y, _f32p), n, num_features, int(normalize), _p(ff, _f32p),
                              _p(tb, _f32p), _p(st, _f32p))
        return ODataset(ff[: nb * num_features * 8], tb[: nb * 8], num_features, n, st)

    def subset(self, target_samples: int) -> "ODataset":
        s = min(target_samples, self.ns)
        nb = lib().orc_dataset_num_batches(s)
        ff = np.zeros(max(nb * self.F * 8, 1), np.float32)
        tb = np.zeros(max(nb * 8, 1), np.float32)
        out_s = C.c_int64()
        lib().orc_dataset_subset(_p(self.ff, _f32p), _p(self.tb, _f32p), self.F, self.ns, target_samples,
                                 _p(ff, _f32p), _p(tb, _f32p), C.byref(out_s))
        return ODataset(ff[: nb * self.F * 8], tb[: nb * 8], self.F, out_s.value)

    def soa(self) -> Tuple[np.ndarray, np.ndarray]:
        """Features as [F][nb*8] and targets [nb*8] (point order)."""
        f = self.ff.reshape(self.nb, self.F, 8).transpose(1, 0, 2).reshape(self.F, self.nb * 8)
        return f, self.tb.copy()

    def dargs(self, with_target=True):
        if with_target:
            return [_p(self.ff, _f32p), _p(self.tb, _f32p), self.F, self.nb, self.ns]
        return [_p(self.ff, _f32p), self.F, self.nb]


def _rc(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")


def eval_points(prog, ds: ODataset) -> np.ndarray:
    op = OProgram(prog)
    out = np.zeros(ds.nb * 8, np.float32)
    _rc(lib().orc_eval_points(*op.args_(), *ds.dargs(False), _p(out, _f32p)))
    return out


def compute_mse(prog, ds: ODataset) -> float:
    op = OProgram(prog)
    out = C.c_float()
    _rc(lib().orc_compute_mse(*op.args_(), *ds.dargs(), C.byref(out)))
    return out.value


def compute_mse_with_gradient(prog, ds: ODataset):
    op = OProgram(prog)
    m = C.c_float()
    g = np.zeros(32, np.float32)
    _rc(lib().orc_compute_mse_with_gradient(*op.args_(), *ds.dargs(), C.byref(m), _p(g, _f32p)))
    return m.value, g


def eval_points_dual(prog, ds: ODataset, active: int):
    op = OProgram(prog)
    v = np.zeros(ds.nb * 8, np.float32)
    g = np.zeros(ds.nb * 8, np.float32)
    _rc(lib().orc_eval_points_dual(*op.args_(), *ds.dargs(False), active, _p(v, _f32p), _p(g, _f32p)))
    return v, g


def eval_points_f64(prog, ds: ODataset) -> np.ndarray:
    op = OProgram(prog)
    out = np.zeros(ds.nb * 8, np.float64)
    _rc(lib().orc_eval_points_f64(*op.args_(), *ds.dargs(False), _p(out, _f64p)))
    return out


def compute_mse_f64(prog, ds: ODataset) -> float:
    op = OProgram(prog)
    out = C.c_double()
    _rc(lib().orc_compute_mse_f64(*op.args_(), *ds.dargs(), C.byref(out)))
    return out.value


def run_lbfgs(prog, ds: ODataset, iters: int = 15, m: int = 6, fp64: bool = False):
    """Returns (fitness, fitted constant payloads) -- lbfgs.rs:9-165 (fp64=False)."""
    op = OProgram(prog)
    fit = C.c_float()
    fn = lib().orc_run_lbfgs_f64 if fp64 else lib().orc_run_lbfgs
    rc = fn(*op.args_(), *ds.dargs(), iters, m, C.byref(fit))
    return (fit.value if rc == 0 else None), op.cvals[: max(len(op.cvals), 0)].copy()


def run_cmaes(prog, ds: ODataset, iters: int, lam: int, mu: int, normals: np.ndarray):
    """cmaes.rs:19-192 with an injected normal stream normals[it][lam][32].
    Returns (fitness, fitted constant payloads, iterations run)."""
    weights, mu_eff, decay = cmaes_weights(mu)
    op = OProgram(prog)
    fit = C.c_float()
    its = C.c_int32()
    nrm = np.ascontiguousarray(normals, np.float32)
    rc = lib().orc_run_cmaes(*op.args_(), *ds.dargs(), iters, lam, _p(weights, _f32p), mu, C.c_float(mu_eff),
                             C.c_float(decay), _p(nrm, _f32p), C.byref(fit), C.byref(its))
    return (fit.value if rc == 0 else None), op.cvals.copy(), its.value


def cmaes_weights(mu: int):
    """Recombination weights, mu_eff and the sigma decay exactly as cmaes.rs:85-98, 178 (libm f32)."""
    w = np.zeros(mu, np.float32)
    me = C.c_float()
    dec = C.c_float()
    lib().orc_cmaes_constants(mu, _p(w, _f32p), C.byref(me), C.byref(dec))
    return w, me.value, dec.value


def run_nelder_mead(prog, ds: ODataset, iters: int):
    op = OProgram(prog)
    fit = C.c_float()
    rc = lib().orc_run_nelder_mead(*op.args_(), *ds.dargs(), iters, C.byref(fit))
    return (fit.value if rc == 0 else None), op.cvals.copy()


def assign_rank_and_crowding(fitness: Sequence[float], age: Sequence[int]):
    n = len(fitness)
    f = np.asarray(fitness, np.float32)
    a = np.asarray(age, np.uint64)
    r = np.zeros(max(n, 1), np.uint32)
    c = np.zeros(max(n, 1), np.float32)
    lib().orc_assign_rank_and_crowding(n, _p(f, _f32p), _p(a, _u64p), _p(r, _u32p), _p(c, _f32p))
    return r[:n], c[:n]


def tournament(cands: Sequence[int], rank: np.ndarray, crowd: np.ndarray) -> int:
    cc = np.asarray(cands, np.int32)
    return lib().orc_tournament(len(cc), _p(cc, _i32p), _p(np.ascontiguousarray(rank, np.uint32), _u32p),
                                _p(np.ascontiguousarray(crowd, np.float32), _f32p))


def fitness(mse: float, complexity: int, parsimony: float) -> float:
    """island.rs:241-255: mse + complexity * parsimony, f32::MAX when mse is not finite."""
    return lib().orc_fitness(mse, complexity, parsimony)


def sinf(x: float) -> float:
    return lib().orc_sinf(x)


def cosf(x: float) -> float:
    return lib().orc_cosf(x)


def expf(x: float) -> float:
    return lib().orc_expf(x)


def logf(x: float) -> float:
    return lib().orc_logf(x)


SWEEP_FNS = {"sin": 0, "cos": 1, "exp": 2, "ln": 3}


def ulp_sweep(fn: str, lo: float, hi: float, stride: int):
    """Every stride-th float of [lo, hi] through the oracle's sin/cos/exp/ln against FP64
    libm: (max ulp error where the result is a normal float away from zeros, max absolute
    error elsewhere, the argument of the max ulp error, points swept)."""
    out = (C.c_double * 4)()
    rc = lib().orc_ulp_sweep(SWEEP_FNS[fn], lo, hi, stride, out)
    assert rc == 0
    return out[0], out[1], out[2], int(out[3])
