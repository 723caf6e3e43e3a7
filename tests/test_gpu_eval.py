"""GPU parity: the HIP evaluator against the CPU oracle, through the C ABI.

Bars: per-point values and dual tangents bit-exact (same IEEE operations, same
order, same transcendental algorithm); per-program MSE within 1e-5 relative of the
oracle's f32 result (the summation orders differ: 8-lane sequential in the
reference, pairwise + Kahan on the GPU); gradients within 1e-5 of the summed term
magnitude sum|2 (v - y) t| / N; known-answer programs exactly 0.
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


def same_bits(a, b):
    a = np.asarray(a, f32)
    b = np.asarray(b, f32)
    ok = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b)) | ((a == 0) & (b == 0))
    return bool(ok.all()), int((~ok).sum())


def make(ctx, X, y, types):
    F = X.shape[1]
    return api.Dataset.new(ctx, X, y, types), O.ODataset.new(X, y, F)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


def check_mse(gpu, ref, tol=1e-5):
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    both_max = (gpu == FMAX) & (ref == FMAX)
    ok = both_max | (rel_err(gpu, ref) <= tol) | ((gpu == 0) & (ref == 0))
    return ok


@pytest.mark.parametrize("ppl", [1, 2, 3, 4])
def test_known_answers_exactly_zero(gpu_ctx, ppl):
    gpu_ctx.set_tuning(ppl, 8)
    X, y, t = W.data_c1(4000, seed=7)
    ds, _ = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, [W.lorentz_program()])
    assert api.compute_mse(gpu_ctx, P, ds)[0] == 0.0
    X, y, t = W.data_c2(3000, seed=8)
    ds, _ = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, [W.transform_program()])
    assert api.compute_mse(gpu_ctx, P, ds)[0] == 0.0
    X, y, t = W.data_c3(1000, seed=9)
    ds, _ = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, [W.inverse_trace_program()])
    assert api.compute_mse(gpu_ctx, P, ds)[0] == 0.0
    gpu_ctx.set_tuning(2, 8)


@pytest.mark.parametrize("n", [400, 403, 4, 1, 9])
def test_padding_lanes_match_reference_semantics(gpu_ctx, n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-3, 3, n).astype(f32)
    y = (f32(2.5) * x) * x - f32(1.2)
    ds, ods = make(gpu_ctx, x.reshape(-1, 1), y, [VT.Float])
    progs = [W.square_program(), Program([(Op.LoadConstF, 0)], [Scalar.f(1.0)])]
    P = api.Programs(gpu_ctx, progs)
    got = api.compute_mse(gpu_ctx, P, ds)
    ref = [O.compute_mse(p, ods) for p in progs]
    assert check_mse(got, ref, 1e-6).all(), (got, ref)
    if n % 8 == 0:
        assert got[0] == 0.0


def _population_parity(ctx, progs, X, y, types, n_points_check, ppl=2):
    ctx.set_tuning(ppl, 8)
    ds, ods = make(ctx, X, y, types)
    P = api.Programs(ctx, progs)
    assert (P.status(ds) == 0).all()
    mse = api.compute_mse(ctx, P, ds)
    ref = np.array([O.compute_mse(p, ods) for p in progs], np.float32)
    ok = check_mse(mse, ref)
    assert ok.all(), (np.flatnonzero(~ok)[:10], mse[~ok][:5], ref[~ok][:5])
    pts = api.eval_points(ctx, P, ds, 0, n_points_check)
    for i in range(n_points_check):
        same, nbad = same_bits(pts[i], O.eval_points(progs[i], ods))
        assert same, (i, nbad, progs[i].code)
    ctx.set_tuning(2, 8)
    return mse, ref


def test_c0_population(gpu_ctx):
    X, y, t = W.data_c0(400, seed=1000)
    progs = W.gen_c0(1000, seed=2000)
    mse, ref = _population_parity(gpu_ctx, progs, X, y, t, 200)
    # FP64 oracle: report the error of both f32 paths against truth
    ods = O.ODataset.new(X, y, 1)
    truth = np.array([O.compute_mse_f64(p, ods) for p in progs[:100]])
    fin = np.isfinite(truth) & (ref[:100] < FMAX)
    assert np.median(rel_err(mse[:100][fin], truth[fin])) < 1e-5


@pytest.mark.parametrize("ppl", [1, 2, 3, 4])
def test_c1_population(gpu_ctx, ppl):
    X, y, t = W.data_c1(3001, seed=1001)
    progs = W.gen_c1(256, seed=2001)
    _population_parity(gpu_ctx, progs, X, y, t, 32, ppl)


def test_c2_c3_populations(gpu_ctx):
    X, y, t = W.data_c2(2000, seed=1002)
    _population_parity(gpu_ctx, W.gen_c2(200, seed=2002), X, y, t, 24)
    X, y, t = W.data_c3(2000, seed=1003)
    _population_parity(gpu_ctx, W.gen_c3(200, seed=2003), X, y, t, 24)


def test_all_tunings_bitwise_identical_mse(gpu_ctx):
    X, y, t = W.data_c1(2000, seed=3)
    progs = W.gen_c1(128, seed=4)
    ds, _ = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    outs = []
    for ppl in (1, 2, 3, 4):
        for g in (1, 8, 32):
            gpu_ctx.set_tuning(ppl, g)
            outs.append(api.compute_mse(gpu_ctx, P, ds))
    gpu_ctx.set_tuning(2, 8)
    # same tile size -> same partial sums -> identical bits; across tile sizes within 1e-6
    for o in outs[1:]:
        assert check_mse(o, outs[0], 2e-6).all()
    # a repeat at the default tuning (2 points per lane, 8 programs per item) is bit-identical
    assert np.array_equal(outs[1 * 3 + 1].view(np.uint32), api.compute_mse(gpu_ctx, P, ds).view(np.uint32))


def test_deterministic_repeat(gpu_ctx):
    X, y, t = W.data_c1(5000, seed=5)
    progs = W.gen_c1(300, seed=6)
    ds, _ = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    a = api.compute_mse(gpu_ctx, P, ds)
    b = api.compute_mse(gpu_ctx, P, ds)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_non_finite_is_flt_max_and_rejected_programs(gpu_ctx):
    x = np.array([0, 1, 2, 3], f32)
    ds, _ = make(gpu_ctx, x.reshape(-1, 1), x, [VT.Float])
    progs = [Program([(Op.LoadConstF, 0), (Op.LoadVarF, 0), (Op.DivF, 0)], [Scalar.f(1.0)]),
             Program([(Op.AddF, 0)], []),                        # stack underflow -> rejected
             Program([(Op.LoadVarF, 0), (Op.SqrF, 0)], [])]
    P = api.Programs(gpu_ctx, progs)
    st = P.status(ds)
    assert list(st) == [0, -1, 0]
    mse = api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)
    assert mse[0] == FMAX and np.isnan(mse[1])
    assert mse[2] == pytest.approx(np.mean((x * x - x) ** 2) * 4 / 4, rel=1e-6)
    with pytest.raises(api.RsrError):
        api.compute_mse(gpu_ctx, P, ds)


def test_gradients_against_oracle(gpu_ctx):
    X, y, t = W.data_c2(1000, seed=11)
    progs = W.gen_c2(120, seed=12) + [Program([(Op.LoadConstF, 0), (Op.LoadVarV3, 9), (Op.LoadVarV3, 9),
                                                (Op.SubV3, 0), (Op.NormV3, 0), (Op.AddF, 0)], [Scalar.f(1.0)])]
    ds, ods = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    mse, grad = api.compute_mse_with_gradient(gpu_ctx, P, ds)
    fv, tv = ods.soa()
    for i, p in enumerate(progs):
        m_ref, g_ref = O.compute_mse_with_gradient(p, ods)
        assert check_mse([mse[i]], [m_ref])[0], (i, mse[i], m_ref)
        if m_ref == FMAX:
            assert (grad[i] == 0).all()
            continue
        for k in range(min(p.param_count(), 32)):
            v, tk = O.eval_points_dual(p, ods, k)
            scale = np.sum(np.abs(2.0 * (v.astype(np.float64) - tv) * tk)) / ods.ns
            if np.isnan(g_ref[k]):
                assert np.isnan(grad[i, k]), (i, k)
                continue
            assert abs(float(grad[i, k]) - float(g_ref[k])) <= 1e-5 * scale + 1e-30, (i, k, grad[i, k], g_ref[k])
        assert (grad[i, p.param_count():] == 0).all()


def test_dual_points_bit_exact(gpu_ctx):
    X, y, t = W.data_c2(300, seed=13)
    progs = W.gen_c2(30, seed=14)
    ds, ods = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    for i, p in enumerate(progs):
        for k in range(min(p.param_count(), 3)):
            v, g = api.eval_points_dual(gpu_ctx, P, ds, i, k)
            rv, rg = O.eval_points_dual(p, ods, k)
            assert same_bits(v, rv)[0] and same_bits(g, rg)[0], (i, k)


def test_subset_and_normalised_dataset(gpu_ctx):
    X, y, t = W.data_c1(1003, seed=15)
    progs = W.gen_c1(64, seed=16)
    ds = api.Dataset.new(gpu_ctx, X, y, t, normalize=True)
    ods = O.ODataset.new(X, y, 9, normalize=True)
    assert np.allclose(ds.stats(), ods.stats, rtol=0, atol=0)
    for n in (400, 64, 1003, 5000):
        sub, osub = ds.subset(n), ods.subset(n)
        assert sub.num_samples == osub.ns
        P = api.Programs(gpu_ctx, progs)
        got = api.compute_mse(gpu_ctx, P, sub)
        ref = [O.compute_mse(p, osub) for p in progs]
        assert check_mse(got, ref).all()


def test_aosoa_dataset_matches(gpu_ctx):
    X, y, t = W.data_c3(777, seed=17)
    ods = O.ODataset.new(X, y, 8)
    ds = api.Dataset.from_aosoa(gpu_ctx, ods.ff, ods.tb, 8, ods.ns)
    progs = W.gen_c3(50, seed=18)
    P = api.Programs(gpu_ctx, progs)
    assert check_mse(api.compute_mse(gpu_ctx, P, ds), [O.compute_mse(p, ods) for p in progs]).all()


def test_set_constants_roundtrip(gpu_ctx):
    X, y, t = W.data_c0(400, seed=19)
    progs = W.gen_c0(50, seed=20)
    ds, ods = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    c = P.get_constants()
    c2 = (c * f32(0.5)).astype(f32)
    P.set_constants(c2)
    assert np.array_equal(P.get_constants(), c2)
    got = api.compute_mse(gpu_ctx, P, ds)
    off = 0
    for i, p in enumerate(progs):
        consts = []
        for s in p.constants:
            consts.append(Scalar(s.vt, tuple(c2[off:off + len(s.vals)])))
            off += len(s.vals)
        assert check_mse([got[i]], [O.compute_mse(Program(p.code, consts), ods)])[0]


def test_full_size_c1_properties(gpu_ctx):
    """The Lorentz workload at full size: 8,192 programs x 100,000 points.  Reference order: every
    program bit-identical to the oracle's f32 evaluator (oracle on 16 host threads).  Kahan:
    within 1e-5 of the FP64 oracle on 64 programs, and within 5e-5 of the f32 evaluator on
    every program (the f32 evaluator's own rounding at 10^5 points).  Plus size-independent
    checks: the known answer is 0, duplicates agree, a permutation permutes."""
    from concurrent.futures import ThreadPoolExecutor
    X, y, t = W.data_c1(100_000, seed=1001)
    progs = W.gen_c1(8192, seed=2001)
    progs[17] = W.lorentz_program()
    progs[4000] = progs[5]
    ds, ods = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    gpu_ctx.set_summation(True)
    try:
        mse = api.compute_mse(gpu_ctx, P, ds)
    finally:
        gpu_ctx.set_summation(False)
    kahan = api.compute_mse(gpu_ctx, P, ds)
    assert mse[17] == 0.0 and kahan[17] == 0.0
    assert mse[4000] == mse[5] and kahan[4000] == kahan[5]
    assert (mse >= 0).all()
    with ThreadPoolExecutor(16) as ex:
        ref = np.array(list(ex.map(lambda p: O.compute_mse(p, ods), progs)), f32)
    ok, nbad = same_bits(mse, ref)
    assert ok, nbad
    assert check_mse(kahan, ref, 5e-5).all()
    rng = np.random.default_rng(0)
    sample = rng.choice(8192, 64, replace=False)
    with ThreadPoolExecutor(16) as ex:
        f64 = np.array(list(ex.map(lambda i: O.compute_mse_f64(progs[i], ods), sample)))
    fin = np.isfinite(f64) & (f64 < np.finfo(f32).max)
    assert (np.abs(kahan[sample][fin] - f64[fin]) <= 1e-5 * np.abs(f64[fin])).all()
    perm = rng.permutation(8192)
    P2 = api.Programs(gpu_ctx, [progs[i] for i in perm])
    mse2 = api.compute_mse(gpu_ctx, P2, ds)
    assert np.array_equal(mse2, kahan[perm])


@pytest.mark.parametrize("ppl", [1, 2, 3, 4])
@pytest.mark.parametrize("gen,data,npts", [(W.gen_c0, W.data_c0, 403), (W.gen_c1, W.data_c1, 1001),
                                            (W.gen_c2, W.data_c2, 500), (W.gen_c3, W.data_c3, 777)])
def test_reference_order_summation_bit_identical(gpu_ctx, gen, data, npts, ppl):
    X, y, t = data(npts, seed=41)
    progs = gen(96, seed=42)
    ds, ods = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    gpu_ctx.set_summation(True)
    gpu_ctx.set_tuning(ppl, 8)
    try:
        mse = api.compute_mse(gpu_ctx, P, ds)
        m2, g2 = api.compute_mse_with_gradient(gpu_ctx, P, ds)
    finally:
        gpu_ctx.set_summation(False)
        gpu_ctx.set_tuning(2, 8)
    for i, p in enumerate(progs):
        ref = O.compute_mse(p, ods)
        assert mse[i] == ref, (i, mse[i], ref)
        rm, rg = O.compute_mse_with_gradient(p, ods)
        assert m2[i] == rm, (i, m2[i], rm)
        ok, nbad = same_bits(g2[i], rg)
        assert ok, (i, g2[i][:4], rg[:4])


def test_reference_order_full_size_c1(gpu_ctx):
    """Reference summation at 100,000 points: bit-identical MSE to the CPU f32 oracle."""
    X, y, t = W.data_c1(100_000, seed=1001)
    progs = W.gen_c1(512, seed=2001)
    ds, ods = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    gpu_ctx.set_summation(True)
    gpu_ctx.set_tuning(4, 4)
    try:
        mse = api.compute_mse(gpu_ctx, P, ds)
    finally:
        gpu_ctx.set_summation(False)
        gpu_ctx.set_tuning(2, 8)
    for i in np.random.default_rng(3).choice(len(progs), 8, replace=False):
        assert mse[i] == O.compute_mse(progs[i], ods), i


@pytest.mark.parametrize("reference", [False, True])
def test_split_rounds_same_as_unsplit(gpu_ctx, reference):
    """A small partials arena splits evaluations into tile ranges; results are unchanged."""
    X, y, t = W.data_c2(5003, seed=51)
    progs = W.gen_c2(200, seed=52)
    ds, _ = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    gpu_ctx.set_summation(reference)
    try:
        m0 = api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)
        gm0, g0 = api.compute_mse_with_gradient(gpu_ctx, P, ds, allow_rejected=True)
        gpu_ctx.set_partials_limit(64 * 1024)
        m1 = api.compute_mse(gpu_ctx, P, ds, allow_rejected=True)
        gm1, g1 = api.compute_mse_with_gradient(gpu_ctx, P, ds, allow_rejected=True)
    finally:
        gpu_ctx.set_partials_limit(4 << 30)
        gpu_ctx.set_summation(False)
    assert same_bits(m0, m1)[0]
    assert same_bits(gm0, gm1)[0]
    assert same_bits(g0, g1)[0]


def test_default_summation_is_reference_order():
    """A fresh context sums in the reference's order: its MSEs equal the oracle's f32
    evaluator bit for bit (full-size Lorentz programs, 100,000 points)."""
    ctx = api.Context(0)
    try:
        X, y, t = W.data_c1(100000, seed=1001)
        progs = W.gen_c1(64, seed=2001)
        ds, ods = make(ctx, X, y, t)
        P = api.Programs(ctx, progs)
        mse = api.compute_mse(ctx, P, ds)
        ref = np.array([O.compute_mse(p, ods) for p in progs], f32)
        assert same_bits(mse, ref)[0]
        P.close()
        ds.close()
    finally:
        ctx.close()


def test_reference_gradients_large_batch_and_subset_agree(gpu_ctx):
    """Reference-order gradients of a 4,200-program batch (per-point terms of every program
    and tangent slot, folded per row) and of its first 160 programs alone give the same bits,
    equal to the oracle on sampled programs."""
    X, y, t = W.data_c2(1500, seed=95)
    progs = W.gen_c2(4200, seed=96)
    ds, ods = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    ok = P.status(ds) == 0
    progs = [p for p, g in zip(progs, ok) if g]
    P.close()
    assert len(progs) >= 4096
    Pw = api.Programs(gpu_ctx, progs)
    Pt = api.Programs(gpu_ctx, progs[:160])
    gpu_ctx.set_summation(True)
    try:
        mw, gw = api.compute_mse_with_gradient(gpu_ctx, Pw, ds)
        mt, gt = api.compute_mse_with_gradient(gpu_ctx, Pt, ds)
    finally:
        gpu_ctx.set_summation(False)
    assert same_bits(mw[:160], mt)[0] and same_bits(gw[:160], gt)[0]
    for i in np.random.default_rng(5).choice(len(progs), 12, replace=False):
        om, og = O.compute_mse_with_gradient(progs[i], ods)
        assert same_bits(mw[i], om)[0] and same_bits(gw[i], og)[0], i


@pytest.mark.parametrize("limit", [None, 700 * 1024])
def test_reference_order_few_programs_many_points(gpu_ctx, limit):
    """Few programs over many points: each program's eight reference accumulators run
    over 37,500 terms in chunks of 2,048, and with a small partials arena over tile ranges
    that cut through the chunks, with the accumulators carried between ranges.  MSEs and
    gradients equal the oracle's bit for bit."""
    X, y, t = W.data_c0(300_003, seed=61)
    progs = W.gen_c0(6, seed=62)
    ds, ods = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    gpu_ctx.set_summation(True)
    if limit:
        gpu_ctx.set_partials_limit(limit)
    try:
        mse = api.compute_mse(gpu_ctx, P, ds)
        m2, g2 = api.compute_mse_with_gradient(gpu_ctx, P, ds)
    finally:
        gpu_ctx.set_partials_limit(4 << 30)
        gpu_ctx.set_summation(False)
    for i, p in enumerate(progs):
        assert mse[i] == O.compute_mse(p, ods), i
        rm, rg = O.compute_mse_with_gradient(p, ods)
        assert m2[i] == rm, (i, m2[i], rm)
        assert same_bits(g2[i], rg)[0], i


@pytest.mark.parametrize("reference", [False, True])
def test_persistent_and_oneshot_launches_agree(gpu_ctx, reference):
    """The persistent evaluation launch (a resident grid draining a per-XCD device work queue;
    every launch here has more work items than the device holds at once) and the one-shot
    launch give the same bits at every points-per-lane; the queue is left empty, so launches
    after one another (and the other form after it) agree too."""
    X, y, t = W.data_c1(20_011, seed=81)
    progs = W.gen_c1(1500, seed=82)
    ds, _ = make(gpu_ctx, X, y, t)
    P = api.Programs(gpu_ctx, progs)
    gpu_ctx.set_summation(reference)
    try:
        outs = []
        for ppl in (1, 2, 3, 4):
            gpu_ctx.set_tuning(ppl, 8)
            for mode in (api.LAUNCH_PERSISTENT, api.LAUNCH_ONESHOT, api.LAUNCH_PERSISTENT):
                gpu_ctx.set_launch_mode(mode)
                outs.append((ppl, api.compute_mse(gpu_ctx, P, ds)))
    finally:
        gpu_ctx.set_launch_mode(api.LAUNCH_AUTO)
        gpu_ctx.set_tuning(2, 8)
        gpu_ctx.set_summation(False)
    for ppl in (1, 2, 3, 4):
        same = [o for p, o in outs if p == ppl]
        assert all(np.array_equal(same[0], o) for o in same[1:]), ppl
    if reference:
        assert all(np.array_equal(outs[0][1], o) for _, o in outs)
