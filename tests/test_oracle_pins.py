"""Pins of the oracle's linear algebra, selection and fitters that the reference's own
known answers leave open.

- Every Vec2 / Vec3 / Mat2 / Mat3 operator against FP64 numpy closed forms on
  asymmetric operands, entry by entry (a transposed or operand-swapped product, a wrong
  index or sign fails by O(1)): matrices are column-major (scalar.rs:108-112), so the
  flat components reshape to the transpose of numpy's row-major layout.
- Crowding distances against a brute-force restatement over each front (neighbours found
  by search, not by sorting).
- Nelder-Mead steps worked by hand on a one-dimensional quadratic whose MSE is exact in
  f32 (expansion, reflection after a failed expansion, inside contraction).
- CMA-ES: convergence on a convex least-squares program, and the doubled dimension of
  cmaes.rs:33-75 (the second copy of the constants takes part; nothing beyond it does).
"""
import numpy as np
import pytest

import oracle as O
from rsr_amd.isa import Op, Program, Scalar, VT

f32 = np.float32
P = Program
L = Op


def _cm(flat, n):
    """Column-major flat components [..., n*n] -> numpy matrices [..., n, n]."""
    return np.swapaxes(flat.reshape(flat.shape[:-1] + (n, n)), -1, -2)


def _well_conditioned(rng, count, n):
    out = []
    while len(out) < count:
        a = np.eye(n) + 0.3 * rng.standard_normal((n, n))
        if np.linalg.cond(a) < 50:
            out.append(a)
    return np.asarray(out)


def _close(got, exp, rtol):
    got = np.asarray(got, np.float64)
    return np.abs(got - exp) <= rtol * (1.0 + np.abs(exp))


# ---------------------------------------------------------------- Mat3 / Vec3
N3 = 96


@pytest.fixture(scope="module")
def mat3_data():
    """Features: A (Mat3, 0..8), B (Mat3, 9..17), u (Vec3, 18..20), s (21), e (Vec3 unit
    selector, 22..24; point i selects column i % 3)."""
    rng = np.random.default_rng(301)
    A = _well_conditioned(rng, N3, 3)
    B = rng.standard_normal((N3, 3, 3))
    u = rng.standard_normal((N3, 3))
    s = rng.uniform(-2, 2, N3)
    col = np.arange(N3) % 3
    e = np.eye(3)[col]
    flatA = np.swapaxes(A, 1, 2).reshape(N3, 9)  # column-major
    flatB = np.swapaxes(B, 1, 2).reshape(N3, 9)
    X = np.c_[flatA, flatB, u, s, e].astype(f32)
    y = np.zeros(N3, f32)
    ds = O.ODataset.new(X, y, X.shape[1])
    # FP64 values of exactly the f32 inputs
    X64 = X.astype(np.float64)
    return ds, dict(A=_cm(X64[:, 0:9], 3), B=_cm(X64[:, 9:18], 3), u=X64[:, 18:21], s=X64[:, 21], col=col)


def _mat3_entry_programs(mat_code, consts=()):
    """Row r of the selected column: GetR(MulM3V3(M, e)) for r = 0, 1, 2."""
    return [P(list(mat_code) + [(L.LoadVarV3, 22), (L.MulM3V3, 0), (g, 0)], list(consts))
            for g in (L.GetXV3, L.GetYV3, L.GetZV3)]


MAT3_OPS = {
    "MulM3": ([(L.LoadVarM3, 0), (L.LoadVarM3, 9), (L.MulM3, 0)], lambda d: d["A"] @ d["B"], 1e-5),
    "MulM3_swapped": ([(L.LoadVarM3, 9), (L.LoadVarM3, 0), (L.MulM3, 0)], lambda d: d["B"] @ d["A"], 1e-5),
    "AddM3": ([(L.LoadVarM3, 0), (L.LoadVarM3, 9), (L.AddM3, 0)], lambda d: d["A"] + d["B"], 1e-6),
    "SubM3": ([(L.LoadVarM3, 0), (L.LoadVarM3, 9), (L.SubM3, 0)], lambda d: d["A"] - d["B"], 1e-6),
    "ScaleM3": ([(L.LoadVarF, 21), (L.LoadVarM3, 9), (L.ScaleM3, 0)], lambda d: d["s"][:, None, None] * d["B"],
                1e-6),
    "TransposeM3": ([(L.LoadVarM3, 9), (L.TransposeM3, 0)], lambda d: np.swapaxes(d["B"], 1, 2), 0.0),
    "InverseM3": ([(L.LoadVarM3, 0), (L.InverseM3, 0)], lambda d: np.linalg.inv(d["A"]), 1e-4),
    "MakeMat3": ([(L.LoadVarV3, 18), (L.LoadVarV3, 9), (L.LoadVarV3, 3), (L.MakeMat3, 0)],
                 lambda d: np.stack([d["u"], d["B"][:, :, 0], d["A"][:, :, 1]], axis=2), 0.0),
}


@pytest.mark.parametrize("name", sorted(MAT3_OPS))
def test_mat3_operator_entries_match_fp64(mat3_data, name):
    ds, d = mat3_data
    code, fn, tol = MAT3_OPS[name]
    exp = fn(d)  # [N, 3, 3]
    rows = np.arange(N3)
    for r, prog in enumerate(_mat3_entry_programs(code)):
        got = O.eval_points(prog, ds)[:N3]
        want = exp[rows, r, d["col"]]
        assert _close(got, want, tol).all(), (name, r, np.max(np.abs(got - want)))


VEC3_FLOAT_OPS = {
    "DetM3": ([(L.LoadVarM3, 0), (L.DetM3, 0)], lambda d: np.linalg.det(d["A"]), 1e-5),
    "DetM3_B": ([(L.LoadVarM3, 9), (L.DetM3, 0)], lambda d: np.linalg.det(d["B"]), 1e-5),
    "TraceM3": ([(L.LoadVarM3, 9), (L.TraceM3, 0)], lambda d: np.trace(d["B"], axis1=1, axis2=2), 1e-6),
    "DotV3": ([(L.LoadVarV3, 18), (L.LoadVarV3, 9), (L.DotV3, 0)], lambda d: np.sum(d["u"] * d["B"][:, :, 0], 1),
              1e-6),
    "NormV3": ([(L.LoadVarV3, 18), (L.NormV3, 0)], lambda d: np.linalg.norm(d["u"], axis=1), 1e-6),
}
VEC3_OPS = {
    "MulM3V3": ([(L.LoadVarM3, 0), (L.LoadVarV3, 18), (L.MulM3V3, 0)],
                lambda d: np.einsum("nij,nj->ni", d["A"], d["u"]), 1e-5),
    "CrossV3": ([(L.LoadVarV3, 18), (L.LoadVarV3, 9), (L.CrossV3, 0)], lambda d: np.cross(d["u"], d["B"][:, :, 0]),
                1e-6),
    "CrossV3_swapped": ([(L.LoadVarV3, 9), (L.LoadVarV3, 18), (L.CrossV3, 0)],
                        lambda d: np.cross(d["B"][:, :, 0], d["u"]), 1e-6),
    "AddV3": ([(L.LoadVarV3, 18), (L.LoadVarV3, 12), (L.AddV3, 0)], lambda d: d["u"] + d["B"][:, :, 1], 1e-6),
    "SubV3": ([(L.LoadVarV3, 18), (L.LoadVarV3, 12), (L.SubV3, 0)], lambda d: d["u"] - d["B"][:, :, 1], 1e-6),
    "ScaleV3": ([(L.LoadVarF, 21), (L.LoadVarV3, 18), (L.ScaleV3, 0)], lambda d: d["s"][:, None] * d["u"], 1e-6),
    "MakeVec3": ([(L.LoadVarF, 21), (L.LoadVarF, 0), (L.LoadVarF, 13), (L.MakeVec3, 0)],
                 lambda d: np.stack([d["s"], d["A"][:, 0, 0], d["B"][:, 1, 1]], 1), 0.0),
}


@pytest.mark.parametrize("name", sorted(VEC3_FLOAT_OPS))
def test_mat3_float_results_match_fp64(mat3_data, name):
    ds, d = mat3_data
    code, fn, tol = VEC3_FLOAT_OPS[name]
    got = O.eval_points(P(code, []), ds)[:N3]
    assert _close(got, fn(d), tol).all(), name


@pytest.mark.parametrize("name", sorted(VEC3_OPS))
def test_vec3_results_match_fp64(mat3_data, name):
    ds, d = mat3_data
    code, fn, tol = VEC3_OPS[name]
    exp = fn(d)
    for k, g in enumerate((L.GetXV3, L.GetYV3, L.GetZV3)):
        got = O.eval_points(P(code + [(g, 0)], []), ds)[:N3]
        assert _close(got, exp[:, k], tol).all(), (name, k)


def test_inverse_m3_of_singular_is_identity():
    """|det| < 1e-9 -> identity (linalg/eval.rs:481-505)."""
    rows = np.zeros((8, 25), f32)
    rows[:, 0:9] = [1, 2, 3, 2, 4, 6, 0, 1, 5]  # columns 1 and 2 proportional: det 0
    rows[:, 22:25] = np.eye(3)[np.arange(8) % 3]
    ds = O.ODataset.new(rows, np.zeros(8, f32), 25)
    col = np.arange(8) % 3
    for r, prog in enumerate(_mat3_entry_programs([(L.LoadVarM3, 0), (L.InverseM3, 0)])):
        got = O.eval_points(prog, ds)[:8]
        assert np.array_equal(got, (col == r).astype(f32)), r


# ---------------------------------------------------------------- Mat2 / Vec2
N2 = 96


@pytest.fixture(scope="module")
def mat2_data():
    """Features: A (Mat2, 0..3), B (Mat2, 4..7), w (Vec2, 8..9), v (Vec2, 10..11), s (12),
    e (Vec2 unit selector, 13..14; point i selects column i % 2)."""
    rng = np.random.default_rng(302)
    A = _well_conditioned(rng, N2, 2)
    B = rng.standard_normal((N2, 2, 2))
    w = rng.standard_normal((N2, 2))
    v = rng.standard_normal((N2, 2))
    s = rng.uniform(-2, 2, N2)
    col = np.arange(N2) % 2
    e = np.eye(2)[col]
    X = np.c_[np.swapaxes(A, 1, 2).reshape(N2, 4), np.swapaxes(B, 1, 2).reshape(N2, 4), w, v, s, e].astype(f32)
    ds = O.ODataset.new(X, np.zeros(N2, f32), X.shape[1])
    X64 = X.astype(np.float64)
    return ds, dict(A=_cm(X64[:, 0:4], 2), B=_cm(X64[:, 4:8], 2), w=X64[:, 8:10], v=X64[:, 10:12], s=X64[:, 12],
                    col=col)


MAT2_OPS = {
    "MulM2": ([(L.LoadVarM2, 0), (L.LoadVarM2, 4), (L.MulM2, 0)], lambda d: d["A"] @ d["B"], 1e-6),
    "MulM2_swapped": ([(L.LoadVarM2, 4), (L.LoadVarM2, 0), (L.MulM2, 0)], lambda d: d["B"] @ d["A"], 1e-6),
    "AddM2": ([(L.LoadVarM2, 0), (L.LoadVarM2, 4), (L.AddM2, 0)], lambda d: d["A"] + d["B"], 1e-6),
    "SubM2": ([(L.LoadVarM2, 0), (L.LoadVarM2, 4), (L.SubM2, 0)], lambda d: d["A"] - d["B"], 1e-6),
    "ScaleM2": ([(L.LoadVarF, 12), (L.LoadVarM2, 4), (L.ScaleM2, 0)], lambda d: d["s"][:, None, None] * d["B"], 1e-6),
    "TransposeM2": ([(L.LoadVarM2, 4), (L.TransposeM2, 0)], lambda d: np.swapaxes(d["B"], 1, 2), 0.0),
    "InverseM2": ([(L.LoadVarM2, 0), (L.InverseM2, 0)], lambda d: np.linalg.inv(d["A"]), 1e-5),
    "MakeMat2": ([(L.LoadVarV2, 8), (L.LoadVarV2, 10), (L.MakeMat2, 0)], lambda d: np.stack([d["w"], d["v"]], 2),
                 0.0),
}


@pytest.mark.parametrize("name", sorted(MAT2_OPS))
def test_mat2_operator_entries_match_fp64(mat2_data, name):
    ds, d = mat2_data
    code, fn, tol = MAT2_OPS[name]
    exp = fn(d)
    rows = np.arange(N2)
    for r, g in enumerate((L.GetXV2, L.GetYV2)):
        prog = P(code + [(L.LoadVarV2, 13), (L.MulM2V2, 0), (g, 0)], [])
        got = O.eval_points(prog, ds)[:N2]
        want = exp[rows, r, d["col"]]
        assert _close(got, want, tol).all(), (name, r, np.max(np.abs(got - want)))


MAT2_FLOAT_OPS = {
    "DetM2": ([(L.LoadVarM2, 4), (L.DetM2, 0)], lambda d: np.linalg.det(d["B"]), 1e-6),
    "TraceM2": ([(L.LoadVarM2, 4), (L.TraceM2, 0)], lambda d: np.trace(d["B"], axis1=1, axis2=2), 1e-6),
    "DotV2": ([(L.LoadVarV2, 8), (L.LoadVarV2, 10), (L.DotV2, 0)], lambda d: np.sum(d["w"] * d["v"], 1), 1e-6),
    "NormV2": ([(L.LoadVarV2, 8), (L.NormV2, 0)], lambda d: np.linalg.norm(d["w"], axis=1), 1e-6),
}
VEC2_OPS = {
    "MulM2V2": ([(L.LoadVarM2, 0), (L.LoadVarV2, 8), (L.MulM2V2, 0)],
                lambda d: np.einsum("nij,nj->ni", d["A"], d["w"]), 1e-6),
    "MulM2V2_B": ([(L.LoadVarM2, 4), (L.LoadVarV2, 10), (L.MulM2V2, 0)],
                  lambda d: np.einsum("nij,nj->ni", d["B"], d["v"]), 1e-6),
    "AddV2": ([(L.LoadVarV2, 8), (L.LoadVarV2, 10), (L.AddV2, 0)], lambda d: d["w"] + d["v"], 1e-6),
    "SubV2": ([(L.LoadVarV2, 8), (L.LoadVarV2, 10), (L.SubV2, 0)], lambda d: d["w"] - d["v"], 1e-6),
    "ScaleV2": ([(L.LoadVarF, 12), (L.LoadVarV2, 8), (L.ScaleV2, 0)], lambda d: d["s"][:, None] * d["w"], 1e-6),
    "MakeVec2": ([(L.LoadVarF, 12), (L.LoadVarF, 0), (L.MakeVec2, 0)], lambda d: np.stack([d["s"], d["A"][:, 0, 0]], 1),
                 0.0),
}


@pytest.mark.parametrize("name", sorted(MAT2_FLOAT_OPS))
def test_mat2_float_results_match_fp64(mat2_data, name):
    ds, d = mat2_data
    code, fn, tol = MAT2_FLOAT_OPS[name]
    got = O.eval_points(P(code, []), ds)[:N2]
    assert _close(got, fn(d), tol).all(), name


@pytest.mark.parametrize("name", sorted(VEC2_OPS))
def test_vec2_results_match_fp64(mat2_data, name):
    ds, d = mat2_data
    code, fn, tol = VEC2_OPS[name]
    exp = fn(d)
    for k, g in enumerate((L.GetXV2, L.GetYV2)):
        got = O.eval_points(P(code + [(g, 0)], []), ds)[:N2]
        assert _close(got, exp[:, k], tol).all(), (name, k)


# ---------------------------------------------------------------- crowding distance
def _brute_crowding(fit, age):
    """assign_rank_and_crowding_distance (selection.rs:80-134) restated per point: fronts by
    exhaustive dominance; within a front of three or more the extremes of each objective
    get f32::MAX and every other point the gap between its nearest neighbours above and
    below, found by search, divided by the front's range (fitness first, then age, in f32)."""
    n = len(fit)
    fmax = np.finfo(f32).max
    dom = lambda i, j: fit[i] <= fit[j] and age[i] <= age[j] and (fit[i] < fit[j] or age[i] < age[j])
    remaining = set(range(n))
    crowd = np.zeros(n, f32)
    while remaining:
        front = sorted(i for i in remaining if not any(dom(j, i) for j in remaining if j != i))
        remaining -= set(front)
        if len(front) < 3:
            crowd[front] = fmax
            continue
        for vals in (np.asarray(fit, f32), np.asarray(age, np.float64).astype(f32)):
            fv = {i: vals[i] for i in front}
            lo_all, hi_all = min(fv.values()), max(fv.values())
            for i in front:
                if fv[i] == lo_all or fv[i] == hi_all:
                    crowd[i] = fmax
            rng_ = f32(hi_all - lo_all)
            if rng_ <= 1e-9:
                continue
            for i in front:
                if crowd[i] == fmax:
                    continue
                below = max(fv[j] for j in front if fv[j] < fv[i])
                above = min(fv[j] for j in front if fv[j] > fv[i])
                crowd[i] = f32(crowd[i] + f32(f32(above - below) / rng_))
    return crowd


def test_crowding_distance_matches_brute_force():
    rng = np.random.default_rng(303)
    checked = 0
    for _ in range(200):
        n = int(rng.integers(3, 13))
        # distinct (fitness, age) pairs with distinct values per objective, so that no two
        # points of one front tie (the order of ties is the sort's)
        fit = rng.choice(np.arange(1, 200, dtype=np.float64) * 0.013, n, replace=False).astype(f32)
        age = rng.choice(np.arange(0, 60), n, replace=False).astype(np.uint64)
        _, c = O.assign_rank_and_crowding(fit, age)
        want = _brute_crowding(list(fit), list(age))
        assert np.array_equal(c.view(np.uint32), want.view(np.uint32)), (fit, age, c, want)
        checked += int(((want != np.finfo(f32).max) & (want > 0)).sum())
    assert checked > 50  # interior points with a finite distance were compared


def test_crowding_distance_hand_worked_front():
    # one front of four: fitness 1, 2, 4, 7 with ages 9, 6, 5, 0 (anti-correlated)
    fit = np.array([2.0, 7.0, 1.0, 4.0], f32)
    age = np.array([6, 0, 9, 5], np.uint64)
    r, c = O.assign_rank_and_crowding(fit, age)
    assert list(r) == [0, 0, 0, 0]
    fmax = np.finfo(f32).max
    # interior: fitness 2 -> (4 - 1) / 6 + (9 - 5) / 9 ; fitness 4 -> (7 - 2) / 6 + (6 - 0) / 9
    assert c[2] == fmax and c[1] == fmax
    assert c[0] == f32(f32(3.0) / f32(6.0)) + f32(f32(4.0) / f32(9.0))
    assert c[3] == f32(f32(5.0) / f32(6.0)) + f32(f32(6.0) / f32(9.0))


# ---------------------------------------------------------------- Nelder-Mead by hand
def _nm_problem(start):
    """c * x with x = 1 and y = 3 at eight points: MSE(c) = (c - 3)^2 exactly (the eight
    equal lanes add to 8 (c - 3)^2 and the division by 8 is exact), fitness (c - 3)^2 +
    0.01 |c| (nelder_mead.rs:7, 44-52)."""
    ds = O.ODataset.new(np.ones((8, 1), f32), np.full(8, 3.0, f32), 1)
    return ds, P([(L.LoadConstF, 0), (L.LoadVarF, 0), (L.MulF, 0)], [Scalar.f(start)])


def test_nelder_mead_expansion_then_reflection():
    # simplex {2.0, 2.2}; iteration 1: reflect to 2.4 (better than the best, 2.2), expand to
    # 2.6, which is better still -> take 2.6.  Iteration 2: centroid 2.6, reflect to 3.0
    # (better than the best), expand to 3.4 (worse than 3.0) -> take the reflection 3.0,
    # whose MSE (below 1e-8) stops iteration 3.
    ds, prog = _nm_problem(2.0)
    f1, c1 = O.run_nelder_mead(prog, ds, 1)
    assert c1[0] == pytest.approx(2.6, abs=1e-6)
    assert f1 == pytest.approx(0.16, abs=1e-5)
    f2, c2 = O.run_nelder_mead(prog, ds, 2)
    assert c2[0] == pytest.approx(3.0, abs=1e-6) and f2 < 1e-8
    f9, c9 = O.run_nelder_mead(prog, ds, 9)
    assert c9[0] == c2[0] and f9 == f2  # stopped at the MSE threshold


def test_nelder_mead_inside_contraction():
    # simplex {2.9, 3.19}: fitness 0.039 and 0.0681.  The reflection 2.61 (fitness 0.178) is
    # worse than the worst, so contract inside: 2.9 + 0.5 (3.19 - 2.9) = 3.045, fitness
    # 0.032475 < 0.0681 -> take it; it is now the best.
    ds, prog = _nm_problem(2.9)
    f1, c1 = O.run_nelder_mead(prog, ds, 1)
    assert c1[0] == pytest.approx(3.045, abs=1e-6)
    assert f1 == pytest.approx(0.045 ** 2, abs=1e-6)


def test_nelder_mead_outside_contraction():
    # simplex {3.5, 3.85}: fitness 0.285 and 0.761.  Reflection 3.15 (fitness 0.054): better
    # than the best, so expand to 2.8 (fitness 0.068), worse than the reflection -> take
    # 3.15.  Iteration 2: simplex {3.15, 3.5}; reflect to 2.8 (0.068): not better than the
    # best (0.054) and not below the second worst (the best) -> it is below the worst
    # (0.285), so contract outside: 3.15 + 0.5 (2.8 - 3.15) = 2.975, fitness 0.030375 < 0.068
    # -> take it.
    ds, prog = _nm_problem(3.5)
    _, c1 = O.run_nelder_mead(prog, ds, 1)
    assert c1[0] == pytest.approx(3.15, abs=1e-6)
    f2, c2 = O.run_nelder_mead(prog, ds, 2)
    assert c2[0] == pytest.approx(2.975, abs=1e-6)
    assert f2 == pytest.approx(0.025 ** 2, abs=1e-6)


# ---------------------------------------------------------------- CMA-ES
def _linear_problem(rng):
    x = rng.uniform(-2, 2, 256).astype(f32)
    y = (f32(1.5) * x + f32(-0.5)).astype(f32)
    ds = O.ODataset.new(x.reshape(-1, 1), y, 1)
    prog = P([(L.LoadConstF, 0), (L.LoadVarF, 0), (L.MulF, 0), (L.LoadConstF, 1), (L.AddF, 0)],
             [Scalar.f(0.2), Scalar.f(0.9)])
    return ds, prog


def test_cmaes_converges_on_convex_least_squares():
    rng = np.random.default_rng(304)
    ds, prog = _linear_problem(rng)
    normals = rng.standard_normal((300, 16, 32)).astype(f32)
    fit, c, its = O.run_cmaes(prog, ds, 300, 16, 8, normals)
    assert fit < O.compute_mse(prog, ds) * 1e-3
    # the 0.01 L1 term pulls the optimum slightly towards zero
    assert abs(c[0] - 1.5) < 0.05 and abs(c[1] + 0.5) < 0.05, c[:2]


def test_cmaes_doubled_dimension():
    """flatten_params fills mean[0..n] and cmaes.rs:33-75 appends the same constants again:
    the search runs in 2n dimensions (the copies enter the L1 term and the covariance),
    and dimensions beyond 2n are never read."""
    rng = np.random.default_rng(305)
    ds, prog = _linear_problem(rng)
    base = rng.standard_normal((12, 16, 32)).astype(f32)
    fit0, c0, _ = O.run_cmaes(prog, ds, 12, 16, 8, base)
    beyond = base.copy()
    beyond[:, :, 4:] = rng.standard_normal(beyond[:, :, 4:].shape)  # dims 4.. unused for n = 2
    fit1, c1, _ = O.run_cmaes(prog, ds, 12, 16, 8, beyond)
    assert fit1 == fit0 and np.array_equal(c1, c0)
    copies = base.copy()
    copies[:, :, 2:4] *= 25.0  # the copies of the two constants: L1 penalties change the ranking
    fit2, c2, _ = O.run_cmaes(prog, ds, 12, 16, 8, copies)
    assert not (fit2 == fit0 and np.array_equal(c2, c0))
