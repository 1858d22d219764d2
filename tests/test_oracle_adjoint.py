"""The oracle's line-integrated adjoint (d out / d(m, A, h, x, y, z)) against
central finite differences of the oracle's own forward projection, which the
other oracle tests pin to the reference.  The table interpolant is piecewise
linear in q, so a difference step is taken small (1e-6 relative) and the
comparison is normalised per parameter by the size of the gradient."""
import numpy as np
import pytest

import pyoracle as O

KERNELS = {"M4": 0, "C4": 4, "M6": 2}


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.random(n) + 0.2
    h = (rng.random(n) * 0.1 + 0.08) * (100.0 / n) ** (1 / 3)
    rho = rng.random(n) + 0.5
    A = 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    return [x, y, z, m, h, rho, A]


def rays(S, seed):
    rng = np.random.default_rng(seed)
    org = rng.random((3, S)) * 0.8 + 0.1
    dirs = rng.normal(size=(3, S))
    dirs /= np.linalg.norm(dirs, axis=0)
    dirs[:, : S // 2] = np.array([[0.0], [0.0], [1.0]])  # image-plane rays along z
    return org, dirs


def forward(p, kernel, shepard, org, dirs, cot):
    x, y, z, m, h, rho, A = p
    o = O.OracleSet(x, y, z, m, h, rho, [rho, A], kernel=kernel)
    out = o.line_samples(dict(sslots=[2], sflags=[int(shepard)]), O.SCATTER, org, dirs)[0]
    out = np.where(np.isnan(out), 0.0, out)
    return float(np.dot(cot, out))


@pytest.mark.parametrize("name", list(KERNELS))
@pytest.mark.parametrize("shepard", [False, True])
def test_line_adjoint_matches_finite_differences(name, shepard):
    kernel = KERNELS[name]
    p = cloud(300, 5 + kernel)
    n = len(p[0])
    org, dirs = rays(160, 9)
    cot = np.random.default_rng(3).normal(size=org.shape[1])
    o = O.OracleSet(*p[:6], [p[5], p[6]], kernel=kernel)
    g = o.line_adjoint(2, shepard, org, dirs, cot)  # rows m, A, h, x, y, z
    assert g.shape == (6, n)
    rng = np.random.default_rng(11)
    picks = rng.choice(n, 24, replace=False)
    # (row of g, index into p, step)
    params = [(0, 3, "rel"), (1, 6, "rel"), (2, 4, "rel"), (3, 0, "h"), (4, 1, "h"), (5, 2, "h")]
    for row, pi, kind in params:
        scale = max(np.sqrt(np.mean(g[row] ** 2)), 1e-300)
        for b in picks:
            base = p[pi][b]
            eps = 1e-6 * (abs(base) if kind == "rel" else p[4][b])
            q_plus = [a.copy() for a in p]
            q_minus = [a.copy() for a in p]
            q_plus[pi][b] = base + eps
            q_minus[pi][b] = base - eps
            fd = (forward(q_plus, kernel, shepard, org, dirs, cot) - forward(q_minus, kernel, shepard, org, dirs,
                                                                            cot)) / (2 * eps)
            assert abs(fd - g[row, b]) <= 2e-4 * scale + 1e-6 * abs(g[row, b]), (row, b, fd, g[row, b], scale)


def test_line_adjoint_is_linear_in_the_cotangent_and_local():
    p = cloud(200, 1)
    org, dirs = rays(40, 2)
    o = O.OracleSet(*p[:6], [p[5], p[6]], kernel=0)
    rng = np.random.default_rng(0)
    c1, c2 = rng.normal(size=40), rng.normal(size=40)
    g1 = o.line_adjoint(2, False, org, dirs, c1)
    g2 = o.line_adjoint(2, False, org, dirs, c2)
    g12 = o.line_adjoint(2, False, org, dirs, 2.0 * c1 - 3.0 * c2)
    np.testing.assert_allclose(g12, 2.0 * g1 - 3.0 * g2, rtol=1e-12, atol=1e-12 * np.abs(g12).max())
    # particles no ray reaches get exact zeros
    far = o.line_adjoint(2, False, [np.full(1, 5.0), np.full(1, 5.0), np.zeros(1)],
                         [np.zeros(1), np.zeros(1), np.ones(1)], np.ones(1))
    assert not far.any()


# ---- scalar outputs of the general interpolation kernel at points ----

STRATS = {"gather": O.GATHER, "scatter": O.SCATTER, "symmetric": O.SYMMETRIC}


def points(S, seed):
    return np.random.default_rng(seed).random((3, S)) * 0.8 + 0.1


def point_forward(p, kernel, strategy, shepard, pts, cot):
    x, y, z, m, h, rho, A = p
    o = O.OracleSet(x, y, z, m, h, rho, [rho, A], kernel=kernel)
    out = o.point_samples(dict(sslots=[2], sflags=[int(shepard)]), strategy, *pts)[0]
    out = np.where(np.isnan(out), 0.0, out)
    return float(np.dot(cot, out))


@pytest.mark.parametrize("strategy", list(STRATS))
@pytest.mark.parametrize("name,shepard", [("M4", False), ("M4", True), ("C4", True), ("M6", False)])
def test_point_adjoint_matches_finite_differences(strategy, name, shepard):
    """d/d(m, A, h, x, y, z) of sum_s cot_s out_s against central differences of
    the oracle's forward; for gather and symmetric the h row includes the
    dependence of h_a on the nearest particle's h."""
    kernel, st = KERNELS[name], STRATS[strategy]
    p = cloud(300, 15 + kernel)
    n = len(p[0])
    pts = points(120, 19)
    cot = np.random.default_rng(4).normal(size=pts.shape[1])
    o = O.OracleSet(*p[:6], [p[5], p[6]], kernel=kernel)
    g = o.point_adjoint(2, shepard, st, *pts, cot)
    assert g.shape == (6, n)
    rng = np.random.default_rng(12)
    # particles the points reach (nonzero rows) plus a few others
    reached = np.flatnonzero(np.abs(g).sum(axis=0))
    assert len(reached) > 20
    picks = np.concatenate([rng.choice(reached, 20, replace=False), rng.choice(n, 4, replace=False)])
    ha, best = o.nearest_h(*pts)
    if st != O.SCATTER:  # include nearest particles, whose h moves h_a
        order = o.tree()["order"].astype(np.int64)  # 1-based input index per leaf
        picks = np.concatenate([picks, order[best[:4] - 1] - 1])
    params = [(0, 3, "rel"), (1, 6, "rel"), (2, 4, "rel"), (3, 0, "h"), (4, 1, "h"), (5, 2, "h")]
    for row, pi, kind in params:
        scale = max(np.sqrt(np.mean(g[row] ** 2)), 1e-300)
        for b in picks:
            base = p[pi][b]
            eps = 1e-6 * (abs(base) if kind == "rel" else p[4][b])
            q_plus = [a.copy() for a in p]
            q_minus = [a.copy() for a in p]
            q_plus[pi][b] = base + eps
            q_minus[pi][b] = base - eps
            fd = (point_forward(q_plus, kernel, st, shepard, pts, cot) -
                  point_forward(q_minus, kernel, st, shepard, pts, cot)) / (2 * eps)
            assert abs(fd - g[row, b]) <= 2e-4 * scale + 1e-6 * abs(g[row, b]), (row, b, fd, g[row, b], scale)


def test_point_adjoint_linear_in_the_cotangent_and_empty_points():
    p = cloud(200, 3)
    o = O.OracleSet(*p[:6], [p[5], p[6]], kernel=0)
    pts = points(50, 4)
    rng = np.random.default_rng(1)
    c1, c2 = rng.normal(size=50), rng.normal(size=50)
    for st in STRATS.values():
        g1 = o.point_adjoint(2, True, st, *pts, c1)
        g2 = o.point_adjoint(2, True, st, *pts, c2)
        g12 = o.point_adjoint(2, True, st, *pts, 2.0 * c1 - 3.0 * c2)
        np.testing.assert_allclose(g12, 2.0 * g1 - 3.0 * g2, rtol=1e-12, atol=1e-12 * np.abs(g12).max())
        # a point far from every particle (NaN output) contributes nothing
        far = o.point_adjoint(2, True, st, [5.0], [5.0], [5.0], [1.0])
        assert not far.any()


# ---- gradient, divergence and curl outputs at points ----

FIELDS = {1: dict(gslots=[2]), 2: dict(dslots=[[2, 3, 4]]), 3: dict(cslots=[[2, 3, 4]])}


def field_cloud(n, seed):
    p = cloud(n, seed)
    x, y, z = p[0], p[1], p[2]
    B = 0.5 + np.cos(2 * np.pi * z) * x
    Cq = y * y - 0.3 * np.sin(2 * np.pi * x * z)
    return p + [B, Cq]  # x y z m h rho A B C


def field_forward(p, kernel, strategy, mode, pts, cot):
    x, y, z, m, h, rho, A, B, Cq = p
    o = O.OracleSet(x, y, z, m, h, rho, [rho, A, B, Cq], kernel=kernel)
    out = o.point_samples(FIELDS[mode], strategy, *pts)  # (1 or 3) x S
    out = np.where(np.isnan(out), 0.0, out)
    return float(np.sum(cot.T * out))


def test_kernel_second_derivative_matches_differences_of_the_first():
    """f'' of every kernel against central differences of f' (kernel.jl's
    derivative functions), away from the knots where f'' jumps."""
    for K in range(6):
        R = O.support(K)
        for q in np.linspace(0.013, R - 0.011, 97):
            e = 1e-6
            fd = (O.kernel_df(K, q + e) - O.kernel_df(K, q - e)) / (2 * e)
            assert abs(fd - O.kernel_d2f(K, q)) <= 1e-6 * (1 + abs(fd)), (K, q, fd, O.kernel_d2f(K, q))
        assert O.kernel_d2f(K, R + 0.1) == 0.0


@pytest.mark.parametrize("strategy", list(STRATS))
@pytest.mark.parametrize("mode,name", [(1, "M4"), (2, "C4"), (3, "M4"), (1, "M6"), (3, "C4")])
def test_field_adjoint_matches_finite_differences(strategy, mode, name):
    """d/d(m, field columns, h, x, y, z) of sum_s cot_s . out_s for the gradient,
    divergence and curl outputs against central differences of the oracle's
    forward interpolation."""
    kernel, st = KERNELS[name], STRATS[strategy]
    p = field_cloud(260, 31 + kernel + 7 * mode)
    n = len(p[0])
    pts = points(90, 23 + mode)
    c = 1 if mode == 2 else 3
    cot = np.random.default_rng(5 + mode).normal(size=(pts.shape[1], c))
    o = O.OracleSet(*p[:6], [p[5], p[6], p[7], p[8]], kernel=kernel)
    slots = [2] if mode == 1 else [2, 3, 4]
    g = o.field_adjoint(mode, slots, False, st, *pts, cot)
    nc = len(slots)
    assert g.shape == (5 + nc, n)
    rng = np.random.default_rng(17)
    reached = np.flatnonzero(np.abs(g).sum(axis=0))
    assert len(reached) > 20
    picks = np.concatenate([rng.choice(reached, 14, replace=False), rng.choice(n, 3, replace=False)])
    if st != O.SCATTER:
        ha, best = o.nearest_h(*pts)
        order = o.tree()["order"].astype(np.int64)
        picks = np.concatenate([picks, order[best[:3] - 1] - 1])
    params = [(0, 3, "rel")] + [(1 + j, 6 + j, "rel") for j in range(nc)] + [(1 + nc, 4, "rel")] + \
        [(2 + nc + d, d, "h") for d in range(3)]
    for row, pi, kind in params:
        scale = max(np.sqrt(np.mean(g[row] ** 2)), 1e-300)
        for b in picks:
            base = p[pi][b]
            eps = 1e-6 * (abs(base) if kind == "rel" else p[4][b])
            q_plus = [a.copy() for a in p]
            q_minus = [a.copy() for a in p]
            q_plus[pi][b] = base + eps
            q_minus[pi][b] = base - eps
            fd = (field_forward(q_plus, kernel, st, mode, pts, cot) -
                  field_forward(q_minus, kernel, st, mode, pts, cot)) / (2 * eps)
            assert abs(fd - g[row, b]) <= 3e-4 * scale + 1e-6 * abs(g[row, b]), (row, b, fd, g[row, b], scale)


def test_field_adjoint_scalar_mode_is_the_point_adjoint():
    p = field_cloud(200, 8)
    o = O.OracleSet(*p[:6], [p[5], p[6], p[7], p[8]], kernel=0)
    pts = points(40, 9)
    cot = np.random.default_rng(2).normal(size=40)
    for st in STRATS.values():
        for shep in (False, True):
            a = o.point_adjoint(2, shep, st, *pts, cot)
            b = o.field_adjoint(0, [2], shep, st, *pts, cot)
            np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-13 * np.abs(a).max())
