"""Forward-mode dual numbers through the oracle's float64 path against its
hand-derived adjoints.

The reference is generic over T <: AbstractFloat (KI/kernel_function/kernel.jl:141),
so ForwardDiff's duals flow through its CPU path; the oracle's interpolation
templates run on a dual type of their own the same way (oracle_dual.hpp): the
value part is the float64 run, bit for bit, and the derivative part is J * delta
for a direction delta of the particle inputs.  The adjoints give J^T * cot, so
for every direction and cotangent

    cot . (J delta) == (J^T cot) . delta

up to rounding (the dot-product test).  Each parameter row (m, the quantity
columns, h, x, y, z) is checked on its own, so a wrong term in any one row of an
adjoint fails; for gather / symmetric the h row includes the nearest particle's
h_a term (d out / d h_a lands on that particle's h)."""
import numpy as np
import pytest

import pyoracle as O

STRATS = {"gather": O.GATHER, "scatter": O.SCATTER, "symmetric": O.SYMMETRIC}
KERNELS = {"M4": 0, "M6": 2, "C4": 4}
ROWS = {"x": 0, "y": 1, "z": 2, "m": 3, "h": 4}  # rows of a tangent (columns: x y z m h rho quantities)
RTOL = 1e-12


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.random(n) + 0.2
    h = (rng.random(n) * 0.1 + 0.08) * (100.0 / n) ** (1 / 3)
    rho = rng.random(n) + 0.5
    A = 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    B = 0.3 + x * y - 0.2 * z
    Cc = np.cos(3 * z) + 0.1 * x
    return [x, y, z, m, h, rho], [rho, A, B, Cc]  # quantity slots 1..4


def points(S, seed):
    return np.random.default_rng(seed).random((3, S)) * 0.8 + 0.1


def check(lhs_terms, rhs_terms):
    """Both sides of the dot-product identity, to RTOL of the sum of |terms|."""
    lhs, rhs = float(np.sum(lhs_terms)), float(np.sum(rhs_terms))
    scale = float(np.sum(np.abs(lhs_terms)) + np.sum(np.abs(rhs_terms)))
    assert scale > 0.0
    assert abs(lhs - rhs) <= RTOL * scale, (lhs, rhs, scale)


def directions(o, rows_of_adjoint, rng):
    """One tangent per adjoint row: (adjoint row, tangent array) with a random
    direction in that parameter only (rho and the unused columns held fixed)."""
    out = []
    for row, col in rows_of_adjoint:
        t = np.zeros((6 + o.ncol, o.n))
        t[col] = rng.normal(size=o.n)
        out.append((row, t))
    return out


def test_dual_values_are_the_float64_run():
    c6, cols = cloud(250, 1)
    o = O.OracleSet(*c6, cols, kernel=4)
    P = points(60, 2)
    cat = dict(sslots=[1, 2], sflags=[0, 1], gslots=[2], dslots=[2, 3, 4], cslots=[2, 3, 4])
    t = np.random.default_rng(3).normal(size=(6 + o.ncol, o.n))
    for st in STRATS.values():
        val, _ = o.point_jvp(cat, st, *P, t)
        ref = o.point_samples(cat, st, *P)
        assert np.array_equal(np.isnan(val), np.isnan(ref))
        assert np.array_equal(val[~np.isnan(val)], ref[~np.isnan(ref)])
    org, dirs = points(40, 4), np.tile([[0.0], [0.0], [1.0]], (1, 40))
    val, _ = o.line_jvp(dict(sslots=[2], sflags=[1]), O.SCATTER, org, dirs, t)
    ref = o.line_samples(dict(sslots=[2], sflags=[1]), O.SCATTER, org, dirs)
    assert np.array_equal(np.isnan(val), np.isnan(ref))
    assert np.array_equal(val[~np.isnan(val)], ref[~np.isnan(ref)])


def test_dual_derivative_of_a_linear_output_is_exact():
    """sum m W(r, h_b) (scatter, unnormalised) is linear in m: its derivative
    along delta m equals the output evaluated with m replaced by delta m."""
    c6, cols = cloud(200, 5)
    o = O.OracleSet(*c6, cols, kernel=0)
    P = points(50, 6)
    dm = np.random.default_rng(7).normal(size=o.n) + 0.1
    t = np.zeros((6 + o.ncol, o.n))
    t[ROWS["m"]] = dm
    # slot 1 is rho: the unnormalised scalar of rho is sum m W
    _, d = o.point_jvp(dict(sslots=[1], sflags=[0]), O.SCATTER, *P, t)
    o2 = O.OracleSet(c6[0], c6[1], c6[2], dm, c6[4], c6[5], [c6[5]] + cols[1:], kernel=0)
    ref = o2.point_samples(dict(sslots=[1], sflags=[0]), O.SCATTER, *P)
    assert np.allclose(d, ref, rtol=1e-13, atol=1e-13 * np.max(np.abs(ref)))


@pytest.mark.parametrize("strategy", list(STRATS))
@pytest.mark.parametrize("name", list(KERNELS))
@pytest.mark.parametrize("shepard", [False, True])
def test_dual_pins_point_adjoint(strategy, name, shepard):
    st, kernel = STRATS[strategy], KERNELS[name]
    c6, cols = cloud(300, 10 + kernel)
    o = O.OracleSet(*c6, cols, kernel=kernel)
    P = points(80, 11)
    rng = np.random.default_rng(12 + kernel + 3 * st)
    cat = dict(sslots=[2], sflags=[int(shepard)])
    val, _ = o.point_jvp(cat, st, *P, np.zeros((6 + o.ncol, o.n)))
    cot = np.where(np.isnan(val[0]), 0.0, rng.normal(size=P.shape[1]))
    g = o.point_adjoint(2, shepard, st, *P, cot)  # rows m, A, h, x, y, z
    rows = [(0, ROWS["m"]), (1, 6 + 1), (2, ROWS["h"]), (3, ROWS["x"]), (4, ROWS["y"]), (5, ROWS["z"])]
    for row, t in directions(o, rows, rng):
        _, d = o.point_jvp(cat, st, *P, t)
        d = np.where(cot == 0.0, 0.0, d[0])
        col = rows[row][1]
        check(cot * d, g[row] * t[col])


@pytest.mark.parametrize("strategy", list(STRATS))
@pytest.mark.parametrize("mode", [1, 2, 3])
def test_dual_pins_field_adjoint(strategy, mode):
    """Gradient (mode 1), divergence (2) and curl (3) outputs of the fused
    kernel: rows m, the field columns, h, x, y, z."""
    st = STRATS[strategy]
    c6, cols = cloud(280, 20 + mode)
    o = O.OracleSet(*c6, cols, kernel=4 if mode != 2 else 0)
    P = points(70, 21 + mode)
    rng = np.random.default_rng(22 + mode + 5 * st)
    slots = [2] if mode == 1 else [2, 3, 4]
    cat = {1: dict(gslots=[2]), 2: dict(dslots=[2, 3, 4]), 3: dict(cslots=[2, 3, 4])}[mode]
    val, _ = o.point_jvp(cat, st, *P, np.zeros((6 + o.ncol, o.n)))
    w = val.shape[0]
    live = ~np.isnan(val).any(axis=0)
    cot = rng.normal(size=(P.shape[1], w)) * live[:, None]
    g = o.field_adjoint(mode, slots, False, st, *P, cot)
    nc = len(slots)
    rows = [(0, ROWS["m"])] + [(1 + j, 6 + slots[j] - 1) for j in range(nc)] + [(1 + nc, ROWS["h"])] + \
        [(2 + nc + d, d) for d in range(3)]
    for row, t in directions(o, rows, rng):
        _, d = o.point_jvp(cat, st, *P, t)
        d = np.where(live[None, :], d, 0.0)
        col = rows[row][1]
        check(cot.T * d, g[row] * t[col])


@pytest.mark.parametrize("name", list(KERNELS))
@pytest.mark.parametrize("shepard", [False, True])
def test_dual_pins_line_adjoint(name, shepard):
    """The line-integrated scatter projection (the table is linear within a
    cell, and both sides differentiate the same cell)."""
    kernel = KERNELS[name]
    c6, cols = cloud(300, 30 + kernel)
    o = O.OracleSet(*c6, cols, kernel=kernel)
    S = 120
    rng = np.random.default_rng(31 + kernel)
    org = rng.random((3, S)) * 0.8 + 0.1
    dirs = rng.normal(size=(3, S))
    dirs /= np.linalg.norm(dirs, axis=0)
    dirs[:, : S // 2] = np.array([[0.0], [0.0], [1.0]])
    cat = dict(sslots=[2], sflags=[int(shepard)])
    val, _ = o.line_jvp(cat, O.SCATTER, org, dirs, np.zeros((6 + o.ncol, o.n)))
    cot = np.where(np.isnan(val[0]), 0.0, rng.normal(size=S))
    g = o.line_adjoint(2, shepard, org, dirs, cot)
    rows = [(0, ROWS["m"]), (1, 6 + 1), (2, ROWS["h"]), (3, ROWS["x"]), (4, ROWS["y"]), (5, ROWS["z"])]
    for row, t in directions(o, rows, rng):
        _, d = o.line_jvp(cat, O.SCATTER, org, dirs, t)
        d = np.where(cot == 0.0, 0.0, d[0])
        col = rows[row][1]
        check(cot * d, g[row] * t[col])


def test_dual_catches_a_planted_adjoint_error():
    """The identity is sharp: scaling one row of the adjoint by 1 + 1e-9, or
    dropping the nearest particle's h_a term, breaks it."""
    c6, cols = cloud(300, 40)
    o = O.OracleSet(*c6, cols, kernel=4)
    P = points(80, 41)
    rng = np.random.default_rng(42)
    cat = dict(sslots=[2], sflags=[1])
    val, _ = o.point_jvp(cat, O.SYMMETRIC, *P, np.zeros((6 + o.ncol, o.n)))
    cot = np.where(np.isnan(val[0]), 0.0, rng.normal(size=P.shape[1]))
    g = o.point_adjoint(2, True, O.SYMMETRIC, *P, cot)
    t = np.zeros((6 + o.ncol, o.n))
    t[ROWS["h"]] = rng.normal(size=o.n)
    _, d = o.point_jvp(cat, O.SYMMETRIC, *P, t)
    d = np.where(cot == 0.0, 0.0, d[0])
    check(cot * d, g[2] * t[ROWS["h"]])
    with pytest.raises(AssertionError):
        check(cot * d, g[2] * (1 + 1e-9) * t[ROWS["h"]])
    # without the h_a term: the adjoint of the h row as if h_a were fixed
    ha, best = o.nearest_h(*P)
    order = o.tree()["order"].astype(np.int64)
    nearest = np.unique(order[best[cot != 0.0] - 1] - 1)
    t2 = np.zeros_like(t)
    t2[ROWS["h"], nearest] = t[ROWS["h"], nearest]
    _, d2 = o.point_jvp(cat, O.SYMMETRIC, *P, t2)
    d2 = np.where(cot == 0.0, 0.0, d2[0])
    assert np.max(np.abs(d2)) > 0.0  # the nearest particles' h moves the outputs through h_a
