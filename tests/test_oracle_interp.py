"""Oracle interpolation (the fused per-sample kernel and the line-integrated
kernel) against O(N) brute force (tests/brute.py, after
test/interpolation_test_common.jl) at the reference tests' tolerances, plus the
reference's empty-input, uniform-field and analytic-field regressions
(test/interpolation_tests.jl, test/point_samples_interpolation_tests.jl,
test/traversal_analytic.jl)."""
import numpy as np
import pytest

import brute
import pyoracle as O

RNG = np.random.default_rng(4242)
FULL = dict(sslots=[1, 2], sflags=[0, 1], gslots=[2], dslots=[3, 4, 5], cslots=[3, 4, 5])


def random_set(n, kernel="M4_spline", hscale=0.12, hmin=0.04):
    """test/interpolation_test_common.jl:355-368 (own generator)."""
    x, y, z = RNG.random(n), RNG.random(n), RNG.random(n)
    m = RNG.random(n) + 0.05
    h = RNG.random(n) * hscale + hmin
    rho = RNG.random(n) + 0.3
    q = [RNG.random(n) for _ in range(3)]
    cols = [rho, q[0], q[1], q[2], -q[0]]
    o = O.OracleSet(x, y, z, m, h, rho, cols, kernel=kernel)
    c = o.columns()
    inp = dict(x=c[0], y=c[1], z=c[2], m=c[3], h=c[4], rho=c[5], cols=[None] + list(c[6:]))
    return o, inp


@pytest.mark.parametrize("strategy", [O.GATHER, O.SCATTER, O.SYMMETRIC])
@pytest.mark.parametrize("kernel", ["M4_spline", "C4_Wendland", "M6_spline"])
def test_fused_kernel_equals_brute_force(strategy, kernel):
    o, inp = random_set(200, kernel)
    K = O.KERNELS[kernel]
    pts = np.array([[0.4, 0.35, 0.25], [0.2, 0.3, 0.4], [0.7, 0.2, 0.1], [0.95, 0.9, 0.05]])
    out = o.point_samples(FULL, strategy, pts[:, 0], pts[:, 1], pts[:, 2])
    for s, p in enumerate(pts):
        ha = brute.nearest_h(inp, p) if strategy != O.SCATTER else 0.0
        rho = brute.density(inp, p, ha, K, strategy)
        exp = [rho, brute.quantity(inp, p, ha, 2, K, strategy)]
        exp += list(brute.gradient_quantity(inp, p, ha, 2, K, strategy))
        exp += [brute.divergence(inp, p, ha, (3, 4, 5), K, strategy)]
        exp += list(brute.curl(inp, p, ha, (3, 4, 5), K, strategy))
        got = out[:, s]
        if np.isnan(exp[1]):
            assert np.all(np.isnan(got))
            continue
        # density / scalars: rtol 1e-8 (interpolation_tests.jl:240-246); derivatives atol 5e-10 (:270-276)
        assert got[0] == pytest.approx(exp[0], rel=1e-8, abs=1e-10)
        assert got[1] == pytest.approx(exp[1], rel=1e-8, abs=1e-10)
        assert np.allclose(got[2:], exp[2:], rtol=1e-8, atol=5e-10)


def test_empty_and_far_point_are_nan(fixtures):
    fx = fixtures["empty_input"]
    o = O.OracleSet(*(np.array(fx[k]) for k in ("x", "y", "z", "m", "h", "rho")), [np.array(fx["q1"])] * 5)
    for strategy in (O.GATHER, O.SCATTER, O.SYMMETRIC):
        out = o.point_samples(FULL, strategy, *([v] for v in fx["reference_point"]))
        assert np.all(np.isnan(out))
    o2, _ = random_set(20)
    for strategy in (O.GATHER, O.SCATTER, O.SYMMETRIC):
        out = o2.point_samples(FULL, strategy, *([v] for v in fx["far_point"]))
        if strategy == O.SCATTER:
            assert np.all(np.isnan(out))
        else:  # nearest-h gather reaches particles only if 2*ha covers the distance
            ha = o2.nearest_h(*([v] for v in fx["far_point"]))[0][0]
            assert np.all(np.isnan(out)) == (ha * 2 < 1.0)


def test_uniform_field_fixture(fixtures):
    """test/interpolation_tests.jl:305-336: constant v on 4 particles -> div = curl = 0 (atol 1e-12)."""
    fx = fixtures["uniform_field"]
    x, y, z = (np.array(fx[k]) for k in "xyz")
    n = len(x)
    vx, vy, vz = (np.full(n, v) for v in fx["v"])
    o = O.OracleSet(x, y, z, np.full(n, fx["m"]), np.full(n, fx["h"]), np.full(n, fx["rho"]), [vx, vy, vz])
    cat = dict(dslots=[1, 2, 3], cslots=[1, 2, 3])
    for strategy in (O.GATHER, O.SCATTER, O.SYMMETRIC):
        # ha = h of the first particle (the sample sits on it), as in the reference test
        p = [[x[0]], [y[0]], [z[0]]]
        out = o.point_samples(cat, strategy, *p)
        assert np.all(np.abs(out) <= fx["atol"])


def test_uniform_field_has_no_divergence_or_curl():
    """Random cloud, constant fields: derivatives vanish to rounding of the O(1e4) sums."""
    n = 400
    x, y, z = RNG.random(n), RNG.random(n), RNG.random(n)
    m, h, rho = RNG.random(n) + 0.1, RNG.random(n) * 0.1 + 0.08, RNG.random(n) + 0.5
    one = np.ones(n)
    cols = [rho, 2.5 * one, 1.3 * one, -0.7 * one, 0.2 * one]
    o = O.OracleSet(x, y, z, m, h, rho, cols)
    pts = RNG.random((3, 50)) * 0.6 + 0.2
    for strategy in (O.GATHER, O.SCATTER, O.SYMMETRIC):
        out = o.point_samples(FULL, strategy, *pts)
        assert np.allclose(out[1], 2.5, rtol=1e-12)
        assert np.max(np.abs(out[2:5])) < 1e-10   # grad of a constant
        assert np.max(np.abs(out[5])) < 1e-10     # div
        assert np.max(np.abs(out[6:9])) < 1e-10   # curl


def uniform_cloud(nx=12, eta=1.2, variable_h=True, kernel="M4_spline"):
    """test/interpolation_analytic_test_common.jl:31-72."""
    dx = 1.0 / nx
    c = np.arange(dx / 2, 1.0, dx)
    g = np.array(np.meshgrid(c, c, c, indexing="ij")).reshape(3, -1)
    x, y, z = g
    n = len(x)
    h = eta * dx * (1.0 + 0.18 * (x - 0.5) - 0.10 * (y - 0.5) + 0.08 * (z - 0.5)) if variable_h else np.full(n, eta * dx)
    q = x + y + z
    cols = [q, x, y, z]
    return O.OracleSet(x, y, z, np.full(n, dx ** 3), h, np.ones(n), cols, kernel=kernel)


def test_linear_field_regression(fixtures):
    """test/point_samples_interpolation_tests.jl:211-239 tolerances."""
    o = uniform_cloud()
    fx = fixtures["analytic_points"]
    px, py, pz = (np.array(fx[k]) for k in "xyz")
    cat = dict(sslots=[1], sflags=[1], gslots=[1], dslots=[2, 3, 4], cslots=[2, 3, 4])
    tol = fixtures["uniform_cloud"]["tol"]
    for strategy in (O.GATHER, O.SCATTER, O.SYMMETRIC):
        out = o.point_samples(cat, strategy, px, py, pz)
        assert np.allclose(out[0], px + py + pz, atol=tol["scalar"][0], rtol=tol["scalar"][1])
        assert np.allclose(out[1:4], 1.0, atol=tol["grad"][0], rtol=tol["grad"][1])
        assert np.allclose(out[4], 3.0, atol=tol["div"][0], rtol=tol["div"][1])
        assert np.allclose(out[5:8], 0.0, atol=tol["curl"][0], rtol=tol["curl"][1])


def test_grid_fixture_symmetric(fixtures):
    """test/point_samples_interpolation_tests.jl:53-79 on the 4-particle fixture."""
    fx = fixtures["grid_fixture"]
    a = {k: np.array(fx[k]) for k in ("x", "y", "z", "m", "h", "rho", "temp", "vx", "vy", "vz")}
    o = O.OracleSet(a["x"], a["y"], a["z"], a["m"], a["h"], a["rho"], [a["temp"], a["vx"], a["vy"], a["vz"]])
    c = o.columns()
    inp = dict(x=c[0], y=c[1], z=c[2], m=c[3], h=c[4], rho=c[5], cols=[None] + list(c[6:]))
    t = fixtures["point_template"]
    px, py, pz = (np.array(t[k]) for k in "xyz")
    out = o.point_samples(dict(sslots=[1], sflags=[1], dslots=[2, 3, 4]), O.SYMMETRIC, px, py, pz)
    for s in range(len(px)):
        p = (px[s], py[s], pz[s])
        ha = brute.nearest_h(inp, p)
        assert out[0, s] == pytest.approx(brute.quantity(inp, p, ha, 1, 0, 2), abs=1e-12, rel=1e-10)
        assert out[1, s] == pytest.approx(brute.divergence(inp, p, ha, (2, 3, 4), 0, 2), abs=1e-12, rel=1e-10)


def test_line_fixture_scatter(fixtures):
    """test/point_samples_interpolation_tests.jl:121-155 on the line fixture."""
    fx = fixtures["grid_fixture"]
    a = {k: np.array(fx[k]) for k in ("x", "y", "z", "m", "h", "rho", "temp", "vx")}
    o = O.OracleSet(a["x"], a["y"], a["z"], a["m"], a["h"], a["rho"], [a["temp"], a["vx"]])
    c = o.columns()
    inp = dict(x=c[0], y=c[1], z=c[2], m=c[3], h=c[4], rho=c[5], cols=[None] + list(c[6:]))
    lf = fixtures["line_fixture"]
    s2 = 1.0 / np.sqrt(2.0)
    conv = lambda v: [s2 if e == "invsqrt2" else e for e in v]  # noqa: E731
    origin = [np.array(v) for v in lf["origin"]]
    direction = [np.array(conv(lf["direction_x"])), np.array(lf["direction_y"]), np.array(conv(lf["direction_z"]))]
    out = o.line_samples(dict(sslots=[1, 2], sflags=[1, 1]), O.SCATTER, origin, direction)
    for s in range(3):
        o3 = [origin[d][s] for d in range(3)]
        m3 = [direction[d][s] for d in range(3)]
        for j, col in enumerate((1, 2)):
            exp = brute.line_quantity(inp, o3, m3, col, 0, O.line_W)
            assert out[j, s] == pytest.approx(exp, abs=1e-12, rel=1e-10) or (np.isnan(exp) and np.isnan(out[j, s]))


def test_column_density_random():
    """Σ = Σ_b m_b I(r_b/h_b)/h_b for an unnormalised rho column (C3 semantics)."""
    n = 3000
    x, y, z = RNG.random(n), RNG.random(n), RNG.random(n)
    m, h, rho = np.full(n, 1.0 / n), RNG.random(n) * 0.03 + 0.02, RNG.random(n) + 0.5
    o = O.OracleSet(x, y, z, m, h, rho, [rho])
    c = o.columns()
    inp = dict(x=c[0], y=c[1], z=c[2], m=c[3], h=c[4], rho=c[5], cols=[None, c[6]])
    S = 40
    ox, oy = RNG.random(S), RNG.random(S)
    out = o.line_samples(dict(sslots=[1], sflags=[0]), O.SCATTER, (ox, oy, np.zeros(S)),
                         (np.zeros(S), np.zeros(S), np.ones(S)))
    for s in range(S):
        exp = brute.line_quantity(inp, (ox[s], oy[s], 0.0), (0, 0, 1.0), 1, 0, O.line_W, shepard=False)
        assert out[0, s] == pytest.approx(exp, rel=1e-10) or (np.isnan(exp) and np.isnan(out[0, s]))


def test_float32_path_close_to_float64():
    """The Float32 instantiation follows the same algorithm: results within float32 rounding."""
    n = 2000
    x, y, z = RNG.random(n), RNG.random(n), RNG.random(n)
    m, h, rho = RNG.random(n) + 0.1, RNG.random(n) * 0.05 + 0.05, RNG.random(n) + 0.5
    A = np.sin(2 * np.pi * x)
    cols = [rho, A, x, y, -x]
    o64 = O.OracleSet(x, y, z, m, h, rho, cols)
    o32 = O.OracleSet(x, y, z, m, h, rho, cols, precision=32)
    pts = RNG.random((3, 64)) * 0.6 + 0.2
    a = o64.point_samples(dict(sslots=[1, 2], sflags=[0, 1]), O.SYMMETRIC, *pts)
    b = o32.point_samples(dict(sslots=[1, 2], sflags=[0, 1]), O.SYMMETRIC, *pts)
    assert np.allclose(a, b, rtol=2e-4, atol=1e-5)
