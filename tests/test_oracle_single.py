"""The oracle's single-quantity kernels (_density_kernel, _number_density_kernel,
_gradient_density_kernel) and _line_integrated_density_kernel / the line
quantities kernel with a caller's h_a, against O(N) brute force written from the
reference tests' own references (test/interpolation_test_common.jl:88-186), at
the reference tests' points, h_a values and tolerances
(test/interpolation_tests.jl:52-110, 220-300)."""
import numpy as np
import pytest

import brute
import pyoracle as O

STRATS = [O.GATHER, O.SCATTER, O.SYMMETRIC]
KERNELS = ["M4_spline", "C4_Wendland", "M6_spline"]


def random_set(n, seed, kernel="M4_spline", hscale=0.12, hmin=0.04):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.random(n) + 0.05
    h = rng.random(n) * hscale + hmin
    rho = rng.random(n) + 0.3
    cols = [rho, rng.random(n), rng.random(n)]
    o = O.OracleSet(x, y, z, m, h, rho, cols, kernel=kernel)
    c = o.columns()
    inp = dict(x=c[0], y=c[1], z=c[2], m=c[3], h=c[4], rho=c[5], cols=[None] + list(c[6:]))
    return o, inp


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("strategy", STRATS)
def test_density_and_number_density_equal_brute_force(kernel, strategy):
    """interpolation_tests.jl:220-245: 200 particles, point (0.4, 0.35, 0.25), h_a = 0.12."""
    o, inp = random_set(200, 11, kernel)
    K = O.KERNELS[kernel]
    p = np.array([0.4, 0.35, 0.25])
    ha = np.array([0.12])
    rho = o.point_single(0, strategy, [p[0]], [p[1]], [p[2]], ha=ha)[0, 0]
    n = o.point_single(1, strategy, [p[0]], [p[1]], [p[2]], ha=ha)[0, 0]
    assert rho == pytest.approx(brute.density(inp, p, 0.12, K, strategy), rel=1e-8, abs=1e-10)
    assert n == pytest.approx(brute.number_density(inp, p, 0.12, K, strategy), rel=1e-8, abs=1e-10)
    assert rho > 0 and n > 0


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("strategy", STRATS)
def test_gradient_density_equals_brute_force(kernel, strategy):
    """interpolation_tests.jl:250-276: 80 particles, two points, h_a 0.05 and 0.12;
    the mean term uses the unnormalised rho = sum m W (gradient_interpolation.jl:47-55)."""
    o, inp = random_set(80, 12, kernel)
    K = O.KERNELS[kernel]
    pts = np.array([[0.2, 0.3, 0.4], [0.7, 0.2, 0.1]])
    for ha in (0.05, 0.12):
        got = o.point_single(2, strategy, pts[:, 0], pts[:, 1], pts[:, 2], ha=np.full(2, ha))
        for s, p in enumerate(pts):
            exp = np.array(brute.gradient_density(inp, p, ha, K, strategy))
            if np.isnan(exp).any():
                assert np.all(np.isnan(got[:, s]))
            else:
                assert np.allclose(got[:, s], exp, rtol=1e-8, atol=5e-10)


def test_gradient_density_differs_from_the_fused_density_gradient():
    """The fused kernel's density gradient (catalog gradient slot 0) subtracts the
    Shepard mean sum m W / S1 (general_interpolation.jl:148-169); the
    single-quantity kernel subtracts the unnormalised sum m W: both match their
    own brute references and differ from each other."""
    o, inp = random_set(300, 13)
    p = np.array([[0.5], [0.45], [0.55]])
    ha = brute.nearest_h(inp, p[:, 0])
    single = o.point_single(2, O.SYMMETRIC, *p)[:, 0]
    fused = o.point_samples(dict(gslots=[0]), O.SYMMETRIC, *p)[:, 0]
    sel, w, g = brute._weights(inp, p[:, 0], ha, 0, O.SYMMETRIC, True)
    m, rb = inp["m"][sel], inp["rho"][sel]
    shepard_mean = np.sum(m * w) / np.sum(m / rb * w)
    exp_fused = [np.sum(m * gd) - shepard_mean * np.sum(m / rb * gd) for gd in g]
    assert np.allclose(fused, exp_fused, rtol=1e-8, atol=5e-10)
    assert np.allclose(single, brute.gradient_density(inp, p[:, 0], ha, 0, O.SYMMETRIC), rtol=1e-8, atol=5e-10)
    assert np.max(np.abs(single - fused)) > 1e-3 * np.max(np.abs(fused))


def test_no_neighbour_gives_zero_density_and_nan_gradient():
    """interpolation_tests.jl:84-110: 20 particles, far point (2, 2, 2), h_a = 1e-4."""
    o, _ = random_set(20, 14)
    for strategy in STRATS:
        args = ([2.0], [2.0], [2.0])
        assert o.point_single(0, strategy, *args, ha=[1e-4])[0, 0] == 0.0
        assert o.point_single(1, strategy, *args, ha=[1e-4])[0, 0] == 0.0
        assert np.all(np.isnan(o.point_single(2, strategy, *args, ha=[1e-4])))


def test_nearest_h_is_the_default_h_a():
    o, inp = random_set(400, 15)
    P = np.random.default_rng(1).random((3, 40))
    ha = np.array([brute.nearest_h(inp, P[:, s]) for s in range(40)])
    for kind in (0, 1, 2):
        a = o.point_single(kind, O.SYMMETRIC, *P)
        b = o.point_single(kind, O.SYMMETRIC, *P, ha=ha)
        assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


@pytest.mark.parametrize("strategy", STRATS)
def test_line_density_and_quantity_with_h_a_equal_brute_force(strategy):
    """interpolation_tests.jl:281-300: 150 particles, origin (0.2, 0.8, 0), +z, h_a = 0.08."""
    o, inp = random_set(150, 16)
    org = ([0.2], [0.8], [0.0])
    dirs = ([0.0], [0.0], [1.0])
    sigma = o.line_density(strategy, org, dirs, ha=[0.08])[0]
    exp = brute.line_density(inp, (0.2, 0.8, 0.0), (0.0, 0.0, 1.0), 0.08, 0, O.line_W, strategy)
    assert sigma == pytest.approx(exp, rel=1e-8, abs=1e-10) and sigma > 0
    q = o.line_samples(dict(sslots=[2], sflags=[1]), strategy, org, dirs, ha=[0.08])[0, 0]
    qe = brute.line_quantity_ha(inp, (0.2, 0.8, 0.0), (0.0, 0.0, 1.0), 0.08, 2, 0, O.line_W, strategy)
    assert q == pytest.approx(qe, rel=1e-8, abs=1e-10)


def test_line_density_without_neighbours_is_zero():
    o, _ = random_set(30, 17)
    for strategy in STRATS:
        assert o.line_density(strategy, ([5.0], [5.0], [0.0]), ([0.0], [0.0], [1.0]), ha=[1e-3])[0] == 0.0
