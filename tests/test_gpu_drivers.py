"""The reference's driver-level tests, run through the Python mirror of its API
on the HIP backend (test/point_samples_interpolation_tests.jl,
test/structured_grid_interpolation_tests.jl, test/interpolation_tests.jl:52-117)."""
import copy

import numpy as np
import pytest

import brute
import pyoracle as O
from partia import (ArgumentError, C4_Wendland, Cartesian, Cylindrical, InterpolationInput, LineSamples,
                    LineSamples_interpolation, LinearBVH_, M4_spline, MethodError, PointSamples,
                    PointSamples_interpolation, Spherical, StructuredGrid, StructuredGrid_interpolation, UInt32,
                    build_input, div_slots, flatten, initialize_interpolation, itpGather, itpScatter, itpSymmetric,
                    matches_lbvh_leaf_order, scalar_index, synthetic)
from partia.catalog import InterpolationCatalog

pytestmark = pytest.mark.gpu


def grid_fixture(fixtures, backend):
    """test/grid_interpolation_test_common.jl:40-64."""
    f = {k: np.array(v) for k, v in fixtures["grid_fixture"].items() if isinstance(v, list)}
    inp = InterpolationInput((f["x"], f["y"], f["z"]), f["m"], f["h"], f["rho"], (f["temp"], f["vx"], f["vy"], f["vz"]),
                             smoothed_kernel=M4_spline)
    cat = InterpolationCatalog.from_columns(("temp", "vx", "vy", "vz"), 3, scalars=("temp",), divergences=("v",))
    lbvh = LinearBVH_(inp, backend=backend)
    return inp, cat, lbvh


def point_template(fixtures):
    t = fixtures["point_template"]
    return PointSamples(np.zeros(3), tuple(np.array(t[k]) for k in "xyz"))


def as_brute(inp):
    return dict(x=inp.coord[0], y=inp.coord[1], z=inp.coord[2], m=inp.m, h=inp.h, rho=inp.ρ,
                cols=[None] + list(inp.quant))


def permute(inp, p):
    q = copy.deepcopy(inp)
    for a in list(q.coord) + [q.m, q.h, q.ρ] + list(q.quant):
        a[:] = a[p]
    return q


def test_leaf_order_matching(fixtures, backend):
    inp, cat, lbvh = grid_fixture(fixtures, backend)
    assert matches_lbvh_leaf_order(inp, lbvh)
    assert not matches_lbvh_leaf_order(permute(inp, [1, 0, 2, 3]), lbvh)


def test_point_samples_cpu_consistency(fixtures, backend):
    inp, cat, _ = grid_fixture(fixtures, backend)
    tpl = point_template(fixtures)
    res = PointSamples_interpolation(backend, tpl, inp, cat, itpSymmetric)
    assert res.names == cat.ordered_names and len(res.grids) == 2
    assert res.grids[0].coor is res.grids[1].coor
    b = as_brute(inp)
    for i in range(3):
        p = tuple(tpl.coor[d][i] for d in range(3))
        ha = brute.nearest_h(b, p)
        exp_s = brute.quantity(b, p, ha, scalar_index(cat, "temp"), 0, 2)
        exp_d = brute.divergence(b, p, ha, div_slots(cat, "v"), 0, 2)
        assert res.grids[0].grid[i] == pytest.approx(exp_s, abs=1e-12, rel=1e-10)
        assert res.grids[1].grid[i] == pytest.approx(exp_d, abs=1e-12, rel=1e-10)


def test_external_lbvh_overload(fixtures, backend):
    inp, cat, lbvh = grid_fixture(fixtures, backend)
    tpl = point_template(fixtures)
    auto = PointSamples_interpolation(backend, tpl, copy.deepcopy(inp), cat, itpSymmetric)
    manual = PointSamples_interpolation(backend, tpl, inp, lbvh, cat, itpSymmetric)
    assert manual.names == auto.names
    for a, m in zip(auto.grids, manual.grids):
        assert np.array_equal(a.grid, m.grid)
    with pytest.raises(ArgumentError):
        PointSamples_interpolation(backend, tpl, permute(inp, [1, 0, 2, 3]), lbvh, cat, itpSymmetric)


def line_fixture(fixtures, backend):
    f = {k: np.array(v) for k, v in fixtures["grid_fixture"].items() if isinstance(v, list)}
    inp = InterpolationInput((f["x"], f["y"], f["z"]), f["m"], f["h"], f["rho"], (f["temp"], f["vx"]),
                             smoothed_kernel=M4_spline)
    cat = InterpolationCatalog.from_columns(("temp", "vx"), 3, scalars=("temp", "vx"))
    lf = fixtures["line_fixture"]
    s2 = 1.0 / np.sqrt(2.0)
    conv = lambda v: np.array([s2 if e == "invsqrt2" else e for e in v])  # noqa: E731
    tpl = LineSamples(np.zeros(3), tuple(np.array(v) for v in lf["origin"]),
                      (conv(lf["direction_x"]), np.array(lf["direction_y"]), conv(lf["direction_z"])))
    return inp, cat, tpl


def test_line_samples_scatter(fixtures, backend):
    inp, cat, tpl = line_fixture(fixtures, backend)
    res = LineSamples_interpolation(backend, tpl, inp, cat, itpScatter)
    assert res.names == cat.ordered_names
    assert res.grids[0].origin is res.grids[1].origin
    b = as_brute(inp)
    for i in range(3):
        o = [tpl.origin[d][i] for d in range(3)]
        m = [tpl.direction[d][i] for d in range(3)]
        for j, col in enumerate(cat.scalar_slots):
            exp = brute.line_quantity(b, o, m, col, 0, O.line_W)
            got = res.grids[j].grid[i]
            assert (np.isnan(exp) and np.isnan(got)) or got == pytest.approx(exp, abs=1e-12, rel=1e-10)


def test_line_samples_external_and_errors(fixtures, backend):
    inp, cat, tpl = line_fixture(fixtures, backend)
    lbvh = LinearBVH_(inp, backend=backend)
    auto = LineSamples_interpolation(backend, tpl, copy.deepcopy(inp), cat, itpScatter)
    manual = LineSamples_interpolation(backend, tpl, inp, lbvh, cat, itpScatter)
    for a, m in zip(auto.grids, manual.grids):
        assert np.array_equal(a.grid, m.grid, equal_nan=True)
    with pytest.raises(ArgumentError):
        LineSamples_interpolation(backend, tpl, permute(inp, [1, 0, 2, 3]), lbvh, cat, itpScatter)
    for strategy in (itpGather, itpSymmetric):
        with pytest.raises(ArgumentError):
            LineSamples_interpolation(backend, tpl, inp, cat, strategy)
    pinp, pcat, _ = grid_fixture(fixtures, backend)
    with pytest.raises(MethodError):
        LineSamples_interpolation(backend, tpl, pinp, pcat, itpScatter)


def uniform_cloud(nx=12, eta=1.2):
    dx = 1.0 / nx
    c = np.arange(dx / 2, 1.0, dx)
    x, y, z = (a.ravel() for a in np.meshgrid(c, c, c, indexing="ij"))
    n = len(x)
    h = eta * dx * (1.0 + 0.18 * (x - 0.5) - 0.10 * (y - 0.5) + 0.08 * (z - 0.5))
    inp = InterpolationInput((x, y, z), np.full(n, dx ** 3), h, np.ones(n), (x + y + z, x.copy(), y.copy(), z.copy()),
                             smoothed_kernel=M4_spline)
    cat = InterpolationCatalog.from_columns(("q", "vx", "vy", "vz"), 3, scalars=("q",), gradients=("q",),
                                            divergences=("v",), curls=("v",))
    return inp, cat


def test_analytic_linear_field(fixtures, backend):
    fx = fixtures["analytic_points"]
    tpl = PointSamples(np.zeros(3), tuple(np.array(fx[k]) for k in "xyz"))
    tol = fixtures["uniform_cloud"]["tol"]
    for strategy in (itpGather, itpScatter, itpSymmetric):
        inp, cat = uniform_cloud()
        res = PointSamples_interpolation(backend, tpl, inp, cat, strategy)
        x, y, z = tpl.coor
        assert np.allclose(res.grids[0].grid, x + y + z, atol=tol["scalar"][0], rtol=tol["scalar"][1])
        for k in (1, 2, 3):
            assert np.allclose(res.grids[k].grid, 1.0, atol=tol["grad"][0], rtol=tol["grad"][1])
        assert np.allclose(res.grids[4].grid, 3.0, atol=tol["div"][0], rtol=tol["div"][1])
        for k in (5, 6, 7):
            assert np.allclose(res.grids[k].grid, 0.0, atol=tol["curl"][0], rtol=tol["curl"][1])


def test_structured_grid_equals_flattened_points(fixtures, backend):
    """test/structured_grid_interpolation_tests.jl:256-321 (here bit-identical)."""
    p = [tuple(q) for q in fixtures["structured_template"]["params"]]
    g = StructuredGrid(Cartesian, *p)
    inp, cat, _ = grid_fixture(fixtures, backend)
    sres = StructuredGrid_interpolation(backend, Cartesian, g, copy.deepcopy(inp), cat, itpSymmetric)
    pres = PointSamples_interpolation(backend, flatten(Cartesian, g), inp, cat, itpSymmetric)
    for sg, pg in zip(sres.grids, pres.grids):
        assert sg.size == g.size and sg.axes[0] is not None
        assert np.array_equal(sg.grid.ravel(order="F"), pg.grid, equal_nan=True)


@pytest.mark.parametrize("coord,params", [
    (Cartesian, [(0.35, 0.55, 3)] * 3),
    (Cylindrical, [(0.35, 0.60, 3), (0.55, 0.95, 3), (0.25, 0.45, 3)]),
    (Spherical, [(0.55, 0.75, 3), (0.45, 0.75, 3), (0.85, 1.05, 3)]),
])
def test_structured_grid_analytic(backend, coord, params):
    """test/structured_grid_interpolation_tests.jl:325-370."""
    g = StructuredGrid(coord, *params)
    inp, cat = uniform_cloud()
    res = StructuredGrid_interpolation(backend, coord, g, inp, cat, itpSymmetric)
    X, Y, Z = flatten(coord, g).coor
    assert np.allclose(res.grids[0].grid.ravel(order="F"), X + Y + Z, atol=2e-2, rtol=1e-2)
    assert np.allclose(res.grids[4].grid.ravel(order="F"), 3.0, atol=1e-1, rtol=2e-2)


@pytest.mark.parametrize("coord,params", [
    (Cylindrical, [(0.0, 0.45, 9), (0.0, 2 * np.pi, 13), (0.2, 0.8, 6)]),
    (Spherical, [(0.05, 0.45, 7), (0.3, 6.0, 11), (0.1, 3.0, 5)]),
])
def test_structured_grid_coords_generated_on_the_device(backend, coord, params):
    """Cylindrical and spherical grids generated on the device
    (ph_structured_grid_coords: _cylin2cart / _sph2cart with the axes' sin /
    cos pairs) equal point samples on the same coordinates computed on the host
    in the reference's order, bit for bit, and match coordinate_grid's
    closed form."""
    from partia import grids as G
    g = StructuredGrid(coord, *params)
    inp, cat = uniform_cloud()
    res = StructuredGrid_interpolation(backend, coord, g, copy.deepcopy(inp), cat, itpSymmetric)
    xyz = G.coordinates_from_pairs(coord, g)  # the part in x, y > 0 lies in the cloud
    for a, b in zip(xyz, G.coordinate_grid(coord, g)):
        np.testing.assert_allclose(a, b, atol=1e-12)
    pts = PointSamples(np.zeros(len(xyz[0])), tuple(np.asarray(c) for c in xyz))
    pres = PointSamples_interpolation(backend, pts, inp, cat, itpSymmetric)
    for sg, pg in zip(res.grids, pres.grids):
        assert sg.size == g.size
        assert np.array_equal(sg.grid.ravel(order="F"), pg.grid, equal_nan=True)
    assert np.isfinite(res.grids[0].grid).mean() > 0.2


def test_empty_input_all_nan(fixtures, backend):
    fx = fixtures["empty_input"]
    inp = InterpolationInput(tuple(np.array(fx[k]) for k in "xyz"), np.array(fx["m"]), np.array(fx["h"]),
                             np.array(fx["rho"]), (np.array(fx["q1"]),) * 4, smoothed_kernel=M4_spline)
    cat = InterpolationCatalog.from_columns(("a", "vx", "vy", "vz"), 3, scalars=("a",), gradients=("a",),
                                            divergences=("v",), curls=("v",))
    tpl = PointSamples(np.zeros(1), tuple(np.array([v]) for v in fx["reference_point"]))
    for strategy in (itpGather, itpScatter, itpSymmetric):
        res = PointSamples_interpolation(backend, tpl, copy.deepcopy(inp), cat, strategy)
        assert all(np.isnan(gr.grid[0]) for gr in res.grids)


def test_linear_bvh_permutes_input_like_the_oracle(backend):
    w = synthetic.plummer_sphere(30000, seed=5)
    raw = {k: w[k].copy() for k in w}
    inp, cat = build_input(backend, w["x"], w["y"], w["z"], w["m"], w["h"], w["rho"], (w["A"], w["vx"]),
                           column_names=("A", "vx"), scalars=("A",), smoothed_kernel=C4_Wendland)
    lb = LinearBVH_(inp, UInt32, backend)
    o = O.OracleSet(raw["x"], raw["y"], raw["z"], raw["m"], raw["h"], raw["rho"], (raw["A"], raw["vx"]),
                    kernel="C4_Wendland", code_bits=30)
    cols = o.columns()
    mine = [inp.coord[0], inp.coord[1], inp.coord[2], inp.m, inp.h, inp.ρ, inp.quant[0], inp.quant[1]]
    for a, b in zip(mine, cols):
        assert np.array_equal(a, b)
    assert np.array_equal(lb.brt.left, o.tree()["left"])
    assert lb.code_bits == 30 and matches_lbvh_leaf_order(inp, lb)


def test_initialize_interpolation_returns_reference_tuple(fixtures, backend):
    inp, cat, _ = grid_fixture(fixtures, backend)
    tpl = point_template(fixtures)
    grids, lbvh, names, conc = initialize_interpolation(backend, tpl, inp, cat)
    assert len(grids) == cat.L and names == cat.ordered_names and conc.scalar_slots == cat.scalar_slots
    assert lbvh.n == 4 and lbvh.brt.root == 1
    grids2, names2, _ = initialize_interpolation(backend, tpl, cat)
    assert len(grids2) == cat.L
