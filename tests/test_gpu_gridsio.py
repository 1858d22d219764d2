"""GridDataset files fed from device results: a StructuredGrid and a
PointSamples interpolation on the GPU, written with write_GridBundle
(IO/gridsIO/write_grids.jl:179-215) and read back with read_GridDataset
(IO/gridsIO/read_grids.jl:30-81); the values read equal the GPU's and the
oracle's, bit for bit."""
import numpy as np
import pytest

import pyoracle as O
from partia import (Cartesian, M4_spline, PointSamples, PointSamples_interpolation, StructuredGrid,
                    StructuredGrid_interpolation, build_input, itpGather, read_GridDataset, synthetic,
                    write_GridBundle)

pytestmark = pytest.mark.gpu

UNITS = dict(umass=1.0, udist=1.0, utime=1.0, umagfd=1.0)


def uniform_input(backend, n=20000):
    w = synthetic.uniform_box(n, seed=11)
    raw = {k: w[k].copy() for k in ("x", "y", "z", "m", "h", "rho")}
    inp, cat = build_input(backend, w["x"], w["y"], w["z"], w["m"], w["h"], w["rho"], (w["rho"], w["A"]),
                           column_names=("rho", "A"), scalars=("rho", "A"), smoothed_kernel=M4_spline)
    o = O.OracleSet(raw["x"], raw["y"], raw["z"], raw["m"], raw["h"], raw["rho"], [raw["rho"], w["A"]],
                    kernel="M4_spline")
    return inp, cat, o


def test_structured_grid_results_written_and_read(backend, tmp_path):
    inp, cat, o = uniform_input(backend)
    g = StructuredGrid(Cartesian, (0.05, 0.95, 20), (0.0, 1.0, 17), (0.1, 0.9, 9))
    bundle = StructuredGrid_interpolation(backend, Cartesian, g, inp, cat, itpGather)
    fn = str(tmp_path / "grid.h5")
    write_GridBundle(bundle, fn, code_units=UNITS, operation_name="gpu")
    gd = read_GridDataset(fn)
    assert gd.params["grid_type"] == "StructuredGrid" and gd.params["dtype"] == "Float64"
    assert gd.data.names == bundle.names
    X, Y, Z = np.meshgrid(*g.axes, indexing="ij")
    ref = o.point_samples(dict(sslots=[1, 2], sflags=[0, 1]), O.GATHER, X.ravel(order="F"), Y.ravel(order="F"),
                          Z.ravel(order="F"))
    for k, (a, b) in enumerate(zip(gd.data.grids, bundle.grids)):
        assert a.size == g.size
        assert np.array_equal(a.grid, b.grid, equal_nan=True)
        assert np.array_equal(a.grid.ravel(order="F"), ref[k], equal_nan=True)
        assert all(np.array_equal(x, y) for x, y in zip(a.axes, g.axes))


def test_point_samples_results_written_and_read(backend, tmp_path):
    inp, cat, o = uniform_input(backend)
    P = np.random.default_rng(12).random((3, 3000))
    tpl = PointSamples(np.zeros(P.shape[1]), tuple(P))
    bundle = PointSamples_interpolation(backend, tpl, inp, cat, itpGather, precision="fp32c")
    fn = str(tmp_path / "points.h5")
    write_GridBundle(bundle, fn, code_units=UNITS)
    gd = read_GridDataset(fn)
    # the host grids hold the input's float type whatever the evaluation precision
    assert gd.params["grid_type"] == "PointSamples" and gd.params["dtype"] == "Float64"
    for a, b in zip(gd.data.grids, bundle.grids):
        assert a.grid.dtype == np.float64 and np.array_equal(a.grid, b.grid, equal_nan=True)
        assert all(np.array_equal(x, y) for x, y in zip(a.coor, b.coor))
