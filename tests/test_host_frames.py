"""The reference's frame-plane constructor tests on the host mirror
(test/structured_grid_interpolation_tests.jl:110-254): PointSamples(Cartesian /
Polar, frame, ...) against the flattened 2D StructuredGrid, box corners, the
polar plane's closed form, LineSamples(..., ParallelBeam, frame, ...) against the
PointSamples plane and the frame's forward vector, and the invalid plane
parameters.  Frames and parameters come from tests/golden/frame_fixtures.json."""
import json
import math
import os

import numpy as np
import pytest

from partia import (ArgumentError, Cartesian, Frame, LineSamples, ParallelBeam, PointSamples, Polar, StructuredGrid,
                    flatten, frame_forward, frame_right, frame_up)

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "frame_fixtures.json")))
TOL = dict(atol=1e-12, rtol=1e-12)


def frame(name):
    f = FIX["frames"][name]
    return Frame(f["x"], f["f0"], f["u0"])


def params(key):
    lo, hi, n = FIX[key]
    return (lo, hi, int(n))


@pytest.mark.parametrize("name", ["origin", "offset"])
def test_frame_basis_matches_the_fixture(name):
    f, fx = frame(name), FIX["frames"][name]
    assert np.allclose(frame_right(f), fx["right"], atol=0, rtol=0)
    assert np.allclose(frame_up(f), fx["up"], atol=0, rtol=0)
    assert np.allclose(frame_forward(f), fx["forward"], atol=0, rtol=0)
    with pytest.raises(ArgumentError, match="orthogonal"):
        Frame((0, 0, 0), (0, 0, 1), (0, 1, 1))
    with pytest.raises(ArgumentError, match="nonzero"):
        Frame((0, 0, 0), (0, 0, 0), (0, 1, 0))


def test_cartesian_plane_and_box_match_the_flattened_grid():
    """:110-135."""
    fr = frame("origin")
    xp, yp, zp = params("xparams"), params("yparams"), params("zparams")
    flat = flatten(Cartesian, StructuredGrid(Cartesian, xp, yp))
    plane = PointSamples.from_frame(Cartesian, fr, xp, yp)
    assert np.array_equal(plane.grid, np.zeros(len(flat.grid)))
    assert np.allclose(plane.coor[0], flat.coor[0], **TOL)
    assert np.allclose(plane.coor[1], flat.coor[1], **TOL)
    assert np.allclose(plane.coor[2], 0.0, **TOL)
    box = PointSamples.from_frame(Cartesian, fr, xp, yp, zp)
    assert len(box.grid) == xp[2] * yp[2] * zp[2]
    assert np.array_equal(box.grid, np.zeros(len(box.grid)))
    for a, lo, hi in ((0, xp[0], xp[1]), (1, yp[0], yp[1]), (2, -zp[0], -zp[1])):
        assert np.isclose(box.coor[a][0], lo, **TOL) and np.isclose(box.coor[a][-1], hi, **TOL)


def test_polar_plane_matches_the_flattened_grid_and_its_closed_form():
    """:137-197."""
    fr = frame("origin")
    sp, pp, qp = params("sparams"), params("phiparams"), params("partial_phiparams")
    smin, smax, ns = sp
    nphi = pp[2]
    flat = flatten(Polar, StructuredGrid(Polar, sp, pp))
    plane = PointSamples.from_frame(Polar, fr, sp, pp)
    assert np.array_equal(plane.grid, np.zeros(len(flat.grid)))
    assert np.allclose(plane.coor[0], flat.coor[0], **TOL)
    assert np.allclose(plane.coor[1], flat.coor[1], **TOL)
    assert np.allclose(plane.coor[2], 0.0, **TOL)
    pflat = flatten(Polar, StructuredGrid(Polar, sp, qp))
    pplane = PointSamples.from_frame(Polar, fr, sp, qp)
    assert np.allclose(pplane.coor[0], pflat.coor[0], **TOL)
    assert np.allclose(pplane.coor[1], pflat.coor[1], **TOL)
    assert np.allclose(pplane.coor[2], 0.0, **TOL)
    s = np.linspace(smin, smax, ns)
    ang = np.linspace(0.0, 2 * math.pi, nphi + 1)[:-1]
    ex = np.array([s[i] * math.cos(ang[j]) for j in range(nphi) for i in range(ns)])
    ey = np.array([s[i] * math.sin(ang[j]) for j in range(nphi) for i in range(ns)])
    assert np.allclose(plane.coor[0], ex, **TOL) and np.allclose(plane.coor[1], ey, **TOL)
    assert np.allclose(plane.coor[2], np.zeros(ns * nphi), **TOL)
    assert len(plane.grid) == ns * nphi
    assert np.allclose(plane.coor[0][:ns], s, **TOL) and np.allclose(plane.coor[1][:ns], 0.0, **TOL)
    outer = np.arange(ns - 1, ns * nphi, ns)
    assert np.allclose(np.hypot(plane.coor[0][outer], plane.coor[1][outer]), smax, **TOL)
    centre = np.arange(0, ns * nphi, ns)
    for a in range(3):
        assert np.allclose(plane.coor[a][centre], 0.0, **TOL)


def test_parallel_beam_lines_match_the_point_planes():
    """:202-239."""
    fr = frame("offset")
    fx, fy, fz = frame_forward(fr)
    for coord, p1, p2 in ((Cartesian, params("xparams"), params("yparams")),
                          (Polar, params("sparams"), params("phiparams"))):
        pts = PointSamples.from_frame(coord, fr, p1, p2)
        lines = LineSamples.from_frame(coord, ParallelBeam, fr, p1, p2)
        assert np.array_equal(lines.grid, np.zeros(len(pts.grid)))
        for a in range(3):
            assert np.allclose(lines.origin[a], pts.coor[a], **TOL)
        for a, v in enumerate((fx, fy, fz)):
            assert np.allclose(lines.direction[a], v, **TOL)
        if coord is Polar:
            n = p1[2] * p2[2]
            assert len(lines.grid) == n and len(lines.origin[0]) == n and len(lines.direction[0]) == n


def test_invalid_plane_parameters():
    """:241-254."""
    fr = frame("origin")
    for xp, yp in FIX["invalid_cartesian"]:
        with pytest.raises(ArgumentError):
            LineSamples.from_frame(Cartesian, ParallelBeam, fr, tuple(xp[:2]) + (int(xp[2]),),
                                   tuple(yp[:2]) + (int(yp[2]),))
    for sp, pp in FIX["invalid_polar"]:
        with pytest.raises(ArgumentError):
            LineSamples.from_frame(Polar, ParallelBeam, fr, tuple(sp[:2]) + (int(sp[2]),),
                                   tuple(pp[:2]) + (int(pp[2]),))
