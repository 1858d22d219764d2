"""Host restatements of the frame-based sample generators
(Grids/grids/coordinate.jl:44-276) that the batch tests compare the device
generator with: the reference test's closed forms
(test/structured_grid_interpolation_tests.jl:136-175: s cos(phi), s sin(phi)
at 1e-12 on the identity frame), sample order, frames and the ArgumentError
rules."""
import math

import numpy as np
import pytest

from partia import ArgumentError
from partia import grids

ID = ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))


def test_polar_plane_closed_form_and_order():
    for sp, pp in (((0.0, 1.0, 4), (0.0, 2 * math.pi, 8)), ((0.0, 1.0, 4), (math.pi / 4, 3 * math.pi / 4, 5)),
                   ((0.5, 2.0, 7), (0.3, 0.4, 1))):
        x, y, z = grids.polar_plane_coordinates(*ID[:3], sp, pp)
        s = np.linspace(sp[0], sp[1], sp[2])
        phi = pp[0] + np.arange(pp[2]) * ((pp[1] - pp[0]) / pp[2])  # half-open
        ex = np.concatenate([s * math.cos(a) for a in phi])  # s fastest
        ey = np.concatenate([s * math.sin(a) for a in phi])
        np.testing.assert_allclose(x, ex, atol=1e-12, rtol=1e-12)
        np.testing.assert_allclose(y, ey, atol=1e-12, rtol=1e-12)
        assert np.all(np.abs(z) <= 1e-12)


def test_cylinder_is_the_polar_plane_stacked_along_forward():
    rng = np.random.default_rng(3)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    pos, r, u, f = rng.random(3), q[:, 0], q[:, 1], q[:, 2]
    sp, pp, zp = (0.1, 0.9, 5), (0.0, 2 * math.pi, 6), (-0.3, 0.4, 3)
    cyl = grids.cylindrical_coordinates(pos, r, u, f, sp, pp, zp)
    n = sp[2] * pp[2]
    zeta = np.linspace(zp[0], zp[1], zp[2])
    for k3 in range(zp[2]):
        plane = grids.polar_plane_coordinates(pos + zeta[k3] * f, r, u, sp, pp)
        for a in range(3):
            np.testing.assert_allclose(cyl[a][k3 * n:(k3 + 1) * n], plane[a], atol=1e-12)
    # every point lies at radius s from the axis and height zeta along it
    d = np.stack(cyl, axis=1) - pos
    h = d @ f
    rad = np.linalg.norm(d - np.outer(h, f), axis=1)
    s = np.tile(np.linspace(sp[0], sp[1], sp[2]), pp[2] * zp[2])
    np.testing.assert_allclose(rad, s, atol=1e-12)
    np.testing.assert_allclose(h, np.repeat(zeta, n), atol=1e-12)


def test_an_explicit_angle_table_is_used_as_given():
    pp = (0.0, 1.0, 3)
    t = np.array([[0.0, 1.0], [1.0, 0.0], [0.0, -1.0]])  # not the angles' sin / cos
    x, y, _ = grids.polar_plane_coordinates(*ID[:3], (1.0, 2.0, 2), pp, trig=t)
    np.testing.assert_array_equal(x, [1.0, 2.0, 0.0, 0.0, -1.0, -2.0])
    np.testing.assert_array_equal(y, [0.0, 0.0, 1.0, 2.0, 0.0, 0.0])


@pytest.mark.parametrize("sp,pp,zp,msg", [
    ((-0.1, 1.0, 3), (0.0, 1.0, 2), (0.0, 1.0, 2), "smin must be nonnegative."),
    ((0.5, 0.5, 3), (0.0, 1.0, 2), (0.0, 1.0, 2), "smax must be greater than smin."),
    ((0.0, 1.0, 3), (0.0, 7.0, 2), (0.0, 1.0, 2), "angular range"),
    ((0.0, 1.0, 3), (1.0, 1.0, 2), (0.0, 1.0, 2), "angular range"),
    ((0.0, 1.0, 1), (0.0, 1.0, 2), (0.0, 1.0, 2), "ns must be at least 2."),
    ((0.0, 1.0, 3), (0.0, 1.0, 0), (0.0, 1.0, 2), "must be at least 1."),
    ((0.0, 1.0, 3), (0.0, 1.0, 2), (1.0, 1.0, 2), "zmax must be greater than zmin."),
    ((0.0, 1.0, 3), (0.0, 1.0, 2), (0.0, 1.0, 1), "nz must be at least 2."),
])
def test_argument_errors(sp, pp, zp, msg):
    with pytest.raises(ArgumentError, match=msg.replace(".", r"\.")):
        grids.cylindrical_coordinates(*ID, sp, pp, zp)


def test_phimax_may_round_to_two_pi():
    x, _, _ = grids.polar_plane_coordinates(*ID[:3], (0.0, 1.0, 2), (0.0, 2 * math.pi * (1 + 1e-10), 4))
    assert len(x) == 8
