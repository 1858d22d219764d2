"""The reference's frame-plane tests (test/structured_grid_interpolation_tests.jl:
110-254) through the device generators: the frames and axis parameters of
tests/golden/frame_fixtures.json as batch requests (PointSamples(Cartesian,
frame, ...) planes and boxes, PointSamples(Polar, frame, ...) planes), as the
image-plane rays of ph_line_grid_device and the polar rays of
ph_line_polar_device (LineSamples(..., ParallelBeam, frame, ...)).  Each is
bit-identical to the explicit-coordinate call on the host mirror's
PointSamples.from_frame / LineSamples.from_frame coordinates (which
test_host_frames.py pins to the reference's expectations), and the invalid
plane parameters are rejected."""
import json
import os

import numpy as np
import pytest

from gpu_util import DevArray as Dev, DeviceSet, assert_same
from partia import ArgumentError, Cartesian, Frame, LineSamples, ParallelBeam, PointSamples, Polar
from partia import _native as N
from partia import batch as PB
from partia import grids

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "frame_fixtures.json")))
CAT = dict(sslots=[1, 2], sflags=[0, 1], gslots=[2], dslots=[3, 4, 5], cslots=[3, 4, 5])


def frame(name):
    f = FIX["frames"][name]
    return Frame(f["x"], f["f0"], f["u0"])


def params(key):
    lo, hi, n = FIX[key]
    return (lo, hi, int(n))


def cloud(n, seed, centre):
    """Particles around the frame planes (|x| <= 1.3, |y| <= 1.1, |z - c| <= 0.8)."""
    rng = np.random.default_rng(seed)
    x = rng.random(n) * 2.6 - 1.3 + centre[0]
    y = rng.random(n) * 2.2 - 1.1 + centre[1]
    z = rng.random(n) * 1.6 - 0.8 + centre[2]
    m = rng.random(n) + 0.05
    h = (rng.random(n) * 0.12 + 0.04) * (200.0 / n) ** (1 / 3) * 2.0
    rho = rng.random(n) + 0.3
    cols = [rho, np.sin(2 * x) * np.cos(2 * y), x + 0.3 * y, y - z, z * x]
    return (x, y, z, m, h, rho), cols


@pytest.mark.parametrize("name", ["origin", "offset"])
@pytest.mark.parametrize("precision", [N.PH_PREC_FP64, N.PH_PREC_FP32C])
def test_frame_requests_match_the_host_planes(backend, name, precision):
    ctx = backend.context
    fr = frame(name)
    c6, cols = cloud(20000, 3, fr.position())
    d = DeviceSet(ctx, *c6, cols, kernel=4)
    ca = N.CatalogArgs(**CAT)
    xp, yp, zp = params("xparams"), params("yparams"), params("zparams")
    sp, pp, qp = params("sparams"), params("phiparams"), params("partial_phiparams")
    pos, r, u, f = fr.position(), fr.right(), fr.up(), fr.forward()
    specs = [(PB.plane(pos, r, u, xp, yp, None), PointSamples.from_frame(Cartesian, fr, xp, yp)),
             (PB.box(pos, r, u, f, xp, yp, zp, None), PointSamples.from_frame(Cartesian, fr, xp, yp, zp)),
             (PB.polar(pos, r, u, sp, pp, None), PointSamples.from_frame(Polar, fr, sp, pp)),
             (PB.polar(pos, r, u, sp, qp, None), PointSamples.from_frame(Polar, fr, sp, qp))]
    esz = 8 if precision == N.PH_PREC_FP64 else 4
    outs = []
    try:
        reqs = []
        for req, ps in specs:
            o = Dev(ctx, nbytes=ca.L * len(ps.grid) * esz)
            req.out = o.p.value
            reqs.append(req)
            outs.append(o)
        for strategy in (N.PH_GATHER, N.PH_SYMMETRIC, N.PH_SCATTER):
            b = PB.Batch(ctx, d.h, ca, strategy, precision, reqs)
            try:
                b.run()
                for (_, ps), o in zip(specs, outs):
                    S = len(ps.grid)
                    got = o.get(np.float64 if esz == 8 else np.float32, ca.L * S).astype(np.float64)
                    ref = d.points(CAT, strategy, *ps.coor, precision=precision)
                    assert_same(got.reshape(ca.L, S), ref)
                    assert np.isfinite(ref).any()
            finally:
                b.close()
    finally:
        for o in outs:
            o.free()


@pytest.mark.parametrize("name", ["origin", "offset"])
def test_parallel_beam_rays_match_the_host_lines(backend, name):
    ctx = backend.context
    fr = frame(name)
    c6, cols = cloud(20000, 5, fr.position())
    d = DeviceSet(ctx, *c6, cols, kernel=0)
    cat = dict(sslots=[1, 2], sflags=[0, 1])
    ca = N.CatalogArgs(**cat)
    f64 = [np.ascontiguousarray(np.asarray(v, dtype=np.float64))
           for v in (fr.position(), fr.right(), fr.up(), fr.forward())]
    # Cartesian: the image-plane entry point with the plane's axis values
    xp, yp = params("xparams"), params("yparams")
    ax = xp[0] + np.arange(xp[2]) * ((xp[1] - xp[0]) / (xp[2] - 1))
    ay = yp[0] + np.arange(yp[2]) * ((yp[1] - yp[0]) / (yp[2] - 1))
    ls = LineSamples.from_frame(Cartesian, ParallelBeam, fr, xp, yp)
    S = len(ls.grid)
    dax, day, dout = Dev(ctx, ax), Dev(ctx, ay), Dev(ctx, np.zeros(ca.L * S))
    try:
        ctx.check(N.lib().ph_line_grid_device(ctx.handle, d.h, ca.ref(), N.PH_SCATTER, 64, *[N.ptr(a) for a in f64],
                                              dax.p, len(ax), day.p, len(ay), dout.p))
        got = dout.get(np.float64, ca.L * S).reshape(ca.L, S)
        ref = d.lines(cat, N.PH_SCATTER, ls.origin, ls.direction)
        assert_same(got, ref)
        assert (ref[0] > 0).sum() > S // 2
    finally:
        for b in (dax, day, dout):
            b.free()
    # Polar: s values and the angle table of the host plane
    sp, pp = params("sparams"), params("phiparams")
    ls = LineSamples.from_frame(Polar, ParallelBeam, fr, sp, pp)
    S = len(ls.grid)
    s_vals = sp[0] + np.arange(sp[2]) * ((sp[1] - sp[0]) / (sp[2] - 1))
    tab = grids.angle_table(pp)
    ds, dt, dout = Dev(ctx, s_vals), Dev(ctx, np.ascontiguousarray(tab.ravel())), Dev(ctx, np.zeros(ca.L * S))
    try:
        ctx.check(N.lib().ph_line_polar_device(ctx.handle, d.h, ca.ref(), N.PH_SCATTER, 64, *[N.ptr(a) for a in f64],
                                               ds.p, sp[2], dt.p, pp[2], dout.p))
        got = dout.get(np.float64, ca.L * S).reshape(ca.L, S)
        assert_same(got, d.lines(cat, N.PH_SCATTER, ls.origin, ls.direction))
    finally:
        for b in (ds, dt, dout):
            b.free()


def test_invalid_plane_parameters_are_rejected(backend):
    """:241-254: the same parameter sets as batch requests."""
    ctx = backend.context
    c6, cols = cloud(500, 7, (0.0, 0.0, 0.0))
    d = DeviceSet(ctx, *c6, cols, kernel=0)
    ca = N.CatalogArgs(sslots=[1, 2], sflags=[0, 1])
    fr = frame("origin")
    o = Dev(ctx, nbytes=4096)
    try:
        bad = [PB.plane(fr.position(), fr.right(), fr.up(), tuple(xp[:2]) + (int(xp[2]),),
                        tuple(yp[:2]) + (int(yp[2]),), o.p) for xp, yp in FIX["invalid_cartesian"]]
        bad += [PB.polar(fr.position(), fr.right(), fr.up(), tuple(sp[:2]) + (int(sp[2]),),
                         tuple(pp[:2]) + (int(pp[2]),), o.p) for sp, pp in FIX["invalid_polar"]]
        for r in bad:
            with pytest.raises(ArgumentError):
                PB.Batch(ctx, d.h, ca, N.PH_SYMMETRIC, N.PH_PREC_FP64, [r])
    finally:
        o.free()
