"""Adjoints of point and grid interpolation on the GPU
(ph_point_samples_adjoint_device, ph_structured_grid_adjoint_device) against
the oracle's point_adjoint, which test_oracle_adjoint.py pins by central
finite differences of the oracle's forward interpolation.

The GPU sums the per-sample terms with float64 atomics, so results match the
oracle to rounding (1e-10 of each gradient row's rms), not bitwise."""
import numpy as np
import pytest

from gpu_util import DevArray, DeviceSet, grid_points, oracle_of
from partia import ArgumentError
from partia import _native as N
from partia import synthetic

import pyoracle as O

pytestmark = pytest.mark.gpu
STRATS = [O.GATHER, O.SCATTER, O.SYMMETRIC]


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.random(n) + 0.2
    h = (rng.random(n) * 0.1 + 0.08) * (400.0 / n) ** (1 / 3)
    rho = rng.random(n) + 0.5
    A = 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    return [x, y, z, m, h, rho], [rho, A]


def gpu_point_adjoint(d, slot, shepard, strategy, pts, cot, precision=N.PH_PREC_FP64):
    ctx = d.ctx
    ca = N.CatalogArgs(sslots=[slot], sflags=[int(shepard)])
    S = len(cot)
    arrs = [DevArray(ctx, np.ascontiguousarray(a, dtype=np.float64)) for a in pts]
    dc = DevArray(ctx, np.asarray(cot, dtype=np.float64 if precision == N.PH_PREC_FP64 else np.float32))
    dg = DevArray(ctx, nbytes=6 * d.n * 8)
    try:
        ctx.check(N.lib().ph_point_samples_adjoint_device(ctx.handle, d.h, ca.ref(), strategy, precision, S,
                                                          *[a.p for a in arrs], dc.p, dg.p))
        return dg.get(np.float64, 6 * d.n).reshape(6, d.n)
    finally:
        for a in arrs + [dc, dg]:
            a.free()


def gpu_grid_adjoint(d, slot, shepard, strategy, ax, ay, az, cot, precision=N.PH_PREC_FP64):
    ctx = d.ctx
    ca = N.CatalogArgs(sslots=[slot], sflags=[int(shepard)])
    axs = [DevArray(ctx, np.asarray(a, np.float64)) for a in (ax, ay, az)]
    dc = DevArray(ctx, np.asarray(cot, dtype=np.float64 if precision == N.PH_PREC_FP64 else np.float32))
    dg = DevArray(ctx, nbytes=6 * d.n * 8)
    try:
        ctx.check(N.lib().ph_structured_grid_adjoint_device(ctx.handle, d.h, ca.ref(), strategy, precision,
                                                            axs[0].p, len(ax), axs[1].p, len(ay), axs[2].p, len(az),
                                                            dc.p, dg.p))
        return dg.get(np.float64, 6 * d.n).reshape(6, d.n)
    finally:
        for a in axs + [dc, dg]:
            a.free()


def assert_rows_close(got, ref, rtol):
    for r in range(6):
        scale = max(np.sqrt(np.mean(ref[r] ** 2)), 1e-300)
        err = np.max(np.abs(got[r] - ref[r])) / scale
        assert err <= rtol, "row %d: max error %.3e of rms" % (r, err)


@pytest.mark.parametrize("strategy", STRATS)
@pytest.mark.parametrize("kernel,shepard", [(0, False), (0, True), (4, True), (2, False), (5, True)])
def test_point_adjoint_matches_oracle(backend, strategy, kernel, shepard):
    c6, cols = cloud(3000, 40 + kernel)
    d = DeviceSet(backend.context, *c6, cols, kernel=kernel)
    o = oracle_of(*c6, cols, kernel=kernel)
    rng = np.random.default_rng(8)
    pts = rng.random((3, 2000)) * 1.1 - 0.05  # some outside the cloud
    pts[:, :5] = 7.0  # no neighbours: NaN outputs, no gradient
    cot = rng.normal(size=2000)
    ref = o.point_adjoint(2, shepard, strategy, *pts, cot)
    got = gpu_point_adjoint(d, 2, shepard, strategy, pts, cot)
    assert np.abs(ref).sum(axis=1).min() > 0
    assert_rows_close(got, ref, 1e-10)


@pytest.mark.parametrize("strategy", STRATS)
def test_grid_adjoint_matches_oracle(backend, strategy):
    """A ragged grid (tiles cut at every face) in both precisions' cotangent
    types; the density (no Shepard) and a Shepard scalar."""
    c6, cols = cloud(4000, 51)
    d = DeviceSet(backend.context, *c6, cols, kernel=4)
    o = oracle_of(*c6, cols, kernel=4)
    ax, ay, az = synthetic.linrange(-0.1, 1.1, 19), synthetic.linrange(0.0, 1.0, 13), synthetic.linrange(0.1, 0.9, 10)
    pts = grid_points(ax, ay, az)
    cot = np.random.default_rng(2).normal(size=len(pts[0]))
    for slot, shepard in ((1, False), (2, True)):
        ref = o.point_adjoint(slot, shepard, strategy, *pts, cot)
        assert_rows_close(gpu_grid_adjoint(d, slot, shepard, strategy, ax, ay, az, cot), ref, 1e-10)
        c32 = cot.astype(np.float32)
        ref32 = o.point_adjoint(slot, shepard, strategy, *pts, c32.astype(np.float64))
        got32 = gpu_grid_adjoint(d, slot, shepard, strategy, ax, ay, az, c32, precision=N.PH_PREC_FP32C)
        assert_rows_close(got32, ref32, 1e-10)


def test_grid_adjoint_equals_point_adjoint_on_the_same_points(backend):
    c6, cols = cloud(5000, 60)
    d = DeviceSet(backend.context, *c6, cols, kernel=0)
    ax = synthetic.linrange(0.0, 1.0, 33)
    pts = grid_points(ax, ax, ax)
    cot = np.random.default_rng(5).normal(size=len(pts[0]))
    for st in STRATS:
        a = gpu_grid_adjoint(d, 2, True, st, ax, ax, ax, cot)
        b = gpu_point_adjoint(d, 2, True, st, pts, cot)
        assert_rows_close(a, b, 1e-11)


def test_point_adjoint_tiny_trees_and_empty(backend):
    """Trees of 1, 3 and 9 particles (one or a few buckets) and zero samples
    (the gradient is cleared).  With one particle a Shepard-normalised output
    equals A_b exactly, so its m, h and x rows are rounding noise: that tree is
    checked without the Shepard flag."""
    for n in (1, 3, 9):
        c6, cols = cloud(n, 70 + n)
        c6[4] = np.full(n, 0.6)
        d = DeviceSet(backend.context, *c6, cols, kernel=0)
        o = oracle_of(*c6, cols, kernel=0)
        rng = np.random.default_rng(n)
        pts = rng.random((3, 300))
        cot = rng.normal(size=300)
        for shepard in ((False,) if n == 1 else (False, True)):
            for st in STRATS:
                ref = o.point_adjoint(2, shepard, st, *pts, cot)
                assert_rows_close(gpu_point_adjoint(d, 2, shepard, st, pts, cot), ref, 1e-10)
    got = gpu_point_adjoint(d, 2, False, O.GATHER, np.zeros((3, 0)), np.zeros(0))
    assert not got.any()


def test_point_adjoint_errors(backend):
    c6, cols = cloud(500, 80)
    d = DeviceSet(backend.context, *c6, cols, kernel=0)
    ctx = d.ctx
    dg = DevArray(ctx, nbytes=6 * d.n * 8)
    dc = DevArray(ctx, np.zeros(4))
    pts = [DevArray(ctx, np.zeros(4)) for _ in range(3)]
    try:
        two = N.CatalogArgs(sslots=[1, 2], sflags=[0, 1])
        with pytest.raises(ArgumentError):  # two scalars
            ctx.check(N.lib().ph_point_samples_adjoint_device(ctx.handle, d.h, two.ref(), O.GATHER,
                                                              N.PH_PREC_FP64, 4, *[a.p for a in pts], dc.p, dg.p))
        one = N.CatalogArgs(sslots=[2], sflags=[1])
        with pytest.raises(ArgumentError):  # Float32 evaluation
            ctx.check(N.lib().ph_point_samples_adjoint_device(ctx.handle, d.h, one.ref(), O.GATHER,
                                                              N.PH_PREC_FP32, 4, *[a.p for a in pts], dc.p, dg.p))
        with pytest.raises(ArgumentError):  # null gradient
            ctx.check(N.lib().ph_point_samples_adjoint_device(ctx.handle, d.h, one.ref(), O.GATHER,
                                                              N.PH_PREC_FP64, 4, *[a.p for a in pts], dc.p, None))
    finally:
        for a in pts + [dc, dg]:
            a.free()
