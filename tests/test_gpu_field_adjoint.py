"""Adjoints of gradient, divergence and curl outputs of point and grid
interpolation on the GPU (ph_point_samples_adjoint_device and
ph_structured_grid_adjoint_device with a catalog of one gradient, divergence
or curl) against the oracle's field_adjoint, which test_oracle_adjoint.py pins
by central finite differences of the oracle's forward interpolation.

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
CATS = {1: dict(gslots=[2]), 2: dict(dslots=[2, 3, 4]), 3: dict(cslots=[2, 3, 4])}
SLOTS = {1: [2], 2: [2, 3, 4], 3: [2, 3, 4]}


def width(mode):
    return 1 if mode == 2 else 3


def rows(mode):
    return 6 if mode == 1 else 8


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.random(n) + 0.2
    h = (rng.random(n) * 0.1 + 0.08) * (400.0 / n) ** (1 / 3)
    rho = rng.random(n) + 0.5
    A = 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    B = 0.5 + np.cos(2 * np.pi * z) * x
    Cq = y * y - 0.3 * np.sin(2 * np.pi * x * z)
    return [x, y, z, m, h, rho], [rho, A, B, Cq]


def cot_dtype(precision):
    return np.float64 if precision == N.PH_PREC_FP64 else np.float32


def gpu_point_field_adjoint(d, mode, strategy, pts, cot, precision=N.PH_PREC_FP64):
    """cot: S x width; the device takes it component-major (width x S)."""
    ctx = d.ctx
    ca = N.CatalogArgs(**CATS[mode])
    S = cot.shape[0]
    R = rows(mode)
    arrs = [DevArray(ctx, np.ascontiguousarray(a, dtype=np.float64)) for a in pts]
    dc = DevArray(ctx, np.ascontiguousarray(cot.T, dtype=cot_dtype(precision)))
    dg = DevArray(ctx, nbytes=R * d.n * 8)
    try:
        ctx.check(N.lib().ph_point_samples_adjoint_device(ctx.handle, d.h, ca.ref(), strategy, precision, S,
                                                          *[a.p for a in arrs], dc.p, dg.p))
        return dg.get(np.float64, R * d.n).reshape(R, d.n)
    finally:
        for a in arrs + [dc, dg]:
            a.free()


def gpu_grid_field_adjoint(d, mode, strategy, ax, ay, az, cot, precision=N.PH_PREC_FP64):
    ctx = d.ctx
    ca = N.CatalogArgs(**CATS[mode])
    R = rows(mode)
    axs = [DevArray(ctx, np.asarray(a, np.float64)) for a in (ax, ay, az)]
    dc = DevArray(ctx, np.ascontiguousarray(cot.T, dtype=cot_dtype(precision)))
    dg = DevArray(ctx, nbytes=R * d.n * 8)
    try:
        ctx.check(N.lib().ph_structured_grid_adjoint_device(ctx.handle, d.h, ca.ref(), strategy, precision,
                                                            axs[0].p, len(ax), axs[1].p, len(ay), axs[2].p, len(az),
                                                            dc.p, dg.p))
        return dg.get(np.float64, R * d.n).reshape(R, d.n)
    finally:
        for a in axs + [dc, dg]:
            a.free()


def assert_rows_close(got, ref, rtol):
    assert got.shape == ref.shape
    for r in range(ref.shape[0]):
        scale = max(np.sqrt(np.mean(ref[r] ** 2)), 1e-300)
        err = np.max(np.abs(got[r] - ref[r])) / scale
        assert err <= rtol, "row %d: max error %.3e of rms" % (r, err)


@pytest.mark.parametrize("strategy", STRATS)
@pytest.mark.parametrize("mode,kernel", [(1, 0), (2, 4), (3, 0), (1, 5), (2, 1), (3, 3)])
def test_field_adjoint_matches_oracle(backend, strategy, mode, kernel):
    c6, cols = cloud(3000, 90 + kernel + 3 * mode)
    d = DeviceSet(backend.context, *c6, cols, kernel=kernel)
    o = oracle_of(*c6, cols, kernel=kernel)
    rng = np.random.default_rng(7 + mode)
    pts = rng.random((3, 2000)) * 1.1 - 0.05
    pts[:, :5] = 7.0  # no neighbours: NaN outputs, no gradient
    pts[:, 5:8] = np.array([c6[0][:3], c6[1][:3], c6[2][:3]])  # on particles (r = 0 terms)
    cot = rng.normal(size=(2000, width(mode)))
    ref = o.field_adjoint(mode, SLOTS[mode], False, strategy, *pts, cot)
    got = gpu_point_field_adjoint(d, mode, strategy, pts, cot)
    assert np.abs(ref).sum(axis=1).min() > 0
    assert_rows_close(got, ref, 1e-10)


@pytest.mark.parametrize("strategy", STRATS)
def test_grid_field_adjoint_matches_oracle(backend, strategy):
    """A ragged grid in both cotangent types, each output kind."""
    c6, cols = cloud(4000, 101)
    d = DeviceSet(backend.context, *c6, cols, kernel=4)
    o = oracle_of(*c6, cols, kernel=4)
    ax, ay, az = synthetic.linrange(-0.1, 1.1, 19), synthetic.linrange(0.0, 1.0, 13), synthetic.linrange(0.1, 0.9, 10)
    pts = grid_points(ax, ay, az)
    S = len(pts[0])
    for mode in (1, 2, 3):
        cot = np.random.default_rng(mode).normal(size=(S, width(mode)))
        ref = o.field_adjoint(mode, SLOTS[mode], False, strategy, *pts, cot)
        assert_rows_close(gpu_grid_field_adjoint(d, mode, strategy, ax, ay, az, cot), ref, 1e-10)
        c32 = cot.astype(np.float32)
        ref32 = o.field_adjoint(mode, SLOTS[mode], False, strategy, *pts, c32.astype(np.float64))
        got32 = gpu_grid_field_adjoint(d, mode, strategy, ax, ay, az, c32, precision=N.PH_PREC_FP32C)
        assert_rows_close(got32, ref32, 1e-10)


def test_grid_field_adjoint_equals_point_field_adjoint(backend):
    c6, cols = cloud(5000, 110)
    d = DeviceSet(backend.context, *c6, cols, kernel=0)
    ax = synthetic.linrange(0.0, 1.0, 21)
    pts = grid_points(ax, ax, ax)
    for mode in (1, 3):
        cot = np.random.default_rng(5).normal(size=(len(pts[0]), width(mode)))
        for st in STRATS:
            a = gpu_grid_field_adjoint(d, mode, st, ax, ax, ax, cot)
            b = gpu_point_field_adjoint(d, mode, st, pts, cot)
            assert_rows_close(a, b, 1e-11)


def test_field_adjoint_tiny_trees_and_empty(backend):
    for n in (3, 9):
        c6, cols = cloud(n, 120 + n)
        c6[4] = np.full(n, 0.6)
        d = DeviceSet(backend.context, *c6, cols, kernel=0)
        o = oracle_of(*c6, cols, kernel=0)
        rng = np.random.default_rng(n)
        pts = rng.random((3, 300))
        for mode in (1, 2, 3):
            cot = rng.normal(size=(300, width(mode)))
            for st in STRATS:
                ref = o.field_adjoint(mode, SLOTS[mode], False, st, *pts, cot)
                assert_rows_close(gpu_point_field_adjoint(d, mode, st, pts, cot), ref, 1e-10)
    got = gpu_point_field_adjoint(d, 3, O.GATHER, np.zeros((3, 0)), np.zeros((0, 3)))
    assert got.shape == (8, d.n) and not got.any()


def test_field_adjoint_errors(backend):
    c6, cols = cloud(500, 130)
    d = DeviceSet(backend.context, *c6, cols, kernel=0)
    ctx = d.ctx
    dg = DevArray(ctx, nbytes=8 * d.n * 8)
    dc = DevArray(ctx, np.zeros(12))
    pts = [DevArray(ctx, np.zeros(4)) for _ in range(3)]
    try:
        for kw in (dict(gslots=[0]),  # the density gradient branch
                   dict(gslots=[2], dslots=[2, 3, 4]),  # two output groups
                   dict(gslots=[2, 3])):
            ca = N.CatalogArgs(**kw)
            with pytest.raises(ArgumentError):
                ctx.check(N.lib().ph_point_samples_adjoint_device(ctx.handle, d.h, ca.ref(), O.GATHER,
                                                                  N.PH_PREC_FP64, 4, *[a.p for a in pts], dc.p, dg.p))
    finally:
        for a in pts + [dc, dg]:
            a.free()
