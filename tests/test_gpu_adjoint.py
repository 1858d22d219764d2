"""Line-projection adjoints on the GPU (ph_*_adjoint_device) against the oracle's
adjoint (itself pinned by finite differences in test_oracle_adjoint.py) and
against central finite differences of the GPU's own forward projection.

The GPU sums the per-ray terms with float64 atomics, so results match the
oracle to rounding (1e-10 of each gradient row's rms in float64), not bitwise;
the Float32 path is compared with the oracle's Float32 instantiation at 1e-4."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import DevArray, DeviceSet, oracle_of
from partia import ArgumentError
from partia import _native as N
from partia import synthetic

import pyoracle as O

pytestmark = pytest.mark.gpu


def cloud(n, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.random(n) + 0.2
    h = (rng.random(n) * 0.1 + 0.08) * (400.0 / n) ** (1 / 3)
    rho = rng.random(n) + 0.5
    A = 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    c6 = [x, y, z, m, h, rho]
    cols = [rho, A]
    if dtype == np.float32:
        c6 = [a.astype(np.float32).astype(np.float64) for a in c6]
        cols = [a.astype(np.float32).astype(np.float64) for a in cols]
    return c6, cols


def gpu_line_adjoint(d, slot, shepard, org, dirs, cot, precision=64):
    ctx = d.ctx
    ca = N.CatalogArgs(sslots=[slot], sflags=[int(shepard)])
    S = len(cot)
    arrs = [DevArray(ctx, np.ascontiguousarray(a, dtype=np.float64)) for a in list(org) + list(dirs)]
    dc = DevArray(ctx, np.asarray(cot, dtype=np.float64 if precision == 64 else np.float32))
    dg = DevArray(ctx, nbytes=6 * d.n * 8)
    try:
        ctx.check(N.lib().ph_line_samples_adjoint_device(ctx.handle, d.h, ca.ref(), N.PH_SCATTER, precision, S,
                                                         *[a.p for a in arrs], dc.p, dg.p))
        return dg.get(np.float64, 6 * d.n).reshape(6, d.n)
    finally:
        for a in arrs + [dc, dg]:
            a.free()


def line_rays(S, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    org = rng.random((3, S)) * 0.8 + 0.1
    dirs = rng.normal(size=(3, S))
    dirs /= np.linalg.norm(dirs, axis=0)
    dirs[:, : S // 2] = np.array([[0.0], [0.0], [1.0]])
    if dtype == np.float32:
        org = org.astype(np.float32).astype(np.float64)
        dirs = dirs.astype(np.float32).astype(np.float64)
    return org, dirs


def assert_rows_close(got, ref, rtol):
    for r in range(6):
        scale = max(np.sqrt(np.mean(ref[r] ** 2)), 1e-300)
        err = np.max(np.abs(got[r] - ref[r])) / scale
        assert err <= rtol, "row %d: max error %.3e of rms" % (r, err)


@pytest.mark.parametrize("kernel", [0, 4])
@pytest.mark.parametrize("shepard", [False, True])
def test_adjoint_float64_matches_oracle(backend, kernel, shepard):
    c6, cols = cloud(4000, 20 + kernel)
    d = DeviceSet(backend.context, *c6, cols, kernel=kernel)
    o = oracle_of(*c6, cols, kernel=kernel)
    org, dirs = line_rays(1500, 3)
    cot = np.random.default_rng(7).normal(size=1500)
    ref = o.line_adjoint(2, shepard, org, dirs, cot)
    got = gpu_line_adjoint(d, 2, shepard, org, dirs, cot)
    assert_rows_close(got, ref, 1e-10)
    # the column-density form: A is the rho column itself (slot 1, no Shepard)
    assert_rows_close(gpu_line_adjoint(d, 1, False, org, dirs, cot), o.line_adjoint(1, False, org, dirs, cot), 1e-10)


@pytest.mark.parametrize("shepard", [False, True])
def test_adjoint_float32_matches_oracle(backend, shepard):
    c6, cols = cloud(4000, 31, np.float32)
    d = DeviceSet(backend.context, *c6, cols, kernel=0, precision=32)
    o = oracle_of(*c6, cols, kernel=0, precision=32)
    org, dirs = line_rays(1500, 5, np.float32)
    cot = np.random.default_rng(8).normal(size=1500).astype(np.float32).astype(np.float64)
    ref = o.line_adjoint(2, shepard, org, dirs, cot)
    got = gpu_line_adjoint(d, 2, shepard, org, dirs, cot, precision=32)
    assert_rows_close(got, ref, 1e-4)


def project(ctx, c6, cols, kernel, frame, ax, precision=64):
    """GPU column map of the scalar in slot 1 (non-Shepard) on an image plane."""
    dtype = np.float64 if precision == 64 else np.float32
    d = DeviceSet(ctx, *c6, cols, kernel=kernel, precision=precision)
    S = len(ax) ** 2
    dax = DevArray(ctx, ax)
    dout = DevArray(ctx, nbytes=S * np.dtype(dtype).itemsize)
    ca = N.CatalogArgs(sslots=[1], sflags=[0])
    fr = [np.ascontiguousarray(np.asarray(v, dtype=np.float64)) for v in frame]
    try:
        ctx.check(N.lib().ph_line_grid_device(ctx.handle, d.h, ca.ref(), N.PH_SCATTER, precision,
                                              *[N.ptr(v) for v in fr], dax.p, len(ax), dax.p, len(ax), dout.p))
        out = dout.get(dtype, S).astype(np.float64)
    finally:
        dax.free()
        dout.free()
    return d, out


def test_adjoint_column_density_finite_differences(backend):
    """Loss 1/2 |Sigma - T|^2 of a Plummer column-density map (A = rho, slot 1):
    the adjoint of the image-plane projection against central differences of
    the GPU forward at 16 particles, for m, h, x, y (z moves along the rays).
    Float64, relative step 1e-6 (positions: 1e-6 h)."""
    ctx = backend.context
    w = synthetic.plummer_sphere(20000, seed=5)
    c6 = [w[k] for k in ("x", "y", "z", "m", "h", "rho")]
    cols = [w["rho"]]
    ax = synthetic.linrange(-0.6, 0.6, 192)
    frame = ([0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0])
    d, sigma = project(ctx, c6, cols, 4, frame, ax)
    rng = np.random.default_rng(2)
    target = sigma * np.exp(rng.normal(0.0, 0.1, sigma.shape))
    cot = np.nan_to_num(sigma - target, nan=0.0)  # rays without neighbours (NaN) are left out
    ca = N.CatalogArgs(sslots=[1], sflags=[0])
    dax, dc, dg = DevArray(ctx, ax), DevArray(ctx, cot), DevArray(ctx, nbytes=6 * d.n * 8)
    fr = [np.ascontiguousarray(np.asarray(v, dtype=np.float64)) for v in frame]
    ctx.check(N.lib().ph_line_grid_adjoint_device(ctx.handle, d.h, ca.ref(), N.PH_SCATTER, 64,
                                                  *[N.ptr(v) for v in fr], dax.p, len(ax), dax.p, len(ax), dc.p,
                                                  dg.p))
    g = dg.get(np.float64, 6 * d.n).reshape(6, d.n)
    for b in (dax, dc, dg):
        b.free()

    def loss(p):
        _, s = project(ctx, p, cols, 4, frame, ax)
        return 0.5 * float(np.nansum((s - target) ** 2))

    r2 = c6[0] ** 2 + c6[1] ** 2
    picks = rng.choice(np.nonzero(r2 < 0.3 ** 2)[0], 16, replace=False)
    for row, pi in ((0, 3), (2, 4), (3, 0), (4, 1)):
        scale = np.sqrt(np.mean(g[row] ** 2))
        for b in picks:
            eps = 1e-6 * (abs(c6[pi][b]) if pi in (3, 4) else c6[4][b])
            pp, pm = [a.copy() for a in c6], [a.copy() for a in c6]
            pp[pi][b] += eps
            pm[pi][b] -= eps
            fd = (loss(pp) - loss(pm)) / (2 * eps)
            assert np.isfinite(fd)
            assert abs(fd - g[row, b]) <= 1e-3 * scale + 1e-4 * abs(g[row, b]), (row, b, fd, g[row, b])


def test_adjoint_errors(backend):
    c6, cols = cloud(300, 1)
    d = DeviceSet(backend.context, *c6, cols)
    org, dirs = line_rays(10, 1)
    with pytest.raises(ArgumentError, match="exactly one scalar"):
        ca = N.CatalogArgs(sslots=[1, 2], sflags=[0, 1])
        d.ctx.check(N.lib().ph_line_samples_adjoint_device(d.ctx.handle, d.h, ca.ref(), N.PH_SCATTER, 64, 0,
                                                           *([None] * 6), None, C.c_void_p(8)))
    with pytest.raises(ArgumentError, match="itpScatter"):
        ca = N.CatalogArgs(sslots=[1], sflags=[0])
        d.ctx.check(N.lib().ph_line_samples_adjoint_device(d.ctx.handle, d.h, ca.ref(), N.PH_GATHER, 64, 0,
                                                           *([None] * 6), None, C.c_void_p(8)))
