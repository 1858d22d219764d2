"""Line-integrated projections on the GPU: the image-plane entry point
(ph_line_grid_device) against the oracle on the same rays, and the wave-walked
traversal against the per-ray one.

Both traversals test, for every ray, exactly the nodes and leaves of the
reference's DFS (LinearBVHTraversalMacros.jl:539-644) in the same order, so all
results are required to be bit-identical (NaN masks equal)."""
import numpy as np
import pytest

from gpu_util import DevArray as Dev, DeviceSet, assert_same, oracle_of, traversal
from partia import _native as N
from partia import synthetic

import pyoracle as O

pytestmark = pytest.mark.gpu
KERNELS = list(range(6))


def cloud(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.random(n) + 0.05
    h = (rng.random(n) * 0.12 + 0.04) * spread * (200.0 / max(n, 1)) ** (1 / 3)
    rho = rng.random(n) + 0.3
    A = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    return (x, y, z, m, h, rho), [rho, A, x + 0.3 * y]


def line_grid(d, cat, frame, ax, ay, precision=64):
    """ph_line_grid_device on host arrays; returns L x (nx*ny) float64."""
    ctx = d.ctx
    ca = N.CatalogArgs(**cat)
    o, u, v, dr = (np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in frame)
    nx, ny = len(ax), len(ay)
    S = nx * ny
    dax, day = Dev(ctx, np.asarray(ax, np.float64)), Dev(ctx, np.asarray(ay, np.float64))
    esz = 8 if precision == 64 else 4
    dout = Dev(ctx, np.zeros(max((ca.L * S * esz + 7) // 8, 1)))
    try:
        ctx.check(N.lib().ph_line_grid_device(ctx.handle, d.h, ca.ref(), O.SCATTER, precision, N.ptr(o), N.ptr(u),
                                              N.ptr(v), N.ptr(dr), dax.p, nx, day.p, ny, dout.p))
        out = dout.get(np.float64 if precision == 64 else np.float32, ca.L * S).astype(np.float64)
    finally:
        for b in (dax, day, dout):
            b.free()
    return out.reshape(ca.L, S)


def frame_rays(frame, ax, ay):
    """Ray (i, j), i fastest: origin (o + u a_i) + v b_j, direction dir, as the
    device generates them (float64, unfused)."""
    o, u, v, dr = (np.asarray(a, dtype=np.float64) for a in frame)
    A, B = np.meshgrid(np.asarray(ax, np.float64), np.asarray(ay, np.float64), indexing="ij")
    a, b = A.ravel(order="F"), B.ravel(order="F")
    org = [(o[k] + u[k] * a) + v[k] * b for k in range(3)]
    dirs = [np.full(a.shape, dr[k]) for k in range(3)]
    return org, dirs


Z_FRAME = ([0.0, 0.0, -0.5], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0])


def oblique_frame():
    d = np.array([0.3, -0.4, 0.8])
    d /= np.linalg.norm(d)
    u = np.cross(d, [0.0, 0.0, 1.0])
    u /= np.linalg.norm(u)
    v = np.cross(d, u)
    return ([0.5, 0.5, 0.5], u, v, d)


@pytest.mark.parametrize("kernel", KERNELS)
def test_line_grid_matches_oracle(backend, kernel):
    """Image-plane rays along z (ragged 8x8 tiles: 37 x 21) against the oracle."""
    c6, cols = cloud(5000, 60 + kernel)
    d = DeviceSet(backend.context, *c6, cols, kernel=kernel)
    o = oracle_of(*c6, cols, kernel=kernel)
    cat = dict(sslots=[1, 2, 3], sflags=[0, 1, 1])
    ax, ay = synthetic.linrange(-0.05, 1.05, 37), synthetic.linrange(0.0, 1.0, 21)
    ref = o.line_samples(cat, O.SCATTER, *frame_rays(Z_FRAME, ax, ay))
    for mode in (N.PH_TRAVERSAL_AUTO, N.PH_TRAVERSAL_LANE):
        with traversal(d.ctx, mode):
            assert_same(line_grid(d, cat, Z_FRAME, ax, ay), ref)


def test_line_grid_oblique_and_tiny_trees(backend):
    """Oblique rays; trees of 1, 2, 3 and 9 particles (root leaf, brute force)."""
    cat = dict(sslots=[1, 2], sflags=[0, 1])
    ax, ay = synthetic.linrange(-0.6, 0.6, 19), synthetic.linrange(-0.6, 0.6, 11)
    fr = oblique_frame()
    for n, spread in ((1, 8.0), (2, 6.0), (3, 6.0), (9, 3.0), (3000, 1.0)):
        c6, cols = cloud(n, 70 + n, spread)
        d = DeviceSet(backend.context, *c6, cols, kernel=0)
        o = oracle_of(*c6, cols, kernel=0)
        ref = o.line_samples(cat, O.SCATTER, *frame_rays(fr, ax, ay))
        assert np.isfinite(ref).any() or n < 3
        for mode in (N.PH_TRAVERSAL_AUTO, N.PH_TRAVERSAL_LANE):
            with traversal(d.ctx, mode):
                assert_same(line_grid(d, cat, fr, ax, ay), ref)


def test_line_grid_float32(backend):
    """Float32 tree and rays: bit-identical to the oracle's Float32 instantiation."""
    c6, cols = cloud(6000, 81)
    c6 = tuple(a.astype(np.float32).astype(np.float64) for a in c6)
    cols = [a.astype(np.float32).astype(np.float64) for a in cols]
    d = DeviceSet(backend.context, *c6, cols, kernel=0, precision=32)
    o = oracle_of(*c6, cols, kernel=0, precision=32)
    cat = dict(sslots=[1], sflags=[0])
    ax, ay = synthetic.linrange(0.0, 1.0, 45), synthetic.linrange(0.0, 1.0, 29)
    org, dirs = frame_rays(Z_FRAME, ax, ay)
    ref = o.line_samples(cat, O.SCATTER, [a.astype(np.float32).astype(np.float64) for a in org], dirs)
    for mode in (N.PH_TRAVERSAL_AUTO, N.PH_TRAVERSAL_LANE):
        with traversal(d.ctx, mode):
            assert_same(line_grid(d, cat, Z_FRAME, ax, ay, precision=32), ref)


AXIS_FRAMES = [
    ([-0.5, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]),
    ([0.0, 1.5, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, -1.0, 0.0]),
    ([0.0, 0.0, 1.5], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, -1.0]),
    Z_FRAME,
]


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("frame", range(len(AXIS_FRAMES)))
def test_line_axis_frames_same_as_lane(backend, precision, frame):
    """Image planes looking along +-x, +-y, +-z take the wave walk's axis-aligned
    distances; the images equal the per-ray walk's bit for bit."""
    c6, cols = cloud(7000, 61 + frame)
    if precision == 32:
        c6 = tuple(a.astype(np.float32).astype(np.float64) for a in c6)
        cols = [a.astype(np.float32).astype(np.float64) for a in cols]
    d = DeviceSet(backend.context, *c6, cols, kernel=frame % 6, precision=precision)
    cat = dict(sslots=[1, 2], sflags=[0, 1])
    ax, ay = synthetic.linrange(-0.1, 1.1, 37), synthetic.linrange(-0.1, 1.1, 23)
    fr = AXIS_FRAMES[frame]
    res = {}
    for mode in (N.PH_TRAVERSAL_LANE, N.PH_TRAVERSAL_WAVE):
        with traversal(d.ctx, mode):
            res[mode] = line_grid(d, cat, fr, ax, ay, precision)
    assert np.isfinite(res[N.PH_TRAVERSAL_LANE]).mean() > 0.5
    assert_same(res[N.PH_TRAVERSAL_WAVE], res[N.PH_TRAVERSAL_LANE])


@pytest.mark.parametrize("precision", [64, 32])
def test_line_axis_frame_huge_coordinates(backend, precision):
    """Coordinates beyond 2^(emax - 2) send the wave walk down the general
    distances: ray origins there (those waves only; the others keep the axis
    path) and a particle there (the whole tree).  The images still equal the
    per-ray walk's."""
    big = 1.0e38 if precision == 32 else 1.0e308
    c6, cols = cloud(3000, 71)
    ax = np.concatenate([synthetic.linrange(0.0, 1.0, 33), [big, -big, 0.25]])
    ay = synthetic.linrange(0.0, 1.0, 9)
    for outlier in (False, True):
        p6, pc = c6, cols
        if outlier:
            p6 = tuple(np.append(v, e) for v, e in zip(c6, (big, 0.5, 0.5, 1.0, 0.05, 1.0)))
            pc = [np.append(v, 1.0) for v in cols]
        if precision == 32:
            p6 = tuple(v.astype(np.float32).astype(np.float64) for v in p6)
            pc = [v.astype(np.float32).astype(np.float64) for v in pc]
        d = DeviceSet(backend.context, *p6, pc, kernel=0, precision=precision)
        cat = dict(sslots=[1], sflags=[0])
        res = {}
        for mode in (N.PH_TRAVERSAL_LANE, N.PH_TRAVERSAL_WAVE):
            with traversal(d.ctx, mode):
                res[mode] = line_grid(d, cat, Z_FRAME, ax, ay, precision)
        assert np.isfinite(res[N.PH_TRAVERSAL_LANE]).sum() > 100
        assert_same(res[N.PH_TRAVERSAL_WAVE], res[N.PH_TRAVERSAL_LANE])


@pytest.mark.parametrize("precision", [64, 32])
def test_line_list_wave_same_as_lane(backend, precision):
    """An explicit list of random rays (some axis-aligned), walked per ray and per
    wave: identical outputs and accepted-neighbour counts."""
    c6, cols = cloud(8000, 90)
    if precision == 32:
        c6 = tuple(a.astype(np.float32).astype(np.float64) for a in c6)
        cols = [a.astype(np.float32).astype(np.float64) for a in cols]
    d = DeviceSet(backend.context, *c6, cols, kernel=4, precision=precision)
    rng = np.random.default_rng(5)
    S = 3000
    org = rng.random((3, S))
    dirs = rng.normal(size=(3, S))
    dirs /= np.linalg.norm(dirs, axis=0)
    dirs[:, :300] = np.array([[1.0], [0.0], [0.0]])
    cat = dict(sslots=[1, 2, 3], sflags=[0, 1, 0])
    ctx = d.ctx
    res = {}
    for mode in (N.PH_TRAVERSAL_LANE, N.PH_TRAVERSAL_WAVE):
        with traversal(ctx, mode):
            ctx.set_count_neighbours(True)
            try:
                res[mode] = (d.lines(cat, O.SCATTER, org, dirs, precision=precision), ctx.stats()["accepted"])
            finally:
                ctx.set_count_neighbours(False)
    assert_same(res[N.PH_TRAVERSAL_WAVE][0], res[N.PH_TRAVERSAL_LANE][0])
    assert res[N.PH_TRAVERSAL_WAVE][1] == res[N.PH_TRAVERSAL_LANE][1] > 0


def test_line_block_orders_give_identical_images(backend):
    """The block and tile orders (ph_set_launch xcd_map, tile_order) only move
    work between XCDs and waves: image-plane and list projections are
    bit-identical for every order, with both traversals, and equal to the
    oracle."""
    c6, cols = cloud(20000, 95)
    d = DeviceSet(backend.context, *c6, cols, kernel=0)
    o = oracle_of(*c6, cols, kernel=0)
    cat = dict(sslots=[1, 2], sflags=[0, 1])
    ax, ay = synthetic.linrange(0.0, 1.0, 203), synthetic.linrange(0.0, 1.0, 77)  # 26 x 10 ragged tiles
    ref = o.line_samples(cat, O.SCATTER, *frame_rays(Z_FRAME, ax, ay))
    rng = np.random.default_rng(8)
    org, dirs = rng.random((3, 2500)), rng.normal(size=(3, 2500))
    dirs /= np.linalg.norm(dirs, axis=0)
    ctx = d.ctx
    try:
        ref_l = d.lines(cat, O.SCATTER, org, dirs)
        for xm, order in ((-1, 0), (0, 0), (1, 0), (2, 0), (5, 0), (7, 0), (-1, 1), (3, 1), (5, 1)):
            ctx.set_launch(1.0, xm, order)
            for mode in (N.PH_TRAVERSAL_AUTO, N.PH_TRAVERSAL_LANE, N.PH_TRAVERSAL_WAVE):
                with traversal(ctx, mode):
                    assert_same(line_grid(d, cat, Z_FRAME, ax, ay), ref)
                    assert_same(d.lines(cat, O.SCATTER, org, dirs), ref_l)
    finally:
        ctx.set_launch()


def test_line_polar_rays_match_the_ray_list(backend):
    """LineSamples(Polar, ParallelBeam, frame, sparams, phiparams): rays from the
    polar sampling of the frame plane along the frame's forward vector, equal
    bit for bit to the same rays as an explicit list (host restatement of the
    origins with the same angle table) and to the oracle; the adjoint equals
    the list adjoint to rounding."""
    from partia import grids
    c6, cols = cloud(6000, 97)
    d = DeviceSet(backend.context, *c6, cols, kernel=0)
    o = oracle_of(*c6, cols, kernel=0)
    ctx = d.ctx
    cat = dict(sslots=[1, 2], sflags=[0, 1])
    pos, u, v, fw = np.array([0.5, 0.5, -0.5]), np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0]), \
        np.array([0.0, 0.0, 1.0])
    sp, pp = (0.0, 0.45, 23), (0.0, 2 * np.pi, 37)
    tab = grids.angle_table(pp)
    org = grids.polar_plane_coordinates(pos, u, v, sp, pp, trig=tab)
    S = len(org[0])
    dirs = [np.full(S, fw[k]) for k in range(3)]
    ref = o.line_samples(cat, O.SCATTER, org, dirs)
    s_vals = sp[0] + np.arange(sp[2]) * ((sp[1] - sp[0]) / (sp[2] - 1))
    ds, dt = Dev(ctx, s_vals), Dev(ctx, np.ascontiguousarray(tab.ravel()))
    ca = N.CatalogArgs(**cat)
    dout = Dev(ctx, np.zeros(ca.L * S))
    frame = [np.ascontiguousarray(a, dtype=np.float64) for a in (pos, u, v, fw)]
    try:
        for mode in (N.PH_TRAVERSAL_AUTO, N.PH_TRAVERSAL_LANE):
            with traversal(ctx, mode):
                ctx.check(N.lib().ph_line_polar_device(ctx.handle, d.h, ca.ref(), O.SCATTER, 64,
                                                       *[N.ptr(f) for f in frame], ds.p, sp[2], dt.p, pp[2], dout.p))
                got = dout.get(np.float64, ca.L * S).reshape(ca.L, S)
                assert_same(got, ref)
                assert_same(got, d.lines(cat, O.SCATTER, org, dirs))
        # adjoint of the Shepard scalar: the polar form against the list form
        one = N.CatalogArgs(sslots=[2], sflags=[1])
        cot = np.random.default_rng(1).normal(size=S)
        dc = Dev(ctx, cot)
        dg1, dg2 = Dev(ctx, np.zeros(6 * d.n)), Dev(ctx, np.zeros(6 * d.n))
        arrs = [Dev(ctx, np.ascontiguousarray(a)) for a in list(org) + dirs]
        try:
            ctx.check(N.lib().ph_line_polar_adjoint_device(ctx.handle, d.h, one.ref(), O.SCATTER, 64,
                                                           *[N.ptr(f) for f in frame], ds.p, sp[2], dt.p, pp[2],
                                                           dc.p, dg1.p))
            ctx.check(N.lib().ph_line_samples_adjoint_device(ctx.handle, d.h, one.ref(), O.SCATTER, 64, S,
                                                             *[a.p for a in arrs], dc.p, dg2.p))
            g1 = dg1.get(np.float64, 6 * d.n).reshape(6, d.n)
            g2 = dg2.get(np.float64, 6 * d.n).reshape(6, d.n)
            for r in range(6):
                scale = max(np.sqrt(np.mean(g2[r] ** 2)), 1e-300)
                assert np.max(np.abs(g1[r] - g2[r])) <= 1e-10 * scale
        finally:
            for a in arrs + [dc, dg1, dg2]:
                a.free()
    finally:
        for b in (ds, dt, dout):
            b.free()


def test_line_grid_empty_image(backend):
    c6, cols = cloud(100, 3)
    d = DeviceSet(backend.context, *c6, cols)
    assert line_grid(d, dict(sslots=[1], sflags=[0]), Z_FRAME, np.zeros(0), np.zeros(4)).shape == (1, 0)
