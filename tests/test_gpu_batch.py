"""Batched sampling (ph_batch_*) against one PointSamples call per request.

Every sample of a batch runs the per-call arithmetic, so each request's outputs
must equal those of ph_point_samples on the same coordinates bit for bit (NaN
masks equal), in float64 and FP32C.  Plane and box coordinates are generated on
the device; the per-request calls use the host restatement of
Grids/grids/coordinate.jl:44-143, so the comparison also pins the generator."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import DevArray, DeviceSet, assert_same, traversal
from partia import ArgumentError
from partia import _native as N
from partia import batch as PB
from partia import grids

pytestmark = pytest.mark.gpu
FULL = dict(sslots=[1, 2], sflags=[0, 1], gslots=[2], dslots=[3, 4, 5], cslots=[3, 4, 5])
SMALL = dict(sslots=[1, 2], sflags=[0, 1])
WIDE = dict(sslots=[1, 2, 3, 4, 5], sflags=[0, 1, 1, 1, 1], gslots=[2, 3], dslots=[3, 4, 5], cslots=[3, 4, 5])
O_SYM = N.PH_SYMMETRIC


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.random(n) + 0.05
    h = (rng.random(n) * 0.12 + 0.04) * (200.0 / n) ** (1 / 3)
    rho = rng.random(n) + 0.3
    A = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    return (x, y, z, m, h, rho), [rho, A, x + 0.3 * y, y - z, z * x]


def random_frame(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    Rm = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                   [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                   [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return rng.random(3) * 0.6 + 0.2, Rm[:, 0], Rm[:, 2], Rm[:, 1]  # position, right, up, forward


def make_requests(ctx, rng, L, precision, kinds):
    """Requests of the given kinds with host copies of their coordinates."""
    esz = 8 if precision == N.PH_PREC_FP64 else 4
    reqs, coords, bufs, outs = [], [], [], []
    for kind in kinds:
        if kind == "points":
            S = int(rng.choice([0, 1, 63, 64, 257, int(rng.integers(300, 3000))]))
            P = [rng.random(S) * 1.2 - 0.1 for _ in range(3)]
            d = [DevArray(ctx, a) for a in P]
            o = DevArray(ctx, nbytes=max(1, L * S) * esz)
            reqs.append(PB.points(d[0].p, d[1].p, d[2].p, S, o.p))
            bufs += d
        else:
            pos, r, u, f = random_frame(rng)
            ext = [(-0.3 * rng.random() - 0.05, 0.3 * rng.random() + 0.05, int(rng.integers(2, 40))) for _ in range(3)]
            if kind == "plane":
                P = grids.cartesian_plane_coordinates(pos, r, u, ext[0], ext[1])
                S = len(P[0])
                o = DevArray(ctx, nbytes=L * S * esz)
                reqs.append(PB.plane(pos, r, u, ext[0], ext[1], o.p))
            else:
                ext[2] = (ext[2][0], ext[2][1], int(rng.integers(2, 12)))
                P = grids.cartesian_box_coordinates(pos, r, u, f, ext[0], ext[1], ext[2])
                S = len(P[0])
                o = DevArray(ctx, nbytes=L * S * esz)
                reqs.append(PB.box(pos, r, u, f, ext[0], ext[1], ext[2], o.p))
        coords.append(P)
        outs.append((o, S))
    return reqs, coords, bufs, outs


def read_out(o, S, L, precision):
    dt = np.float64 if precision == N.PH_PREC_FP64 else np.float32
    return o.get(dt, L * S).astype(np.float64).reshape(L, S)


@pytest.mark.parametrize("precision", [N.PH_PREC_FP64, N.PH_PREC_FP32C])
@pytest.mark.parametrize("strategy", [N.PH_GATHER, N.PH_SCATTER, N.PH_SYMMETRIC])
def test_batch_matches_single_calls(backend, precision, strategy):
    ctx = backend.context
    c6, cols = cloud(20000, 7 + strategy)
    d = DeviceSet(ctx, *c6, cols, kernel=4)
    rng = np.random.default_rng(100 + 10 * strategy + precision)
    for cat in (FULL, SMALL, WIDE):
        ca = N.CatalogArgs(**cat)
        kinds = ["points", "plane", "box"] * 4 + ["points"] * 3
        reqs, coords, bufs, outs = make_requests(ctx, rng, ca.L, precision, kinds)
        b = PB.Batch(ctx, d.h, ca, strategy, precision, reqs)
        try:
            b.run()
            for r, (P, (o, S)) in enumerate(zip(coords, outs)):
                got = read_out(o, S, ca.L, precision)
                ref = d.points(cat, strategy, *P, precision=precision)
                assert got.shape == ref.shape
                assert_same(got, ref)
        finally:
            b.close()
            for x in bufs + [o for o, _ in outs]:
                x.free()


@pytest.mark.parametrize("precision", [N.PH_PREC_FP64, N.PH_PREC_FP32C])
@pytest.mark.parametrize("mode", [N.PH_TRAVERSAL_AUTO, N.PH_TRAVERSAL_WAVE, N.PH_TRAVERSAL_BUCKET,
                                  N.PH_TRAVERSAL_LANE])
def test_batch_walks_under_every_traversal(backend, precision, mode):
    """By default each wave takes the per-wave union walk of the 8-leaf bucket
    tree or the per-lane bucket walk, by its samples' spread; PH_TRAVERSAL_WAVE
    puts every wave on the union walk (with the per-wave nearest-h search),
    BUCKET and LANE on the per-lane walk.  Every choice gives the per-call
    results bit for bit (gather and symmetric)."""
    ctx = backend.context
    c6, cols = cloud(12000, 31)
    d = DeviceSet(ctx, *c6, cols, kernel=1)
    rng = np.random.default_rng(41 + mode + 7 * precision)
    for cat, strategy in ((FULL, N.PH_GATHER), (WIDE, O_SYM), (SMALL, O_SYM)):
        ca = N.CatalogArgs(**cat)
        reqs, coords, bufs, outs = make_requests(ctx, rng, ca.L, precision, ["points", "plane", "box", "points"] * 2)
        with traversal(ctx, mode):
            b = PB.Batch(ctx, d.h, ca, strategy, precision, reqs)
        try:
            b.run()
            for P, (o, S) in zip(coords, outs):
                assert_same(read_out(o, S, ca.L, precision), d.points(cat, strategy, *P, precision=precision))
        finally:
            b.close()
            for x in bufs + [o for o, _ in outs]:
                x.free()


def test_batch_point_lists_in_walk_order(backend):
    """Point lists long enough to be walked in Morton order: a list of one
    repeated point (zero-extent box), a list with NaN coordinates, a list in
    reverse Morton order and a random list; each output lands at its own index,
    equal to the single call's."""
    ctx = backend.context
    c6, cols = cloud(20000, 17)
    d = DeviceSet(ctx, *c6, cols, kernel=4)
    rng = np.random.default_rng(5)
    lists = [[np.full(1000, 0.4), np.full(1000, 0.6), np.full(1000, 0.5)]]
    P = [rng.random(700) for _ in range(3)]
    P[0][::7] = np.nan
    P[2][3] = np.inf
    lists.append(P)
    g = np.linspace(0.0, 1.0, 12)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    lists.append([X.ravel()[::-1].copy(), Y.ravel()[::-1].copy(), Z.ravel()[::-1].copy()])
    lists.append([rng.random(5000) for _ in range(3)])
    ca = N.CatalogArgs(**FULL)
    for precision in (N.PH_PREC_FP64, N.PH_PREC_FP32C):
        esz = 8 if precision == N.PH_PREC_FP64 else 4
        bufs, outs, reqs = [], [], []
        for P in lists:
            dp = [DevArray(ctx, np.ascontiguousarray(a)) for a in P]
            o = DevArray(ctx, nbytes=ca.L * len(P[0]) * esz)
            reqs.append(PB.points(dp[0].p, dp[1].p, dp[2].p, len(P[0]), o.p))
            bufs += dp
            outs.append(o)
        b = PB.Batch(ctx, d.h, ca, O_SYM, precision, reqs)
        try:
            b.run()
            for P, o in zip(lists, outs):
                assert_same(read_out(o, len(P[0]), ca.L, precision), d.points(FULL, O_SYM, *P, precision=precision))
        finally:
            b.close()
            for x in bufs + outs:
                x.free()


@pytest.mark.parametrize("precision", [N.PH_PREC_FP64, N.PH_PREC_FP32C])
def test_batch_polar_and_cylindrical_requests(backend, precision):
    """PointSamples(Polar, ...) and PointSamples(Cylindrical, ...) generated on
    the device (coordinate.jl:153-276): equal, bit for bit, to single calls on
    the host restatement's coordinates with the same angle table (the C
    library's, or one the caller passes)."""
    ctx = backend.context
    c6, cols = cloud(20000, 27)
    d = DeviceSet(ctx, *c6, cols, kernel=4)
    rng = np.random.default_rng(8)
    ca = N.CatalogArgs(**FULL)
    esz = 8 if precision == N.PH_PREC_FP64 else 4
    specs = [("polar", (0.0, 0.45, 13), (0.0, 2 * np.pi, 21), None, None),
             ("polar", (0.05, 0.3, 2), (np.pi / 4, 3 * np.pi / 4, 1), None, None),
             ("cylinder", (0.0, 0.4, 9), (0.5, 6.0, 11), (-0.2, 0.3, 5), None),
             ("polar", (0.1, 0.4, 30), (0.0, 2 * np.pi, 17), None, "table"),
             ("cylinder", (0.0, 0.35, 6), (0.0, 2 * np.pi * (1 + 1e-10), 7), (-0.1, 0.1, 2), "table")]
    reqs, coords, outs = [], [], []
    for kind, sp, pp, zp, tab in specs:
        pos, r, u, f = random_frame(rng)
        trig = None
        if tab:
            phi = pp[0] + np.arange(pp[2]) * ((pp[1] - pp[0]) / pp[2])
            trig = np.stack([np.sin(phi), np.cos(phi)], axis=1)  # numpy's, not the C library's
        if kind == "polar":
            P = grids.polar_plane_coordinates(pos, r, u, sp, pp, trig=trig)
        else:
            P = grids.cylindrical_coordinates(pos, r, u, f, sp, pp, zp, trig=trig)
        o = DevArray(ctx, nbytes=ca.L * len(P[0]) * esz)
        if kind == "polar":
            reqs.append(PB.polar(pos, r, u, sp, pp, o.p, trig=trig))
        else:
            reqs.append(PB.cylinder(pos, r, u, f, sp, pp, zp, o.p, trig=trig))
        assert PB.sample_count(reqs[-1]) == len(P[0])
        coords.append(P)
        outs.append(o)
    b = PB.Batch(ctx, d.h, ca, O_SYM, precision, reqs)
    try:
        b.run()
        for P, o in zip(coords, outs):
            assert_same(read_out(o, len(P[0]), ca.L, precision), d.points(FULL, O_SYM, *P, precision=precision))
    finally:
        b.close()
        for o in outs:
            o.free()


def test_batch_polar_argument_errors(backend):
    ctx = backend.context
    c6, cols = cloud(500, 2)
    d = DeviceSet(ctx, *c6, cols, kernel=0)
    ca = N.CatalogArgs(**SMALL)
    o = DevArray(ctx, nbytes=4096)
    fr = ((0.5, 0.5, 0.5), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))
    try:
        bad = [PB.polar(*fr[:3], (-0.1, 0.3, 3), (0.0, 1.0, 2), o.p),
               PB.polar(*fr[:3], (0.0, 0.3, 3), (0.0, 7.0, 2), o.p),
               PB.polar(*fr[:3], (0.0, 0.3, 1), (0.0, 1.0, 2), o.p),
               PB.polar(*fr[:3], (0.0, 0.3, 3), (0.0, 1.0, 0), o.p),
               PB.cylinder(*fr, (0.0, 0.3, 3), (0.0, 1.0, 2), (0.2, 0.2, 3), o.p),
               PB.cylinder(*fr, (0.0, 0.3, 3), (0.0, 1.0, 2), (0.0, 0.2, 1), o.p)]
        for r in bad:
            with pytest.raises(ArgumentError):
                PB.Batch(ctx, d.h, ca, O_SYM, N.PH_PREC_FP64, [r])
    finally:
        o.free()


def test_batch_replays_and_times(backend):
    """Graph replays with eager calls in between give identical outputs; per-
    request times are ordered and empty requests report NaN."""
    ctx = backend.context
    c6, cols = cloud(15000, 3)
    d = DeviceSet(ctx, *c6, cols, kernel=0)
    rng = np.random.default_rng(4)
    ca = N.CatalogArgs(**FULL)
    reqs, coords, bufs, outs = make_requests(ctx, rng, ca.L, N.PH_PREC_FP32C,
                                             ["plane", "points", "box", "points", "plane"])
    empty = DevArray(ctx, nbytes=8)
    reqs.append(PB.points(None, None, None, 0, empty.p))
    b = PB.Batch(ctx, d.h, ca, N.PH_SYMMETRIC, N.PH_PREC_FP32C, reqs)
    try:
        first = None
        for rep in range(4):
            b.run()
            res = [read_out(o, S, ca.L, N.PH_PREC_FP32C) for o, S in outs]
            if first is None:
                first = res
            else:
                for a, c in zip(first, res):
                    assert_same(a, c)
            # eager work between replays
            d.points(FULL, N.PH_SYMMETRIC, *coords[0], precision=N.PH_PREC_FP32C)
            t0, t1 = b.times()
            S = [s for _, s in outs] + [0]
            for r in range(len(reqs)):
                if S[r] == 0:
                    assert np.isnan(t0[r]) and np.isnan(t1[r])
                else:
                    assert 0.0 <= t0[r] <= t1[r] < 1e7
    finally:
        b.close()
        for x in bufs + [o for o, _ in outs] + [empty]:
            x.free()


def test_batch_after_column_change(backend):
    """Runs read the current columns: same-shape updates and a reallocation
    (another column count, which re-records the graph)."""
    ctx = backend.context
    c6, cols = cloud(8000, 9)
    d = DeviceSet(ctx, *c6, cols, kernel=4)
    rng = np.random.default_rng(12)
    cat = dict(sslots=[1, 2], sflags=[0, 1], gslots=[2])
    ca = N.CatalogArgs(**cat)
    reqs, coords, bufs, outs = make_requests(ctx, rng, ca.L, N.PH_PREC_FP64, ["points", "plane", "box"])
    b = PB.Batch(ctx, d.h, ca, N.PH_GATHER, N.PH_PREC_FP64, reqs)
    try:
        order = d.get("order", np.uint64, d.n).astype(np.int64) - 1  # Morton position -> source index
        for ncol in (5, 7):
            new = [np.cos(3.0 * c6[0] + k) for k in range(ncol)]
            m_s, rho_s = c6[3][order], c6[5][order]
            cs = [c[order] for c in new]
            arr = (C.c_void_p * ncol)(*[a.ctypes.data for a in cs])
            ctx.check(N.lib().ph_lbvh_set_columns(d.h, N.ptr(m_s), N.ptr(rho_s), ncol, arr))
            b.run()
            for P, (o, S) in zip(coords, outs):
                assert_same(read_out(o, S, ca.L, N.PH_PREC_FP64), d.points(cat, N.PH_GATHER, *P))
    finally:
        b.close()
        for x in bufs + [o for o, _ in outs]:
            x.free()


def test_batch_errors(backend):
    ctx = backend.context
    c6, cols = cloud(500, 1)
    d = DeviceSet(ctx, *c6, cols, kernel=0)
    ca = N.CatalogArgs(**SMALL)
    o = DevArray(ctx, nbytes=1 << 16)
    pos, r, u, f = [0.5] * 3, [1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]
    with pytest.raises(ArgumentError, match="at least 2"):
        PB.Batch(ctx, d.h, ca, N.PH_GATHER, N.PH_PREC_FP64, [PB.plane(pos, r, u, (0.0, 1.0, 1), (0.0, 1.0, 4), o.p)])
    with pytest.raises(ArgumentError, match="greater than"):
        PB.Batch(ctx, d.h, ca, N.PH_GATHER, N.PH_PREC_FP64,
                 [PB.box(pos, r, u, f, (0.0, 1.0, 3), (0.0, 1.0, 4), (2.0, 2.0, 3), o.p)])
    bad = PB.points(None, None, None, 0, o.p)
    bad.kind = 7
    with pytest.raises(ArgumentError, match="unknown kind"):
        PB.Batch(ctx, d.h, ca, N.PH_GATHER, N.PH_PREC_FP64, [bad])
    with pytest.raises(ArgumentError):
        PB.Batch(ctx, d.h, ca, N.PH_GATHER, N.PH_PREC_FP32, [PB.plane(pos, r, u, (0.0, 1.0, 2), (0.0, 1.0, 2), o.p)])
    b = PB.Batch(ctx, d.h, ca, N.PH_GATHER, N.PH_PREC_FP64, [PB.plane(pos, r, u, (0.0, 1.0, 2), (0.0, 1.0, 2), o.p)])
    b.run()
    d.close()  # the tree goes away first
    with pytest.raises(ArgumentError, match="destroyed"):
        b.run()
    b.close()
    o.free()
