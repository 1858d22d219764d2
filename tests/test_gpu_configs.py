"""Parity at the full sizes of the measured workloads:

* column density: 8e6 particles in a log-normal density field, M4
  line-integrated, 2048^2 rays along z, Float32
  (line_integrated_scalar_interpolation.jl:123-166 through LineSamples_driver.jl:39-57);
* batched serving: 4e6 uniform-plus-clump particles and 512 uneven requests
  (point lists, oriented planes, oriented boxes) in one graph launch
  (PointSamples_driver.jl:98-118 per request);
* differentiation: 1e6 Plummer particles, the adjoint of a 1024^2 column
  density map's L2 loss.

Every workload is generated by the same seeded generators as bench.py.  The
oracle checks a strided subset of the samples (its CPU walk of every sample
would take minutes); the GPU's own traversal variants and single calls cover
the rest of each workload bit for bit."""
import ctypes as C
import time

import numpy as np
import pytest

from gpu_util import DevArray, DeviceSet, assert_same, catalog_fields, field_errors, oracle_of, traversal
from partia import _native as N
from partia import batch as PB
from partia import grids, synthetic

import pyoracle as O

pytestmark = pytest.mark.gpu
Z_FRAME = ([0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0])
FULL = dict(sslots=[1, 2], sflags=[0, 1], gslots=[2], dslots=[3, 4, 5], cslots=[3, 4, 5])
FP32C_NORMWISE = 1e-6


def log(msg):
    print("[configs] %s" % msg, flush=True)


def line_grid(d, cat, frame, ax, ay, precision):
    """ph_line_grid_device; returns (L x nx*ny float64, accepted neighbours)."""
    ctx = d.ctx
    ca = N.CatalogArgs(**cat)
    fr = [np.ascontiguousarray(np.asarray(v, dtype=np.float64)) for v in frame]
    S = len(ax) * len(ay)
    dt = np.float64 if precision == 64 else np.float32
    dax, day = DevArray(ctx, np.asarray(ax, np.float64)), DevArray(ctx, np.asarray(ay, np.float64))
    dout = DevArray(ctx, nbytes=ca.L * S * np.dtype(dt).itemsize)
    ctx.set_count_neighbours(True)
    try:
        ctx.check(N.lib().ph_line_grid_device(ctx.handle, d.h, ca.ref(), N.PH_SCATTER, precision,
                                              *[N.ptr(v) for v in fr], dax.p, len(ax), day.p, len(ay), dout.p))
        acc = ctx.stats()["accepted"]
        out = dout.get(dt, ca.L * S).astype(np.float64).reshape(ca.L, S)
    finally:
        ctx.set_count_neighbours(False)
        for b in (dax, day, dout):
            b.free()
    return out, acc


def image_rays(ax, ay, idx, dtype):
    """Origins and directions of rays idx (i fastest) of the z frame."""
    i, j = idx % len(ax), idx // len(ax)
    org = [np.asarray(ax, np.float64)[i], np.asarray(ay, np.float64)[j], np.zeros(len(idx))]
    org = [a.astype(dtype).astype(np.float64) for a in org]
    z = np.zeros(len(idx))
    return org, [z, z, z + 1.0]


@pytest.mark.timeout(600)
def test_lognormal_8e6_column_density_2048_float32(backend):
    ctx = backend.context
    t0 = time.time()
    w = synthetic.lognormal_field(8_000_000, seed=3)
    c6 = [w[k] for k in ("x", "y", "z", "m", "h", "rho")]
    cols = [w["rho"]]
    log("log-normal particles generated (%.1f s)" % (time.time() - t0))
    d = DeviceSet(ctx, *c6, cols, kernel=0, precision=32)
    cat = dict(sslots=[1], sflags=[0])
    nimg = 2048
    ax = synthetic.linrange(0.0, 1.0, nimg)
    S = nimg * nimg
    img, acc = line_grid(d, cat, Z_FRAME, ax, ax, 32)
    log("GPU image done: %d accepted neighbours (%.1f per ray)" % (acc, acc / S))
    assert 2000 < acc / S < 4000  # about 2.8e3 particles per ray
    assert np.isfinite(img).mean() > 0.99
    # one DFS per ray: the same bits and the same neighbour count
    with traversal(ctx, N.PH_TRAVERSAL_LANE):
        lane, acc_lane = line_grid(d, cat, Z_FRAME, ax, ax, 32)
    assert acc_lane == acc
    assert_same(lane, img)
    log("per-ray walk identical")
    # the oracle (Float32 instantiation) on every 1021st ray
    o = oracle_of(*c6, cols, kernel=0, precision=32)
    idx = np.arange(0, S, 1021)
    org, dirs = image_rays(ax, ax, idx, np.float32)
    ref = o.line_samples(cat, O.SCATTER, org, dirs, nthreads=16)
    assert_same(img[:, idx], ref)
    log("%d rays bit-identical to the oracle" % len(idx))
    # accepted neighbours of 256 of those rays: the oracle's neighbour lists
    # against the GPU's count over the same rays as an explicit ray list
    sub = idx[::16][:256]
    org, dirs = image_rays(ax, ax, sub, np.float32)
    n_ref = sum(len(o.neighbors([org[0][k], org[1][k], org[2][k]], O.SCATTER, direction=[0.0, 0.0, 1.0])[0])
                for k in range(len(sub)))
    ctx.set_count_neighbours(True)
    try:
        got = d.lines(cat, O.SCATTER, org, dirs, precision=32)
        n_gpu = ctx.stats()["accepted"]
    finally:
        ctx.set_count_neighbours(False)
    assert n_gpu == n_ref
    assert_same(got, img[:, sub])
    d.close()


def _request_coords(q):
    if q["kind"] == "points":
        st = synthetic.Stream(q["seed"])
        return [st.uniform(q["count"]) for _ in range(3)]
    if q["kind"] == "plane":
        return grids.cartesian_plane_coordinates(q["position"], q["right"], q["up"], q["axis"], q["axis"])
    return grids.cartesian_box_coordinates(q["position"], q["right"], q["up"], q["forward"], q["axis"], q["axis"],
                                           q["axis"])


@pytest.mark.timeout(900)
def test_uniform_clump_4e6_512_requests_batch(backend):
    """FP32C batch of all 512 requests: every request bit-identical to its own
    ph_point_samples_device call; float64 batch of the same requests: every
    32nd request's strided samples bit-identical to the oracle, and the FP32C
    outputs there within the normwise tolerance of the oracle."""
    ctx = backend.context
    w = synthetic.uniform_plus_clump(4_000_000, seed=4)
    mix = synthetic.request_mix(512, seed=4)
    c6 = [w[k] for k in ("x", "y", "z", "m", "h", "rho")]
    cols = [w["rho"], w["A"], w["vx"], w["vy"], w["vz"]]
    d = DeviceSet(ctx, *c6, cols, kernel=4)
    ca = N.CatalogArgs(**FULL)
    L = ca.L
    reqs32, reqs64, outs, pts, sizes = [], [], [], [], []
    for q in mix:
        S = (q["count"] if q["kind"] == "points" else
             q["axis"][2] ** 2 if q["kind"] == "plane" else q["axis"][2] ** 3)
        o32, o64 = DevArray(ctx, nbytes=4 * L * S), DevArray(ctx, nbytes=8 * L * S)
        outs.append((o32, o64))
        sizes.append(S)
        if q["kind"] == "points":
            P = _request_coords(q)
            dp = [DevArray(ctx, a) for a in P]
            pts.append(dp)
            reqs32.append(PB.points(dp[0].p, dp[1].p, dp[2].p, S, o32.p))
            reqs64.append(PB.points(dp[0].p, dp[1].p, dp[2].p, S, o64.p))
        elif q["kind"] == "plane":
            pts.append(None)
            reqs32.append(PB.plane(q["position"], q["right"], q["up"], q["axis"], q["axis"], o32.p))
            reqs64.append(PB.plane(q["position"], q["right"], q["up"], q["axis"], q["axis"], o64.p))
        else:
            pts.append(None)
            for reqs, o in ((reqs32, o32), (reqs64, o64)):
                reqs.append(PB.box(q["position"], q["right"], q["up"], q["forward"], q["axis"], q["axis"],
                                   q["axis"], o.p))
    log("%d requests, %d samples" % (len(mix), sum(sizes)))
    b32 = PB.Batch(ctx, d.h, ca, N.PH_SYMMETRIC, N.PH_PREC_FP32C, reqs32)
    b32.run()
    b64 = PB.Batch(ctx, d.h, ca, N.PH_SYMMETRIC, N.PH_PREC_FP64, reqs64)
    b64.run()
    ctx.check(N.lib().ph_synchronize(ctx.handle))
    log("both batches ran")
    o = oracle_of(*c6, cols, kernel=4)
    fields = catalog_fields(FULL)
    tmp = DevArray(ctx, nbytes=4 * L * max(sizes))
    for r, (q, S) in enumerate(zip(mix, sizes)):
        P = _request_coords(q)
        dp = pts[r] if pts[r] is not None else [DevArray(ctx, a) for a in P]
        ctx.check(N.lib().ph_point_samples_device(ctx.handle, d.h, ca.ref(), N.PH_SYMMETRIC, N.PH_PREC_FP32C, S,
                                                  dp[0].p, dp[1].p, dp[2].p, tmp.p))
        single = tmp.get(np.float32, L * S).reshape(L, S)
        got32 = outs[r][0].get(np.float32, L * S).reshape(L, S)
        assert np.array_equal(got32.view(np.uint32), single.view(np.uint32)), "request %d (%s)" % (r, q["kind"])
        if pts[r] is None:
            for a in dp:
                a.free()
        if r % 32 == 0:
            k = np.arange(0, S, max(1, S // 1500))
            ref = o.point_samples(FULL, O.SYMMETRIC, P[0][k], P[1][k], P[2][k], nthreads=16)
            got64 = outs[r][1].get(np.float64, L * S).reshape(L, S)[:, k]
            assert_same(got64, ref)
            err = field_errors(got32[:, k].astype(np.float64), ref, fields=fields)
            assert max(e[0] for e in err) <= FP32C_NORMWISE, (r, err)
            log("request %d (%s, %d samples): float64 bit-identical to the oracle on %d samples, FP32C %.2e"
                % (r, q["kind"], S, len(k), max(e[0] for e in err)))
    log("every request equals its single call")
    b32.close()
    b64.close()
    tmp.free()
    for o32, o64 in outs:
        o32.free()
        o64.free()
    for dp in pts:
        for a in dp or ():
            a.free()
    d.close()


def _column_map(ctx, c6, cols, ax):
    d = DeviceSet(ctx, *c6, cols, kernel=0)
    img, _ = line_grid(d, dict(sslots=[1], sflags=[0]), Z_FRAME, ax, ax, 64)
    return d, img[0]


@pytest.mark.timeout(900)
def test_plummer_1e6_column_density_1024_adjoint(backend):
    """The adjoint of the 1024^2 column-density map of 1e6 Plummer particles
    (M4, A = rho in slot 1, float64): against the oracle's adjoint for a
    cotangent supported on every 97th ray (1e-10 of each row's rms), and the
    whole-map loss against central differences of the GPU forward at 64
    particles."""
    ctx = backend.context
    w = synthetic.plummer_sphere(1_000_000, seed=5)
    st = synthetic.Stream(55)
    m_target = w["m"] * np.exp(0.1 * st.normal(len(w["m"])))
    c6 = [w[k] for k in ("x", "y", "z", "m", "h", "rho")]
    cols = [w["rho"]]
    nimg = 1024
    ax = synthetic.linrange(-1.0, 1.0, nimg)
    S = nimg * nimg
    d, sigma = _column_map(ctx, c6, cols, ax)
    dT, target = _column_map(ctx, [c6[0], c6[1], c6[2], m_target, c6[4], c6[5]], cols, ax)
    dT.close()
    cot_full = np.nan_to_num(sigma - target, nan=0.0)
    ca = N.CatalogArgs(sslots=[1], sflags=[0])
    fr = [np.ascontiguousarray(np.asarray(v, dtype=np.float64)) for v in Z_FRAME]
    dax = DevArray(ctx, ax)

    def adjoint(cot):
        dc, dg = DevArray(ctx, cot), DevArray(ctx, nbytes=6 * d.n * 8)
        ctx.check(N.lib().ph_line_grid_adjoint_device(ctx.handle, d.h, ca.ref(), N.PH_SCATTER, 64,
                                                      *[N.ptr(v) for v in fr], dax.p, nimg, dax.p, nimg, dc.p, dg.p))
        g = dg.get(np.float64, 6 * d.n).reshape(6, d.n)
        dc.free()
        dg.free()
        return g

    # cotangent on a strided ray subset against the oracle
    idx = np.arange(0, S, 97)
    cot = np.zeros(S)
    cot[idx] = cot_full[idx]
    g = adjoint(cot)
    o = oracle_of(*c6, cols, kernel=0)
    org, dirs = image_rays(ax, ax, idx, np.float64)
    ref = o.line_adjoint(1, False, org, dirs, cot[idx])
    for r in range(6):
        scale = max(np.sqrt(np.mean(ref[r] ** 2)), 1e-300)
        err = np.max(np.abs(g[r] - ref[r])) / scale
        assert err <= 1e-10, "row %d: max error %.3e of rms" % (r, err)
    log("adjoint of %d rays matches the oracle to 1e-10 of the row rms" % len(idx))
    # whole-map loss: central differences at 64 particles (m, h, x, y)
    g = adjoint(cot_full)

    def loss(p):
        dd, s = _column_map(ctx, p, cols, ax)
        dd.close()
        return 0.5 * float(np.nansum((s - target) ** 2))

    rng = np.random.default_rng(64)
    inside = np.nonzero((np.abs(c6[0]) < 0.9) & (np.abs(c6[1]) < 0.9))[0]
    picks = rng.choice(inside, 64, replace=False)
    worst = {}
    for row, pi in ((0, 3), (2, 4), (3, 0), (4, 1)):
        scale = np.sqrt(np.mean(g[row] ** 2))
        worst[row] = 0.0
        for b in picks:
            eps = 1e-6 * (abs(c6[pi][b]) if pi in (3, 4) else c6[4][b])
            pp, pm = [a.copy() for a in c6], [a.copy() for a in c6]
            pp[pi][b] += eps
            pm[pi][b] -= eps
            fd = (loss(pp) - loss(pm)) / (2 * eps)
            assert np.isfinite(fd)
            err = abs(fd - g[row, b])
            worst[row] = max(worst[row], err / scale)
            assert err <= 1e-3 * scale + 1e-4 * abs(g[row, b]), (row, b, fd, g[row, b])
        log("row %d: central differences at 64 particles, worst %.2e of rms" % (row, worst[row]))
    dax.free()
    d.close()
