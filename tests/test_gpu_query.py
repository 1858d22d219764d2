"""LBVH queries on the GPU (ph_query_device: LBVH_probe_neighbors and
LBVH_query!; ph_find_nearest_device: LBVH_find_nearest) against the oracle's
restatement, which test_oracle_lbvh.py pins to brute force and to the
reference's neighbour-query fixture.  Counts, pools (DFS order), closest
indices and distances are compared bit for bit, on 3D and 2D trees."""
import numpy as np
import pytest

from gpu_util import DevArray, DeviceSet, oracle_of
from partia import ArgumentError
from partia import _native as N

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(99)


def gpu_query(d, Q, r, with_pool=True, dims=3):
    ctx = d.ctx
    S = Q.shape[1]
    pts = [DevArray(ctx, N.f64(Q[k])) for k in range(3 if dims == 3 else 2)]
    dr = DevArray(ctx, N.f64(r))
    dc, db, dd = (DevArray(ctx, nbytes=8 * S) for _ in range(3))
    bufs = pts + [dr, dc, db, dd]
    try:
        pz = pts[2].p if dims == 3 else None
        ctx.check(N.lib().ph_query_device(ctx.handle, d.h, S, pts[0].p, pts[1].p, pz, dr.p, dc.p, db.p, dd.p, None,
                                          None))
        cnt = dc.get(np.int64, S)
        pool = None
        if with_pool:
            offs = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
            do = DevArray(ctx, offs)
            dp = DevArray(ctx, nbytes=8 * max(int(cnt.sum()), 1))
            bufs += [do, dp]
            ctx.check(N.lib().ph_query_device(ctx.handle, d.h, S, pts[0].p, pts[1].p, pz, dr.p, None, None, None,
                                              do.p, dp.p))
            pool = dp.get(np.int64, int(cnt.sum()))
        return cnt, db.get(np.int64, S), dd.get(np.float64, S), pool
    finally:
        for b in bufs:
            b.free()


def gpu_find_nearest(d, Q, dims=3):
    ctx = d.ctx
    S = Q.shape[1]
    pts = [DevArray(ctx, N.f64(Q[k])) for k in range(3 if dims == 3 else 2)]
    di, dd = DevArray(ctx, nbytes=8 * S), DevArray(ctx, nbytes=8 * S)
    try:
        ctx.check(N.lib().ph_find_nearest_device(ctx.handle, d.h, S, pts[0].p, pts[1].p,
                                                 pts[2].p if dims == 3 else None, di.p, dd.p))
        return di.get(np.int64, S), dd.get(np.float64, S)
    finally:
        for b in pts + [di, dd]:
            b.free()


@pytest.mark.parametrize("dims", [3, 2])
@pytest.mark.parametrize("n", [1, 5, 3000, 200_000])
def test_query_and_find_nearest_match_oracle(backend, dims, n):
    P = RNG.random((3, n))
    if n > 1000:
        P[:, : n // 4] = 0.3 + 1e-3 * RNG.random((3, n // 4))  # a dense clump
    if dims == 2:
        P[2] = 0.0
    h = np.full(n, 0.01)
    d = DeviceSet(backend.context, P[0], P[1], P[2], np.ones(n), h, np.ones(n), dims=dims)
    o = oracle_of(P[0], P[1], P[2], np.ones(n), h, np.ones(n), dims=dims)
    S = 3000
    Q = RNG.random((3, S)) * 1.2 - 0.1
    if dims == 2:
        Q[2] = 0.0
    r = RNG.random(S) * (0.02 if n > 1000 else 0.5)
    r[:3] = 0.0
    ref = o.query(*Q, r, with_pool=True)
    got = gpu_query(d, Q, r, dims=dims)
    for a, b, name in zip(got, ref, ("count", "closest", "closest_d2", "pool")):
        assert np.array_equal(a, b), name
    idx, d2 = gpu_find_nearest(d, Q, dims=dims)
    _, best = o.nearest_h(*Q)
    assert np.array_equal(idx, best)
    c = o.columns()
    exp = ((Q[0] - c[0][best - 1]) * (Q[0] - c[0][best - 1]) + (Q[1] - c[1][best - 1]) * (Q[1] - c[1][best - 1])) + \
        (Q[2] - c[2][best - 1]) * (Q[2] - c[2][best - 1])
    assert np.array_equal(d2, exp)


def test_query_errors(backend):
    n = 100
    P = RNG.random((3, n))
    d = DeviceSet(backend.context, *P, np.ones(n), np.full(n, 0.1), np.ones(n))
    ctx = d.ctx
    a = DevArray(ctx, np.zeros(4))
    try:
        with pytest.raises(ArgumentError):  # a 3D tree needs z
            ctx.check(N.lib().ph_find_nearest_device(ctx.handle, d.h, 4, a.p, a.p, None, a.p, None))
        with pytest.raises(ArgumentError):  # a pool needs offsets
            ctx.check(N.lib().ph_query_device(ctx.handle, d.h, 4, a.p, a.p, a.p, a.p, None, None, None, None, a.p))
        d32 = DeviceSet(ctx, *P, np.ones(n), np.full(n, 0.1), np.ones(n), precision=32)
        with pytest.raises(ArgumentError):  # float64 trees only
            ctx.check(N.lib().ph_find_nearest_device(ctx.handle, d32.h, 4, a.p, a.p, a.p, a.p, None))
    finally:
        a.free()
