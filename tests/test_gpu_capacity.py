"""Trees past the old 2^26-leaf bucket limit, toward the LinearBVH's 2^30
(LinearBoundingVolumeHierarchy/BinaryRadixTree/BinaryRadixTree.jl:86-91).

2^27 + 4099 uniform particles: bucket records then hold first-leaf indices
above 2^27, the leaves (32 bytes each) and the FP32C records pass 4 GiB, so the
per-lane FP32C walk takes its 64-bit-address variant and batches take the
union walk.  Every path is checked against an independent one on the same
tree: the float64 bucket walk against the float64 per-leaf LinearBVH walk (no
bucket records), the per-lane FP32C walk against the batch's union walk, the
FP32C grid walks over the 8- and 16-leaf trees against each other and against
float64."""
import numpy as np
import pytest

from gpu_util import DevArray, DeviceSet, assert_same, traversal
from partia import _native as N
from partia import batch as PB
from partia import synthetic

pytestmark = pytest.mark.gpu
CAT = dict(sslots=[1, 2], sflags=[0, 1])


@pytest.fixture(scope="module")
def big(backend):
    n = (1 << 27) + 4099
    print("capacity: n=%d" % n, flush=True)
    rng = np.random.default_rng(2027)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = np.full(n, 1.0 / n)
    h = np.full(n, 1.3 * n ** (-1.0 / 3.0))
    rho = 1.0 + 0.1 * rng.random(n)
    A = np.sin(2 * np.pi * x) + y
    d = DeviceSet(backend.context, x, y, z, m, h, rho, [rho, A], kernel=4)
    del x, y, z, m, h, rho, A
    yield d
    d.close()


def test_float64_bucket_walk_equals_the_leaf_walk(backend, big):
    ctx = backend.context
    P = np.random.default_rng(1).random((3, 4096))
    for st in (N.PH_GATHER, N.PH_SYMMETRIC):
        got = big.points(CAT, st, *P, precision=N.PH_PREC_FP64)
        with traversal(ctx, N.PH_TRAVERSAL_LANE):
            ref = big.points(CAT, st, *P, precision=N.PH_PREC_FP64)
        assert_same(got, ref)
        assert np.isfinite(ref).all()


def test_fp32c_wide_lane_walk_equals_the_union_walk(backend, big):
    ctx = backend.context
    P = np.random.default_rng(2).random((3, 3000))
    ca = N.CatalogArgs(**CAT)
    got = big.points(CAT, N.PH_SYMMETRIC, *P, precision=N.PH_PREC_FP32C)
    dp = [DevArray(ctx, a) for a in P]
    o = DevArray(ctx, nbytes=ca.L * P.shape[1] * 4)
    try:
        for mode in (N.PH_TRAVERSAL_AUTO, N.PH_TRAVERSAL_WAVE):
            with traversal(ctx, mode):
                b = PB.Batch(ctx, big.h, ca, N.PH_SYMMETRIC, N.PH_PREC_FP32C,
                             [PB.points(dp[0].p, dp[1].p, dp[2].p, P.shape[1], o.p)])
            try:
                b.run()
                assert_same(o.get(np.float32, ca.L * P.shape[1]).astype(np.float64).reshape(ca.L, -1), got)
            finally:
                b.close()
    finally:
        for a in dp + [o]:
            a.free()
    ref = big.points(CAT, N.PH_SYMMETRIC, *P, precision=N.PH_PREC_FP64)
    for k in range(ref.shape[0]):
        assert np.linalg.norm(got[k] - ref[k]) <= 1e-6 * np.linalg.norm(ref[k])


def test_fp32c_grids_over_every_bucket_tree(backend, big):
    ctx = backend.context
    ax = synthetic.linrange(0.02, 0.98, 24)
    try:
        ref = big.grid(CAT, N.PH_SYMMETRIC, ax, ax, ax, precision=N.PH_PREC_FP32C)
        for leaves in (4, 16):
            ctx.set_tile_shape(leaves, 0)
            assert_same(big.grid(CAT, N.PH_SYMMETRIC, ax, ax, ax, precision=N.PH_PREC_FP32C), ref)
        ctx.set_tile_shape()
        with traversal(ctx, N.PH_TRAVERSAL_BUCKET):  # grids keep the tile walk past 32-bit offsets
            assert_same(big.grid(CAT, N.PH_SYMMETRIC, ax, ax, ax, precision=N.PH_PREC_FP32C), ref)
    finally:
        ctx.set_tile_shape()
    f64 = big.grid(CAT, N.PH_SYMMETRIC, ax, ax, ax, precision=N.PH_PREC_FP64)
    for k in range(f64.shape[0]):
        assert np.linalg.norm(ref[k] - f64[k]) <= 1e-6 * np.linalg.norm(f64[k])
