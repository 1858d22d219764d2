"""The bucket traversal: float64 results bit-identical to the oracle (and to the
per-leaf traversal) for every kernel and strategy, on point lists, ragged grids
and particle counts around the bucket size; the FP32-compensated evaluation
within the stated tolerance of the float64 oracle with identical neighbour sets."""
import numpy as np
import pytest

from gpu_util import DeviceSet, assert_same, catalog_fields, field_errors, grid_points, oracle_of, traversal
from partia import _native as N
from partia import synthetic

import pyoracle as O

pytestmark = pytest.mark.gpu
FULL = dict(sslots=[1, 2], sflags=[0, 1], gslots=[2], dslots=[3, 4, 5], cslots=[3, 4, 5])
STRATS = [O.GATHER, O.SCATTER, O.SYMMETRIC]
# FP32-compensated tolerance: normwise relative L2 error per output field (the
# three components of a gradient or curl form one field)
FP32C_NORMWISE = 1e-6


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.random(n) + 0.05
    h = (rng.random(n) * 0.12 + 0.04) * (200.0 / n) ** (1 / 3)
    rho = rng.random(n) + 0.3
    A = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    return (x, y, z, m, h, rho), [rho, A, x + 0.3 * y, y - z, z * x]


def pair(ctx, n, kernel, seed):
    c6, cols = cloud(n, seed)
    return DeviceSet(ctx, *c6, cols, kernel=kernel), oracle_of(*c6, cols, kernel=kernel)


@pytest.mark.parametrize("kernel", range(6))
@pytest.mark.parametrize("strategy", STRATS)
def test_bucket_f64_points_bitwise(backend, kernel, strategy):
    ctx = backend.context
    d, o = pair(ctx, 3000, kernel, 100 + kernel)
    P = np.random.default_rng(kernel).random((3, 2000)) * 1.2 - 0.1
    with traversal(ctx, N.PH_TRAVERSAL_BUCKET):
        assert_same(d.points(FULL, strategy, *P), o.point_samples(FULL, strategy, *P))


@pytest.mark.parametrize("strategy", STRATS)
def test_bucket_and_lane_grids_bitwise(backend, strategy):
    ctx = backend.context
    d, o = pair(ctx, 7000, 4, 7)
    axes = (synthetic.linrange(-0.1, 1.1, 13), synthetic.linrange(0.0, 1.0, 9), synthetic.linrange(0.2, 0.9, 6))
    ref = o.point_samples(FULL, strategy, *grid_points(*axes))
    for mode in (N.PH_TRAVERSAL_AUTO, N.PH_TRAVERSAL_BUCKET, N.PH_TRAVERSAL_LANE):
        with traversal(ctx, mode):
            assert_same(d.grid(FULL, strategy, *axes), ref)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 15, 16, 17, 33, 257])
def test_bucket_boundaries(backend, n):
    """Particle counts around the bucket size (4 leaves) and tiny trees (n = 1)."""
    ctx = backend.context
    c6, cols = cloud(n, 50 + n)
    c6 = (c6[0], c6[1], c6[2], c6[3], c6[4] * 0 + 0.3, c6[5])
    d, o = DeviceSet(ctx, *c6, cols, kernel=0), oracle_of(*c6, cols, kernel=0)
    P = np.random.default_rng(n).random((3, 300)) * 1.4 - 0.2
    for strategy in STRATS:
        ref = o.point_samples(FULL, strategy, *P)
        with traversal(ctx, N.PH_TRAVERSAL_BUCKET):
            assert_same(d.points(FULL, strategy, *P), ref)
        got = d.points(FULL, strategy, *P, precision=N.PH_PREC_FP32C)
        # with a handful of particles the derivative fields are close to zero;
        # measure them against the scale max|A| / h of their terms
        scale = max(np.max(np.abs(c)) for c in cols) / 0.3
        fields = catalog_fields(FULL)
        floors = [0.0, 0.0] + [scale] * (len(fields) - 2)
        for k, (nw, _) in enumerate(field_errors(got, ref, floor=floors, fields=fields)):
            assert nw <= FP32C_NORMWISE, (strategy, k, nw)


def test_bucket_many_outputs_and_density_gradient(backend):
    ctx = backend.context
    d, o = pair(ctx, 5000, 3, 9)
    cat = dict(sslots=[1, 2, 3, 4, 5, 2], sflags=[0, 1, 1, 1, 1, 0], gslots=[0, 2, 3], dslots=[3, 4, 5, 5, 4, 3],
               cslots=[3, 4, 5, 2, 2, 2])
    P = np.random.default_rng(3).random((3, 1500))
    ref = o.point_samples(cat, O.SYMMETRIC, *P)
    with traversal(ctx, N.PH_TRAVERSAL_BUCKET):
        assert_same(d.points(cat, O.SYMMETRIC, *P), ref)
    got = d.points(cat, O.SYMMETRIC, *P, precision=N.PH_PREC_FP32C)
    for k, (nw, _) in enumerate(field_errors(got, ref, fields=catalog_fields(cat))):
        assert nw <= FP32C_NORMWISE, (k, nw)


@pytest.mark.parametrize("kernel", range(6))
@pytest.mark.parametrize("strategy", STRATS)
def test_fp32c_points(backend, kernel, strategy):
    ctx = backend.context
    d, o = pair(ctx, 4000, kernel, 200 + kernel)
    P = np.random.default_rng(7 + kernel).random((3, 2500)) * 1.2 - 0.1
    ref = o.point_samples(FULL, strategy, *P)
    got = d.points(FULL, strategy, *P, precision=N.PH_PREC_FP32C)
    errs = field_errors(got, ref, fields=catalog_fields(FULL))
    print("FP32C kernel %d strategy %d (normwise, pointwise):" % (kernel, strategy),
          ["%.2e/%.2e" % e for e in errs])
    assert max(e[0] for e in errs) <= FP32C_NORMWISE, errs


def test_fp32c_same_neighbour_sets(backend):
    """The guard band makes the FP32C accepted set the float64 one (equal counts,
    identical NaN masks), also for samples placed exactly on a support boundary."""
    ctx = backend.context
    n = 4096
    g = (np.arange(16) + 0.5) / 16.0
    X = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1)
    h = np.full(n, 1.0 / 32.0)  # support radius 2h = one lattice spacing
    cols = [np.ones(n), X[0].copy()]
    d = DeviceSet(ctx, *X, np.ones(n), h, np.ones(n), cols, kernel=0)
    o = oracle_of(*X, np.ones(n), h, np.ones(n), cols, kernel=0)
    # samples on lattice points (distance exactly 1/16 to 6 neighbours) and between
    P = np.concatenate([X[:, ::7], X[:, ::11] + 1.0 / 32.0], axis=1)
    cat = dict(sslots=[1, 2], sflags=[0, 1], gslots=[2])
    counts = []
    for prec in (N.PH_PREC_FP64, N.PH_PREC_FP32C):
        ctx.set_count_neighbours(True)
        try:
            got = d.points(cat, O.SCATTER, *P, precision=prec)
            counts.append(ctx.stats()["accepted"])
        finally:
            ctx.set_count_neighbours(False)
        ref = o.point_samples(cat, O.SCATTER, *P)
        errs = field_errors(got, ref, fields=catalog_fields(cat))
        assert max(e[0] for e in errs) <= FP32C_NORMWISE, errs
    assert counts[0] == counts[1]


def test_fp32c_plummer_C2_full_grid(backend):
    """2e6 Plummer particles, C4, symmetric, L = 9 on the 256^3 grid: FP32C against
    the float64 GPU result of the whole grid (itself bit-identical to the oracle,
    checked on a subset), per-field normwise error <= 1e-6 and identical NaN masks."""
    ctx = backend.context
    w = synthetic.plummer_sphere(2_000_000, seed=2)
    c6 = (w["x"], w["y"], w["z"], w["m"], w["h"], w["rho"])
    cols = [w["rho"], w["A"], w["vx"], w["vy"], w["vz"]]
    d = DeviceSet(ctx, *c6, cols, kernel=4)
    ax = synthetic.linrange(-1.0, 1.0, 256)
    ref = d.grid(FULL, O.SYMMETRIC, ax, ax, ax)
    X, Y, Z = grid_points(ax, ax, ax)
    idx = np.arange(0, 256 ** 3, 4099)
    o = oracle_of(*c6, cols, kernel=4)
    assert_same(ref[:, idx], o.point_samples(FULL, O.SYMMETRIC, X[idx], Y[idx], Z[idx]))
    got = d.grid(FULL, O.SYMMETRIC, ax, ax, ax, precision=N.PH_PREC_FP32C)
    errs = field_errors(got, ref, fields=catalog_fields(FULL))
    print("FP32C vs FP64 (normwise, pointwise) per field:", ["%.2e/%.2e" % e for e in errs])
    assert max(e[0] for e in errs) <= FP32C_NORMWISE, errs


@pytest.mark.parametrize("strategy", [O.GATHER, O.SYMMETRIC])
def test_samples_beyond_every_support_are_nan(backend, strategy):
    """Samples with no particle within K max(h) skip the exact nearest-h search
    (their outputs are NaN whatever h_a is): float64 outputs stay bit-identical
    to the oracle's full search, FP32C has the same NaN mask, and samples just
    inside the bound keep their neighbours."""
    ctx = backend.context
    for kernel in (0, 2, 4):
        d, o = pair(ctx, 3000, kernel, 70 + kernel)
        rng = np.random.default_rng(kernel)
        far = rng.random((3, 3000)) * 7.0 - 3.0  # mostly outside [0,1]^3
        hmax = d.bvh()["leaf_scale"].max()
        K = O.support(kernel)
        edge = np.array([[1.0 + K * hmax * f for f in (0.2, 0.6, 0.95, 0.999)],
                         [0.5] * 4, [0.5] * 4])  # beside the cloud, within the bound
        P = np.concatenate([far, edge], axis=1)
        ref = o.point_samples(FULL, strategy, *P)
        assert np.isnan(ref[0, :3000]).mean() > 0.5
        with traversal(ctx, N.PH_TRAVERSAL_BUCKET):
            assert_same(d.points(FULL, strategy, *P), ref)
            got32 = d.points(FULL, strategy, *P, precision=N.PH_PREC_FP32C)
        assert np.array_equal(np.isnan(got32), np.isnan(ref))


WIDE = dict(sslots=[1, 2, 3, 4, 5], sflags=[0, 1, 1, 1, 0], gslots=[0, 3], dslots=[3, 4, 5], cslots=[5, 4, 2])
SMALL = dict(sslots=[1, 2], sflags=[0, 1])


@pytest.mark.parametrize("kernel", range(6))
@pytest.mark.parametrize("strategy", STRATS)
def test_tile_walk_equals_lane_walk_fp32c_grids(backend, kernel, strategy):
    """FP32C grids: the per-wave walk (default) and the per-lane walk give
    identical outputs and neighbour counts, for the C2 catalog (shared sums),
    a two-scalar catalog and a catalog split into chunks; ragged tile edges."""
    ctx = backend.context
    d, o = pair(ctx, 9000, kernel, 300 + kernel)
    axes = (synthetic.linrange(-0.15, 1.1, 19), synthetic.linrange(0.0, 1.0, 11), synthetic.linrange(0.1, 0.95, 7))
    for cat in (FULL, SMALL, WIDE):
        res = {}
        for mode in (N.PH_TRAVERSAL_AUTO, N.PH_TRAVERSAL_BUCKET):
            with traversal(ctx, mode):
                ctx.set_count_neighbours(True)
                try:
                    res[mode] = (d.grid(cat, strategy, *axes, precision=N.PH_PREC_FP32C), ctx.stats()["accepted"])
                finally:
                    ctx.set_count_neighbours(False)
        assert_same(res[N.PH_TRAVERSAL_AUTO][0], res[N.PH_TRAVERSAL_BUCKET][0])
        assert res[N.PH_TRAVERSAL_AUTO][1] == res[N.PH_TRAVERSAL_BUCKET][1]
    # and against the float64 oracle
    ref = o.point_samples(FULL, strategy, *grid_points(*axes))
    got = d.grid(FULL, strategy, *axes, precision=N.PH_PREC_FP32C)
    assert max(e[0] for e in field_errors(got, ref, fields=catalog_fields(FULL))) <= FP32C_NORMWISE


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 16, 17, 65, 300])
def test_tile_walk_small_trees_and_far_grids(backend, n):
    """Tiny trees (a single bucket, n = 1; n = 8 is one full 8-leaf bucket and
    n = 9 the first split of the tile kernel's tree), more staged leaves than
    one batch holds (large h), and grids reaching far beyond the particles."""
    ctx = backend.context
    c6, cols = cloud(n, 90 + n)
    c6 = (c6[0], c6[1], c6[2], c6[3], c6[4] * 0 + (0.9 if n > 60 else 0.3), c6[5])
    d = DeviceSet(ctx, *c6, cols, kernel=4)
    axes = (synthetic.linrange(-2.0, 3.0, 9), synthetic.linrange(-0.2, 1.2, 10), synthetic.linrange(0.0, 1.0, 5))
    for strategy in STRATS:
        res = {}
        for mode in (N.PH_TRAVERSAL_AUTO, N.PH_TRAVERSAL_BUCKET):
            with traversal(ctx, mode):
                res[mode] = d.grid(FULL, strategy, *axes, precision=N.PH_PREC_FP32C)
        assert_same(res[N.PH_TRAVERSAL_AUTO], res[N.PH_TRAVERSAL_BUCKET])
        # float64: the per-wave nearest-h search (default) against the per-leaf walk
        with traversal(ctx, N.PH_TRAVERSAL_LANE):
            lane = d.grid(FULL, strategy, *axes)
        assert_same(d.grid(FULL, strategy, *axes), lane)


def test_tree_outliving_its_context_after_fp32c_grid():
    """A tree whose context was destroyed after an FP32C grid call (which uses
    both bucket trees) can still be destroyed: the context released every
    device buffer of the tree, the 8-leaf bucket tree included."""
    ctx = N.Context(0)
    c6, cols = cloud(3000, 77)
    d = DeviceSet(ctx, *c6, cols, kernel=4)
    axes = (synthetic.linrange(0.0, 1.0, 8),) * 3
    got = d.grid(FULL, O.SYMMETRIC, *axes, precision=N.PH_PREC_FP32C)
    assert np.isfinite(got).any()
    ctx.close()
    d.close()  # ph_lbvh_destroy of an orphaned tree
    ctx2 = N.Context(0)  # the device is still usable
    d2 = DeviceSet(ctx2, *c6, cols, kernel=4)
    assert_same(d2.grid(FULL, O.SYMMETRIC, *axes, precision=N.PH_PREC_FP32C), got)
    d2.close()
    ctx2.close()


@pytest.mark.parametrize("kernel", range(6))
@pytest.mark.parametrize("strategy", STRATS)
def test_tile_walk_f64_grids_bitwise(backend, kernel, strategy):
    """Float64 grids: the per-wave walk (k_tile_f64, default) equals the
    per-lane walk and the oracle bit for bit with the same neighbour counts,
    for the C2 catalog, a two-scalar catalog and a catalog split into chunks
    with a density gradient; ragged tile edges."""
    ctx = backend.context
    d, o = pair(ctx, 9000, kernel, 400 + kernel)
    axes = (synthetic.linrange(-0.15, 1.1, 19), synthetic.linrange(0.0, 1.0, 11), synthetic.linrange(0.1, 0.95, 7))
    pts = grid_points(*axes)
    for cat in (FULL, SMALL, WIDE):
        res = {}
        for mode in (N.PH_TRAVERSAL_AUTO, N.PH_TRAVERSAL_BUCKET):
            with traversal(ctx, mode):
                ctx.set_count_neighbours(True)
                try:
                    res[mode] = (d.grid(cat, strategy, *axes), ctx.stats()["accepted"])
                finally:
                    ctx.set_count_neighbours(False)
        assert_same(res[N.PH_TRAVERSAL_AUTO][0], res[N.PH_TRAVERSAL_BUCKET][0])
        assert res[N.PH_TRAVERSAL_AUTO][1] == res[N.PH_TRAVERSAL_BUCKET][1]
        assert_same(res[N.PH_TRAVERSAL_AUTO][0], o.point_samples(cat, strategy, *pts))


@pytest.mark.parametrize("strategy", STRATS)
def test_nan_column_values_stay_local(backend, strategy):
    """A NaN in a column reaches only the samples that have that particle as a
    neighbour, as in the reference: lanes stepping in lockstep without a
    pending neighbour must not read it (the Morton-last particle, which
    per-lane kernels once used as a stand-in, is among the NaN ones)."""
    ctx = backend.context
    c6, cols = cloud(6000, 91)
    order = DeviceSet(ctx, *c6, cols, kernel=4).get("order", np.uint64, 6000)
    bad = np.concatenate([[int(order[-1]) - 1, int(order[0]) - 1], np.random.default_rng(3).integers(0, 6000, 6)])
    cols = [c.copy() for c in cols]
    cols[1][bad] = np.nan
    d, o = DeviceSet(ctx, *c6, cols, kernel=4), oracle_of(*c6, cols, kernel=4)
    axes = (synthetic.linrange(-0.05, 1.05, 21), synthetic.linrange(0.0, 1.0, 13), synthetic.linrange(0.0, 1.0, 9))
    P = np.random.default_rng(4).random((3, 3000)) * 1.1 - 0.05
    fields = catalog_fields(FULL)
    ref_g, ref_p = o.point_samples(FULL, strategy, *grid_points(*axes)), o.point_samples(FULL, strategy, *P)
    assert np.isnan(ref_g[1]).any() and np.isfinite(ref_g[1]).sum() > ref_g.shape[1] // 2
    for mode in (N.PH_TRAVERSAL_AUTO, N.PH_TRAVERSAL_BUCKET):
        with traversal(ctx, mode):
            assert_same(d.grid(FULL, strategy, *axes), ref_g)
            assert_same(d.points(FULL, strategy, *P), ref_p)
            for got, ref in ((d.grid(FULL, strategy, *axes, precision=N.PH_PREC_FP32C), ref_g),
                             (d.points(FULL, strategy, *P, precision=N.PH_PREC_FP32C), ref_p)):
                assert max(e[0] for e in field_errors(got, ref, fields=fields)) <= FP32C_NORMWISE
