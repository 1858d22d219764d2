"""GPU interpolation against the oracle on the same seeded inputs.

The float64 path evaluates the reference's arithmetic in the reference's order,
so every field is required to be bit-identical to the oracle (NaN masks equal);
the float32 path likewise against the oracle's Float32 instantiation."""
import numpy as np
import pytest

from gpu_util import DeviceSet, assert_same, grid_points, oracle_of
from partia import synthetic

import pyoracle as O

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(99)
FULL = dict(sslots=[1, 2], sflags=[0, 1], gslots=[2], dslots=[3, 4, 5], cslots=[3, 4, 5])
KERNELS = list(range(6))


def random_cloud(n, seed=0):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.random(n) + 0.05
    h = (rng.random(n) * 0.12 + 0.04) * (200.0 / n) ** (1 / 3)
    rho = rng.random(n) + 0.3
    A = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    cols = [rho, A, x + 0.3 * y, y - z, z * x]
    return (x, y, z, m, h, rho), cols


def both(ctx, n=5000, kernel=0, seed=0, code_bits=63, precision=64):
    c6, cols = random_cloud(n, seed)
    d = DeviceSet(ctx, *c6, cols, kernel=kernel, code_bits=code_bits, precision=precision)
    o = oracle_of(*c6, cols, kernel=kernel, code_bits=code_bits, precision=precision)
    return d, o


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("strategy", [O.GATHER, O.SCATTER, O.SYMMETRIC])
def test_points_bitwise(backend, kernel, strategy):
    d, o = both(backend.context, 4000, kernel, seed=kernel)
    S = 3000
    P = RNG.random((3, S)) * 1.2 - 0.1  # includes samples outside the cloud (NaN outputs)
    assert_same(d.points(FULL, strategy, *P), o.point_samples(FULL, strategy, *P))


@pytest.mark.parametrize("strategy", [O.GATHER, O.SCATTER, O.SYMMETRIC])
def test_ragged_grid_bitwise(backend, strategy):
    d, o = both(backend.context, 8000, 0, seed=3)
    axes = (synthetic.linrange(-0.1, 1.1, 13), synthetic.linrange(0.0, 1.0, 7), synthetic.linrange(0.2, 0.9, 5))
    assert_same(d.grid(FULL, strategy, *axes), o.point_samples(FULL, strategy, *grid_points(*axes)))


def test_many_outputs_split_into_chunks(backend):
    """Catalogs larger than one kernel pass (6 scalars, 2 gradients, 2 divergences, 2 curls)."""
    d, o = both(backend.context, 6000, 4, seed=5)
    cat = dict(sslots=[1, 2, 3, 4, 5, 2], sflags=[0, 1, 1, 1, 1, 0], gslots=[2, 3], dslots=[3, 4, 5, 5, 4, 3],
               cslots=[3, 4, 5, 2, 2, 2])
    P = RNG.random((3, 2000))
    assert_same(d.points(cat, O.SYMMETRIC, *P), o.point_samples(cat, O.SYMMETRIC, *P))


def test_density_gradient_slot_zero(backend):
    """grad_slots entry 0 selects the density-gradient branch (general_interpolation.jl:76-82)."""
    d, o = both(backend.context, 5000, 0, seed=8)
    cat = dict(sslots=[1], sflags=[0], gslots=[0, 2])
    P = RNG.random((3, 1500))
    for strategy in (O.GATHER, O.SCATTER, O.SYMMETRIC):
        assert_same(d.points(cat, strategy, *P), o.point_samples(cat, strategy, *P))


def test_nearest_h_bitwise(backend):
    d, o = both(backend.context, 20000, 0, seed=11)
    P = RNG.random((3, 5000)) * 1.5 - 0.25
    assert np.array_equal(d.nearest_h(*P), o.nearest_h(*P)[0])
    # lattice with exact ties: the first leaf in Morton order wins
    c = np.linspace(0.0, 1.0, 9)
    g = np.array(np.meshgrid(c, c, c, indexing="ij")).reshape(3, -1)
    n = g.shape[1]
    hl = np.arange(1.0, n + 1)
    dl = DeviceSet(d.ctx, *g, np.ones(n), hl, np.ones(n))
    ol = oracle_of(*g, np.ones(n), hl, np.ones(n))
    q = (np.arange(1, 8) / 8.0 - 1.0 / 16.0)
    Q = np.array(np.meshgrid(q, q, q, indexing="ij")).reshape(3, -1)
    assert np.array_equal(dl.nearest_h(*Q), ol.nearest_h(*Q)[0])


@pytest.mark.parametrize("kernel", KERNELS)
def test_lines_bitwise(backend, kernel):
    d, o = both(backend.context, 6000, kernel, seed=20 + kernel)
    S = 2000
    org = RNG.random((3, S))
    dirs = RNG.normal(size=(3, S))
    dirs /= np.linalg.norm(dirs, axis=0)
    dirs[:, :200] = np.array([[0.0], [0.0], [1.0]])  # axis-aligned rays (zero direction components)
    cat = dict(sslots=[1, 2, 3], sflags=[0, 1, 1])
    assert_same(d.lines(cat, O.SCATTER, org, dirs), o.line_samples(cat, O.SCATTER, org, dirs))


@pytest.mark.parametrize("strategy", [O.GATHER, O.SYMMETRIC])
def test_float32_path_bitwise(backend, strategy):
    c6, cols = random_cloud(6000, 31)
    c6 = tuple(a.astype(np.float32).astype(np.float64) for a in c6)
    cols = [a.astype(np.float32).astype(np.float64) for a in cols]
    d = DeviceSet(backend.context, *c6, cols, kernel=0, precision=32)
    o = oracle_of(*c6, cols, kernel=0, precision=32)
    P = (RNG.random((3, 2000)) * 1.1 - 0.05).astype(np.float32).astype(np.float64)
    assert_same(d.points(FULL, strategy, *P, precision=32), o.point_samples(FULL, strategy, *P))
    org = RNG.random((3, 500)).astype(np.float32).astype(np.float64)
    z = np.zeros(500)
    dirs = (z, z, z + 1.0)
    cat = dict(sslots=[1], sflags=[0])
    assert_same(d.lines(cat, O.SCATTER, org, dirs, precision=32), o.line_samples(cat, O.SCATTER, org, dirs))


def test_thirty_bit_codes_same_fields(backend):
    """30-bit codes change the tree but not the neighbour sets: fields equal the
    63-bit oracle's up to summation order."""
    d30, o30 = both(backend.context, 8000, 0, seed=40, code_bits=30)
    P = RNG.random((3, 2000))
    assert_same(d30.points(FULL, O.SYMMETRIC, *P), o30.point_samples(FULL, O.SYMMETRIC, *P))
    o63 = both(backend.context, 8000, 0, seed=40, code_bits=63)[1]
    assert_same(d30.points(FULL, O.SYMMETRIC, *P), o63.point_samples(FULL, O.SYMMETRIC, *P), bitwise=False,
                rtol=1e-9)


def test_uniform_box_C1_full_size(backend):
    """1e5 uniform particles, M4, gather, density + scalar on the full 64^3 grid."""
    w = synthetic.uniform_box(100_000, seed=1)
    c6 = (w["x"], w["y"], w["z"], w["m"], w["h"], w["rho"])
    cols = [w["rho"], w["A"]]
    d = DeviceSet(backend.context, *c6, cols, kernel=0)
    o = oracle_of(*c6, cols, kernel=0)
    ax = synthetic.linrange(0.0, 1.0, 64)
    cat = dict(sslots=[1, 2], sflags=[0, 1])
    got = d.grid(cat, O.GATHER, ax, ax, ax)
    assert_same(got, o.point_samples(cat, O.GATHER, *grid_points(ax, ax, ax)))
    # physics sanity: interior density ~ 1, A follows sin(2 pi x) cos(2 pi y)
    X, Y, Z = grid_points(ax, ax, ax)
    inner = (X > 0.1) & (X < 0.9) & (Y > 0.1) & (Y < 0.9) & (Z > 0.1) & (Z < 0.9)
    assert abs(np.mean(got[0][inner]) - 1.0) < 0.02
    assert np.corrcoef(got[1][inner], synthetic.scalar_field(X, Y)[inner])[0, 1] > 0.99


def test_plummer_C2_full_size_subset(backend):
    """2e6 Plummer particles, C4, symmetric, L = 9 on the 256^3 grid: the full grid
    on the GPU, a strided subset of it against the oracle (bit-identical), and
    size-independent properties of the whole grid."""
    w = synthetic.plummer_sphere(2_000_000, seed=2)
    c6 = (w["x"], w["y"], w["z"], w["m"], w["h"], w["rho"])
    cols = [w["rho"], w["A"], w["vx"], w["vy"], w["vz"]]
    d = DeviceSet(backend.context, *c6, cols, kernel=4)
    ax = synthetic.linrange(-1.0, 1.0, 256)
    got = d.grid(FULL, O.SYMMETRIC, ax, ax, ax)
    X, Y, Z = grid_points(ax, ax, ax)
    idx = np.arange(0, 256 ** 3, 1031)
    o = oracle_of(*c6, cols, kernel=4)
    ref = o.point_samples(FULL, O.SYMMETRIC, X[idx], Y[idx], Z[idx])
    assert_same(got[:, idx], ref)
    nan = np.isnan(got[0])
    assert np.array_equal(np.isnan(got), np.broadcast_to(nan, got.shape))  # all-or-nothing per sample
    assert 0.2 < nan.mean() < 0.45
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    sel = (~nan) & (r > 0.1) & (r < 0.6)
    # solid-body rotation plus N(0, 0.01^2) noise: the noise dominates single
    # samples where h is small, but averages out: <div v> ~ 0, curl v ~ (0, 0, 2)
    assert abs(np.mean(got[5][sel])) < 0.1
    assert abs(np.median(got[8][sel]) - 2.0) < 0.3
