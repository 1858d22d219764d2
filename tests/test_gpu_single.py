"""The single-quantity kernels on the GPU (ph_point_single, ph_line_samples_ha,
ph_line_density) against the oracle's restatement of _density_kernel,
_number_density_kernel, _gradient_density_kernel
(scalar_interpolation.jl:1-166, gradient_interpolation.jl:1-170),
_line_integrated_quantities_interpolate_kernel in every strategy and
_line_integrated_density_kernel (line_integrated_scalar_interpolation.jl:1-214):
bit-identical in float64 and in Float32, NaN masks equal."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import DeviceSet, assert_same, oracle_of
from partia import ArgumentError
from partia import _native as N

import pyoracle as O

pytestmark = pytest.mark.gpu
STRATS = [O.GATHER, O.SCATTER, O.SYMMETRIC]


def cloud(n, seed, f32=False):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.random(n) + 0.05
    h = (rng.random(n) * 0.12 + 0.04) * (200.0 / n) ** (1 / 3)
    rho = rng.random(n) + 0.3
    cols = [rho, np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y), x + 0.3 * y]
    c6 = [x, y, z, m, h, rho]
    if f32:
        c6 = [a.astype(np.float32).astype(np.float64) for a in c6]
        cols = [a.astype(np.float32).astype(np.float64) for a in cols]
    return c6, cols


def point_single(d, kind, strategy, P, ha=None, precision=64):
    S = P.shape[1]
    L = 3 if kind == N.PH_SINGLE_GRADIENT_DENSITY else 1
    out = np.zeros(L * S)
    x, y, z = (N.f64(P[k]) for k in range(3))
    hv = N.f64(ha) if ha is not None else None
    d.ctx.check(N.lib().ph_point_single(d.ctx.handle, d.h, kind, strategy, precision, S, N.ptr(x), N.ptr(y),
                                       N.ptr(z), N.ptr(hv), N.ptr(out)))
    return out.reshape(L, S)


def line_call(d, cat, strategy, org, dirs, ha, precision=64):
    S = len(org[0])
    arrs = [N.f64(a) for a in list(org) + list(dirs)]
    hv = N.f64(ha) if ha is not None else None
    if cat is None:
        out = np.zeros(S)
        d.ctx.check(N.lib().ph_line_density(d.ctx.handle, d.h, strategy, precision, S, *[N.ptr(a) for a in arrs],
                                            N.ptr(hv), N.ptr(out)))
        return out.reshape(1, S)
    ca = N.CatalogArgs(**cat)
    out = np.zeros(ca.L * S)
    d.ctx.check(N.lib().ph_line_samples_ha(d.ctx.handle, d.h, ca.ref(), strategy, precision, S,
                                           *[N.ptr(a) for a in arrs], N.ptr(hv), N.ptr(out)))
    return out.reshape(ca.L, S)


def rays(S, seed, f32=False):
    rng = np.random.default_rng(seed)
    org = rng.random((3, S)) * 1.2 - 0.1
    dirs = rng.normal(size=(3, S))
    dirs /= np.linalg.norm(dirs, axis=0)
    dirs[:, : S // 3] = np.array([[0.0], [0.0], [1.0]])
    if f32:
        org, dirs = (a.astype(np.float32).astype(np.float64) for a in (org, dirs))
    return org, dirs


@pytest.mark.parametrize("kernel", range(6))
@pytest.mark.parametrize("strategy", STRATS)
def test_point_single_kernels_bitwise(backend, kernel, strategy):
    c6, cols = cloud(4000, 500 + kernel)
    d, o = DeviceSet(backend.context, *c6, cols, kernel=kernel), oracle_of(*c6, cols, kernel=kernel)
    P = np.random.default_rng(kernel).random((3, 1500)) * 1.3 - 0.15
    ha = np.random.default_rng(7).random(1500) * 0.05 + 0.01
    for kind in (N.PH_SINGLE_DENSITY, N.PH_SINGLE_NUMBER_DENSITY, N.PH_SINGLE_GRADIENT_DENSITY):
        assert_same(point_single(d, kind, strategy, P), o.point_single(kind, strategy, *P))
        assert_same(point_single(d, kind, strategy, P, ha=ha), o.point_single(kind, strategy, *P, ha=ha))


@pytest.mark.parametrize("strategy", STRATS)
def test_point_single_float32_tree(backend, strategy):
    c6, cols = cloud(3000, 61, f32=True)
    d = DeviceSet(backend.context, *c6, cols, kernel=0, precision=32)
    o = oracle_of(*c6, cols, kernel=0, precision=32)
    P = (np.random.default_rng(3).random((3, 1000)) * 1.2 - 0.1).astype(np.float32).astype(np.float64)
    for kind in (0, 1, 2):
        assert_same(point_single(d, kind, strategy, P, precision=32), o.point_single(kind, strategy, *P))


@pytest.mark.parametrize("kernel", [0, 1, 4])
@pytest.mark.parametrize("strategy", STRATS)
def test_line_kernels_with_h_a_bitwise(backend, kernel, strategy):
    c6, cols = cloud(4000, 600 + kernel)
    d, o = DeviceSet(backend.context, *c6, cols, kernel=kernel), oracle_of(*c6, cols, kernel=kernel)
    org, dirs = rays(800, kernel)
    ha = np.random.default_rng(9).random(800) * 0.04 + 0.01
    cat = dict(sslots=[1, 2, 3], sflags=[0, 1, 1])
    assert_same(line_call(d, cat, strategy, org, dirs, ha), o.line_samples(cat, strategy, org, dirs, ha=ha))
    got = line_call(d, None, strategy, org, dirs, ha)
    ref = o.line_density(strategy, org, dirs, ha=ha)[None, :]
    assert_same(got, ref)
    assert (got > 0).sum() > 100 and (got == 0).any()


@pytest.mark.parametrize("strategy", STRATS)
def test_line_kernels_float32_and_tiny_trees(backend, strategy):
    c6, cols = cloud(2500, 71, f32=True)
    d = DeviceSet(backend.context, *c6, cols, kernel=0, precision=32)
    o = oracle_of(*c6, cols, kernel=0, precision=32)
    org, dirs = rays(600, 5, f32=True)
    ha = (np.random.default_rng(2).random(600) * 0.05 + 0.01).astype(np.float32).astype(np.float64)
    cat = dict(sslots=[2], sflags=[1])
    assert_same(line_call(d, cat, strategy, org, dirs, ha, 32), o.line_samples(cat, strategy, org, dirs, ha=ha))
    assert_same(line_call(d, None, strategy, org, dirs, ha, 32), o.line_density(strategy, org, dirs, ha=ha)[None, :])
    for n in (1, 2, 5):
        c6, cols = cloud(n, 80 + n)
        c6[4] = c6[4] * 0 + 0.4
        d, o = DeviceSet(backend.context, *c6, cols), oracle_of(*c6, cols)
        org, dirs = rays(50, n)
        ha = np.full(50, 0.3)
        assert_same(line_call(d, None, strategy, org, dirs, ha), o.line_density(strategy, org, dirs, ha=ha)[None, :])
        P = np.random.default_rng(n).random((3, 50))
        assert_same(point_single(d, 2, strategy, P, ha=ha), o.point_single(2, strategy, *P, ha=ha))


def test_single_kernel_errors(backend):
    c6, cols = cloud(300, 1)
    d = DeviceSet(backend.context, *c6, cols)
    P = np.random.default_rng(0).random((3, 10))
    with pytest.raises(ArgumentError, match="PH_PREC_FP64 or PH_PREC_FP32"):
        point_single(d, 0, O.GATHER, P, precision=N.PH_PREC_FP32C)
    with pytest.raises(ArgumentError, match="unknown single-quantity kind"):
        point_single(d, 7, O.GATHER, P)
    org, dirs = rays(10, 1)
    with pytest.raises(ArgumentError, match="smoothing length ha"):
        line_call(d, None, O.GATHER, org, dirs, None)
    with pytest.raises(ArgumentError, match="scalar quantities only"):
        line_call(d, dict(gslots=[2]), O.SCATTER, org, dirs, None)
    # scatter needs no h_a and equals the driver's line projection
    cat = dict(sslots=[1, 2], sflags=[0, 1])
    assert_same(line_call(d, cat, O.SCATTER, org, dirs, None), d.lines(cat, O.SCATTER, org, dirs))
