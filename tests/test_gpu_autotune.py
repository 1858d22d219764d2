"""Autotuner and launch configurations (ph_set_launch, ph_set_autotune): every
configuration must give bit-identical outputs, since neither knob changes the
neighbour order or any arithmetic; only the speed may change."""
import numpy as np
import pytest

from gpu_util import DeviceSet, assert_same
from partia import ArgumentError
from partia import _native as N
from partia import synthetic

pytestmark = pytest.mark.gpu
FULL = dict(sslots=[1, 2], sflags=[0, 1], gslots=[2], dslots=[3, 4, 5], cslots=[3, 4, 5])
CONFIGS = ([(occ, xm, o) for occ in (1.0, 0.75, 0.5) for xm in (-1, 2, 1, 0) for o in (-1, 0)]
           + [(1.0, xm, o) for xm in range(8) for o in (0, 1) if (xm, o) not in ((2, 0), (1, 0), (0, 0))])


def plummer_set(ctx, n=60000):
    w = synthetic.plummer_sphere(n, seed=21)
    c6 = [w[k] for k in ("x", "y", "z", "m", "h", "rho")]
    cols = [w["rho"], w["A"], w["vx"], w["vy"], w["vz"]]
    return DeviceSet(ctx, *c6, cols, kernel=4)


@pytest.mark.parametrize("precision", [N.PH_PREC_FP64, N.PH_PREC_FP32C])
def test_every_launch_configuration_gives_identical_outputs(backend, precision):
    ctx = backend.context
    d = plummer_set(ctx)
    ax = synthetic.linrange(-0.5, 0.5, 37)
    rng = np.random.default_rng(3)
    P = rng.random((3, 5000)) - 0.5
    try:
        ref_g = d.grid(FULL, N.PH_SYMMETRIC, ax, ax, ax, precision=precision)
        ref_p = d.points(FULL, N.PH_GATHER, *P, precision=precision)
        for occ, xm, order in CONFIGS:
            ctx.set_launch(occ, xm, order)
            assert_same(d.grid(FULL, N.PH_SYMMETRIC, ax, ax, ax, precision=precision), ref_g)
            assert_same(d.points(FULL, N.PH_GATHER, *P, precision=precision), ref_p)
    finally:
        ctx.set_launch()


SHAPES = [(4, 256), (16, 256), (8, 64), (8, 128), (8, 512), (16, 512), (4, 64), (16, 128)]


@pytest.mark.parametrize("precision", [N.PH_PREC_FP64, N.PH_PREC_FP32C])
def test_every_tile_shape_gives_identical_outputs(backend, precision):
    """The grid tile walks' bucket size (4, 8, 16 leaves: the 16-leaf tree is
    built on first use) and block size change only which candidates are
    staged and how waves are grouped: outputs bit-identical, ragged grids
    included (partial tiles, a last block with fewer tiles than waves)."""
    ctx = backend.context
    d = plummer_set(ctx, 50000)
    axes = [(synthetic.linrange(-0.5, 0.5, 37),) * 3,
            (synthetic.linrange(-0.4, 0.45, 13), synthetic.linrange(-0.3, 0.3, 6), synthetic.linrange(-0.2, 0.5, 9))]
    try:
        for ax in axes:
            for st in (N.PH_GATHER, N.PH_SYMMETRIC, N.PH_SCATTER):
                ctx.set_tile_shape()
                ref = d.grid(FULL, st, *ax, precision=precision)
                for leaves, block in SHAPES:
                    ctx.set_tile_shape(leaves, block)
                    assert_same(d.grid(FULL, st, *ax, precision=precision), ref)
    finally:
        ctx.set_tile_shape()


def test_set_tile_shape_rejects_bad_values(backend):
    ctx = backend.context
    for bad in ((2, 256), (32, 256), (8, 96), (8, 1024), (-1, 0)):
        with pytest.raises(ArgumentError):
            ctx.set_tile_shape(*bad)
    ctx.set_tile_shape()


def test_autotuner_keeps_results_and_records_one_choice_per_key(backend):
    ctx = backend.context
    d = plummer_set(ctx, 40000)
    ax = synthetic.linrange(-0.5, 0.5, 41)
    ref = d.grid(FULL, N.PH_SYMMETRIC, ax, ax, ax, precision=N.PH_PREC_FP32C)
    before = {e["key"] for e in ctx.autotune_report()}
    ctx.set_autotune(True)
    try:
        first = d.grid(FULL, N.PH_SYMMETRIC, ax, ax, ax, precision=N.PH_PREC_FP32C)  # tunes
        rep = [e for e in ctx.autotune_report() if e["key"] not in before]
        assert len(rep) == 1
        e = rep[0]
        assert (float(e["occupancy"]), int(e["xcd_map"]), int(e["tile_order"])) in CONFIGS
        assert int(e["bucket_leaves"]) in (4, 8, 16) and int(e["block_size"]) in (128, 256, 512)
        assert 0.0 < e["best_ms"] <= e["default_ms"]
        second = d.grid(FULL, N.PH_SYMMETRIC, ax, ax, ax, precision=N.PH_PREC_FP32C)  # uses the choice
        assert len([e for e in ctx.autotune_report() if e["key"] not in before]) == 1
        assert_same(first, ref)
        assert_same(second, ref)
    finally:
        ctx.set_autotune(False)


def test_set_launch_rejects_bad_values(backend):
    ctx = backend.context
    with pytest.raises(ArgumentError):
        ctx.set_launch(0.0, 1)
    with pytest.raises(ArgumentError):
        ctx.set_launch(1.0, 8)
    with pytest.raises(ArgumentError):
        ctx.set_launch(1.0, -2)
    with pytest.raises(ArgumentError):
        ctx.set_launch(1.0, 1, 2)
    with pytest.raises(ArgumentError):
        ctx.set_launch(1.0, 1, -2)
