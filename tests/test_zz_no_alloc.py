"""ph_point_samples_device_noalloc: the per-call buffers live in the library's
static device workspace.  Its outputs must equal ph_point_samples_device's bit
for bit (float64 and FP32-compensated, every strategy, catalogs of one and of
two chunks), on two calls in a row, the second with a smaller tree and fewer
samples, and sizes beyond the workspace must be rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import DevArray, DeviceSet
from partia import _native as N

import pyoracle as O

pytestmark = pytest.mark.gpu
FULL = dict(sslots=[1, 2], sflags=[0, 1], gslots=[2], dslots=[3, 4, 5], cslots=[3, 4, 5])
# six scalars: two catalog chunks (at most four scalars per chunk)
WIDE = dict(sslots=[1, 2, 3, 4, 5, 2], sflags=[0, 1, 1, 1, 1, 0])
WORKSPACE_BYTES = 64 << 20


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.random(n) + 0.05
    h = (rng.random(n) * 0.12 + 0.04) * (200.0 / n) ** (1 / 3)
    rho = rng.random(n) + 0.3
    A = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    return (x, y, z, m, h, rho), [rho, A, x + 0.3 * y, y - z, z * x]


def run(ctx, d, cat, strategy, precision, P, noalloc):
    ca = N.CatalogArgs(**cat)
    S = P.shape[1]
    pts = [DevArray(ctx, P[k]) for k in range(3)]
    es = 8 if precision == N.PH_PREC_FP64 else 4
    out = DevArray(ctx, nbytes=ca.L * S * es)
    fn = N.lib().ph_point_samples_device_noalloc if noalloc else N.lib().ph_point_samples_device
    print("noalloc=%d strategy=%d precision=%d n=%d S=%d L=%d" % (noalloc, strategy, precision, d.n, S, ca.L),
          flush=True)
    ctx.check(fn(ctx.handle, d.h, ca.ref(), strategy, precision, S, pts[0].p, pts[1].p, pts[2].p, out.p))
    ctx.check(N.lib().ph_synchronize(ctx.handle))
    got = out.get(np.float64 if es == 8 else np.float32, ca.L * S).reshape(ca.L, S)
    for a in pts + [out]:
        a.free()
    return got


def same_bits(a, b):
    assert a.shape == b.shape
    assert np.array_equal(a.view(np.uint64 if a.dtype == np.float64 else np.uint32),
                          b.view(np.uint64 if b.dtype == np.float64 else np.uint32))


@pytest.mark.parametrize("precision", [N.PH_PREC_FP64, N.PH_PREC_FP32C])
def test_noalloc_equals_allocating_entry_point(backend, precision):
    ctx = backend.context
    rng = np.random.default_rng(5)
    # first call: larger tree and sample set; second call: smaller ones
    for n, S, seed in ((6000, 5000, 11), (700, 900, 12)):
        c6, cols = cloud(n, seed)
        d = DeviceSet(ctx, *c6, cols, kernel=4)
        P = rng.random((3, S)) * 1.2 - 0.1
        for strategy in (O.GATHER, O.SCATTER, O.SYMMETRIC):
            for cat in (FULL, WIDE):
                ref = run(ctx, d, cat, strategy, precision, P, noalloc=False)
                got = run(ctx, d, cat, strategy, precision, P, noalloc=True)
                same_bits(got, ref)
                assert np.isfinite(got).any()
        d.close()


def test_noalloc_rejects_sizes_beyond_the_workspace(backend):
    ctx = backend.context
    c6, cols = cloud(500, 3)
    d = DeviceSet(ctx, *c6, cols, kernel=0)
    ca = N.CatalogArgs(**FULL)
    dummy = DevArray(ctx, nbytes=64)
    S = WORKSPACE_BYTES // 8 + 1  # the h_a of S samples alone exceeds the workspace
    rc = N.lib().ph_point_samples_device_noalloc(ctx.handle, d.h, ca.ref(), O.SYMMETRIC, N.PH_PREC_FP64, S,
                                                 dummy.p, dummy.p, dummy.p, dummy.p)
    assert rc == N.PH_ERR_CAPACITY
    assert b"workspace" in N.lib().ph_last_error(ctx.handle)
    # a Float32 evaluation is not offered by this entry point
    rc = N.lib().ph_point_samples_device_noalloc(ctx.handle, d.h, ca.ref(), O.SYMMETRIC, N.PH_PREC_FP32, 10,
                                                 dummy.p, dummy.p, dummy.p, dummy.p)
    assert rc == N.PH_ERR_ARGUMENT
    dummy.free()
    d.close()
