"""The 2D path on the GPU: LinearBVH! of an InterpolationInput{2}
(ph_lbvh_build2d[_device]) against the oracle's 2D build, which
test_oracle_lbvh.py pins to an independent bit-by-bit Morton pipeline and
top-down tree (test/neighbor_search_tests.jl runs its invariants for D = 2 and
3).  Codes, order, topology and escapes are compared bit for bit, boxes and
scales exactly; nearest-h queries in the plane match the oracle bit for bit."""
import ctypes as C

import numpy as np
import pytest

import lbvh_ref as R
from gpu_util import DevArray, DeviceSet, oracle_of
from partia import ArgumentError
from partia import _native as N
from partia import synthetic

import pyoracle as O

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(2024)


def compare2d(ctx, cols5, code_bits=63, precision=64):
    x, y, m, h, rho = cols5
    d = DeviceSet(ctx, x, y, None, m, h, rho, code_bits=code_bits, precision=precision, dims=2)
    o = oracle_of(x, y, None, m, h, rho, code_bits=code_bits, precision=precision, dims=2)
    td, to = d.tree(), o.tree()
    for k in ("order", "codes", "left", "right", "escape", "parent"):
        assert np.array_equal(td[k], to[k]), k
    assert td["root"] == to["root"]
    bd, bo = d.bvh(), o.bvh()
    for k in bo:
        assert np.array_equal(bd[k], bo[k]), k
    cd = d.columns()
    assert np.array_equal(cd, o.columns()) and not cd[2].any()
    return d, o


@pytest.mark.parametrize("code_bits", [30, 63])
def test_2d_prime_hash_all_cases(backend, code_bits):
    for offset in range(5):
        for n in range(1, 21):
            x, y = R.prime_hash_coords(2, n, offset)
            compare2d(backend.context, (x, y, np.ones(n), np.full(n, 0.1), np.ones(n)), code_bits)


@pytest.mark.parametrize("n", [1, 2, 8, 16, 1000])
def test_2d_identical_codes(backend, n):
    compare2d(backend.context, (np.full(n, 0.5), np.full(n, 0.5), np.ones(n), np.full(n, 0.1), np.ones(n)))


@pytest.mark.parametrize("code_bits", [30, 63])
def test_2d_random_clustered_and_float32(backend, code_bits):
    for n in (3, 100, 4097, 100_000):
        P = RNG.random((2, n))
        P[:, : n // 3] = 0.25 + 1e-9 * RNG.random((2, n // 3))  # duplicate codes even at 32 bits per axis
        compare2d(backend.context, (*P, np.ones(n), RNG.random(n) * 0.1, np.ones(n)), code_bits)
    n = 30_000
    P = RNG.random((2, n)).astype(np.float32).astype(np.float64)
    h = (RNG.random(n) * 0.05).astype(np.float32).astype(np.float64)
    compare2d(backend.context, (*P, np.ones(n), h, np.ones(n)), code_bits, precision=32)


def test_2d_full_size_tree(backend):
    d = synthetic.plummer_sphere(2_000_000, seed=2)
    compare2d(backend.context, (d["x"], d["y"], d["m"], d["h"], d["rho"]))


def test_2d_device_build_nearest_h_and_dims(backend):
    ctx = backend.context
    n = 20_000
    P = RNG.random((2, n))
    h = RNG.random(n) * 0.02 + 0.001
    d, o = compare2d(ctx, (*P, np.ones(n), h, np.ones(n)))
    dims = C.c_int32()
    assert N.lib().ph_lbvh_dims(d.h, C.byref(dims)) == 0 and dims.value == 2
    # the device-pointer build gives the same tree
    arrs = [DevArray(ctx, N.f64(a)) for a in (P[0], P[1], np.ones(n), h, np.ones(n))]
    hd = C.c_void_p()
    try:
        ctx.check(N.lib().ph_lbvh_build2d_device(ctx.handle, n, *[a.p for a in arrs], 0, None, 0, 63, 64,
                                                 C.byref(hd)))
        for which, dt, size in (("order", np.uint64, n), ("codes", np.uint64, n), ("left", np.int32, 2 * n - 1),
                                ("escape", np.int32, 2 * n - 1), ("node_min", np.float64, 3 * (n - 1))):
            out = np.zeros(size, dt)
            ctx.check(N.lib().ph_lbvh_get(hd, N.GET[which], N.ptr(out)))
            assert np.array_equal(out, d.get(which, dt, size)), which
    finally:
        N.lib().ph_lbvh_destroy(hd)
        for a in arrs:
            a.free()
    # LBVH_find_nearest_h in the plane (pz null), bit for bit with the oracle
    Q = RNG.random((2, 4000)) * 1.3 - 0.15
    ha = np.zeros(4000)
    ctx.check(N.lib().ph_nearest_h(ctx.handle, d.h, 4000, N.ptr(N.f64(Q[0])), N.ptr(N.f64(Q[1])), None, N.ptr(ha)))
    ref, _ = o.nearest_h(Q[0], Q[1], np.zeros(4000))
    assert np.array_equal(ha, ref)


def test_2d_tree_rejected_by_interpolation(backend):
    ctx = backend.context
    n = 100
    P = RNG.random((2, n))
    d = DeviceSet(ctx, P[0], P[1], None, np.ones(n), np.full(n, 0.1), np.ones(n), cols=[np.ones(n)], dims=2)
    ca = N.CatalogArgs(sslots=[1], sflags=[1])
    pts = [DevArray(ctx, np.zeros(4)) for _ in range(3)]
    out = DevArray(ctx, np.zeros(4))
    try:
        with pytest.raises(ArgumentError):
            ctx.check(N.lib().ph_point_samples_device(ctx.handle, d.h, ca.ref(), O.GATHER, N.PH_PREC_FP64, 4,
                                                      *[a.p for a in pts], out.p))
        assert "InterpolationInput{2}" in N.lib().ph_last_error(ctx.handle).decode()
    finally:
        for a in pts + [out]:
            a.free()


def test_2d_host_linear_bvh(backend):
    """LinearBVH!(InterpolationInput{2}) through the host mirror: the input's
    columns are permuted into Morton order and the tree exposes 2D arrays."""
    from partia.input import InterpolationInput, LinearBVH_
    from partia.kernels import M4_spline
    n = 500
    P = RNG.random((2, n))
    m, h, rho, A = RNG.random(n) + 0.5, np.full(n, 0.05), RNG.random(n) + 1.0, RNG.random(n)
    inp = InterpolationInput((P[0].copy(), P[1].copy()), m.copy(), h.copy(), rho.copy(), (A.copy(),),
                             smoothed_kernel=M4_spline)
    lb = LinearBVH_(inp, backend=backend)
    o = oracle_of(P[0], P[1], None, m, h, rho, [A], kernel=0, dims=2)
    oc = o.columns()
    assert np.array_equal(inp.coord[0], oc[0]) and np.array_equal(inp.coord[1], oc[1])
    assert np.array_equal(inp.quant[0], oc[6])
    assert len(lb.leaf_coor) == 2 and len(lb.node_aabb.min) == 2
    assert np.array_equal(lb.order, o.tree()["order"])
    lb.close()


def test_2d_duplicate_codes_near_the_top_of_the_code_range(backend):
    """2D UInt64 codes use all 64 bits.  Duplicates within 2^32 of the top are
    rewritten by the sequential tail pass; here three particles share the code
    2^64 - 4 (rewritten to -4, -3, -2) below one at the corner (2^64 - 1)."""
    e = 1e-10  # floor((2^32 - 1)(1 - e)) = 2^32 - 2 on both axes
    x = np.array([0.0, 0.3, 1.0 - e, 1.0 - e, 1.0 - e, 0.7, 1.0])
    y = np.array([0.0, 0.6, 1.0 - e, 1.0 - e, 1.0 - e, 0.2, 1.0])
    n = len(x)
    _, o = compare2d(backend.context, (x, y, np.ones(n), np.full(n, 0.1), np.ones(n)))
    top = [int(c) - 2 ** 64 for c in o.tree()["codes"][-4:]]
    assert top == [-4, -3, -2, -1]


def test_2d_duplicate_codes_at_the_top_are_rejected(backend):
    """Particles sharing the (xmax, ymax) corner all get the code 2^64 - 1; the
    reference's duplicate rewrite (codes[i-1] + 1) wraps to 0 and the codes are
    no longer sorted.  The refit's check rejects the tree, in the oracle and on
    the GPU alike."""
    from partia.errors import JuliaAssertionError
    x = np.array([0.0, 0.3, 1.0, 1.0, 1.0, 0.7])
    y = np.array([0.0, 0.6, 1.0, 1.0, 1.0, 0.2])
    n = len(x)
    with pytest.raises(ValueError, match="refit"):
        oracle_of(x, y, None, np.ones(n), np.full(n, 0.1), np.ones(n), dims=2)
    with pytest.raises(JuliaAssertionError, match="refit"):
        DeviceSet(backend.context, x, y, None, np.ones(n), np.full(n, 0.1), np.ones(n), dims=2)
    compare2d(backend.context, (x[:3], y[:3], np.ones(3), np.full(3, 0.1), np.ones(3)))  # context still usable
