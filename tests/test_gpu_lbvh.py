"""GPU LBVH build (ph_lbvh_build) against the oracle: Morton codes, sorted order,
tree topology and escapes are compared bit for bit (memcmp), boxes and scales
exactly, at small (every prime-hash case of test/neighbor_search_tests.jl) and
full (2e6-particle Plummer) sizes, for 30- and 63-bit codes and float32 trees."""
import ctypes as C

import numpy as np
import pytest

import lbvh_ref as R
from gpu_util import DeviceSet, oracle_of
from partia import _native as N
from partia import synthetic

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(777)


def compare(ctx, cols6, code_bits=63, precision=64, check_bvh=True):
    x, y, z, m, h, rho = cols6
    d = DeviceSet(ctx, x, y, z, m, h, rho, code_bits=code_bits, precision=precision)
    o = oracle_of(x, y, z, m, h, rho, code_bits=code_bits, precision=precision)
    td, to = d.tree(), o.tree()
    for k in ("order", "codes", "left", "right", "escape", "parent"):
        assert np.array_equal(td[k], to[k]), k
    assert td["root"] == to["root"]
    if check_bvh:
        bd, bo = d.bvh(), o.bvh()
        for k in bo:
            assert np.array_equal(bd[k], bo[k]), k
    assert np.array_equal(d.columns(), o.columns())
    return d, o


@pytest.mark.parametrize("code_bits", [30, 63])
def test_prime_hash_all_cases(backend, code_bits):
    ctx = backend.context
    for offset in range(5):
        for n in range(1, 21):
            P = R.prime_hash_coords(3, n, offset)
            compare(ctx, (*P, np.ones(n), np.full(n, 0.1), np.ones(n)), code_bits)


@pytest.mark.parametrize("n", [1, 2, 8, 16, 1000])
def test_identical_codes(backend, n):
    compare(backend.context, (np.full(n, 0.5),) * 3 + (np.ones(n), np.full(n, 0.1), np.ones(n)))


@pytest.mark.parametrize("code_bits", [30, 63])
def test_random_and_clustered(backend, code_bits):
    ctx = backend.context
    for n in (3, 100, 4097, 100_000):
        P = RNG.random((3, n))
        P[:, : n // 3] = 0.25 + 1e-6 * RNG.random((3, n // 3))  # a tight clump forces duplicate codes
        compare(ctx, (*P, np.ones(n), RNG.random(n) * 0.1, np.ones(n)), code_bits)


@pytest.mark.parametrize("code_bits", [30, 63])
def test_full_size_plummer_tree(backend, code_bits):
    d = synthetic.plummer_sphere(2_000_000, seed=2)
    compare(backend.context, (d["x"], d["y"], d["z"], d["m"], d["h"], d["rho"]), code_bits)


def test_float32_tree(backend):
    n = 50_000
    P = RNG.random((3, n)).astype(np.float32).astype(np.float64)
    h = (RNG.random(n) * 0.05).astype(np.float32).astype(np.float64)
    for bits in (30, 63):
        compare(backend.context, (*P, np.ones(n), h, np.ones(n)), bits, precision=32)


def test_topology_against_independent_construction(backend):
    n = 3000
    P = RNG.random((3, n))
    d = DeviceSet(backend.context, *P, np.ones(n), np.full(n, 0.1), np.ones(n))
    t = d.tree()
    left, right, parent, root = R.topdown_tree(t["codes"])
    assert np.array_equal(t["left"], left) and np.array_equal(t["right"], right)
    assert np.array_equal(t["escape"], R.expected_escape(left, right, n))


def test_build_errors_leave_context_usable(backend):
    ctx = backend.context
    L = N.lib()
    h = C.c_void_p()
    z = np.zeros(1)
    assert L.ph_lbvh_build(ctx.handle, 0, *[N.ptr(z)] * 6, 0, None, 0, 63, 64, C.byref(h)) == -5  # capacity
    assert L.ph_lbvh_build(ctx.handle, 1, *[N.ptr(z)] * 6, 0, None, 9, 63, 64, C.byref(h)) == -1  # kernel
    assert L.ph_lbvh_build(ctx.handle, 1, *[N.ptr(z)] * 6, 0, None, 0, 40, 64, C.byref(h)) == -1  # code bits
    assert L.ph_lbvh_build(ctx.handle, 1 << 31, *[N.ptr(z)] * 6, 0, None, 0, 63, 64, C.byref(h)) == -5
    assert "capacity" in L.ph_last_error(ctx.handle).decode()
    compare(ctx, (np.array([0.1, 0.9]),) * 3 + (np.ones(2), np.ones(2), np.ones(2)))
