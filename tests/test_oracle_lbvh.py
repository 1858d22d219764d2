"""Oracle LBVH (Morton encoding, sort, duplicate rewrite, radix tree, escapes,
refit, traversals, nearest h) against independent references (tests/lbvh_ref.py)
and the reference's own invariant tests (test/neighbor_search_tests.jl)."""
import numpy as np
import pytest

import lbvh_ref as R
import pyoracle as O

RNG = np.random.default_rng(20260512)


def test_expand_bits_is_bit_interleave():
    for _ in range(200):
        v = int(RNG.integers(0, 2 ** 21))
        e64 = O.lib().oracle_expand_bits3_u64(v)
        assert e64 == R.interleave_bits(v, 0, 0, 21) >> 2
        v10 = v & 1023
        assert O.lib().oracle_expand_bits3_u32(v10) == R.interleave_bits(v10, 0, 0, 10) >> 2


@pytest.mark.parametrize("code_bits", [30, 63])
@pytest.mark.parametrize("n", [1, 2, 7, 300, 5000])
def test_codes_and_order_match_independent_pipeline(code_bits, n):
    # clustered data so 30-bit codes collide and the duplicate rewrite is exercised
    P = np.concatenate([RNG.random((3, n // 2)), 0.5 + 1e-4 * RNG.random((3, n - n // 2))], axis=1)
    o = O.OracleSet(*P, np.ones(n), np.full(n, 0.01), np.ones(n), code_bits=code_bits)
    t = o.tree()
    codes, order = R.morton_pipeline(*P, code_bits)
    assert np.array_equal(t["order"], order)
    assert np.array_equal(t["codes"], codes)
    assert np.all(np.diff(t["codes"].astype(object)) > 0)  # strictly increasing after the rewrite


def test_dedup_prefix_max_equals_serial_loop():
    for _ in range(50):
        c = np.sort(RNG.integers(0, 40, size=int(RNG.integers(1, 60))))
        serial = [int(c[0])]
        for v in c[1:]:
            serial.append(int(v) if v > serial[-1] else serial[-1] + 1)
        i = np.arange(len(c))
        formula = i + np.maximum.accumulate(c - i)
        assert list(formula) == serial


def _tree_invariants(P, code_bits=63):
    n = len(P[0])
    o = O.OracleSet(*P, np.ones(n), np.full(n, 0.1), np.ones(n), code_bits=code_bits)
    t = o.tree()
    left, right, parent, root = R.topdown_tree(t["codes"])
    assert t["root"] == root
    assert np.array_equal(t["left"], left)
    assert np.array_equal(t["right"], right)
    assert np.array_equal(t["parent"], parent)
    assert np.array_equal(t["escape"], R.expected_escape(left, right, n))
    assert np.count_nonzero(t["parent"] == 0) == 1
    b = o.bvh()
    mn, mx, sc = R.recursive_boxes(left, right, b["leaf_coor"], b["leaf_scale"], n)
    assert np.array_equal(b["node_min"], mn) and np.array_equal(b["node_max"], mx)
    assert np.array_equal(b["node_scale"], sc)
    return o, t


@pytest.mark.parametrize("offset", [0, 1, 2, 3, 4])
def test_prime_hash_trees(offset, fixtures):
    for n in range(1, fixtures["prime_hash"]["n_max"] + 1):
        for bits in (30, 63):
            _tree_invariants(R.prime_hash_coords(3, n, offset), bits)


@pytest.mark.parametrize("n", [1, 2, 8, 16])
def test_identical_codes(n):
    o, t = _tree_invariants([np.full(n, 0.5)] * 3)
    assert t["root"] == (1 if n >= 2 else 0)


def test_random_trees_scale_and_boxes():
    for n in (64, 1000, 20000):
        o, t = _tree_invariants(list(RNG.random((3, n))))


def test_lattice_boxes_contain_children(fixtures):
    for n in fixtures["aabb_lattice"]["n"]:
        c = np.linspace(0.0, 1.0, n)
        _tree_invariants([c, c, c])


def test_escape_walk_visits_every_node_once():
    n = 777
    P = RNG.random((3, n))
    o = O.OracleSet(*P, np.ones(n), np.full(n, 0.1), np.ones(n))
    t = o.tree()
    seen, node = [], t["root"]
    while node != 0:  # unpruned stackless walk: left on internal nodes, escape on leaves
        seen.append(node)
        node = t["left"][node - 1] if node < n else t["escape"][node - 1]
    assert sorted(seen) == list(range(1, 2 * n))


def test_neighbour_query_fixture(fixtures):
    fx = fixtures["neighbour_query"]
    x, y, z = (np.array(fx[k]) for k in "xyz")
    o = O.OracleSet(x, y, z, np.ones(4), np.full(4, fx["h"]), np.ones(4))
    cols = o.columns()
    for p, r in fx["cases_3d"]:
        got, _ = o.neighbors(p, O.GATHER, ha=r / 2.0)  # Kvalid(M4) = 2 -> radius r
        exp = R.brute_point_neighbours(cols[:3], cols[4], p, 0, 2.0, r / 2.0)
        assert sorted(got) == sorted(exp) and len(exp) >= 1


@pytest.mark.parametrize("strategy", [O.GATHER, O.SCATTER, O.SYMMETRIC])
def test_point_traversal_equals_brute_force(strategy):
    n = 3000
    P = RNG.random((3, n))
    h = RNG.random(n) * 0.05 + 0.01
    o = O.OracleSet(*P, np.ones(n), h, np.ones(n))
    cols = o.columns()
    for _ in range(40):
        p = RNG.random(3) * 1.2 - 0.1
        ha = float(RNG.random() * 0.05 + 0.005)
        got, d2 = o.neighbors(p, strategy, ha=ha)
        exp = R.brute_point_neighbours(cols[:3], cols[4], p, strategy, 2.0, ha)
        assert np.array_equal(got, exp)  # DFS order is increasing Morton index
        assert np.all(np.diff(got) > 0)


@pytest.mark.parametrize("strategy", [O.GATHER, O.SCATTER, O.SYMMETRIC])
def test_line_traversal_equals_brute_force(strategy, fixtures):
    fx = fixtures["line_traversal"]
    x, y, h = np.array(fx["x"]), np.array(fx["y"]), np.array(fx["h"])
    o = O.OracleSet(x, y, np.zeros(5), np.ones(5), h, np.ones(5))
    cols = o.columns()
    # Kvalid(M4) = 2: ha such that (2 ha)^2 = radius2, and hb rescaled so 2 hb = Kvalid * h
    ha = np.sqrt(fx["radius2"]) / 2.0
    got, _ = o.neighbors(fx["origin"], strategy, ha=ha, direction=fx["direction"])
    exp = R.brute_line_neighbours(cols[:3], cols[4], fx["origin"], fx["direction"], strategy, 2.0, ha)
    assert np.array_equal(got, exp)
    n = 4000
    P = RNG.random((3, n))
    hh = RNG.random(n) * 0.03 + 0.005
    o = O.OracleSet(*P, np.ones(n), hh, np.ones(n))
    cols = o.columns()
    for _ in range(30):
        org = RNG.random(3)
        d = RNG.normal(size=3)
        d /= np.linalg.norm(d)
        got, _ = o.neighbors(org, strategy, ha=0.01, direction=d)
        exp = R.brute_line_neighbours(cols[:3], cols[4], org, d, strategy, 2.0, 0.01)
        assert np.array_equal(got, exp)


def test_nearest_h_brute_force_and_ties():
    n = 5000
    P = RNG.random((3, n))
    h = RNG.random(n)
    o = O.OracleSet(*P, np.ones(n), h, np.ones(n))
    cols = o.columns()
    Q = RNG.random((3, 500)) * 1.4 - 0.2
    ha, best = o.nearest_h(*Q)
    for s in range(Q.shape[1]):
        d2 = (Q[0, s] - cols[0]) ** 2 + (Q[1, s] - cols[1]) ** 2 + (Q[2, s] - cols[2]) ** 2
        k = int(np.argmin(d2))  # first minimum in Morton order = the reference's tie rule
        assert best[s] == k + 1 and ha[s] == cols[4][k]
    # exact ties on a lattice: the first leaf in DFS (Morton) order wins
    c = np.linspace(0.0, 1.0, 5)
    g = np.array(np.meshgrid(c, c, c, indexing="ij")).reshape(3, -1)
    hl = np.arange(1, g.shape[1] + 1, dtype=float)
    o = O.OracleSet(*g, np.ones(len(hl)), hl, np.ones(len(hl)))
    cols = o.columns()
    q = np.array([[0.125], [0.125], [0.125]])  # equidistant from 8 lattice points
    ha, best = o.nearest_h(*q)
    d2 = (q[0] - cols[0]) ** 2 + (q[1] - cols[1]) ** 2 + (q[2] - cols[2]) ** 2
    ties = np.nonzero(d2 == d2.min())[0]
    assert len(ties) == 8 and best[0] == ties[0] + 1


def test_empty_tree_rejected():
    with pytest.raises(ValueError):
        O.OracleSet(np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0))


# ---- the 2D path: MortonEncoding(x, y) and LinearBVH! of an InterpolationInput{2} ----

def test_expand_bits2_is_bit_interleave():
    for _ in range(200):
        v = int(RNG.integers(0, 2 ** 32))
        assert O.lib().oracle_expand_bits2_u64(v) == R.interleave_bits2(v, 0, 32) >> 1
        v16 = v & 0xFFFF
        assert O.lib().oracle_expand_bits2_u32(v16) == R.interleave_bits2(v16, 0, 16) >> 1


@pytest.mark.parametrize("code_bits", [30, 63])
@pytest.mark.parametrize("n", [1, 2, 7, 300, 5000])
def test_2d_codes_and_order_match_independent_pipeline(code_bits, n):
    P = np.concatenate([RNG.random((2, n // 2)), 0.5 + 1e-6 * RNG.random((2, n - n // 2))], axis=1)
    o = O.OracleSet(P[0], P[1], None, np.ones(n), np.full(n, 0.01), np.ones(n), code_bits=code_bits, dims=2)
    t = o.tree()
    codes, order = R.morton_pipeline2(P[0], P[1], code_bits)
    assert np.array_equal(t["order"], order)
    assert np.array_equal(t["codes"], codes)
    assert not o.columns()[2].any()  # z stays zero


def _tree_invariants_2d(P, code_bits=63):
    n = len(P[0])
    o = O.OracleSet(P[0], P[1], None, np.ones(n), np.full(n, 0.1), np.ones(n), code_bits=code_bits, dims=2)
    t = o.tree()
    left, right, parent, root = R.topdown_tree(t["codes"])
    assert t["root"] == root
    assert np.array_equal(t["left"], left) and np.array_equal(t["right"], right)
    assert np.array_equal(t["parent"], parent)
    assert np.array_equal(t["escape"], R.expected_escape(left, right, n))
    b = o.bvh()
    mn, mx, sc = R.recursive_boxes(left, right, b["leaf_coor"], b["leaf_scale"], n)
    assert np.array_equal(b["node_min"], mn) and np.array_equal(b["node_max"], mx)
    assert np.array_equal(b["node_scale"], sc)
    return o, t


@pytest.mark.parametrize("offset", [0, 1, 2, 3, 4])
def test_2d_prime_hash_trees(offset, fixtures):
    """neighbor_search_tests.jl:177-232 with D = 2."""
    for n in range(1, fixtures["prime_hash"]["n_max"] + 1):
        for bits in (30, 63):
            _tree_invariants_2d(R.prime_hash_coords(2, n, offset), bits)


@pytest.mark.parametrize("n", [1, 2, 8, 16])
def test_2d_identical_codes(n):
    o, t = _tree_invariants_2d([np.full(n, 0.5)] * 2)
    assert t["root"] == (1 if n >= 2 else 0)
    assert np.count_nonzero(o.tree()["parent"] == 0) == 1


def test_2d_lattice_boxes_contain_children(fixtures):
    for n in fixtures["aabb_lattice"]["n"]:
        c = np.linspace(0.0, 1.0, n)
        _tree_invariants_2d([c, c])


@pytest.mark.parametrize("strategy", [O.GATHER, O.SCATTER, O.SYMMETRIC])
def test_2d_point_traversal_and_nearest_h_equal_brute_force(strategy):
    n = 3000
    P = RNG.random((2, n))
    h = RNG.random(n) * 0.05 + 0.01
    o = O.OracleSet(P[0], P[1], None, np.ones(n), h, np.ones(n), dims=2)
    cols = o.columns()
    for _ in range(40):
        p = np.array([*(RNG.random(2) * 1.2 - 0.1), 0.0])
        ha = float(RNG.random() * 0.05 + 0.005)
        got, _ = o.neighbors(p, strategy, ha=ha)
        exp = R.brute_point_neighbours(cols[:3], cols[4], p, strategy, 2.0, ha)
        assert np.array_equal(got, exp)
    Q = RNG.random((2, 300)) * 1.4 - 0.2
    ha, best = o.nearest_h(Q[0], Q[1], np.zeros(300))
    for s in range(Q.shape[1]):
        d2 = (Q[0, s] - cols[0]) ** 2 + (Q[1, s] - cols[1]) ** 2
        k = int(np.argmin(d2))
        assert best[s] == k + 1 and ha[s] == cols[4][k]


# ---- LBVH_probe_neighbors / LBVH_query! / LBVH_find_nearest ----

@pytest.mark.parametrize("dims", [3, 2])
def test_query_equals_brute_force(dims):
    """Counts, DFS-ordered pools (increasing Morton index) and the closest
    accepted leaf (first index on ties) against brute force over all leaves;
    radius 0 and radii past every leaf included."""
    n = 4000
    P = RNG.random((3, n))
    if dims == 2:
        P[2] = 0.0
    o = O.OracleSet(*P, np.ones(n), np.full(n, 0.01), np.ones(n), dims=dims)
    c = o.columns()
    S = 300
    Q = RNG.random((3, S)) * 1.2 - 0.1
    if dims == 2:
        Q[2] = 0.0
    Q[:, :5] = c[:3, :5]  # on particles: d2 = 0
    r = RNG.random(S) * 0.08
    r[5], r[6] = 0.0, 10.0
    cnt, best, cd2, pool = o.query(*Q, r, with_pool=True)
    off = 0
    for s in range(S):
        d2 = ((Q[0, s] - c[0]) * (Q[0, s] - c[0]) + (Q[1, s] - c[1]) * (Q[1, s] - c[1])) + \
            (Q[2, s] - c[2]) * (Q[2, s] - c[2])
        acc = np.flatnonzero(d2 <= r[s] * r[s])
        assert cnt[s] == len(acc)
        assert np.array_equal(pool[off:off + cnt[s]], acc + 1)
        off += cnt[s]
        if len(acc):
            k = acc[np.argmin(d2[acc])]
            assert best[s] == k + 1 and cd2[s] == d2[k]
        else:
            assert best[s] == 0 and cd2[s] == np.inf
    assert cnt[6] == n


def test_query_on_the_reference_fixture(fixtures):
    """LBVH_query! on test/neighbor_search_tests.jl's point set and radii."""
    fx = fixtures["neighbour_query"]
    x, y, z = (np.array(fx[k]) for k in "xyz")
    o = O.OracleSet(x, y, z, np.ones(4), np.full(4, fx["h"]), np.ones(4))
    cols = o.columns()
    for p, r in fx["cases_3d"]:
        cnt, best, cd2, pool = o.query([p[0]], [p[1]], [p[2]], [r], with_pool=True)
        exp = R.brute_point_neighbours(cols[:3], cols[4], p, 0, 2.0, r / 2.0)
        assert sorted(pool) == sorted(exp) and cnt[0] == len(exp) >= 1
