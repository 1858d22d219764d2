"""Independent tree references used to pin the oracle and the GPU build.

Nothing here follows the reference's construction code: Morton codes are
interleaved bit by bit, the radix tree is built top-down by splitting every
range at its highest differing bit (the binary radix tree over distinct keys is
unique, so any correct construction must agree), escapes are the preorder
successors of each subtree, and boxes/scales are computed recursively.
"""
import numpy as np

PRIMES = (37, 61, 97, 131, 197, 263)


def prime_hash_coords(dim, n, offset, modulus=1021):
    """test/neighbor_search_tests.jl:42-54."""
    return [np.array([((PRIMES[d] * i + PRIMES[d + dim] * offset) % modulus) / modulus for i in range(1, n + 1)])
            for d in range(dim)]


def interleave_bits(ix, iy, iz, bits):
    """x at bit 3k+2, y at 3k+1, z at 3k, for k = 0..bits-1 (bit by bit)."""
    code = 0
    for k in range(bits):
        code |= ((int(ix) >> k) & 1) << (3 * k + 2)
        code |= ((int(iy) >> k) & 1) << (3 * k + 1)
        code |= ((int(iz) >> k) & 1) << (3 * k)
    return code


def interleave_bits2(ix, iy, bits):
    """2D: x at bit 2k+1, y at 2k, for k = 0..bits-1 (bit by bit)."""
    code = 0
    for k in range(bits):
        code |= ((int(ix) >> k) & 1) << (2 * k + 1)
        code |= ((int(iy) >> k) & 1) << (2 * k)
    return code


def quantise(v, bits):
    v = np.asarray(v, dtype=np.float64)
    scale = float(2 ** bits - 1)
    lo, hi = v.min(), v.max()
    d = hi - lo
    if d == 0:
        return np.full(len(v), int(np.floor(scale * 0.5)), dtype=np.int64)
    inv = 1.0 / d
    return np.floor(scale * ((v - lo) * inv)).astype(np.int64)


def morton_pipeline(x, y, z, code_bits):
    """(codes after dedup, 1-based order) by an independent route."""
    bits = 10 if code_bits == 30 else 21
    ix, iy, iz = quantise(x, bits), quantise(y, bits), quantise(z, bits)
    codes = np.array([interleave_bits(a, b, c, bits) for a, b, c in zip(ix, iy, iz)], dtype=object)
    order = np.argsort(np.array(codes, dtype=np.uint64), kind="stable")
    c = [int(codes[i]) for i in order]
    # dedup as a prefix maximum: c'[i] = i + max_{j<=i}(c[j] - j)
    out, run = [], None
    for i, cj in enumerate(c):
        run = cj - i if run is None else max(run, cj - i)
        out.append(run + i)
    return np.array(out, dtype=np.uint64), (order + 1).astype(np.uint64)


def morton_pipeline2(x, y, code_bits):
    """The same for a 2D input: 16 (UInt32) or 32 (UInt64) bits per axis."""
    bits = 16 if code_bits == 30 else 32
    ix, iy = quantise(x, bits), quantise(y, bits)
    codes = np.array([interleave_bits2(a, b, bits) for a, b in zip(ix, iy)], dtype=object)
    order = np.argsort(np.array(codes, dtype=np.uint64), kind="stable")
    c = [int(codes[i]) for i in order]
    out, run = [], None
    for i, cj in enumerate(c):
        run = cj - i if run is None else max(run, cj - i)
        out.append(run + i)
    return np.array(out, dtype=np.uint64), (order + 1).astype(np.uint64)


def topdown_tree(codes):
    """left/right/parent (unified IDs) of the radix tree over sorted distinct codes."""
    n = len(codes)
    tot = 2 * n - 1
    left = np.zeros(tot, np.int32)
    right = np.zeros(tot, np.int32)
    parent = np.zeros(tot, np.int32)
    if n < 2:
        return left, right, parent, 0
    c = [int(v) for v in codes]
    off = n - 1
    stack = [(1, n, 1)]  # (lo, hi, node id) with 1-based leaf indices
    while stack:
        lo, hi, node = stack.pop()
        b = (c[lo - 1] ^ c[hi - 1]).bit_length() - 1  # highest differing bit
        s = lo
        while s + 1 <= hi and not ((c[s] >> b) & 1):  # last index whose bit b is 0
            s += 1
        lid = lo + off if s == lo else s
        rid = hi + off if s + 1 == hi else s + 1
        left[node - 1], right[node - 1] = lid, rid
        parent[lid - 1] = parent[rid - 1] = node
        if s != lo:
            stack.append((lo, s, s))
        if s + 1 != hi:
            stack.append((s + 1, hi, s + 1))
    return left, right, parent, 1


def preorder(left, right, n):
    """Recursive preorder of unified IDs starting at the root (1), or [leaf] for n = 1."""
    if n == 1:
        return [1]
    out, stack = [], [1]
    while stack:
        v = stack.pop()
        out.append(v)
        if v < n:
            stack.append(int(right[v - 1]))
            stack.append(int(left[v - 1]))
    return out


def expected_escape(left, right, n):
    """escape[v] = the node visited after the whole subtree of v in preorder (0 at the end)."""
    order = preorder(left, right, n)
    size = {}

    def subtree(v):
        if v >= n:
            return 1
        if v not in size:
            size[v] = 1 + subtree(int(left[v - 1])) + subtree(int(right[v - 1]))
        return size[v]

    import sys
    sys.setrecursionlimit(100000)
    pos = {v: i for i, v in enumerate(order)}
    esc = np.zeros(2 * n - 1, np.int32)
    for v in order:
        j = pos[v] + subtree(v)
        esc[v - 1] = order[j] if j < len(order) else 0
    return esc


def recursive_boxes(left, right, leaf_xyz, leaf_h, n):
    """node min/max (3 x n-1) and node scale (n-1) by recursion from the root."""
    ni = n - 1
    mn, mx, sc = np.zeros((3, ni)), np.zeros((3, ni)), np.zeros(ni)
    if n < 2:
        return mn, mx, sc
    order = preorder(left, right, n)
    for v in reversed(order):  # children before parents
        if v >= n:
            continue
        vals = []
        for ch in (int(left[v - 1]), int(right[v - 1])):
            if ch >= n:
                k = ch - n
                vals.append((leaf_xyz[:, k], leaf_xyz[:, k], leaf_h[k]))
            else:
                vals.append((mn[:, ch - 1], mx[:, ch - 1], sc[ch - 1]))
        mn[:, v - 1] = np.minimum(vals[0][0], vals[1][0])
        mx[:, v - 1] = np.maximum(vals[0][1], vals[1][1])
        sc[v - 1] = max(vals[0][2], vals[1][2])
    return mn, mx, sc


def brute_point_neighbours(xyz, h, p, strategy, Kv, ha):
    """1-based leaf indices accepted by the reference's leaf test (d2 <= r2)."""
    dx, dy, dz = p[0] - xyz[0], p[1] - xyz[1], p[2] - xyz[2]
    d2 = dx * dx + dy * dy + dz * dz
    ra = Kv * ha
    r2a = ra * ra
    rb = Kv * h
    r2b = rb * rb
    r2 = {0: np.full_like(h, r2a), 1: r2b, 2: np.maximum(r2a, r2b)}[strategy]
    return np.nonzero(d2 <= r2)[0] + 1


def brute_line_neighbours(xyz, h, o, m, strategy, Kv, ha):
    D = [xyz[d] - o[d] for d in range(3)]
    D2 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2]
    Dm = D[0] * m[0] + D[1] * m[1] + D[2] * m[2]
    d2 = D2 - Dm * Dm
    ra = Kv * ha
    r2a = ra * ra
    rb = Kv * h
    r2b = rb * rb
    r2 = {0: np.full_like(h, r2a), 1: r2b, 2: np.maximum(r2a, r2b)}[strategy]
    return np.nonzero(d2 <= r2)[0] + 1
