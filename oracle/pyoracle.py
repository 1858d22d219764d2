"""ctypes wrapper of the parity oracle (liboracle.so, built from partia_oracle.hpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

KERNELS = {"M4_spline": 0, "M5_spline": 1, "M6_spline": 2, "C2_Wendland": 3, "C4_Wendland": 4, "C6_Wendland": 5}
GATHER, SCATTER, SYMMETRIC = 0, 1, 2

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d, i64, i32, vp = C.c_double, C.c_int64, C.c_int, C.c_void_p
        L.oracle_build.restype = vp
        L.oracle_build.argtypes = [i32, i64] + [vp] * 6 + [i32, vp, i32, i32]
        L.oracle_free.argtypes = [vp]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_get_tree.argtypes = [vp] * 8
        L.oracle_get_bvh.argtypes = [vp] * 6
        L.oracle_get_columns.argtypes = [vp, vp]
        L.oracle_point_samples.argtypes = [vp, i32, vp, vp, i32, vp, i32, vp, i32, vp, i32, i64, vp, vp, vp, vp, i32]
        L.oracle_line_samples.argtypes = [vp, i32, vp, vp, i32, i64] + [vp] * 7 + [vp, i32]
        L.oracle_point_single.argtypes = [vp, i32, i32, i64, vp, vp, vp, vp, vp, i32]
        L.oracle_line_density.argtypes = [vp, i32, i64] + [vp] * 6 + [vp, vp, i32]
        L.oracle_nearest_h.argtypes = [vp, i64, vp, vp, vp, vp, vp]
        L.oracle_line_adjoint.argtypes = [vp, i32, i32, i64] + [vp] * 6 + [vp, vp]
        L.oracle_point_adjoint.argtypes = [vp, i32, i32, i32, i64] + [vp] * 3 + [vp, vp]
        L.oracle_field_adjoint.argtypes = [vp, i32, vp, i32, i32, i64] + [vp] * 3 + [vp, vp]
        L.oracle_point_jvp.argtypes = [vp, i32, vp, vp, i32, vp, i32, vp, i32, vp, i32, i64] + [vp] * 6 + [i32]
        L.oracle_line_jvp.argtypes = [vp, i32, vp, vp, i32, i64] + [vp] * 10 + [i32]
        L.oracle_kernel_d2f.argtypes = [i32, d]
        L.oracle_query.argtypes = [vp, i64] + [vp] * 8
        L.oracle_kernel_d2f.restype = d
        L.oracle_neighbors.restype = i64
        L.oracle_neighbors.argtypes = [vp, i32, i32, vp, vp, d, vp, vp, i64]
        for name in ("oracle_kernel_f", "oracle_kernel_df"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [i32, d]
        for name in ("oracle_kernel_sigma3", "oracle_kernel_support"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [i32]
        L.oracle_kernel_W.restype = d
        L.oracle_kernel_W.argtypes = [i32, d, d]
        L.oracle_kernel_gradW.argtypes = [i32, d, d, d, d, vp]
        L.oracle_line_W.restype = d
        L.oracle_line_W.argtypes = [i32, d, d]
        L.oracle_line_W32.restype = C.c_float
        L.oracle_line_W32.argtypes = [i32, C.c_float, C.c_float]
        L.oracle_expand_bits3_u32.restype = C.c_uint32
        L.oracle_expand_bits3_u32.argtypes = [C.c_uint32]
        L.oracle_expand_bits3_u64.restype = C.c_uint64
        L.oracle_expand_bits3_u64.argtypes = [C.c_uint64]
        L.oracle_expand_bits2_u32.restype = C.c_uint32
        L.oracle_expand_bits2_u32.argtypes = [C.c_uint32]
        L.oracle_expand_bits2_u64.restype = C.c_uint64
        L.oracle_expand_bits2_u64.argtypes = [C.c_uint64]
        L.oracle_build2d.restype = vp
        L.oracle_build2d.argtypes = [i32, i64] + [vp] * 5 + [i32, vp, i32, i32]
        L.oracle_find_range_u64.argtypes = [vp, i64, i64, vp, vp]
        L.oracle_split_position_u64.restype = i64
        L.oracle_split_position_u64.argtypes = [vp, i64, i64, i64]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


class OracleSet:
    """One InterpolationInput after LinearBVH! on the CPU oracle."""

    def __init__(self, x, y, z, m, h, rho, cols=(), kernel="M4_spline", code_bits=63, precision=64, dims=3):
        """dims=2: an InterpolationInput{2} (z is ignored; the tree's z is 0)."""
        L = lib()
        self.n = len(x)
        self.ncol = len(cols)
        self.kernel = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        self.precision = precision
        self.dims = dims
        cl = [_f64(c) for c in cols]
        ptrs = (C.c_void_p * max(1, len(cl)))(*[c.ctypes.data for c in cl]) if cl else None
        if dims == 2:
            arrs = [_f64(a) for a in (x, y, m, h, rho)]
            self._keep = arrs + cl
            self.h = L.oracle_build2d(precision, self.n, *[_p(a) for a in arrs], len(cl), ptrs, self.kernel,
                                      code_bits)
        else:
            arrs = [_f64(a) for a in (x, y, z, m, h, rho)]
            self._keep = arrs + cl
            self.h = L.oracle_build(precision, self.n, *[_p(a) for a in arrs], len(cl), ptrs, self.kernel,
                                    code_bits)
        if not self.h:
            raise ValueError(L.oracle_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_free(self.h)
            self.h = None

    def tree(self):
        n = self.n
        t = dict(order=np.zeros(n, np.uint64), codes=np.zeros(n, np.uint64))
        for k in ("left", "right", "escape", "parent"):
            t[k] = np.zeros(2 * n - 1, np.int32)
        root = C.c_int32(0)
        lib().oracle_get_tree(self.h, _p(t["order"]), _p(t["codes"]), _p(t["left"]), _p(t["right"]),
                              _p(t["escape"]), _p(t["parent"]), C.byref(root))
        t["root"] = root.value
        return t

    def bvh(self):
        n, ni = self.n, max(self.n - 1, 0)
        leaf = np.zeros((3, n))
        ls = np.zeros(n)
        nmin, nmax, ns = np.zeros((3, ni)), np.zeros((3, ni)), np.zeros(ni)
        lib().oracle_get_bvh(self.h, _p(leaf), _p(ls), _p(nmin), _p(nmax), _p(ns))
        return dict(leaf_coor=leaf, leaf_scale=ls, node_min=nmin, node_max=nmax, node_scale=ns)

    def columns(self):
        out = np.zeros((6 + self.ncol, self.n))
        lib().oracle_get_columns(self.h, _p(out))
        return out

    def point_samples(self, cat, strategy, px, py, pz, nthreads=0):
        """cat: dict(sslots, sflags, gslots, dslots, cslots) with 1-based slots."""
        px, py, pz = _f64(px), _f64(py), _f64(pz)
        S = len(px)
        ss, gs = _i32(cat.get("sslots", ())), _i32(cat.get("gslots", ()))
        sf = np.ascontiguousarray(np.asarray(cat.get("sflags", ()), dtype=np.uint8))
        ds, cs = _i32(np.ravel(cat.get("dslots", ()))), _i32(np.ravel(cat.get("cslots", ())))
        L = len(ss) + 3 * len(gs) + len(ds) // 3 + len(cs)
        out = np.zeros((L, S))
        lib().oracle_point_samples(self.h, len(ss), _p(ss), _p(sf), len(gs), _p(gs), len(ds) // 3, _p(ds),
                                   len(cs) // 3, _p(cs), strategy, S, _p(px), _p(py), _p(pz), _p(out), nthreads)
        return out

    def line_samples(self, cat, strategy, origin, direction, ha=None, nthreads=0):
        o = [_f64(a) for a in origin]
        m = [_f64(a) for a in direction]
        S = len(o[0])
        ss = _i32(cat.get("sslots", ()))
        sf = np.ascontiguousarray(np.asarray(cat.get("sflags", ()), dtype=np.uint8))
        hv = _f64(ha) if ha is not None else None
        out = np.zeros((len(ss), S))
        lib().oracle_line_samples(self.h, len(ss), _p(ss), _p(sf), strategy, S, *[_p(a) for a in o + m],
                                  _p(hv) if hv is not None else None, _p(out), nthreads)
        return out

    def point_single(self, kind, strategy, px, py, pz, ha=None, nthreads=0):
        """Single-quantity kernels: kind 0 density, 1 number density (1 x S), 2
        density gradient (3 x S); ha None: each point's nearest-particle h."""
        px, py, pz = _f64(px), _f64(py), _f64(pz)
        S = len(px)
        hv = _f64(ha) if ha is not None else None
        out = np.zeros((3 if kind == 2 else 1, S))
        lib().oracle_point_single(self.h, int(kind), int(strategy), S, _p(px), _p(py), _p(pz),
                                  _p(hv) if hv is not None else None, _p(out), nthreads)
        return out

    def line_density(self, strategy, origin, direction, ha=None, nthreads=0):
        """_line_integrated_density_kernel per line (S values)."""
        o = [_f64(a) for a in origin]
        m = [_f64(a) for a in direction]
        S = len(o[0])
        hv = _f64(ha) if ha is not None else None
        out = np.zeros(S)
        lib().oracle_line_density(self.h, int(strategy), S, *[_p(a) for a in o + m],
                                  _p(hv) if hv is not None else None, _p(out), nthreads)
        return out

    def line_adjoint(self, slot, shepard, origin, direction, cot):
        """6 x n gradients (input order) of sum_s cot[s] * out_s for catalog scalar
        `slot` along the lines (scatter)."""
        o = [_f64(a) for a in origin]
        m = [_f64(a) for a in direction]
        c = _f64(cot)
        S = len(o[0])
        g = np.zeros((6, self.n))
        lib().oracle_line_adjoint(self.h, int(slot), int(bool(shepard)), S, *[_p(a) for a in o + m], _p(c), _p(g))
        return g

    def point_adjoint(self, slot, shepard, strategy, px, py, pz, cot):
        """6 x n gradients (input order) of sum_s cot[s] * out_s for catalog scalar
        `slot` at the points (rows m, A, h, x, y, z; rho held fixed)."""
        px, py, pz, c = _f64(px), _f64(py), _f64(pz), _f64(cot)
        S = len(px)
        g = np.zeros((6, self.n))
        lib().oracle_point_adjoint(self.h, int(slot), int(bool(shepard)), int(strategy), S, _p(px), _p(py), _p(pz),
                                   _p(c), _p(g))
        return g

    def field_adjoint(self, mode, slots, shepard, strategy, px, py, pz, cot):
        """(5 + c) x n gradients (input order; rows m, the c field columns, h, x,
        y, z) of sum_s cot_s . out_s for one output group: mode 0 scalar, 1
        gradient, 2 divergence, 3 curl of the 1-based column slots."""
        px, py, pz = _f64(px), _f64(py), _f64(pz)
        c = 3 if mode >= 2 else 1
        S = len(px)
        ct = _f64(np.asarray(cot, dtype=np.float64).reshape(S, -1))
        sl = _i32(list(slots) + [0] * (3 - len(slots)))
        g = np.zeros((5 + c, self.n))
        lib().oracle_field_adjoint(self.h, int(mode), _p(sl), int(bool(shepard)), int(strategy), S, _p(px), _p(py),
                                   _p(pz), _p(ct), _p(g))
        return g

    def point_jvp(self, cat, strategy, px, py, pz, tan, nthreads=0):
        """Forward mode (dual numbers) through the float64 point path: the
        outputs (L x S) and their derivatives along tan ((6 + ncol) x n, input
        order: x, y, z, m, h, rho, the quantity columns)."""
        px, py, pz = _f64(px), _f64(py), _f64(pz)
        S = len(px)
        ss, gs = _i32(cat.get("sslots", ())), _i32(cat.get("gslots", ()))
        sf = np.ascontiguousarray(np.asarray(cat.get("sflags", ()), dtype=np.uint8))
        ds, cs = _i32(np.ravel(cat.get("dslots", ()))), _i32(np.ravel(cat.get("cslots", ())))
        L = len(ss) + 3 * len(gs) + len(ds) // 3 + len(cs)
        t = _f64(np.asarray(tan, dtype=np.float64).reshape(6 + self.ncol, self.n))
        out, dout = np.zeros((L, S)), np.zeros((L, S))
        rc = lib().oracle_point_jvp(self.h, len(ss), _p(ss), _p(sf), len(gs), _p(gs), len(ds) // 3, _p(ds),
                                    len(cs) // 3, _p(cs), int(strategy), S, _p(px), _p(py), _p(pz), _p(t), _p(out),
                                    _p(dout), nthreads)
        if rc != 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        return out, dout

    def line_jvp(self, cat, strategy, origin, direction, tan, ha=None, nthreads=0):
        """Forward mode through the float64 line path: (outputs, derivatives
        along tan), each nS x S."""
        o = [_f64(a) for a in origin]
        m = [_f64(a) for a in direction]
        S = len(o[0])
        ss = _i32(cat.get("sslots", ()))
        sf = np.ascontiguousarray(np.asarray(cat.get("sflags", ()), dtype=np.uint8))
        hv = _f64(ha) if ha is not None else None
        t = _f64(np.asarray(tan, dtype=np.float64).reshape(6 + self.ncol, self.n))
        out, dout = np.zeros((len(ss), S)), np.zeros((len(ss), S))
        rc = lib().oracle_line_jvp(self.h, len(ss), _p(ss), _p(sf), int(strategy), S, *[_p(a) for a in o + m],
                                   _p(hv) if hv is not None else None, _p(t), _p(out), _p(dout), nthreads)
        if rc != 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        return out, dout

    def query(self, px, py, pz, radius, with_pool=False):
        """LBVH_probe_neighbors / LBVH_query! per point: (counts, closest
        1-based index or 0, closest d2 or inf, pool or None)."""
        px, py, pz, r = _f64(px), _f64(py), _f64(pz), _f64(radius)
        S = len(px)
        cnt, best, cd2 = np.zeros(S, np.int64), np.zeros(S, np.int64), np.zeros(S)
        lib().oracle_query(self.h, S, _p(px), _p(py), _p(pz), _p(r), _p(cnt), _p(best), _p(cd2), None)
        pool = None
        if with_pool:
            pool = np.zeros(max(int(cnt.sum()), 1), np.int64)
            lib().oracle_query(self.h, S, _p(px), _p(py), _p(pz), _p(r), _p(cnt), _p(best), _p(cd2), _p(pool))
            pool = pool[: int(cnt.sum())]
        return cnt, best, cd2, pool

    def nearest_h(self, px, py, pz):
        px, py, pz = _f64(px), _f64(py), _f64(pz)
        S = len(px)
        ha, best = np.zeros(S), np.zeros(S, np.int64)
        lib().oracle_nearest_h(self.h, S, _p(px), _p(py), _p(pz), _p(ha), _p(best))
        return ha, best

    def neighbors(self, point, strategy, ha=0.0, direction=None):
        a = _f64(point)
        mode = 0 if direction is None else 1
        dv = _f64(direction) if direction is not None else _f64([0, 0, 0])
        cap = self.n
        idx, d2 = np.zeros(cap, np.int64), np.zeros(cap)
        cnt = lib().oracle_neighbors(self.h, mode, strategy, _p(a), _p(dv), float(ha), _p(idx), _p(d2), cap)
        return idx[:cnt], d2[:cnt]


def kernel_d2f(K, q):
    return lib().oracle_kernel_d2f(KERNELS.get(K, K) if isinstance(K, str) else K, q)


def kernel_f(K, q):
    return lib().oracle_kernel_f(KERNELS.get(K, K) if isinstance(K, str) else K, q)


def kernel_df(K, q):
    return lib().oracle_kernel_df(KERNELS.get(K, K) if isinstance(K, str) else K, q)


def sigma3(K):
    return lib().oracle_kernel_sigma3(KERNELS[K] if isinstance(K, str) else K)


def support(K):
    return lib().oracle_kernel_support(KERNELS[K] if isinstance(K, str) else K)


def kernel_W(K, r, h):
    return lib().oracle_kernel_W(KERNELS[K] if isinstance(K, str) else K, r, h)


def kernel_gradW(K, dx, dy, dz, h):
    g = np.zeros(3)
    lib().oracle_kernel_gradW(KERNELS[K] if isinstance(K, str) else K, dx, dy, dz, h, _p(g))
    return g


def line_W(K, r, h, precision=64):
    k = KERNELS[K] if isinstance(K, str) else K
    if precision == 32:
        return lib().oracle_line_W32(k, r, h)
    return lib().oracle_line_W(k, r, h)
