"""Helpers for the GPU parity tests: build a device tree through the C ABI and
read its arrays, side by side with the oracle."""
import ctypes as C

import numpy as np

from partia import _native as N

import pyoracle as O


class DeviceSet:
    """One ph_lbvh built from host float64 columns (x, y, z, m, h, rho, cols...)."""

    def __init__(self, ctx, x, y, z, m, h, rho, cols=(), kernel=0, code_bits=63, precision=64, dims=3):
        """dims=2: ph_lbvh_build2d (z is ignored)."""
        self.ctx = ctx
        self.n = len(x)
        self.ncol = len(cols)
        arrs = [N.f64(a) for a in ((x, y, z, m, h, rho) if dims == 3 else (x, y, m, h, rho))]
        cl = [N.f64(c) for c in cols]
        cptr = (C.c_void_p * max(1, len(cl)))(*[c.ctypes.data for c in cl])
        self.h = C.c_void_p()
        N.register_tree(self)
        build = N.lib().ph_lbvh_build if dims == 3 else N.lib().ph_lbvh_build2d
        ctx.check(build(ctx.handle, self.n, *[N.ptr(a) for a in arrs], len(cl), cptr, kernel, code_bits, precision,
                        C.byref(self.h)))

    def get(self, which, dtype, size):
        out = np.zeros(size, dtype)
        self.ctx.check(N.lib().ph_lbvh_get(self.h, N.GET[which], N.ptr(out)))
        return out

    def tree(self):
        n = self.n
        t = {k: self.get(k, np.int32, 2 * n - 1) for k in ("left", "right", "escape", "parent")}
        t["order"] = self.get("order", np.uint64, n)
        t["codes"] = self.get("codes", np.uint64, n)
        root = C.c_int32()
        N.lib().ph_lbvh_info(self.h, None, C.byref(root), None, None, None)
        t["root"] = root.value
        return t

    def bvh(self):
        n, ni = self.n, max(self.n - 1, 0)
        return dict(leaf_coor=self.get("leaf_coor", np.float64, 3 * n).reshape(3, n),
                    leaf_scale=self.get("leaf_scale", np.float64, n),
                    node_min=self.get("node_min", np.float64, 3 * ni).reshape(3, ni),
                    node_max=self.get("node_max", np.float64, 3 * ni).reshape(3, ni),
                    node_scale=self.get("node_scale", np.float64, ni))

    def columns(self):
        return self.get("columns", np.float64, (6 + self.ncol) * self.n).reshape(6 + self.ncol, self.n)

    def points(self, cat, strategy, px, py, pz, precision=64):
        ca = N.CatalogArgs(**cat)
        S = len(px)
        out = np.zeros(max(ca.L, 1) * S)
        x, y, z = N.f64(px), N.f64(py), N.f64(pz)
        self.ctx.check(N.lib().ph_point_samples(self.ctx.handle, self.h, ca.ref(), strategy, precision, S, N.ptr(x),
                                                N.ptr(y), N.ptr(z), N.ptr(out)))
        return out[:ca.L * S].reshape(ca.L, S)

    def grid(self, cat, strategy, ax, ay, az, precision=64):
        ca = N.CatalogArgs(**cat)
        S = len(ax) * len(ay) * len(az)
        out = np.zeros(max(ca.L, 1) * S)
        a, b, c = N.f64(ax), N.f64(ay), N.f64(az)
        self.ctx.check(N.lib().ph_structured_grid(self.ctx.handle, self.h, ca.ref(), strategy, precision, N.ptr(a),
                                                  len(a), N.ptr(b), len(b), N.ptr(c), len(c), N.ptr(out)))
        return out[:ca.L * S].reshape(ca.L, S)

    def lines(self, cat, strategy, origin, direction, precision=64):
        ca = N.CatalogArgs(**cat)
        S = len(origin[0])
        out = np.zeros(max(ca.L, 1) * S)
        arrs = [N.f64(a) for a in list(origin) + list(direction)]
        self.ctx.check(N.lib().ph_line_samples(self.ctx.handle, self.h, ca.ref(), strategy, precision, S,
                                               *[N.ptr(a) for a in arrs], N.ptr(out)))
        return out[:ca.L * S].reshape(ca.L, S)

    def nearest_h(self, px, py, pz):
        x, y, z = N.f64(px), N.f64(py), N.f64(pz)
        ha = np.zeros(len(x))
        self.ctx.check(N.lib().ph_nearest_h(self.ctx.handle, self.h, len(x), N.ptr(x), N.ptr(y), N.ptr(z), N.ptr(ha)))
        return ha

    def close(self):
        if self.h:
            N.lib().ph_lbvh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DevArray:
    """Device buffer from ph_device_alloc, optionally filled from a host array."""

    def __init__(self, ctx, arr=None, nbytes=None):
        self.ctx = ctx
        a = None if arr is None else np.ascontiguousarray(arr)
        size = a.nbytes if a is not None else int(nbytes)
        self.p = C.c_void_p()
        ctx.check(N.lib().ph_device_alloc(ctx.handle, max(size, 8), C.byref(self.p)))
        if a is not None and a.nbytes:
            ctx.check(N.lib().ph_memcpy_htod(ctx.handle, self.p, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def get(self, dtype, count):
        out = np.zeros(count, dtype)
        if count:
            self.ctx.check(N.lib().ph_memcpy_dtoh(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self.p,
                                                  out.nbytes))
        return out

    def free(self):
        if self.p:
            N.lib().ph_device_free(self.ctx.handle, self.p)
            self.p = C.c_void_p()


def grid_points(ax, ay, az):
    g = np.meshgrid(ax, ay, az, indexing="ij")
    return [a.ravel(order="F") for a in g]


def assert_same(got, ref, bitwise=True, rtol=1e-12):
    """NaN masks identical; values bit-identical (float64 path) or within rtol."""
    assert got.shape == ref.shape
    assert np.array_equal(np.isnan(got), np.isnan(ref)), "NaN masks differ"
    fin = ~np.isnan(ref)
    if bitwise:
        bad = np.nonzero(got[fin] != ref[fin])[0]
        assert len(bad) == 0, "%d values differ, max rel %.3e" % (
            len(bad), np.max(np.abs(got[fin] - ref[fin]) / np.maximum(np.abs(ref[fin]), 1e-300)))
    else:
        # normwise per output row: max |got - ref| / max |ref|
        for k in range(ref.shape[0]):
            f = fin[k]
            if f.any():
                scale = max(np.max(np.abs(ref[k][f])), 1e-300)
                assert np.max(np.abs(got[k][f] - ref[k][f])) <= rtol * scale, "output %d" % k


def oracle_of(x, y, z, m, h, rho, cols=(), kernel=0, code_bits=63, precision=64, dims=3):
    return O.OracleSet(x, y, z, m, h, rho, cols, kernel=kernel, code_bits=code_bits, precision=precision, dims=dims)


def catalog_fields(cat):
    """Output rows of each field of a catalog, in output order: one row per scalar
    and divergence, three per gradient and curl (a vector field is one field)."""
    ns, ng = len(cat.get("sslots", ())), len(cat.get("gslots", ()))
    nd, nc = len(cat.get("dslots", ())) // 3, len(cat.get("cslots", ())) // 3
    groups, r = [], 0
    for size, count in ((1, ns), (3, ng), (1, nd), (3, nc)):
        for _ in range(count):
            groups.append(list(range(r, r + size)))
            r += size
    return groups


def field_errors(got, ref, floor=0.0, fields=None):
    """Per field (a list of output rows; default one row each): normwise
    ||g - f||_2 / max(||f||_2, floor sqrt(count)) and the pointwise maximum of
    |g - f| / max(|f|, 1e-3 rms f, floor) over the samples where ref is finite
    (NaN masks must be identical).  floor is an absolute scale for fields that
    vanish identically (e.g. the gradient of one particle), where the reference
    holds only rounding noise; a scalar or one value per field."""
    assert got.shape == ref.shape
    assert np.array_equal(np.isnan(got), np.isnan(ref)), "NaN masks differ"
    if fields is None:
        fields = [[k] for k in range(ref.shape[0])]
    floors = np.broadcast_to(np.asarray(floor, dtype=float), (len(fields),))
    out = []
    for fl, rows in zip(floors, fields):
        floor = float(fl)
        f = ~np.isnan(ref[rows[0]])
        if not f.any():
            out.append((0.0, 0.0))
            continue
        g, r = got[rows][:, f], ref[rows][:, f]
        nr = max(np.linalg.norm(r), floor * np.sqrt(r.size), 1e-300)
        norm = np.linalg.norm(g - r) / nr
        rms = np.sqrt(np.mean(r * r))
        point = np.max(np.abs(g - r) / np.maximum(np.maximum(np.abs(r), 1e-3 * rms), max(floor, 1e-300)))
        out.append((float(norm), float(point)))
    return out


class traversal:
    """Context manager selecting the traversal of a context, restoring AUTO."""

    def __init__(self, ctx, mode):
        self.ctx, self.mode = ctx, mode

    def __enter__(self):
        self.ctx.set_traversal(self.mode)
        return self

    def __exit__(self, *a):
        self.ctx.set_traversal(N.PH_TRAVERSAL_AUTO)
