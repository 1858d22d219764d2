"""InterpolationInput, build_input and LinearBVH_ (the in-place LinearBVH!).

Mirrors KI/interpolation_setup/InterpolationInput.jl (container and length
checks :83-107, LinearBVH! :201-221, matches_lbvh_leaf_order :265-267) and
KI/interpolation_setup/constructor.jl:68-136 (build_input: promote every column
to one float type, copy, build the catalog).  The tree itself is built on the
GPU (ph_lbvh_build); LinearBVH_ permutes the host arrays in place with the
permutation the device computed, exactly as LinearBVH! does.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .catalog import InterpolationCatalog
from .errors import DimensionMismatch, MethodError
from .kernels import M5_spline, AbstractSPHKernel, HIPComputeBackend, CPUComputeBackend, default_backend


class InterpolationInput:
    """InterpolationInput{D, T, V, K, NCOLUMN}: SoA particle columns, D = 3 or 2.

    A 2D input supports LinearBVH! and the tree queries; the interpolation
    kernels are 3D only (README.md:152) and reject its tree."""

    def __init__(self, coord, m, h, rho, quant=(), smoothed_kernel=M5_spline):
        coord = tuple(coord)
        if len(coord) not in (2, 3):
            raise MethodError("InterpolationInput takes 2D or 3D coordinates")
        if not (isinstance(smoothed_kernel, type) and issubclass(smoothed_kernel, AbstractSPHKernel)):
            raise MethodError("smoothed_kernel must be an SPH kernel type")
        npart = len(m)
        for d, c in enumerate(coord):
            if len(c) != npart:
                raise DimensionMismatch("coord[%d] length %d != Nparticles %d" % (d + 1, len(c), npart))
        if len(h) != npart:
            raise DimensionMismatch("h length %d != Nparticles %d" % (len(h), npart))
        if len(rho) != npart:
            raise DimensionMismatch("ρ length %d != Nparticles %d" % (len(rho), npart))
        for j, q in enumerate(quant):
            if len(q) != npart:
                raise DimensionMismatch("quant[%d] length %d != Nparticles %d" % (j + 1, len(q), npart))
        self.Npart = npart
        self.D = len(coord)
        self.smoothed_kernel = smoothed_kernel
        self.coord = coord
        self.m = m
        self.h = h
        self.ρ = rho
        self.quant = tuple(quant)

    rho = property(lambda self: self.ρ)

    @property
    def dtype(self):
        return self.m.dtype

    def __len__(self):
        return self.Npart


def get_coord(inp):
    return inp.coord


def get_xcoord(inp):
    return inp.coord[0]


def get_ycoord(inp):
    return inp.coord[1]


def get_zcoord(inp):
    return inp.coord[2]


def _promote(arrays):
    dt = np.result_type(*[np.asarray(a).dtype for a in arrays])
    if not np.issubdtype(dt, np.floating):
        dt = np.dtype(np.float64)
    if dt not in (np.float32, np.float64):
        dt = np.dtype(np.float64)
    return dt


def build_input(backend, x, y, z, m, h, rho, quantity_columns=(), column_names=(), scalars=(), gradients=(),
                divergences=(), curls=(), smoothed_kernel=M5_spline):
    """build_input(backend, x, y, z, m, h, ρ, quantity_columns; column_names, ...) (constructor.jl:68-136)."""
    if not isinstance(backend, (HIPComputeBackend, CPUComputeBackend)):
        raise MethodError("no build_input method for backend %r" % (backend,))
    quantity_columns = tuple(quantity_columns)
    if len(column_names) != len(quantity_columns):
        raise MethodError("column_names must name every quantity column")
    T = _promote([x, y, z, m, h, rho] + list(quantity_columns))
    cp = lambda a: np.array(a, dtype=T, copy=True)  # noqa: E731
    inp = InterpolationInput((cp(x), cp(y), cp(z)), cp(m), cp(h), cp(rho), tuple(cp(q) for q in quantity_columns),
                             smoothed_kernel=smoothed_kernel)
    cat = InterpolationCatalog.from_columns(column_names, 3, scalars=scalars, gradients=gradients,
                                            divergences=divergences, curls=curls)
    return inp, cat


class BinaryRadixTree:
    """Host view of the device tree: root, nleaf and the unified-ID arrays."""

    def __init__(self, lbvh):
        self._lbvh = lbvh
        self.root = lbvh._root
        self.nleaf = lbvh.n

    def _arr(self, which):
        return self._lbvh._get(which, np.int32, 2 * self.nleaf - 1)

    left = property(lambda self: self._arr("left"))
    right = property(lambda self: self._arr("right"))
    escape = property(lambda self: self._arr("escape"))
    parent = property(lambda self: self._arr("parent"))


class AABB:
    def __init__(self, mn, mx):
        self.min = mn
        self.max = mx


class LinearBVH:
    """Device-resident LinearBVH{3, T} (LBVH/LinearBVH/LinearBVH.jl:8-14).

    Owns the ph_lbvh handle (tree + particle columns in Morton order on the GPU).
    The arrays of the reference struct are available as properties, copied from
    the device on first access.
    """

    def __init__(self, backend, handle, n, ncol, kernel, code_bits, dtype, dims=3):
        self.backend = backend
        self.D = dims
        self.handle = handle
        N.register_tree(self)
        self.n = n
        self.ncol = ncol
        self.kernel = kernel
        self.code_bits = code_bits
        self.dtype = np.dtype(dtype)
        root = C.c_int32()
        N.lib().ph_lbvh_info(handle, None, C.byref(root), None, None, None)
        self._root = root.value
        self._cache = {}
        self.brt = BinaryRadixTree(self)

    @property
    def ctx(self):
        return self.backend.context

    def _get(self, which, dtype, size):
        if which not in self._cache:
            out = np.zeros(size, dtype=dtype)
            self.ctx.check(N.lib().ph_lbvh_get(self.handle, N.GET[which], N.ptr(out)))
            self._cache[which] = out
        return self._cache[which]

    @property
    def leaf_coor(self):
        a = self._get("leaf_coor", np.float64, 3 * self.n).reshape(3, self.n).astype(self.dtype)
        return tuple(a[d] for d in range(self.D))

    @property
    def leaf_scale(self):
        return self._get("leaf_scale", np.float64, self.n).astype(self.dtype)

    @property
    def node_aabb(self):
        ni = max(self.n - 1, 0)
        mn = self._get("node_min", np.float64, 3 * ni).reshape(3, ni).astype(self.dtype)
        mx = self._get("node_max", np.float64, 3 * ni).reshape(3, ni).astype(self.dtype)
        return AABB(tuple(mn[d] for d in range(self.D)), tuple(mx[d] for d in range(self.D)))

    @property
    def node_scale(self):
        return self._get("node_scale", np.float64, max(self.n - 1, 0)).astype(self.dtype)

    @property
    def order(self):
        """MortonEncoding.order: original 1-based index of each leaf."""
        return self._get("order", np.uint64, self.n)

    @property
    def codes(self):
        return self._get("codes", np.uint64, self.n)

    def refresh_columns(self, inp):
        """Upload the input's current m, ρ and quantity columns (Morton order)."""
        cols = [N.f64(q) for q in inp.quant]
        cptr = (C.c_void_p * max(1, len(cols)))(*[c.ctypes.data for c in cols])
        m, rho = N.f64(inp.m), N.f64(inp.ρ)
        self.ctx.check(N.lib().ph_lbvh_set_columns(self.handle, N.ptr(m), N.ptr(rho), len(cols), cptr))
        self.ncol = len(cols)

    def close(self):
        if getattr(self, "handle", None):
            N.lib().ph_lbvh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


UInt64 = np.uint64
UInt32 = np.uint32


def LinearBVH_(inp, CodeType=np.uint64, backend=None):
    """LinearBVH!(input; CodeType) (InterpolationInput.jl:201-221; the 2D
    method :224-242) on the GPU.

    Builds the tree from the input's columns, then permutes the coordinates, m,
    h, ρ and every quantity column of `inp` in place into Morton order.
    """
    backend = backend or default_backend()
    if not isinstance(backend, HIPComputeBackend):
        raise MethodError("LinearBVH_ runs on a HIPComputeBackend")
    if CodeType in (np.uint64, "UInt64", 63):
        bits = 63
    elif CodeType in (np.uint32, "UInt32", 30):
        bits = 30
    else:
        raise MethodError("CodeType must be UInt32 or UInt64")
    ctx = backend.context
    dt = inp.m.dtype
    prec = N.PH_PREC_FP32 if dt == np.float32 else N.PH_PREC_FP64
    dims = len(inp.coord)
    base = [N.f64(a) for a in tuple(inp.coord) + (inp.m, inp.h, inp.ρ)]
    cols = [N.f64(q) for q in inp.quant]
    cptr = (C.c_void_p * max(1, len(cols)))(*[c.ctypes.data for c in cols])
    h = C.c_void_p()
    build = N.lib().ph_lbvh_build if dims == 3 else N.lib().ph_lbvh_build2d
    ctx.check(build(ctx.handle, inp.Npart, *[N.ptr(a) for a in base], len(cols), cptr, inp.smoothed_kernel.id, bits,
                    prec, C.byref(h)))
    lbvh = LinearBVH(backend, h, inp.Npart, len(cols), inp.smoothed_kernel, bits, dt, dims)
    p = lbvh.order.astype(np.int64) - 1
    for arr in list(inp.coord) + [inp.m, inp.h, inp.ρ] + list(inp.quant):
        arr[:] = arr[p]
    return lbvh


def matches_lbvh_leaf_order(inp, lbvh):
    """input.coord[d] == lbvh.leaf_coor[d] for every d and input.h == lbvh.leaf_scale (:265-267)."""
    lc = lbvh.leaf_coor
    return len(inp.coord) == lbvh.D and \
        all(np.array_equal(np.asarray(inp.coord[d]), lc[d]) for d in range(lbvh.D)) and \
        np.array_equal(np.asarray(inp.h), lbvh.leaf_scale)
