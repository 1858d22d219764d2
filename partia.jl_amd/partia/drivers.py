"""Interpolation drivers on the HIP backend.

Mirrors KI/grid_interpolation/setup/initialize_interpolation.jl:39-111 and the
drivers KI/grid_interpolation/drivers/{PointSamples,LineSamples,StructuredGrid}_driver.jl
with the same signatures, overloads, defaults and errors:
  * PointSamples_interpolation(backend, template, input, catalog, strategy=itpSymmetric)
    and (backend, template, input, lbvh, catalog, strategy) -- the external-LBVH
    overload checks matches_lbvh_leaf_order first (ArgumentError);
  * LineSamples_interpolation(...) likewise, scatter only (ArgumentError for gather
    and symmetric), scalar catalogs only (MethodError otherwise);
  * StructuredGrid_interpolation(backend, COORD, template, input, catalog, strategy);
    on Cartesian grids the sample coordinates are generated on the device.
Each returns a GridBundle whose grids share the template's geometry.  As in the
reference, the auto-LBVH overloads permute the input's arrays into Morton order.
"""
import numpy as np

from . import _native as N
from .catalog import InterpolationCatalog, to_concise_catalog
from .errors import ArgumentError, MethodError
from .grids import (Cartesian, Cylindrical, GridBundle, LineSamples, PointSamples, Spherical, StructuredGrid, device_axes,
                    flatten, restore_struct)
from .input import InterpolationInput, LinearBVH, LinearBVH_, matches_lbvh_leaf_order
from .kernels import AbstractInterpolationStrategy, HIPComputeBackend, itpScatter, itpSymmetric

PRECISIONS = {"fp64": N.PH_PREC_FP64, "fp32": N.PH_PREC_FP32, "fp32c": N.PH_PREC_FP32C}


def _check_backend(backend):
    if not isinstance(backend, HIPComputeBackend):
        raise MethodError("no method of this driver for backend %r (this package implements HIPComputeBackend)"
                          % (backend,))


def _check_strategy(strategy):
    if not (isinstance(strategy, type) and issubclass(strategy, AbstractInterpolationStrategy)):
        raise MethodError("strategy must be itpGather, itpScatter or itpSymmetric")


def _precision_of(inp, precision):
    if precision is None:
        return N.PH_PREC_FP32 if inp.m.dtype == np.float32 else N.PH_PREC_FP64
    return PRECISIONS[precision] if isinstance(precision, str) else int(precision)


def initialize_interpolation(backend, grid_template, *args):
    """initialize_interpolation(backend, template, input, catalog) -> (grids, LBVH, names, concise)
    initialize_interpolation(backend, template, catalog)        -> (grids, names, concise)"""
    _check_backend(backend)
    if len(args) == 2:
        inp, catalog = args
        grids = tuple(grid_template.similar() for _ in range(catalog.L))
        lbvh = LinearBVH_(inp, np.uint64, backend)
        return grids, lbvh, catalog.ordered_names, to_concise_catalog(catalog)
    if len(args) == 1:
        (catalog,) = args
        grids = tuple(grid_template.similar() for _ in range(catalog.L))
        return grids, catalog.ordered_names, to_concise_catalog(catalog)
    raise MethodError("initialize_interpolation: unexpected arguments")


def _split_overload(args):
    """(input, catalog[, strategy]) or (input, lbvh, catalog[, strategy])."""
    args = list(args)
    if len(args) >= 2 and isinstance(args[1], LinearBVH):
        inp, lbvh, catalog = args[0], args[1], args[2]
        rest = args[3:]
    elif len(args) >= 2:
        inp, lbvh, catalog = args[0], None, args[1]
        rest = args[2:]
    else:
        raise MethodError("missing arguments")
    if not isinstance(inp, InterpolationInput) or not isinstance(catalog, InterpolationCatalog):
        raise MethodError("expected an InterpolationInput and an InterpolationCatalog")
    return inp, lbvh, catalog, rest


def _run_points(backend, lbvh, catalog, strategy, precision, coords):
    ctx = backend.context
    cn = to_concise_catalog(catalog).native()
    S = len(coords[0])
    out = np.zeros(cn.L * S, dtype=np.float64)
    x, y, z = (N.f64(c) for c in coords)
    ctx.check(N.lib().ph_point_samples(ctx.handle, lbvh.handle, cn.ref(), strategy.id, precision, S, N.ptr(x),
                                       N.ptr(y), N.ptr(z), N.ptr(out)))
    return out.reshape(cn.L, S) if cn.L else out.reshape(0, S)


def PointSamples_interpolation(backend, grid_template, *args, precision=None):
    _check_backend(backend)
    inp, lbvh, catalog, rest = _split_overload(args)
    strategy = rest[0] if rest else itpSymmetric
    _check_strategy(strategy)
    if not isinstance(grid_template, PointSamples) or len(grid_template.coor) != 3:
        raise MethodError("PointSamples_interpolation needs a 3D PointSamples template")
    if lbvh is None:
        grids, lbvh, names, _ = initialize_interpolation(backend, grid_template, inp, catalog)
    else:
        if not matches_lbvh_leaf_order(inp, lbvh):
            raise ArgumentError("Provided LBVH leaf order does not match the current input ordering. "
                                "Ensure the LBVH was built from the same Morton-reordered input.")
        lbvh.refresh_columns(inp)
        grids, names, _ = initialize_interpolation(backend, grid_template, catalog)
    vals = _run_points(backend, lbvh, catalog, strategy, _precision_of(inp, precision), grid_template.coor)
    for k, g in enumerate(grids):
        g.grid[:] = vals[k].astype(g.grid.dtype)
    return GridBundle(grids, names)


def _line_catalog_ok(catalog):
    N_, G, D, C = catalog.shape
    if G or D or C:
        raise MethodError("LineSamples_interpolation accepts catalogs of scalar quantities only")


def LineSamples_interpolation(backend, grid_template, *args, precision=None):
    _check_backend(backend)
    inp, lbvh, catalog, rest = _split_overload(args)
    strategy = rest[0] if rest else itpScatter
    _check_strategy(strategy)
    _line_catalog_ok(catalog)
    if not isinstance(grid_template, LineSamples) or len(grid_template.origin) != 3:
        raise MethodError("LineSamples_interpolation needs a 3D LineSamples template")
    if strategy is not itpScatter:
        raise ArgumentError("LineSamples_interpolation only supports itpScatter. Line-integrated samples do not "
                            "have a well-defined query smoothing length ha, so itpGather and itpSymmetric are not "
                            "supported.")
    if lbvh is None:
        grids, lbvh, names, _ = initialize_interpolation(backend, grid_template, inp, catalog)
    else:
        if not matches_lbvh_leaf_order(inp, lbvh):
            raise ArgumentError("Provided LBVH leaf order does not match the current input ordering. "
                                "Ensure the LBVH was built from the same Morton-reordered input.")
        lbvh.refresh_columns(inp)
        grids, names, _ = initialize_interpolation(backend, grid_template, catalog)
    ctx = backend.context
    cn = to_concise_catalog(catalog).native()
    S = len(grid_template)
    out = np.zeros(max(cn.L, 1) * S)
    o = [N.f64(a) for a in grid_template.origin]
    d = [N.f64(a) for a in grid_template.direction]
    ctx.check(N.lib().ph_line_samples(ctx.handle, lbvh.handle, cn.ref(), strategy.id, _precision_of(inp, precision),
                                      S, *[N.ptr(a) for a in o + d], N.ptr(out)))
    out = out[:cn.L * S].reshape(cn.L, S)
    for k, g in enumerate(grids):
        g.grid[:] = out[k].astype(g.grid.dtype)
    return GridBundle(grids, names)


def StructuredGrid_interpolation(backend, coord, grid_template, inp, catalog, strategy=itpSymmetric, precision=None):
    """StructuredGrid_interpolation (StructuredGrid_driver.jl:46-64)."""
    _check_backend(backend)
    _check_strategy(strategy)
    if not isinstance(grid_template, StructuredGrid) or len(grid_template.axes) != 3:
        raise MethodError("StructuredGrid_interpolation needs a 3D StructuredGrid template")
    if coord not in (Cartesian, Cylindrical, Spherical):
        raise MethodError("StructuredGrid_interpolation takes Cartesian, Cylindrical or Spherical grids")
    # the flattened coordinates are generated on the device from the axes
    # (angle axes as sin / cos pairs)
    lbvh = LinearBVH_(inp, np.uint64, backend)
    ctx = backend.context
    cn = to_concise_catalog(catalog).native()
    nx, ny, nz = grid_template.size
    S = nx * ny * nz
    out = np.zeros(max(cn.L, 1) * S)
    cid, a1, a2, a3 = device_axes(coord, grid_template)
    ax, ay, az = N.f64(a1), N.f64(a2), N.f64(a3)
    ctx.check(N.lib().ph_structured_grid_coords(ctx.handle, lbvh.handle, cn.ref(), strategy.id,
                                                _precision_of(inp, precision), cid, N.ptr(ax), nx, N.ptr(ay), ny,
                                                N.ptr(az), nz, N.ptr(out)))
    lbvh.close()
    T = grid_template.dtype
    axes = tuple(np.asarray(a, dtype=T) for a in grid_template.axes)
    grids = tuple(StructuredGrid(np.reshape(out[k * S:(k + 1) * S].astype(T), grid_template.size, order="F"), axes,
                                 grid_template.size) for k in range(cn.L))
    return GridBundle(grids, catalog.ordered_names)


def LBVH_find_nearest_h(lbvh, points):
    """LBVH_find_nearest_h (LBVH/LinearBVH/LinearBVHQueries.jl:118-140) for many points:
    points is (x, y, z); returns the h of the nearest leaf of each point."""
    ctx = lbvh.ctx
    x, y, z = (N.f64(c) for c in points)
    ha = np.zeros(len(x))
    ctx.check(N.lib().ph_nearest_h(ctx.handle, lbvh.handle, len(x), N.ptr(x), N.ptr(y), N.ptr(z), N.ptr(ha)))
    return ha
