"""partia -- MI355X (gfx950) implementation of Partia.jl's SPH interpolation path.

The public names mirror the reference's Julia API (Partia.jl exports) for this
path; the numerical work runs in libpartia_hip.so through its C ABI:

    from partia import *
    backend = HIPComputeBackend()
    input, catalog = build_input(backend, x, y, z, m, h, rho, (A, vx, vy, vz),
                                 column_names=("A", "vx", "vy", "vz"),
                                 scalars=("A",), gradients=("A",), divergences=("v",))
    grid = StructuredGrid(Cartesian, (0.0, 1.0, 64), (0.0, 1.0, 64), (0.0, 1.0, 64))
    bundle = StructuredGrid_interpolation(backend, Cartesian, grid, input, catalog, itpSymmetric)

Julia's `LinearBVH!` is `LinearBVH_` here (trailing underscore = in place).
"""
from .catalog import (InterpolationCatalog, InterpolationCatalogConcise, curl_slots, div_slots, grad_slot,
                      ordered_quantity_names, scalar_index, to_concise_catalog)
from .drivers import (LBVH_find_nearest_h, LineSamples_interpolation, PointSamples_interpolation,
                      StructuredGrid_interpolation, initialize_interpolation)
from .errors import ArgumentError, DeviceError, DimensionMismatch, JuliaAssertionError, MethodError
from .grids import (AbstractGrid, AbstractSamples, Cartesian, Cylindrical, Frame, GridBundle, LineSamples,
                    ParallelBeam, PointSamples, frame_forward, frame_position, frame_right, frame_up,
                    Polar, Spherical, StructuredGrid, coordinate_grid, flatten, restore_struct)
from .gridsio import GridDataset, file_identifier, read_GridDataset, write_GridBundle, write_GridDataset
from .input import (AABB, BinaryRadixTree, InterpolationInput, LinearBVH, LinearBVH_, UInt32, UInt64, build_input,
                    get_coord, get_xcoord, get_ycoord, get_zcoord, matches_lbvh_leaf_order)
from .kernels import (ALL_KERNELS, AbstractExecutionBackend, AbstractInterpolationStrategy, AbstractSPHKernel,
                      C2_Wendland, C4_Wendland, C6_Wendland, CPUComputeBackend, HIPComputeBackend, KernelFunctionNneigh,
                      KernelFunctionnorm, KernelFunctionValid, M4_spline, M5_spline, M6_spline, itpGather, itpScatter,
                      itpSymmetric)
from ._native import NativeLibraryError

__all__ = [n for n in dir() if not n.startswith("_")]
