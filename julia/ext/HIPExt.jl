# HIPExt.jl -- the AMD Instinct (gfx950) backend of Partia.jl's interpolation
# hot path: methods of the reference's driver functions for
# `HIPComputeBackend` (declared in HIPExtDummy.jlh), each calling
# libpartia_hip through `ccall`.  The library's C ABI is
# include/partia_hip.h; every call below passes exactly the C parameter types
# declared there (tests/test_julia_binding.py checks each ccall's arity and
# types against the header, and each field access against the reference's
# structs).
#
# The reference's CUDA extension builds the LBVH on the host and runs one
# thread per sample (ext/CUDAExt/KernelInterpolation/grid_interpolation/
# drivers/PointSamples_driver.jl:12-35); here `LinearBVH!` runs on the GPU and
# the host input is permuted with the device's Morton order afterwards, as
# LinearBVH! permutes it (src/julia/KernelInterpolation/interpolation_setup/
# InterpolationInput.jl:201-221).  The library path is ENV["PARTIA_HIP_LIB"].
module HIPExt

using Libdl
using Partia
using Partia: AbstractInterpolationStrategy, C2_Wendland, C4_Wendland, C6_Wendland, Cartesian, Cylindrical,
              GridBundle, HIPComputeBackend, InterpolationCatalog, InterpolationCatalogConcise, InterpolationInput,
              LineSamples, LinearBVH, M4_spline, M5_spline, M6_spline, PointSamples, Spherical, StructuredGrid,
              itpGather, itpScatter, itpSymmetric

const libpartia_hip = get(ENV, "PARTIA_HIP_LIB", "libpartia_hip")

# ---- constants of include/partia_hip.h ----
const PH_OK = Cint(0)
const PH_ERR_ARGUMENT = Cint(-1)
const PH_ERR_DIMENSION = Cint(-2)
const PH_ERR_STRATEGY = Cint(-3)
const PH_ERR_LEAF_ORDER = Cint(-4)
const PH_ERR_CAPACITY = Cint(-5)
const PH_ERR_INTERNAL = Cint(-8)
const PH_PREC_FP64 = Int32(64)
const PH_PREC_FP32 = Int32(32)
const PH_PREC_FP32C = Int32(33)
const PH_GET_ORDER = Int32(0)
const PH_COORDS_CARTESIAN = Int32(0)
const PH_COORDS_CYLINDRICAL = Int32(1)
const PH_COORDS_SPHERICAL = Int32(2)

# kernel_function/kernels/<name>.jl -> PH_KERNEL_*
kernel_id(:: M4_spline) = Int32(0)
kernel_id(:: M5_spline) = Int32(1)
kernel_id(:: M6_spline) = Int32(2)
kernel_id(:: C2_Wendland) = Int32(3)
kernel_id(:: C4_Wendland) = Int32(4)
kernel_id(:: C6_Wendland) = Int32(5)
# InterpolationStrategy.jl:1-5 -> PH_GATHER / PH_SCATTER / PH_SYMMETRIC
strategy_id(:: Type{itpGather}) = Int32(0)
strategy_id(:: Type{itpScatter}) = Int32(1)
strategy_id(:: Type{itpSymmetric}) = Int32(2)
# tree precision of an input float type, and the evaluation precision of a call
tree_precision(:: Type{Float64}) = PH_PREC_FP64
tree_precision(:: Type{Float32}) = PH_PREC_FP32
call_precision(b :: HIPComputeBackend, :: Type{Float64}) = b.compensated ? PH_PREC_FP32C : PH_PREC_FP64
call_precision(b :: HIPComputeBackend, :: Type{Float32}) = PH_PREC_FP32
coords_id(:: Type{Cartesian}) = PH_COORDS_CARTESIAN
coords_id(:: Type{Cylindrical}) = PH_COORDS_CYLINDRICAL
coords_id(:: Type{Spherical}) = PH_COORDS_SPHERICAL

const LINE_STRATEGY_MESSAGE =
    "LineSamples_interpolation only supports itpScatter. " *
    "Line-integrated samples do not have a well-defined query smoothing length ha, " *
    "so itpGather and itpSymmetric are not supported."
const LEAF_ORDER_MESSAGE =
    "Provided LBVH leaf order does not match the current input ordering. " *
    "Ensure the LBVH was built from the same Morton-reordered input."

# ---- status codes -> the exceptions the reference throws ----
function check(rc :: Cint, ctx :: Ptr{Cvoid})
    rc == PH_OK && return nothing
    msg = unsafe_string(ccall((:ph_last_error, libpartia_hip), Cstring, (Ptr{Cvoid},), ctx))
    rc == PH_ERR_DIMENSION && throw(DimensionMismatch(msg))
    rc in (PH_ERR_ARGUMENT, PH_ERR_STRATEGY, PH_ERR_LEAF_ORDER, PH_ERR_CAPACITY) && throw(ArgumentError(msg))
    rc == PH_ERR_INTERNAL && throw(AssertionError(msg))
    error("libpartia_hip status $(rc): $(msg)")  # out of memory, HIP error, no device
end

# ---- one context (GPU + stream) per device, for the process's lifetime ----
const CONTEXTS = Dict{Int, Ptr{Cvoid}}()
const CONTEXT_LOCK = ReentrantLock()

function context(b :: HIPComputeBackend)
    lock(CONTEXT_LOCK) do
        get!(CONTEXTS, b.device) do
            out = Ref{Ptr{Cvoid}}(C_NULL)
            check(ccall((:ph_ctx_create, libpartia_hip), Cint, (Cint, Ref{Ptr{Cvoid}}), b.device, out), C_NULL)
            out[]
        end
    end
end

function Partia.Greeting_HIP(b :: HIPComputeBackend = HIPComputeBackend())
    ctx = context(b)
    @info "Hello from HIPExt: libpartia_hip context on device $(b.device)" libpartia_hip ctx
    return nothing
end

# ---- the concise catalog as the C ABI reads it ----
# ph_catalog (include/partia_hip.h) mirrors InterpolationCatalogConcise
# (InterpolationCatalog.jl:26-32); its slots are Int tuples, the C ABI reads
# int32 / uint8 arrays, so the tuples are copied into vectors kept alive by
# CatalogArrays while a call runs.
struct PhCatalog
    n_scalar :: Int32
    scalar_slots :: Ptr{Int32}
    scalar_shepard :: Ptr{UInt8}
    n_grad :: Int32
    grad_slots :: Ptr{Int32}
    n_div :: Int32
    div_slots :: Ptr{Int32}
    n_curl :: Int32
    curl_slots :: Ptr{Int32}
end

struct CatalogArrays
    scalars :: Vector{Int32}
    shepard :: Vector{UInt8}
    grads :: Vector{Int32}
    divs :: Vector{Int32}
    curls :: Vector{Int32}
end

function CatalogArrays(cc :: InterpolationCatalogConcise)
    flat(t) = Int32[s for tup in t for s in tup]
    return CatalogArrays(Int32[s for s in cc.scalar_slots], UInt8[f ? 0x01 : 0x00 for f in cc.scalar_snormalization],
                         Int32[s for s in cc.grad_slots], flat(cc.div_slots), flat(cc.curl_slots))
end

# Only valid while `a` is preserved (GC.@preserve a ...).
PhCatalog(a :: CatalogArrays) =
    PhCatalog(Int32(length(a.scalars)), pointer(a.scalars), pointer(a.shepard), Int32(length(a.grads)),
              pointer(a.grads), Int32(length(a.divs) ÷ 3), pointer(a.divs), Int32(length(a.curls) ÷ 3),
              pointer(a.curls))

# ---- device trees ----
# A ph_lbvh: the device LinearBVH of one InterpolationInput, destroyed with
# the Julia object.
mutable struct DeviceLBVH
    handle :: Ptr{Cvoid}
    function DeviceLBVH(handle :: Ptr{Cvoid})
        t = new(handle)
        finalizer(t) do x
            if x.handle != C_NULL
                ccall((:ph_lbvh_destroy, libpartia_hip), Cint, (Ptr{Cvoid},), x.handle)
                x.handle = C_NULL
            end
        end
        return t
    end
end

# The C ABI reads float64 columns (Float32 inputs are widened exactly and the
# library rounds them back for a Float32 tree).
as_f64(v :: Vector{Float64}) = v
as_f64(v :: AbstractVector) = Vector{Float64}(v)

# LinearBVH!(input; CodeType) on the GPU: code_bits 63 (UInt64, the drivers'
# default) or 30 (UInt32).
function lbvh_build(ctx :: Ptr{Cvoid}, input :: InterpolationInput{3, T}; code_bits :: Integer = 63) where {T}
    x, y, z = as_f64(input.coord[1]), as_f64(input.coord[2]), as_f64(input.coord[3])
    m, h, rho = as_f64(input.m), as_f64(input.h), as_f64(input.ρ)
    quant = Vector{Float64}[as_f64(q) for q in input.quant]
    qptr = Ptr{Float64}[pointer(q) for q in quant]
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve x y z m h rho quant qptr begin
        check(ccall((:ph_lbvh_build, libpartia_hip), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                     Ptr{Float64}, Int32, Ptr{Ptr{Float64}}, Int32, Int32, Int32, Ref{Ptr{Cvoid}}),
                    ctx, input.Npart, x, y, z, m, h, rho, length(qptr), qptr, kernel_id(input.smoothed_kernel),
                    code_bits, tree_precision(T), out), ctx)
    end
    return DeviceLBVH(out[])
end

# Permutes every host column of `input` into the device tree's Morton order,
# exactly as LinearBVH! permutes them (InterpolationInput.jl:209-217).
function permute_like_lbvh!(ctx :: Ptr{Cvoid}, input :: InterpolationInput{3}, t :: DeviceLBVH)
    order = Vector{UInt64}(undef, input.Npart)
    check(ccall((:ph_lbvh_get, libpartia_hip), Cint, (Ptr{Cvoid}, Int32, Ptr{UInt64}), t.handle, PH_GET_ORDER,
                order), ctx)
    for column in (input.coord..., input.m, input.h, input.ρ, input.quant...)
        Base.permute!(column, order)
    end
    return input
end

# Output k of sample i at out[(k - 1) S + i] -> grids[k].grid[i] (linear
# index; column-major for StructuredGrid arrays).
function scatter_outputs!(grids, out :: Vector{Float64}, S :: Integer)
    for k in eachindex(grids)
        g = grids[k].grid
        @inbounds for i in 1:S
            g[i] = out[(k - 1) * S + i]
        end
    end
    return grids
end

# ---- initialize_interpolation: the LBVH is built on the device ----
# Returns (grids, DeviceLBVH, names, concise catalog), the reference's 4-tuple
# (initialize_interpolation.jl:39-55) with the device tree in place of the
# host LinearBVH; the host input is left in Morton order, as LinearBVH! leaves it.
function Partia.initialize_interpolation(b :: HIPComputeBackend, grid_template :: AS,
                                         input :: InterpolationInput{3, T},
                                         catalog :: InterpolationCatalog{3, N, G, Div, C, L}) where {N, G, Div, C, L, T <: AbstractFloat, AS}
    grids, names, cc = Partia.initialize_interpolation(Partia.CPUComputeBackend(), grid_template, catalog)
    ctx = context(b)
    t = lbvh_build(ctx, input)
    permute_like_lbvh!(ctx, input, t)
    return grids, t, names, cc
end

# ---- PointSamples_interpolation (PointSamples_driver.jl:34-48, 98-118) ----
function point_samples!(b :: HIPComputeBackend, t :: DeviceLBVH, grids, cc :: InterpolationCatalogConcise,
                        grid_template :: PointSamples{3, TF}, :: Type{S}) where {TF, S}
    ctx = context(b)
    npts = length(grid_template.grid)
    out = Vector{Float64}(undef, length(grids) * npts)
    px, py, pz = as_f64(grid_template.coor[1]), as_f64(grid_template.coor[2]), as_f64(grid_template.coor[3])
    a = CatalogArrays(cc)
    GC.@preserve a px py pz out begin
        cat = Ref(PhCatalog(a))
        check(ccall((:ph_point_samples, libpartia_hip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ref{PhCatalog}, Int32, Int32, Int64, Ptr{Float64}, Ptr{Float64},
                     Ptr{Float64}, Ptr{Float64}),
                    ctx, t.handle, cat, strategy_id(S), call_precision(b, TF), npts, px, py, pz, out), ctx)
    end
    return scatter_outputs!(grids, out, npts)
end

function Partia.PointSamples_interpolation(b :: HIPComputeBackend, grid_template :: PointSamples{3, TF},
                                           input :: InterpolationInput{3, TF},
                                           catalog :: InterpolationCatalog{3, N, G, Div, C, L},
                                           itp_strategy :: Type{ITPSTRATEGY} = itpSymmetric) where {N, G, Div, C, L, TF <: AbstractFloat, ITPSTRATEGY <: AbstractInterpolationStrategy}
    grids, t, names, cc = Partia.initialize_interpolation(b, grid_template, input, catalog)
    point_samples!(b, t, grids, cc, grid_template, itp_strategy)
    return GridBundle(grids, names)
end

# External-LBVH overload: the input is already in the given tree's Morton
# order, so the device build sorts nothing and reproduces that tree.
function Partia.PointSamples_interpolation(b :: HIPComputeBackend, grid_template :: PointSamples{3, TF},
                                           input :: InterpolationInput{3, TF}, lbvh :: LinearBVH{3, TF},
                                           catalog :: InterpolationCatalog{3, N, G, Div, C, L},
                                           itp_strategy :: Type{ITPSTRATEGY} = itpSymmetric) where {N, G, Div, C, L, TF <: AbstractFloat, ITPSTRATEGY <: AbstractInterpolationStrategy}
    Partia.matches_lbvh_leaf_order(input, lbvh) || throw(ArgumentError(LEAF_ORDER_MESSAGE))
    grids, names, cc = Partia.initialize_interpolation(Partia.CPUComputeBackend(), grid_template, catalog)
    t = lbvh_build(context(b), input)
    point_samples!(b, t, grids, cc, grid_template, itp_strategy)
    return GridBundle(grids, names)
end

# ---- LineSamples_interpolation (LineSamples_driver.jl:39-57, 116-139) ----
function line_samples!(b :: HIPComputeBackend, t :: DeviceLBVH, grids, cc :: InterpolationCatalogConcise,
                       grid_template :: LineSamples{3, TF}) where {TF}
    ctx = context(b)
    nlines = length(grid_template.grid)
    out = Vector{Float64}(undef, length(grids) * nlines)
    ox, oy, oz = (as_f64(grid_template.origin[d]) for d in 1:3)
    dx, dy, dz = (as_f64(grid_template.direction[d]) for d in 1:3)
    a = CatalogArrays(cc)
    GC.@preserve a ox oy oz dx dy dz out begin
        cat = Ref(PhCatalog(a))
        check(ccall((:ph_line_samples, libpartia_hip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ref{PhCatalog}, Int32, Int32, Int64, Ptr{Float64}, Ptr{Float64},
                     Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    ctx, t.handle, cat, strategy_id(itpScatter), call_precision(b, TF), nlines, ox, oy, oz, dx, dy,
                    dz, out), ctx)
    end
    return scatter_outputs!(grids, out, nlines)
end

function Partia.LineSamples_interpolation(b :: HIPComputeBackend, grid_template :: LineSamples{3, TF},
                                          input :: InterpolationInput{3, TF},
                                          catalog :: InterpolationCatalog{3, N, 0, 0, 0, N},
                                          itp_strategy :: Type{ITPSTRATEGY} = itpScatter) where {N, TF <: AbstractFloat, ITPSTRATEGY <: AbstractInterpolationStrategy}
    itp_strategy === itpScatter || throw(ArgumentError(LINE_STRATEGY_MESSAGE))
    grids, t, names, cc = Partia.initialize_interpolation(b, grid_template, input, catalog)
    line_samples!(b, t, grids, cc, grid_template)
    return GridBundle(grids, names)
end

function Partia.LineSamples_interpolation(b :: HIPComputeBackend, grid_template :: LineSamples{3, TF},
                                          input :: InterpolationInput{3, TF}, lbvh :: LinearBVH{3, TF},
                                          catalog :: InterpolationCatalog{3, N, 0, 0, 0, N},
                                          itp_strategy :: Type{ITPSTRATEGY} = itpScatter) where {N, TF <: AbstractFloat, ITPSTRATEGY <: AbstractInterpolationStrategy}
    itp_strategy === itpScatter || throw(ArgumentError(LINE_STRATEGY_MESSAGE))
    Partia.matches_lbvh_leaf_order(input, lbvh) || throw(ArgumentError(LEAF_ORDER_MESSAGE))
    grids, names, cc = Partia.initialize_interpolation(Partia.CPUComputeBackend(), grid_template, catalog)
    t = lbvh_build(context(b), input)
    line_samples!(b, t, grids, cc, grid_template)
    return GridBundle(grids, names)
end

# ---- StructuredGrid_interpolation (StructuredGrid_driver.jl:46-64) ----
# The grid points are generated on the device from the axes (first axis
# fastest, as flatten's coordinate_grid orders them), so no flattened
# coordinate array exists; angle axes go over as Julia's own (sin, cos) pairs.
function sincos_pairs(ax :: AbstractVector)
    out = Vector{Float64}(undef, 2 * length(ax))
    for (j, angle) in enumerate(ax)
        s, c = sincos(angle)
        out[2j - 1] = s
        out[2j] = c
    end
    return out
end
axis_arrays(:: Type{Cartesian}, ax) = (as_f64(ax[1]), as_f64(ax[2]), as_f64(ax[3]))
axis_arrays(:: Type{Cylindrical}, ax) = (as_f64(ax[1]), sincos_pairs(ax[2]), as_f64(ax[3]))
axis_arrays(:: Type{Spherical}, ax) = (as_f64(ax[1]), sincos_pairs(ax[2]), sincos_pairs(ax[3]))

function Partia.StructuredGrid_interpolation(b :: HIPComputeBackend, :: Type{COORD}, grid_template :: StructuredGrid{3},
                                             input :: InterpolationInput{3, T},
                                             catalog :: InterpolationCatalog{3, N, G, Div, C, L},
                                             itp_strategy :: Type{ITPSTRATEGY} = itpSymmetric) where {COORD <: Union{Cartesian, Cylindrical, Spherical}, N, G, Div, C, L, T <: AbstractFloat, ITPSTRATEGY <: AbstractInterpolationStrategy}
    ctx = context(b)
    t = lbvh_build(ctx, input)
    permute_like_lbvh!(ctx, input, t)
    n1, n2, n3 = grid_template.size
    S = n1 * n2 * n3
    a1, a2, a3 = axis_arrays(COORD, grid_template.axes)
    out = Vector{Float64}(undef, L * S)
    a = CatalogArrays(Partia.to_concise_catalog(catalog))
    GC.@preserve a a1 a2 a3 out begin
        cat = Ref(PhCatalog(a))
        check(ccall((:ph_structured_grid_coords, libpartia_hip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ref{PhCatalog}, Int32, Int32, Int32, Ptr{Float64}, Int64, Ptr{Float64},
                     Int64, Ptr{Float64}, Int64, Ptr{Float64}),
                    ctx, t.handle, cat, strategy_id(itp_strategy), call_precision(b, T), coords_id(COORD), a1, n1,
                    a2, n2, a3, n3, out), ctx)
    end
    # outputs in the input's float type on the template's axes (restore_struct)
    axes = ntuple(d -> T.(grid_template.axes[d]), Val(3))
    grids = ntuple(k -> StructuredGrid(reshape(T.(out[(k - 1) * S + 1:k * S]), n1, n2, n3), axes, (n1, n2, n3)),
                   Val(L))
    return GridBundle(grids, catalog.ordered_names)
end

# ---- batched serving: many PointSamples requests over one tree in one
# graph launch (ph_batch_*; no reference counterpart) ----
# ph_request (include/partia_hip.h); kind 0 = device point list, 1 = plane,
# 2 = box of a frame (PointSamples(Cartesian, frame, ...), PointSamples.jl:328-356).
struct PhRequest
    kind :: Int32
    reserved :: Int32
    count :: Int64
    px :: Ptr{Float64}
    py :: Ptr{Float64}
    pz :: Ptr{Float64}
    position :: NTuple{3, Float64}
    right :: NTuple{3, Float64}
    up :: NTuple{3, Float64}
    forward :: NTuple{3, Float64}
    amin :: NTuple{3, Float64}
    amax :: NTuple{3, Float64}
    n :: NTuple{3, Int64}
    out :: Ptr{Cvoid}
    trig :: Ptr{Float64}
end

# PointSamples(Cartesian, frame, xparams, yparams) as a request writing L x nx ny
# outputs to device memory `out`; position / right / up are the frame's
# frame_position / frame_right / frame_up (Frames/struct/Frame.jl:145-186).
plane_request(position, right, up, xparams, yparams, out :: Ptr{Cvoid}) =
    PhRequest(Int32(1), Int32(0), Int64(0), C_NULL, C_NULL, C_NULL, Tuple(position), Tuple(right), Tuple(up),
              (0.0, 0.0, 0.0), (xparams[1], yparams[1], 0.0), (xparams[2], yparams[2], 1.0),
              (Int64(xparams[3]), Int64(yparams[3]), Int64(2)), out, C_NULL)

mutable struct DeviceBatch
    handle :: Ptr{Cvoid}
    tree :: DeviceLBVH  # kept alive while the batch is
    function DeviceBatch(handle :: Ptr{Cvoid}, tree :: DeviceLBVH)
        x = new(handle, tree)
        finalizer(x) do y
            if y.handle != C_NULL
                ccall((:ph_batch_destroy, libpartia_hip), Cint, (Ptr{Cvoid},), y.handle)
                y.handle = C_NULL
            end
        end
        return x
    end
end

function batch_create(b :: HIPComputeBackend, t :: DeviceLBVH, catalog :: InterpolationCatalog,
                      :: Type{S}, reqs :: Vector{PhRequest}, :: Type{TF} = Float64) where {S, TF}
    ctx = context(b)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    a = CatalogArrays(Partia.to_concise_catalog(catalog))
    GC.@preserve a reqs begin
        cat = Ref(PhCatalog(a))
        check(ccall((:ph_batch_create, libpartia_hip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ref{PhCatalog}, Int32, Int32, Int32, Ptr{PhRequest}, Ref{Ptr{Cvoid}}),
                    ctx, t.handle, cat, strategy_id(S), call_precision(b, TF), length(reqs), reqs, out), ctx)
    end
    return DeviceBatch(out[], t)
end

batch_run(b :: HIPComputeBackend, x :: DeviceBatch) =
    check(ccall((:ph_batch_run, libpartia_hip), Cint, (Ptr{Cvoid},), x.handle), context(b))

# ---- column density (LineSamples, one scalar, scatter): the differentiable
# form used by HIPChainRulesExt ----
"""
    column_density(b, grid_template::LineSamples{3,Float64}, x, y, z, m, h, ρ, A, kernel; shepard = false)

Line-integrated `A` along the lines of `grid_template`
(`_line_integrated_quantities_interpolate_kernel`, scatter), from particle
arrays in any order; with `A = ρ` and `shepard = false` this is the column
density Σ m I(r/h)/h.  Differentiable through its ChainRulesCore `rrule`.
"""
function Partia.column_density(b :: HIPComputeBackend, grid_template :: LineSamples{3, Float64}, x, y, z, m, h, ρ,
                               A, kernel; shepard :: Bool = false)
    ctx = context(b)
    input = InterpolationInput(length(x), kernel, (copy(x), copy(y), copy(z)), copy(m), copy(h), copy(ρ),
                               (copy(A),))
    t = lbvh_build(ctx, input)
    nlines = length(grid_template.grid)
    out = Vector{Float64}(undef, nlines)
    cc = InterpolationCatalogConcise{3, 1, 0, 0, 0}((1,), (shepard,), (), (), ())
    line_samples!(b, t, (PointSamplesLike(out),), cc, grid_template)
    return out
end
# A one-output grid holder for scatter_outputs!.
struct PointSamplesLike
    grid :: Vector{Float64}
end

# Device buffer from the library's allocator, filled from a host vector.
function device_copy(ctx :: Ptr{Cvoid}, v :: Vector{Float64})
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:ph_device_alloc, libpartia_hip), Cint, (Ptr{Cvoid}, Int64, Ref{Ptr{Cvoid}}), ctx, sizeof(v), p),
          ctx)
    check(ccall((:ph_memcpy_htod, libpartia_hip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Int64), ctx, p[], v,
                sizeof(v)), ctx)
    return p[]
end
device_free(ctx :: Ptr{Cvoid}, p :: Ptr{Cvoid}) =
    check(ccall((:ph_device_free, libpartia_hip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx, p), ctx)

# Gradient of sum_s cot[s] column_density(...)[s] with respect to (m, A, h, x,
# y, z) of every particle, in the order the arrays were given (6 x n, rho held
# fixed): ph_line_samples_adjoint_device.
function column_density_adjoint(b :: HIPComputeBackend, grid_template :: LineSamples{3, Float64}, x, y, z, m, h, ρ,
                                A, kernel, cot :: Vector{Float64}; shepard :: Bool = false)
    ctx = context(b)
    input = InterpolationInput(length(x), kernel, (copy(x), copy(y), copy(z)), copy(m), copy(h), copy(ρ),
                               (copy(A),))
    t = lbvh_build(ctx, input)
    n = length(x)
    nlines = length(grid_template.grid)
    rays = [device_copy(ctx, as_f64(v)) for v in (grid_template.origin..., grid_template.direction...)]
    dcot = device_copy(ctx, cot)
    dgrad = device_copy(ctx, zeros(6 * n))
    grad = Matrix{Float64}(undef, n, 6)
    a = CatalogArrays(InterpolationCatalogConcise{3, 1, 0, 0, 0}((1,), (shepard,), (), (), ()))
    GC.@preserve a grad begin
        cat = Ref(PhCatalog(a))
        check(ccall((:ph_line_samples_adjoint_device, libpartia_hip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ref{PhCatalog}, Int32, Int32, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                     Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                    ctx, t.handle, cat, strategy_id(itpScatter), PH_PREC_FP64, nlines, rays[1], rays[2], rays[3],
                    rays[4], rays[5], rays[6], dcot, dgrad), ctx)
        check(ccall((:ph_memcpy_dtoh, libpartia_hip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}, Int64), ctx, grad,
                    dgrad, sizeof(grad)), ctx)
    end
    for p in (rays..., dcot, dgrad)
        device_free(ctx, p)
    end
    return grad  # columns: d/dm, d/dA, d/dh, d/dx, d/dy, d/dz
end

end
