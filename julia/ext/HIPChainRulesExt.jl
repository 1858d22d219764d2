# HIPChainRulesExt.jl -- reverse-mode rule of the GPU column density, loaded
# when ChainRulesCore is present next to HIPExt (a weak dependency, like
# CUDA.jl for ext/CUDAExt.jl: Project.toml:19-25 of the reference).  The
# pullback calls ph_line_samples_adjoint_device through
# HIPExt.column_density_adjoint: the library recomputes each ray's neighbours
# instead of storing them.  ρ is held fixed by the adjoint, so its tangent is
# reported as not implemented rather than as zero.
module HIPChainRulesExt

using ChainRulesCore
using Partia
using Partia: HIPComputeBackend, LineSamples

const HIPExt = Base.get_extension(Partia, :HIPExt)

function ChainRulesCore.rrule(:: typeof(Partia.column_density), b :: HIPComputeBackend,
                              grid_template :: LineSamples{3, Float64}, x, y, z, m, h, ρ, A, kernel;
                              shepard :: Bool = false)
    out = Partia.column_density(b, grid_template, x, y, z, m, h, ρ, A, kernel; shepard = shepard)
    function column_density_pullback(Δ)
        cot = Vector{Float64}(unthunk(Δ))
        # NaN outputs (rays without neighbours) carry no gradient
        replace!(v -> isnan(v) ? 0.0 : v, cot)
        g = HIPExt.column_density_adjoint(b, grid_template, x, y, z, m, h, ρ, A, kernel, cot; shepard = shepard)
        dm, dA, dh, dx, dy, dz = (g[:, k] for k in 1:6)
        return (NoTangent(), NoTangent(), NoTangent(), dx, dy, dz, dm, dh,
                @not_implemented("the adjoint holds ρ fixed"), dA, NoTangent())
    end
    return out, column_density_pullback
end

end
