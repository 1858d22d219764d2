"""partia -- MI355X (gfx950) implementation of Partia.jl's SPH interpolation path."""
