// ph_build.h -- LBVH construction on the device.
#pragma once

#include "ph_common.h"

namespace ph {

// Scratch buffers of one build (freed with the object).
struct BuildScratch {
    DevBuf ptrs, mm, keys, dk, adj, visited, tmp, stage, stage32, flag, aos;
};

// Builds the tree of B (n, ncol, kernel, code_bits, precision set by the caller)
// from 6+ncol float64 columns x, y, z, m, h, rho, quant... (host or device
// pointers).  Synchronises the stream before returning (the refit check), unless
// deferred_bad is given: then *deferred_bad is set to the device count of the
// check (in W, or left null for trees of fewer than 2 leaves) for the caller to
// read back and test for 0.
void build_lbvh(ph_lbvh* B, const double* const* src64, bool host_src, hipStream_t s, BuildScratch& W,
                u32** deferred_bad = nullptr);

}  // namespace ph
