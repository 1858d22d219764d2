// ph_primitives.h -- device-wide scan, reduction and radix sort used by the LBVH build.
#pragma once

#include "ph_common.h"

namespace ph {

// Exclusive prefix sum of a u32 array, in place.
void exclusive_sum_u32(u32* data, i64 m, DevBuf& tmp, hipStream_t s);
// Inclusive prefix maximum of a u64 array, in place.
void inclusive_max_u64(u64* data, i64 m, DevBuf& tmp, hipStream_t s);
// Component-wise min and max of three arrays: out6 = {xmin, ymin, zmin, xmax, ymax, zmax} (device).
template <class T> void minmax3(const T* x, const T* y, const T* z, i64 n, T* out6, DevBuf& tmp, hipStream_t s);
// Stable LSD radix sort of (key, value) pairs over the low `bits` bits of the key.
// Ties keep their input order, so sorting (code, index) pairs with the index as
// value reproduces a stable sort by code.
void radix_sort_pairs(u64* keys, u32* vals, i64 n, int bits, DevBuf& tmp, hipStream_t s);
void radix_sort_pairs(u32* keys, u32* vals, i64 n, int bits, DevBuf& tmp, hipStream_t s);

// Blocks of `block` threads covering n items, one item per thread: kernels
// launched on it without a grid-stride loop need the default cap, which covers
// 2^31 items at 256 threads (the LinearBVH's 2^30 leaves, 2^31 - 1 nodes).
inline int grid_for(i64 n, int block = kBlock, i64 cap = i64(1) << 23) {
    i64 g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return int(g);
}

}  // namespace ph
