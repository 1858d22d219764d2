// ph_primitives.hip -- scans, min/max reduction and the stable LSD radix sort.
//
// All kernels use 256-thread workgroups (4 wave64 wavefronts).  Scans are the
// classic reduce / scan-partials / downsweep triple; the radix sort uses 8-bit
// digits, a per-tile LDS histogram, a device-wide exclusive scan of the
// digit-major count table, and a scatter pass whose in-tile ranks are stable:
// each 256-key chunk is ranked with wave64 ballots (lanes holding the same digit
// are found by AND-ing eight bit ballots), wave offsets are combined in LDS, and
// chunks are processed in input order.
#include "ph_primitives.h"

namespace ph {
namespace {

constexpr int kItems = 8;                 // items per thread in scans
constexpr int kScanTile = kBlock * kItems;  // 2048
constexpr int kSortChunks = 8;            // 256-key chunks per sort tile
constexpr int kSortTile = kBlock * kSortChunks;
constexpr int kRadix = 256;

struct SumOp {
    template <class T> __device__ T operator()(T a, T b) const { return a + b; }
};
struct MaxOp {
    template <class T> __device__ T operator()(T a, T b) const { return a > b ? a : b; }
};

template <class T, class Op> __device__ T wave_inclusive_scan(T v, Op op) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        T u = __shfl_up(v, d, kWave);
        if (lane >= d) v = op(v, u);
    }
    return v;
}

// Inclusive scan across the 256 threads of a block; returns the block total too.
template <class T, class Op> __device__ T block_inclusive_scan(T v, Op op, T identity, T* lds4, T& total) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    v = wave_inclusive_scan(v, op);
    if (lane == kWave - 1) lds4[w] = v;
    __syncthreads();
    T prefix = identity, tot = identity;
#pragma unroll
    for (int i = 0; i < kBlock / kWave; ++i) {
        T li = lds4[i];
        if (i < w) prefix = op(prefix, li);
        tot = op(tot, li);
    }
    __syncthreads();
    total = tot;
    return op(prefix, v);
}

template <class T, class Op> __global__ void __launch_bounds__(kBlock) k_scan_reduce(const T* data, i64 m, T identity, T* partial) {
    __shared__ T lds[kBlock / kWave];
    const i64 base = (i64)blockIdx.x * kScanTile + (i64)threadIdx.x * kItems;
    Op op;
    T acc = identity;
    for (int k = 0; k < kItems; ++k)
        if (base + k < m) acc = op(acc, data[base + k]);
    T total;
    block_inclusive_scan(acc, op, identity, lds, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Scan `data` in place over [0, m) in tiles of kScanTile, starting from the
// carry partial[blockIdx.x] (exclusive tile prefixes) or from `identity`.
template <class T, class Op>
__global__ void __launch_bounds__(kBlock) k_scan_tiles(T* data, i64 m, T identity, const T* partial, int tiles_per_block,
                                                       bool inclusive) {
    __shared__ T lds[kBlock / kWave];
    __shared__ T shift[kBlock];
    Op op;
    T carry = partial ? partial[blockIdx.x] : identity;
    for (int tt = 0; tt < tiles_per_block; ++tt) {
        const i64 tile = (i64)blockIdx.x * tiles_per_block + tt;
        const i64 base = tile * kScanTile + (i64)threadIdx.x * kItems;
        T v[kItems];
        T acc = identity;
        for (int k = 0; k < kItems; ++k) {
            v[k] = (base + k < m) ? data[base + k] : identity;
            acc = op(acc, v[k]);
        }
        T total;
        const T incl = block_inclusive_scan(acc, op, identity, lds, total);
        shift[threadIdx.x] = incl;  // exclusive thread prefix = inclusive prefix of the previous thread
        __syncthreads();
        T run = op(carry, threadIdx.x ? shift[threadIdx.x - 1] : identity);
        __syncthreads();
        for (int k = 0; k < kItems; ++k) {
            if (base + k < m) {
                if (inclusive) {
                    run = op(run, v[k]);
                    data[base + k] = run;
                } else {
                    data[base + k] = run;
                    run = op(run, v[k]);
                }
            }
        }
        carry = op(carry, total);
    }
}

inline i64 scan_scratch_elems(i64 m) { return (m + kScanTile - 1) / kScanTile + 1; }

// `scratch` holds at least scan_scratch_elems(m) elements of T.
template <class T, class Op> void device_scan(T* data, i64 m, T identity, bool inclusive, T* scratch, hipStream_t s) {
    if (m <= 0) return;
    const i64 ntiles = (m + kScanTile - 1) / kScanTile;
    if (ntiles <= 1) {
        k_scan_tiles<T, Op><<<1, kBlock, 0, s>>>(data, m, identity, nullptr, 1, inclusive);
        PH_HIP(hipGetLastError());
        return;
    }
    k_scan_reduce<T, Op><<<(unsigned)ntiles, kBlock, 0, s>>>(data, m, identity, scratch);
    PH_HIP(hipGetLastError());
    const i64 ptiles = (ntiles + kScanTile - 1) / kScanTile;  // tile totals scanned by one block
    k_scan_tiles<T, Op><<<1, kBlock, 0, s>>>(scratch, ntiles, identity, nullptr, (int)ptiles, false);
    PH_HIP(hipGetLastError());
    k_scan_tiles<T, Op><<<(unsigned)ntiles, kBlock, 0, s>>>(data, m, identity, scratch, 1, inclusive);
    PH_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// min / max of three coordinate arrays
// ---------------------------------------------------------------------------
template <class T> __device__ T wave_min(T v) {
    for (int d = kWave / 2; d > 0; d >>= 1) v = fmin(v, __shfl_xor(v, d, kWave));
    return v;
}
template <class T> __device__ T wave_max(T v) {
    for (int d = kWave / 2; d > 0; d >>= 1) v = fmax(v, __shfl_xor(v, d, kWave));
    return v;
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_minmax3(const T* x, const T* y, const T* z, i64 n, T* partial /*[grid][6]*/) {
    __shared__ T lds[6][kBlock / kWave];
    const T inf = T(INFINITY);
    T a[6] = {inf, inf, inf, -inf, -inf, -inf};
    for (i64 i = (i64)blockIdx.x * kBlock + threadIdx.x; i < n; i += (i64)gridDim.x * kBlock) {
        T v[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; ++d) {
            a[d] = fmin(a[d], v[d]);
            a[3 + d] = fmax(a[3 + d], v[d]);
        }
    }
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    for (int d = 0; d < 3; ++d) {
        T mn = wave_min(a[d]), mx = wave_max(a[3 + d]);
        if (lane == 0) {
            lds[d][w] = mn;
            lds[3 + d][w] = mx;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int d = threadIdx.x;
        T r = lds[d][0];
        for (int i = 1; i < kBlock / kWave; ++i) r = d < 3 ? fmin(r, lds[d][i]) : fmax(r, lds[d][i]);
        partial[(i64)blockIdx.x * 6 + d] = r;
    }
}

// One block: every thread folds a strided share of the partials, then the
// wave and block reductions of k_minmax3.
template <class T> __global__ void __launch_bounds__(kBlock) k_minmax_final(const T* partial, int nparts, T* out6) {
    __shared__ T lds[6][kBlock / kWave];
    const T inf = T(INFINITY);
    T a[6] = {inf, inf, inf, -inf, -inf, -inf};
    for (int i = threadIdx.x; i < nparts; i += kBlock)
        for (int d = 0; d < 3; ++d) {
            a[d] = fmin(a[d], partial[(i64)i * 6 + d]);
            a[3 + d] = fmax(a[3 + d], partial[(i64)i * 6 + 3 + d]);
        }
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    for (int d = 0; d < 3; ++d) {
        T mn = wave_min(a[d]), mx = wave_max(a[3 + d]);
        if (lane == 0) {
            lds[d][w] = mn;
            lds[3 + d][w] = mx;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int d = threadIdx.x;
        T r = lds[d][0];
        for (int i = 1; i < kBlock / kWave; ++i) r = d < 3 ? fmin(r, lds[d][i]) : fmax(r, lds[d][i]);
        out6[d] = r;
    }
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------
template <class K>
__global__ void __launch_bounds__(kBlock) k_radix_hist(const K* keys, i64 n, int shift, u32* counts, int nblocks) {
    __shared__ u32 h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const i64 base = (i64)blockIdx.x * kSortTile;
    for (int c = 0; c < kSortChunks; ++c) {
        const i64 i = base + (i64)c * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(u32)(keys[i] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    counts[(i64)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Scatter of one tile of kSortTile keys.  Each key's rank among the tile's
// keys of its digit is found chunk by chunk (256 keys; stable: chunk, wave and
// lane order); the tile is then sorted by digit in LDS and written out from
// LDS, so consecutive threads write consecutive addresses of one digit's run.
// The positions equal those of a direct stable scatter.
template <class K>
__global__ void __launch_bounds__(kBlock) k_radix_scatter(const K* __restrict__ keys, const u32* __restrict__ vals,
                                                          K* __restrict__ okeys, u32* __restrict__ ovals, i64 n,
                                                          int shift, const u32* __restrict__ offsets, int nblocks) {
    __shared__ u32 wave_cnt[kBlock / kWave][kRadix];
    __shared__ u32 tcnt[kRadix];   // keys of each digit in the chunks so far, then the tile's digit starts
    __shared__ u32 base_of[kRadix];
    __shared__ u32 wsum[kBlock / kWave];
    __shared__ K skey[kSortTile];
    __shared__ u32 sval[kSortTile];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    base_of[tid] = offsets[(i64)tid * nblocks + blockIdx.x];
    tcnt[tid] = 0;
    for (int i = 0; i < kBlock / kWave; ++i) wave_cnt[i][tid] = 0;
    __syncthreads();
    const u64 lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (kWave - lane));
    const i64 tile = (i64)blockIdx.x * kSortTile;
    K key[kSortChunks];
    u32 val[kSortChunks], rnk[kSortChunks];
#pragma unroll
    for (int c = 0; c < kSortChunks; ++c) {
        const i64 i = tile + (i64)c * kBlock + tid;
        const bool active = i < n;
        key[c] = active ? keys[i] : K(0);
        val[c] = active ? vals[i] : 0u;
        const u32 digit = (u32)(key[c] >> shift) & (kRadix - 1);
        u64 match = __ballot(active);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int bit = (digit >> b) & 1;
            const u64 bb = __ballot(bit);
            match &= bit ? bb : ~bb;
        }
        const u32 rank = (u32)__popcll(match & lt_mask);
        if (active && rank == 0) wave_cnt[w][digit] = (u32)__popcll(match);
        __syncthreads();
        {  // per digit: earlier waves' counts, on top of the earlier chunks'
            u32 run = tcnt[tid];
            for (int q = 0; q < kBlock / kWave; ++q) {
                const u32 cq = wave_cnt[q][tid];
                wave_cnt[q][tid] = run;
                run += cq;
            }
            tcnt[tid] = run;
        }
        __syncthreads();
        rnk[c] = active ? wave_cnt[w][digit] + rank : ~0u;
        __syncthreads();
        for (int q = 0; q < kBlock / kWave; ++q) wave_cnt[q][tid] = 0;
        __syncthreads();
    }
    // the tile's digit starts: exclusive scan of the digit totals
    const u32 tot = tcnt[tid];
    u32 incl = tot;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const u32 u = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += u;
    }
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    u32 pre = 0;
    for (int q = 0; q < w; ++q) pre += wsum[q];
    tcnt[tid] = pre + incl - tot;
    __syncthreads();
    // the tile sorted by digit, in LDS
#pragma unroll
    for (int c = 0; c < kSortChunks; ++c) {
        if (rnk[c] != ~0u) {
            const u32 digit = (u32)(key[c] >> shift) & (kRadix - 1);
            const u32 at = tcnt[digit] + rnk[c];
            skey[at] = key[c];
            sval[at] = val[c];
        }
    }
    __syncthreads();
    const i64 m = n - tile < kSortTile ? n - tile : kSortTile;
#pragma unroll
    for (int c = 0; c < kSortChunks; ++c) {
        const int e = c * kBlock + tid;
        if (e < m) {
            const K k = skey[e];
            const u32 digit = (u32)(k >> shift) & (kRadix - 1);
            const u32 pos = base_of[digit] + (u32)e - tcnt[digit];
            okeys[pos] = k;
            ovals[pos] = sval[e];
        }
    }
}

template <class K> void radix_sort_impl(K* keys, u32* vals, i64 n, int bits, DevBuf& tmp, hipStream_t s) {
    if (n <= 1) return;
    const int nblocks = (int)((n + kSortTile - 1) / kSortTile);
    const i64 ncount = (i64)kRadix * nblocks;
    auto align = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t cnt_bytes = align(sizeof(u32) * (size_t)ncount);
    const size_t key_bytes = sizeof(K) * (size_t)n, val_bytes = sizeof(u32) * (size_t)n;
    const size_t scan_bytes = align(sizeof(u32) * (size_t)scan_scratch_elems(ncount));
    tmp.alloc(cnt_bytes + align(key_bytes) + align(val_bytes) + scan_bytes);
    char* p = tmp.as<char>();
    u32* counts = reinterpret_cast<u32*>(p);
    K* k2 = reinterpret_cast<K*>(p + cnt_bytes);
    u32* v2 = reinterpret_cast<u32*>(p + cnt_bytes + align(key_bytes));
    u32* scan_scratch = reinterpret_cast<u32*>(p + cnt_bytes + align(key_bytes) + align(val_bytes));
    K* src_k = keys;
    u32* src_v = vals;
    K* dst_k = k2;
    u32* dst_v = v2;
    int passes = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        k_radix_hist<K><<<nblocks, kBlock, 0, s>>>(src_k, n, shift, counts, nblocks);
        PH_HIP(hipGetLastError());
        device_scan<u32, SumOp>(counts, ncount, 0u, false, scan_scratch, s);
        k_radix_scatter<K><<<nblocks, kBlock, 0, s>>>(src_k, src_v, dst_k, dst_v, n, shift, counts, nblocks);
        PH_HIP(hipGetLastError());
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
        ++passes;
    }
    if (passes & 1) {
        PH_HIP(hipMemcpyAsync(keys, src_k, key_bytes, hipMemcpyDeviceToDevice, s));
        PH_HIP(hipMemcpyAsync(vals, src_v, val_bytes, hipMemcpyDeviceToDevice, s));
    }
}

}  // namespace

void exclusive_sum_u32(u32* data, i64 m, DevBuf& tmp, hipStream_t s) {
    tmp.alloc(sizeof(u32) * (size_t)scan_scratch_elems(m));
    device_scan<u32, SumOp>(data, m, 0u, false, tmp.as<u32>(), s);
}
void inclusive_max_u64(u64* data, i64 m, DevBuf& tmp, hipStream_t s) {
    tmp.alloc(sizeof(u64) * (size_t)scan_scratch_elems(m));
    device_scan<u64, MaxOp>(data, m, 0ull, true, tmp.as<u64>(), s);
}

template <class T> void minmax3(const T* x, const T* y, const T* z, i64 n, T* out6, DevBuf& tmp, hipStream_t s) {
    const int g = grid_for(n, kBlock, 1024);
    tmp.alloc(sizeof(T) * 6 * (size_t)g);
    k_minmax3<T><<<g, kBlock, 0, s>>>(x, y, z, n, tmp.as<T>());
    PH_HIP(hipGetLastError());
    k_minmax_final<T><<<1, kBlock, 0, s>>>(tmp.as<T>(), g, out6);
    PH_HIP(hipGetLastError());
}
template void minmax3<double>(const double*, const double*, const double*, i64, double*, DevBuf&, hipStream_t);
template void minmax3<float>(const float*, const float*, const float*, i64, float*, DevBuf&, hipStream_t);

void radix_sort_pairs(u64* keys, u32* vals, i64 n, int bits, DevBuf& tmp, hipStream_t s) {
    radix_sort_impl<u64>(keys, vals, n, bits, tmp, s);
}
void radix_sort_pairs(u32* keys, u32* vals, i64 n, int bits, DevBuf& tmp, hipStream_t s) {
    radix_sort_impl<u32>(keys, vals, n, bits, tmp, s);
}

}  // namespace ph
