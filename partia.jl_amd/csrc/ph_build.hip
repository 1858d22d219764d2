// ph_build.hip -- LinearBVH! on the GPU.
//
// Pipeline (one HIP stream, all device-resident):
//   1. per-axis min/max                    LBVH/MortonEncoding/toolbox.jl:281-294
//   2. quantise + 3D interleave            toolbox.jl:311-336, 38-55, 113-118
//   3. stable LSD radix sort (code, index) toolbox.jl:20-26 (sortperm, ties by index)
//   4. permute every column                KI/interpolation_setup/InterpolationInput.jl:209-217
//   5. duplicate rewrite as a prefix max   toolbox.jl:27-31:  c'[i] = i + max_{j<=i}(c[j] - j)
//   6. Karras range/split + parents        LBVH/BinaryRadixTree/toolbox.jl:14-117
//   7. adjacent LCP, internal escapes      BinaryRadixTree.jl:111-121, toolbox.jl:119-134
//   8. leaf escapes                        BinaryRadixTree.jl:124-126, toolbox.jl:136-156
//   9. bottom-up AABB / max-h refit        LBVH/LinearBVH/LinearBVH.jl:77-91, toolbox.jl:130-188
//      ("second arrival" combines): nodes inside a workgroup's 2048-leaf chunk
//      count arrivals in LDS; the few nodes spanning chunks are finished by a
//      second pass with agent-scope acq_rel counters; then a visit-count check.
// Tree arrays keep the reference's unified node IDs so they compare with memcmp.
#include <algorithm>

#include "ph_build.h"
#include "ph_primitives.h"

namespace ph {
namespace {

template <class TI> struct CodeTraits;
// expand / expand2: the 3D and 2D bit spreads of toolbox.jl:38-55, 77-92.
template <> struct CodeTraits<u32> {
    static constexpr int axis_bits = 10;
    static constexpr int axis_bits2 = 16;
    static constexpr int sort_bits = 32;
    __device__ static u32 expand2(u32 x) {
        x = (x | (x << 8)) & 0x00FF00FFu;
        x = (x | (x << 4)) & 0x0F0F0F0Fu;
        x = (x | (x << 2)) & 0x33333333u;
        x = (x | (x << 1)) & 0x55555555u;
        return x;
    }
    __device__ static u32 expand(u32 x) {
        x = (x | (x << 16)) & 0x30000ffu;
        x = (x | (x << 8)) & 0x300f00fu;
        x = (x | (x << 4)) & 0x30c30c3u;
        x = (x | (x << 2)) & 0x9249249u;
        return x;
    }
    __device__ static int clz(u32 v) { return v ? __clz((int)v) : 32; }
};
template <> struct CodeTraits<u64> {
    static constexpr int axis_bits = 21;
    static constexpr int axis_bits2 = 32;
    static constexpr int sort_bits = 64;
    __device__ static u64 expand2(u64 x) {
        x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
        x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
        x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
        x = (x | (x << 2)) & 0x3333333333333333ull;
        x = (x | (x << 1)) & 0x5555555555555555ull;
        return x;
    }
    __device__ static u64 expand(u64 x) {
        x = (x | (x << 32)) & 0x1f00000000ffffull;
        x = (x | (x << 16)) & 0x1f0000ff0000ffull;
        x = (x | (x << 8)) & 0x100f00f00f00f00full;
        x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
        x = (x | (x << 2)) & 0x1249249249249249ull;
        return x;
    }
    __device__ static int clz(u64 v) { return v ? __clzll((long long)v) : 64; }
};

__device__ inline int clz64(u64 v) { return v ? __clzll((long long)v) : 64; }

// Quantise one coordinate: floor(scale * ((v - vmin) * inv(vmax - vmin))), 0.5
// when flat; scale = exp2(b) - 1 rounded to T (as Julia's exp2(T(b)) - one(T):
// 2^32 - 1 becomes 2^32 in Float32).
template <class T, class TI, int BITS = CodeTraits<TI>::axis_bits>
__device__ inline TI quantise(T v, T vmin, T vmax) {
    const T scale = T((1ull << BITS) - 1);
    T dv = vmax - vmin;
    T f;
    if (dv == T(0)) {
        f = T(0.5);
    } else {
        T inv = T(1) / dv;
        f = (v - vmin) * inv;
    }
    return (TI)floor(scale * f);
}

template <class T, class TI>
__global__ void __launch_bounds__(kBlock) k_encode(const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z, i64 n,
                                           const T* __restrict__ mm6, TI* __restrict__ codes, u32* __restrict__ idx) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    TI ix = quantise<T, TI>(x[i], mm6[0], mm6[3]);
    TI iy = quantise<T, TI>(y[i], mm6[1], mm6[4]);
    TI iz = quantise<T, TI>(z[i], mm6[2], mm6[5]);
    codes[i] = (CodeTraits<TI>::expand(ix) << 2) | (CodeTraits<TI>::expand(iy) << 1) | CodeTraits<TI>::expand(iz);
    idx[i] = (u32)i;
}

// 2D codes (MortonEncoding(x, y), MortonEncoding.jl:80-86; toolbox.jl:166-168).
template <class T, class TI>
__global__ void __launch_bounds__(kBlock) k_encode2(const T* __restrict__ x, const T* __restrict__ y, i64 n,
                                                    const T* __restrict__ mm6, TI* __restrict__ codes,
                                                    u32* __restrict__ idx) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    constexpr int b = CodeTraits<TI>::axis_bits2;
    TI ix = quantise<T, TI, b>(x[i], mm6[0], mm6[3]);
    TI iy = quantise<T, TI, b>(y[i], mm6[1], mm6[4]);
    codes[i] = (CodeTraits<TI>::expand2(ix) << 1) | CodeTraits<TI>::expand2(iy);
    idx[i] = (u32)i;
}

// Gather every column into Morton order, dst[c][i] = src[c][idx[i]], in two
// passes: the columns are first interleaved into rows of rs >= ncols values
// (rs even: 16-byte aligned rows), in input order, through an LDS tile of
// `rows` rows (coalesced reads of each column, coalesced writes of the rows);
// then each sorted position reads its particle's row (one or two cache lines
// instead of one line per column) and writes the columns coalesced.
inline int interleave_stride(int ncols) { return (ncols + 1) & ~1; }
inline int interleave_rows(int rs) { return std::min(kBlock, 4096 / rs > 0 ? 4096 / rs : 1); }

template <class T>
__global__ void __launch_bounds__(kBlock) k_interleave(const T* const* __restrict__ src, T* __restrict__ aos, int ncols,
                                                       int rs, int rows, i64 n) {
    extern __shared__ double interleave_lds[];
    T* tile = reinterpret_cast<T*>(interleave_lds);
    const i64 i0 = (i64)blockIdx.x * rows;
    const int m = (int)(n - i0 < rows ? n - i0 : rows);
    for (int c = 0; c < rs; ++c)
        for (int r = threadIdx.x; r < m; r += kBlock) tile[r * rs + c] = c < ncols ? src[c][i0 + r] : T(0);
    __syncthreads();
    for (int e = threadIdx.x; e < m * rs; e += kBlock) aos[i0 * rs + e] = tile[e];
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_permute(const T* __restrict__ aos, T* __restrict__ dst, int ncols, int rs,
                                                    i64 n, const u32* __restrict__ idx) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const T* row = aos + (i64)idx[i] * rs;
    if constexpr (sizeof(T) == 8) {  // rows of doubles are 16-byte aligned
        int c = 0;
        for (; c + 1 < ncols; c += 2) {
            const double2 v = *reinterpret_cast<const double2*>(row + c);
            dst[(i64)c * n + i] = v.x;
            dst[(i64)(c + 1) * n + i] = v.y;
        }
        if (c < ncols) dst[(i64)c * n + i] = row[c];
    } else {
        for (int c = 0; c < ncols; ++c) dst[(i64)c * n + i] = row[c];
    }
}

// Dedup as a prefix max: key_i = c_i + 2^31 - i (never negative for i < 2^31).
// Codes from here up (2D UInt64 codes use all 64 bits) may make c + 2^31 - i
// overflow; they are rewritten by k_dedup_tail instead.
constexpr u64 kTailCode = ~0ull - (1ull << 32) + 1;  // 2^64 - 2^32

template <class TI> __global__ void k_dedup_keys(const TI* codes, u64* keys, i64 n, u32* first) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    keys[i] = (u64)codes[i] + (1ull << 31) - (u64)i;
    if (sizeof(TI) == 8 && (u64)codes[i] >= kTailCode) atomicMin(first, (u32)i);
}
// The reference's sequential rewrite (codes[i] = codes[i-1] + 1, wrapping at
// 2^64, when codes[i] <= codes[i-1]) from the first tail code on, seeded with
// the scan's exact value before it; writes keys in k_dedup_apply's form.  One
// thread: the tail holds only the codes within 2^32 of the top.
template <class TI> __global__ void k_dedup_tail(const TI* codes, u64* keys, i64 n, const u32* first) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const i64 f = (i64)*first;
    if (f >= n) return;
    i64 i = f;
    u64 prev;
    if (i == 0) {
        prev = (u64)codes[0];
        keys[0] = prev + (1ull << 31);
        i = 1;
    } else {
        prev = keys[i - 1] - (1ull << 31) + (u64)(i - 1);
    }
    for (; i < n; ++i) {
        const u64 c = (u64)codes[i];
        const u64 v = c <= prev ? prev + 1ull : c;
        keys[i] = v + (1ull << 31) - (u64)i;
        prev = v;
    }
}
template <class TI> __global__ void k_dedup_apply(TI* codes, const u64* keys, u64* codes64, i64 n) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const u64 c = keys[i] - (1ull << 31) + (u64)i;
    codes[i] = (TI)c;
    codes64[i] = c;
}

// delta(i, j) on 1-based indices; -1 outside [1, n].
template <class TI> __device__ inline i64 lcp(const TI* __restrict__ c, i64 n, i64 i, i64 j) {
    if (j < 1 || j > n) return -1;
    const TI a = c[i - 1], b = c[j - 1];
    if (a == b) return clz64((u64)i ^ (u64)j);
    return CodeTraits<TI>::clz(a ^ b);
}

// _build_child! for every internal node i (toolbox.jl:83-117, with _find_range
// :22-53 and _split_position :55-76).
template <class TI>
__global__ void __launch_bounds__(kBlock) k_karras(const TI* __restrict__ c, i64 n, i32* __restrict__ left, i32* __restrict__ right,
                                           i32* __restrict__ parent, i32* __restrict__ range) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x + 1;
    if (i > n - 1) return;
    // _range_direction
    const i64 dl = lcp(c, n, i, i - 1), dr = lcp(c, n, i, i + 1);
    const i64 sd = dr - dl;
    const i64 d = sd >= 0 ? 1 : -1;  // sign, with 0 treated as +1
    const i64 dmin = lcp(c, n, i, i - d);
    i64 lmax = 2;
    while (true) {
        const i64 j = i + lmax * d;
        if (j < 1 || j > n) break;
        if (lcp(c, n, i, j) <= dmin) break;
        lmax *= 2;
    }
    i64 l = 0;
    for (i64 t = lmax >> 1; t > 0; t >>= 1) {
        const i64 j = i + (l + t) * d;
        if (j >= 1 && j <= n && lcp(c, n, i, j) > dmin) l += t;
    }
    const i64 jj = i + l * d;
    const i64 lo = i < jj ? i : jj, hi = i < jj ? jj : i;
    i64 s;
    if (c[lo - 1] == c[hi - 1]) {
        s = (lo + hi) >> 1;
        s = s < lo ? lo : (s > hi - 1 ? hi - 1 : s);
    } else {
        const i64 pfl = lcp(c, n, lo, hi);
        s = lo;
        i64 step = hi - lo;
        while (step > 1) {
            step = (step + 1) >> 1;
            const i64 ns = s + step;
            if (ns < hi && lcp(c, n, lo, ns) > pfl) s = ns;
        }
    }
    const i64 off = n - 1;
    range[2 * (i - 1)] = (i32)lo;
    range[2 * (i - 1) + 1] = (i32)hi;
    i64 idx = (s == lo) ? lo + off : s;
    left[i - 1] = (i32)idx;
    parent[idx - 1] = (i32)i;
    idx = (s + 1 == hi) ? hi + off : s + 1;
    right[i - 1] = (i32)idx;
    parent[idx - 1] = (i32)i;
}

// adj[i] = delta(i, i+1) for i < n, adj[n] = typemax(Int32).
template <class TI> __global__ void k_adjacent(const TI* c, i64 n, i32* adj) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x + 1;
    if (i > n) return;
    adj[i - 1] = (i < n) ? (i32)lcp(c, n, i, i + 1) : 0x7fffffff;
}

__global__ void k_internal_escape(const i32* __restrict__ range, const i32* __restrict__ adj, i64 n, i32* __restrict__ escape) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x + 1;
    if (i > n - 1) return;
    const i64 off = n - 1, hi = range[2 * (i - 1) + 1];
    i32 e;
    if (hi == n) e = 0;
    else if (hi == n - 1) e = (i32)(off + n);
    else {
        const i64 nx = hi + 1;
        e = (adj[nx - 1] < adj[nx - 2]) ? (i32)(off + nx) : (i32)nx;
    }
    escape[i - 1] = e;
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_leaf_escape(const i32* __restrict__ parent, const i32* __restrict__ left, const i32* __restrict__ right,
                                                i64 n, i32* __restrict__ escape, i32* __restrict__ leaf_escape,
                                                NodeRec<T>* __restrict__ nodes) {
    const i64 k = (i64)blockIdx.x * kBlock + threadIdx.x + 1;  // leaf 1..n, also internal node k < n
    if (k > n) return;
    const i64 id = n - 1 + k;
    const i32 p = parent[id - 1];
    i32 e;
    if (p == 0) e = 0;
    else if (left[p - 1] == (i32)id) e = right[p - 1];
    else e = escape[p - 1];
    escape[id - 1] = e;
    leaf_escape[k - 1] = e;
    if (k < n) {
        nodes[k - 1].left = left[k - 1];
        nodes[k - 1].escape = escape[k - 1];
    }
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_leaves(const T* __restrict__ cols, i64 n, LeafRec<T>* __restrict__ leaves) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    LeafRec<T> L;
    L.x = cols[i];
    L.y = cols[n + i];
    L.z = cols[2 * n + i];
    L.h = cols[4 * n + i];
    leaves[i] = L;
}

__global__ void k_fill_u32(u32* p, i64 n, u32 v) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

// Box of unified id u: an internal node's refitted box, or a leaf's point
// (its h as the scale).
template <class T>
__device__ inline void node_or_leaf_box(const LeafRec<T>* __restrict__ leaves, const NodeRec<T>* nodes, i64 n, i32 u,
                                        T mn[3], T mx[3], T& h) {
    if (u >= (i32)n) {
        const LeafRec<T> L = leaves[u - n];
        mn[0] = mx[0] = L.x;
        mn[1] = mx[1] = L.y;
        mn[2] = mx[2] = L.z;
        h = L.h;
    } else {
        const NodeRec<T>* N = &nodes[u - 1];
        for (int d = 0; d < 3; ++d) {
            mn[d] = N->lo[d];
            mx[d] = N->hi[d];
        }
        h = N->scale;
    }
}

constexpr int kRefitChunk = 512;              // leaves (and threads) per workgroup, level 1
constexpr int kChunkStarts = 128;             // listed nodes per chunk at most
constexpr int kSpanChunks = 64;               // chunks per workgroup, level 2
constexpr i64 kRefitSpan = (i64)kRefitChunk * kSpanChunks;
constexpr int kRefitSpanBlock = 1024;         // threads per workgroup, levels 2 and 3

template <class T>
__global__ void __launch_bounds__(kRefitChunk) k_refit_chunk(const LeafRec<T>* __restrict__ leaves,
                                                             const i32* __restrict__ parent,
                                                             const i32* __restrict__ left,
                                                             const i32* __restrict__ right,
                                                             const i32* __restrict__ range, i64 n, NodeRec<T>* nodes,
                                                             u32* visited, i32* starts, u32* nstarts, u32* bad) {
    __shared__ T lf[kRefitChunk][4];  // leaf c0 + j: x, y, z, h
    __shared__ T nb[kRefitChunk][7];  // internal node c0 + j: lo[3], hi[3], scale
    __shared__ i32 tp[kRefitChunk][5];  // internal node c0 + j: parent, left, right, range lo, range hi
    __shared__ i32 lp[kRefitChunk];  // parent of leaf c0 + j
    __shared__ u32 cnt[kRefitChunk];
    __shared__ u32 nlisted;
    const i64 c0 = (i64)blockIdx.x * kRefitChunk + 1;
    const i64 c1 = c0 + kRefitChunk - 1 < n ? c0 + kRefitChunk - 1 : n;
    const int j = threadIdx.x;
    const i64 k = c0 + j;
    if (k <= c1) {
        const LeafRec<T> L = leaves[k - 1];
        lf[j][0] = L.x;
        lf[j][1] = L.y;
        lf[j][2] = L.z;
        lf[j][3] = L.h;
        lp[j] = parent[n - 1 + k - 1];
    }
    if (k <= n - 1) {
        tp[j][0] = parent[k - 1];
        tp[j][1] = left[k - 1];
        tp[j][2] = right[k - 1];
        tp[j][3] = range[2 * (k - 1)];
        tp[j][4] = range[2 * (k - 1) + 1];
    }
    cnt[j] = 0u;
    if (j == 0) nlisted = 0u;
    __syncthreads();
    // the box of unified id u, which lies inside the chunk (a leaf, or an
    // internal node refitted here)
    auto box = [&](i32 u, T mn[3], T mx[3], T& h) {
        if (u >= (i32)n) {
            const int q = (int)((i64)u - n + 1 - c0);
            for (int d = 0; d < 3; ++d) mn[d] = mx[d] = lf[q][d];
            h = lf[q][3];
        } else {
            const int q = (int)((i64)u - c0);
            for (int d = 0; d < 3; ++d) {
                mn[d] = nb[q][d];
                mx[d] = nb[q][3 + d];
            }
            h = nb[q][6];
        }
    };
    if (k <= c1) {
        i32 child = (i32)(n - 1 + k);
        i32 p = lp[j];
        for (i64 depth = n; p != 0 && depth > 0; --depth) {
            const i64 q = (i64)p - c0;
            const bool inside = q >= 0 && q < kRefitChunk && (i64)p <= n - 1 && tp[q][3] >= c0 && tp[q][4] <= c1;
            if (!inside) {  // p spans chunks: level 2 climbs from `child`
                const u32 slot = atomicAdd(&nlisted, 1u);
                if (slot < (u32)kChunkStarts) starts[(i64)blockIdx.x * kChunkStarts + slot] = child;
                else atomicAdd(bad, 1u);
                break;
            }
            // the children's boxes are read from LDS: the arrival needs only
            // LDS ordering (the global node stores are read by later kernels)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            const u32 old = __hip_atomic_fetch_add(&cnt[q], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
            if (old != 1u) break;  // first arrival
            T lmin[3], lmax[3], rmin[3], rmax[3], hl, hr;
            box(tp[q][1], lmin, lmax, hl);
            box(tp[q][2], rmin, rmax, hr);
            NodeRec<T>* P = &nodes[p - 1];
            for (int d = 0; d < 3; ++d) {
                const T lo = (lmin[d] < rmin[d]) ? lmin[d] : rmin[d];
                const T hi = (lmax[d] > rmax[d]) ? lmax[d] : rmax[d];
                nb[q][d] = lo;
                nb[q][3 + d] = hi;
                P->lo[d] = lo;
                P->hi[d] = hi;
            }
            const T sc = (hl > hr) ? hl : hr;
            nb[q][6] = sc;
            P->scale = sc;
            visited[p - 1] = 2u;
            child = p;
            p = tp[q][0];
        }
    }
    __syncthreads();
    if (j == 0) nstarts[blockIdx.x] = nlisted < (u32)kChunkStarts ? nlisted : (u32)kChunkStarts;
}

// Levels 2 and 3.  Level 2 (out != null): workgroup b owns the leaves
// [b kRefitSpan + 1, (b + 1) kRefitSpan] and climbs from the nodes listed by
// its chunks (starts / nstarts as k_refit_chunk writes them, nchunks chunks in
// all) through the nodes whose range lies there.  Level 3 (out == null, one
// workgroup): climbs from the nin[0] nodes of `in` to the root.
template <class T>
__global__ void __launch_bounds__(kRefitSpanBlock) k_refit_span(const LeafRec<T>* __restrict__ leaves,
                                                                const i32* __restrict__ parent,
                                                                const i32* __restrict__ left,
                                                                const i32* __restrict__ right,
                                                                const i32* __restrict__ range, i64 n, i64 nchunks,
                                                                NodeRec<T>* nodes, u32* visited, const i32* in,
                                                                const u32* nin, i32* out, u32* nout) {
    const i64 C0 = (i64)blockIdx.x * kRefitSpan + 1, C1 = C0 + kRefitSpan - 1;
    const i64 chunk0 = (i64)blockIdx.x * kSpanChunks;
    // level 2: the span's listed nodes numbered consecutively (exclusive
    // prefix of its chunks' list lengths), so that each thread climbs from at
    // most a few of them
    __shared__ u32 pre[kSpanChunks + 1];
    if (out && threadIdx.x < kWave) {
        const i64 ch = chunk0 + threadIdx.x;
        u32 v = threadIdx.x < kSpanChunks && ch < nchunks ? nin[ch] : 0u;
        for (int d = 1; d < kWave; d <<= 1) {  // inclusive scan over the wave
            const u32 o = __shfl_up(v, d, kWave);
            if ((int)threadIdx.x >= d) v += o;
        }
        if (threadIdx.x < kSpanChunks) pre[threadIdx.x + 1] = v;
        if (threadIdx.x == 0) pre[0] = 0u;
    }
    __syncthreads();
    const u32 total = out ? pre[kSpanChunks] : *nin;
    for (u32 t = threadIdx.x; t < total; t += kRefitSpanBlock) {
        i32 child;
        if (out) {  // the chunk holding entry t: the last c with pre[c] <= t
            int lo = 0, hi = kSpanChunks - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pre[mid] <= t) lo = mid;
                else hi = mid - 1;
            }
            child = in[(chunk0 + lo) * kChunkStarts + (t - pre[lo])];
        } else {
            child = in[t];
        }
        // the climbing thread carries its node's box; per level it loads the
        // parent's topology in one round, takes the arrival, then reads the
        // sibling's box (three dependent rounds)
        T mn[3], mx[3], h;
        node_or_leaf_box<T>(leaves, nodes, n, child, mn, mx, h);
        i32 p = parent[child - 1];
        for (i64 depth = n; p != 0 && depth > 0; --depth) {
            const i32 plo = range[2 * (p - 1)], phi = range[2 * (p - 1) + 1];
            const i32 pl = left[p - 1], pr = right[p - 1], pp = parent[p - 1];
            if (out && (plo < C0 || phi > C1)) {
                const u32 slot = atomicAdd(nout, 1u);
                out[slot] = child;
                break;
            }
            const u32 old = __hip_atomic_fetch_add(&visited[p - 1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old != 1u) break;
            T smn[3], smx[3], sh;
            node_or_leaf_box<T>(leaves, nodes, n, pl == child ? pr : pl, smn, smx, sh);
            NodeRec<T>* P = &nodes[p - 1];
            for (int d = 0; d < 3; ++d) {
                // the reference's order of the operands: left child first
                const T lmin = pl == child ? mn[d] : smn[d], rmin = pl == child ? smn[d] : mn[d];
                const T lmax = pl == child ? mx[d] : smx[d], rmax = pl == child ? smx[d] : mx[d];
                mn[d] = (lmin < rmin) ? lmin : rmin;
                mx[d] = (lmax > rmax) ? lmax : rmax;
                P->lo[d] = mn[d];
                P->hi[d] = mx[d];
            }
            const T hl = pl == child ? h : sh, hr = pl == child ? sh : h;
            h = (hl > hr) ? hl : hr;
            P->scale = h;
            child = p;
            p = pp;
        }
    }
}

__global__ void k_check_visited(const u32* visited, i64 m, u32* bad) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (i < m && visited[i] < 2u) atomicAdd(bad, 1u);
}

__global__ void k_d2f(const double* __restrict__ s, float* __restrict__ d, i64 m) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (i < m) d[i] = (float)s[i];
}

template <class T, class TI>
void build_typed(ph_lbvh* B, const T* const* src_cols /* device, 6+ncol */, hipStream_t s, BuildScratch& W,
                 u32** deferred_bad) {
    const i64 n = B->n;
    const int ncols = 6 + B->ncol;
    const int g = grid_for(n);
    auto align = [](size_t b) { return (b + 255) & ~size_t(255); };

    // every buffer first, then the kernels: allocations come from the stream-
    // ordered pool and wait for the stream, so none sits between two launches
    const i64 tot = 2 * n - 1;
    W.ptrs.alloc(sizeof(void*) * ncols);
    W.mm.alloc(sizeof(T) * 8);
    W.keys.alloc(align(sizeof(TI) * n));
    W.dk.alloc(sizeof(u64) * n);
    W.flag.alloc(sizeof(u32) * 64);
    W.adj.alloc(sizeof(i32) * n);
    // refit: arrival counters and check count, the chunks' listed nodes, the
    // level-3 list and the chunks' list lengths
    W.visited.alloc(sizeof(u32) * (n - 1) + sizeof(u32) * 64 +
                    sizeof(i32) * ((n + kRefitChunk - 1) / kRefitChunk * kChunkStarts + n) +
                    sizeof(u32) * ((n + kRefitChunk - 1) / kRefitChunk));
    W.tmp.alloc(align(16 * (size_t)n) + (size_t(1) << 20));  // radix sort, scans, min/max
    const int rs = interleave_stride(ncols), rows = interleave_rows(rs);
    W.aos.alloc(sizeof(T) * (size_t)rs * n);  // the interleaved columns
    B->order.alloc(sizeof(u32) * n);
    B->cols.alloc(sizeof(T) * (size_t)ncols * n);
    B->codes.alloc(sizeof(u64) * n);
    for (DevBuf* b : {&B->left, &B->right, &B->escape, &B->parent}) b->alloc(sizeof(i32) * tot);
    B->leaf_escape.alloc(sizeof(i32) * n);
    B->leaves.alloc(sizeof(LeafRec<T>) * n);
    B->nodes.alloc(sizeof(NodeRec<T>) * (size_t)(n > 1 ? n - 1 : 1));
    B->range.alloc(sizeof(i32) * 2 * (size_t)(n > 1 ? n - 1 : 1));

    // column pointer table on the device
    PH_HIP(hipMemcpyAsync(W.ptrs.p, src_cols, sizeof(void*) * ncols, hipMemcpyHostToDevice, s));

    // 1-2: min/max, encode
    minmax3<T>(src_cols[0], src_cols[1], src_cols[2], n, W.mm.as<T>(), W.tmp, s);
    if (B->dims == 2)
        k_encode2<T, TI><<<g, kBlock, 0, s>>>(src_cols[0], src_cols[1], n, W.mm.as<T>(), W.keys.as<TI>(),
                                              B->order.as<u32>());
    else
        k_encode<T, TI><<<g, kBlock, 0, s>>>(src_cols[0], src_cols[1], src_cols[2], n, W.mm.as<T>(),
                                             W.keys.as<TI>(), B->order.as<u32>());
    PH_HIP(hipGetLastError());
    k_interleave<T><<<(unsigned)((n + rows - 1) / rows), kBlock, sizeof(T) * (size_t)rows * rs, s>>>(
        W.ptrs.as<const T*>(), W.aos.as<T>(), ncols, rs, rows, n);
    PH_HIP(hipGetLastError());
    // 3: stable sort by code
    radix_sort_pairs(W.keys.as<TI>(), B->order.as<u32>(), n, CodeTraits<TI>::sort_bits, W.tmp, s);
    // 4: permute the columns
    k_permute<T><<<g, kBlock, 0, s>>>(W.aos.as<T>(), B->cols.as<T>(), ncols, rs, n, B->order.as<u32>());
    PH_HIP(hipGetLastError());
    // 5: duplicate rewrite
    u32* first = W.flag.as<u32>();
    k_fill_u32<<<1, kBlock, 0, s>>>(first, 1, 0xffffffffu);
    PH_HIP(hipGetLastError());
    k_dedup_keys<TI><<<g, kBlock, 0, s>>>(W.keys.as<TI>(), W.dk.as<u64>(), n, first);
    PH_HIP(hipGetLastError());
    inclusive_max_u64(W.dk.as<u64>(), n, W.tmp, s);
    k_dedup_tail<TI><<<1, 64, 0, s>>>(W.keys.as<TI>(), W.dk.as<u64>(), n, first);
    PH_HIP(hipGetLastError());
    k_dedup_apply<TI><<<g, kBlock, 0, s>>>(W.keys.as<TI>(), W.dk.as<u64>(), B->codes.as<u64>(), n);
    PH_HIP(hipGetLastError());

    // 6-8: topology
    for (DevBuf* b : {&B->left, &B->right, &B->escape, &B->parent})
        PH_HIP(hipMemsetAsync(b->p, 0, sizeof(i32) * tot, s));
    B->root = n >= 2 ? 1 : 0;
    k_leaves<T><<<g, kBlock, 0, s>>>(B->cols.as<T>(), n, B->leaves.as<LeafRec<T>>());
    PH_HIP(hipGetLastError());
    if (n >= 2) {
        const int gi = grid_for(n - 1);
        k_karras<TI><<<gi, kBlock, 0, s>>>(W.keys.as<TI>(), n, B->left.as<i32>(), B->right.as<i32>(),
                                           B->parent.as<i32>(), B->range.as<i32>());
        PH_HIP(hipGetLastError());
        k_adjacent<TI><<<g, kBlock, 0, s>>>(W.keys.as<TI>(), n, W.adj.as<i32>());
        PH_HIP(hipGetLastError());
        k_internal_escape<<<gi, kBlock, 0, s>>>(B->range.as<i32>(), W.adj.as<i32>(), n, B->escape.as<i32>());
        PH_HIP(hipGetLastError());
    }
    k_leaf_escape<T><<<g, kBlock, 0, s>>>(B->parent.as<i32>(), B->left.as<i32>(), B->right.as<i32>(), n,
                                          B->escape.as<i32>(), B->leaf_escape.as<i32>(), B->nodes.as<NodeRec<T>>());
    PH_HIP(hipGetLastError());
    // 9: refit (three levels: chunks in LDS, spans, the top)
    if (n >= 2) {
        const i64 nchunks = (n + kRefitChunk - 1) / kRefitChunk;
        u32* visited = W.visited.as<u32>();
        u32* bad = visited + (n - 1);
        u32* nstarts2 = bad + 1;  // level-3 list length
        i32* starts = reinterpret_cast<i32*>(visited + (n - 1) + 64);  // nchunks segments of kChunkStarts
        i32* starts2 = starts + nchunks * kChunkStarts;                  // level-3 list
        u32* nstarts = reinterpret_cast<u32*>(starts2 + n);             // per chunk
        k_fill_u32<<<grid_for(n - 1 + 2), kBlock, 0, s>>>(visited, n - 1 + 2, 0u);
        PH_HIP(hipGetLastError());
        k_refit_chunk<T><<<(unsigned)nchunks, kRefitChunk, 0, s>>>(B->leaves.as<LeafRec<T>>(), B->parent.as<i32>(),
                                                                  B->left.as<i32>(), B->right.as<i32>(),
                                                                  B->range.as<i32>(), n, B->nodes.as<NodeRec<T>>(),
                                                                  visited, starts, nstarts, bad);
        PH_HIP(hipGetLastError());
        const unsigned nspans = (unsigned)((n + kRefitSpan - 1) / kRefitSpan);
        k_refit_span<T><<<nspans, kRefitSpanBlock, 0, s>>>(B->leaves.as<LeafRec<T>>(), B->parent.as<i32>(),
                                                          B->left.as<i32>(), B->right.as<i32>(), B->range.as<i32>(),
                                                          n, nchunks, B->nodes.as<NodeRec<T>>(), visited, starts,
                                                          nstarts, starts2, nstarts2);
        PH_HIP(hipGetLastError());
        k_refit_span<T><<<1, kRefitSpanBlock, 0, s>>>(B->leaves.as<LeafRec<T>>(), B->parent.as<i32>(),
                                                     B->left.as<i32>(), B->right.as<i32>(), B->range.as<i32>(), n,
                                                     nchunks, B->nodes.as<NodeRec<T>>(), visited, starts2, nstarts2,
                                                     nullptr, nullptr);
        PH_HIP(hipGetLastError());
        k_check_visited<<<grid_for(n - 1), kBlock, 0, s>>>(visited, n - 1, bad);
        PH_HIP(hipGetLastError());
        if (deferred_bad) {  // the caller reads the count back with its own results
            *deferred_bad = bad;
            return;
        }
        u32 hbad = 0;
        PH_HIP(hipMemcpyAsync(&hbad, bad, sizeof(u32), hipMemcpyDeviceToHost, s));
        PH_HIP(hipStreamSynchronize(s));
        if (hbad != 0)
            throw Error{PH_ERR_INTERNAL,
                        "LinearBVH refit: " + std::to_string(hbad) + " internal nodes were not reached twice"};
    }
}

}  // namespace

void build_lbvh(ph_lbvh* B, const double* const* src64, bool host_src, hipStream_t s, BuildScratch& W,
                u32** deferred_bad) {
    const i64 n = B->n;
    const int ncols = 6 + B->ncol;
    std::vector<const double*> dptr(ncols);
    // a 2D input has no z column: the tree's z is a column of zeros
    const bool zero_z = B->dims == 2;
    if (host_src) {
        // stage host columns on the device (float64, original order)
        W.stage.alloc(sizeof(double) * (size_t)ncols * n);
        for (int c = 0; c < ncols; ++c) {
            double* d = W.stage.as<double>() + (size_t)c * n;
            if (c == 2 && zero_z) PH_HIP(hipMemsetAsync(d, 0, sizeof(double) * n, s));
            else PH_HIP(hipMemcpyAsync(d, src64[c], sizeof(double) * n, hipMemcpyHostToDevice, s));
            dptr[c] = d;
        }
    } else {
        for (int c = 0; c < ncols; ++c) dptr[c] = src64[c];
        if (zero_z) {
            W.stage.alloc(sizeof(double) * n);
            PH_HIP(hipMemsetAsync(W.stage.p, 0, sizeof(double) * n, s));
            dptr[2] = W.stage.as<double>();
        }
    }
    if (B->precision == PH_PREC_FP32) {
        W.stage32.alloc(sizeof(float) * (size_t)ncols * n);
        std::vector<const float*> fptr(ncols);
        for (int c = 0; c < ncols; ++c) {
            float* d = W.stage32.as<float>() + (size_t)c * n;
            k_d2f<<<grid_for(n), kBlock, 0, s>>>(dptr[c], d, n);
            PH_HIP(hipGetLastError());
            fptr[c] = d;
        }
        if (B->code_bits == 30) build_typed<float, u32>(B, fptr.data(), s, W, deferred_bad);
        else build_typed<float, u64>(B, fptr.data(), s, W, deferred_bad);
    } else {
        if (B->code_bits == 30) build_typed<double, u32>(B, dptr.data(), s, W, deferred_bad);
        else build_typed<double, u64>(B, dptr.data(), s, W, deferred_bad);
    }
}

}  // namespace ph
