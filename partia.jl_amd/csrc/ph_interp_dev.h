// ph_interp_dev.h -- device-side building blocks shared by the interpolation kernels:
// tree views, the geometry of the traversal tests, the strategy radius rule, the
// float64 accumulator of the reference's general interpolation kernel and the
// sample generators.  Included by the .hip sources only.
#pragma once

#include <type_traits>

#include "ph_interp.h"
#include "ph_sph.h"

namespace ph {
namespace {

// The lanes of the wave where b holds.  HIP's __ballot takes an int, and the
// bool -> int -> bool round trip costs two vector instructions per vote.
__device__ __attribute__((always_inline)) inline u64 wave_ballot(bool b) {
    return __builtin_amdgcn_ballot_w64(b);
}
// Comparisons straight into lane masks (one v_cmp each, no per-lane bool in
// between; LLVM's predicate codes) and a lane's bit of a mask as a bool.
// Lanes outside exec compare false.
enum : int { kCmpOGT = 2, kCmpOGE = 3, kCmpOLT = 4, kCmpOLE = 5, kCmpUGT = 10, kCmpUNE = 14 };
template <int P> __device__ __attribute__((always_inline)) inline u64 lanes_cmp(float a, float b) {
    return __builtin_amdgcn_fcmpf(a, b, P);
}
template <int P> __device__ __attribute__((always_inline)) inline u64 lanes_cmp(double a, double b) {
    return __builtin_amdgcn_fcmp(a, b, P);
}
__device__ __attribute__((always_inline)) inline bool lane_in(u64 m) {
    return __builtin_amdgcn_inverse_ballot_w64(m);
}

template <class T> struct TreeView {
    i64 n;
    i32 root;
    const NodeRec<T>* __restrict__ nodes;
    const LeafRec<T>* __restrict__ leaves;
    const i32* __restrict__ leaf_escape;
};

template <class T> TreeView<T> view_of(const ph_lbvh* B) {
    return TreeView<T>{B->n, B->root, B->nodes.as<NodeRec<T>>(), B->leaves.as<LeafRec<T>>(),
                       B->leaf_escape.as<i32>()};
}

// XCD-aware workgroup index: hardware dispatches workgroup b to XCD b % 8.
//   mode 0: hardware order (logical block b);
//   mode 1: each XCD takes one contiguous range of logical blocks, so
//           neighbouring tiles share an L2 (uneven work per range is not
//           balanced);
//   mode k in 2..7: each XCD takes runs of 2^(k+2) consecutive logical blocks,
//           run j going to XCD j % 8: the blocks one XCD holds at a time are
//           neighbours, and dense regions spread over every XCD.  Blocks past
//           the last whole round of 8 runs keep the hardware order.
// Every mode is a bijection of [0, gridDim.x).
__device__ inline i64 xcd_block(int mode = 2) {
    const i64 nb = gridDim.x, b = blockIdx.x;
    if (mode == 0) return b;
    const i64 x = b % kNumXcd, pos = b / kNumXcd;
    if (mode == 1) {
        const i64 q = nb / kNumXcd, r = nb % kNumXcd;
        return x * q + (x < r ? x : r) + pos;
    }
    const i64 run = i64(1) << (mode + 2), span = run * kNumXcd;
    if (b >= nb / span * span) return b;
    return ((pos / run) * kNumXcd + x) * run + pos % run;
}

// ---- geometry (LinearBVHTraversalMacros.jl:800-953) ----
template <class T> __device__ inline T d2_point(T px, T py, T pz, const LeafRec<T>& L) {
    T dx = px - L.x, dy = py - L.y, dz = pz - L.z;
    return dx * dx + dy * dy + dz * dz;
}
template <class T> __device__ inline T d2_aabb(T px, T py, T pz, const NodeRec<T>& N) {
    const T p[3] = {px, py, pz};
    T s = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T a = N.lo[d], b = N.hi[d];
        const T t = p[d] < a ? (a - p[d]) : (p[d] > b ? (p[d] - b) : T(0));
        s += t * t;
    }
    return s;
}
template <class T> struct Line {
    T o[3], m[3];
};
template <class T> __device__ inline T d2_line_leaf(const Line<T>& l, const LeafRec<T>& L) {
    const T c[3] = {L.x, L.y, L.z};
    T D2 = T(0), Dm = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T D = c[d] - l.o[d];
        D2 += D * D;
        Dm += D * l.m[d];
    }
    return D2 - Dm * Dm;
}
// The terms D2 = |D|^2 and Dm = D.m of the line distance d2 = D2 - Dm^2 with
// D = c - o, in the reference's order.  AX = 0, 1, 2: m is +-e_AX and every
// component of D is finite; then D_d m_d = +-0 for d != AX, those zeros leave
// the sum unchanged (up to the sign of a zero, which Dm^2 drops), and
// Dm = +-D_AX gives the same Dm^2 with two products and two sums fewer.
template <int AX, class T> __device__ inline T d2_line_point(const Line<T>& l, const T (&c)[3]) {
    T D[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) D[d] = c[d] - l.o[d];
    T D2 = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) D2 += D[d] * D[d];
    T Dm;
    if constexpr (AX >= 0) {
        Dm = D[AX];
    } else {
        Dm = T(0);
#pragma unroll
        for (int d = 0; d < 3; ++d) Dm += D[d] * l.m[d];
    }
    return D2 - Dm * Dm;
}
template <int AX, class T> __device__ inline T d2_line_leaf_ax(const Line<T>& l, const LeafRec<T>& L) {
    const T c[3] = {L.x, L.y, L.z};
    return d2_line_point<AX>(l, c);
}
template <class T> __device__ inline T d2_line_aabb_lb(const Line<T>& l, const NodeRec<T>& N) {
    // _line_intersects_aabb (:878-916)
    T tmin = -T(INFINITY), tmax = T(INFINITY);
    bool hit = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T a = N.lo[d], b = N.hi[d], o = l.o[d], m = l.m[d];
        if (m == T(0)) {
            if (o < a || o > b) hit = false;
        } else {
            // x / 1 == x and x / -1 == -x exactly (signed zeros included): axis-
            // aligned rays skip the two divisions with identical results
            T t0, t1;
            if (m == T(1)) {
                t0 = a - o;
                t1 = b - o;
            } else if (m == T(-1)) {
                t0 = -(a - o);
                t1 = -(b - o);
            } else {
                t0 = (a - o) / m;
                t1 = (b - o) / m;
            }
            if (t0 > t1) {
                const T tt = t0;
                t0 = t1;
                t1 = tt;
            }
            tmin = fmax(tmin, t0);
            tmax = fmin(tmax, t1);
            if (tmin > tmax) hit = false;
        }
    }
    if (hit) return T(0);
    // bounding-sphere lower bound (:918-953)
    const T half = T(1) / T(2);
    T c[3], r2 = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] = (N.lo[d] + N.hi[d]) * half;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T hh = (N.hi[d] - N.lo[d]) * half;
        r2 += hh * hh;
    }
    T D2 = T(0), Dm = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T D = c[d] - l.o[d];
        D2 += D * D;
        Dm += D * l.m[d];
    }
    const T dc2 = D2 - Dm * Dm;
    if (dc2 <= r2) return T(0);
    const T dl = sqrt(dc2) - sqrt(r2);
    return dl * dl;
}

// Record of the per-wave line walks: the LBVH in DFS preorder, internal nodes
// and leaves alike, so that a node's left child is the next record and a
// pruned subtree is skipped with one link.  An internal node holds its box,
// the node-only terms of d2_line_aabb_lb's bounding-sphere bound (centre c and
// sr = sqrt(r2), r2 the squared half diagonal) and the scatter radius rad2 =
// (K scale)^2, each computed with the operations d2_line_aabb_lb and radius2
// use, so the walk's pruning decisions are the per-ray DFS's, bit for bit; and
// thresholds on the squared centre distance dc2 that decide most lanes without
// the square root (line_node_prune): keep2 = max(r2, thr_lo) (NaN operands
// ignored) and thr_hi.  A leaf holds its position in lo, its h in hi[0] and
// its scatter radius (K h)^2 (radius2) in hi[1].  link: the record after the
// subtree (a leaf: the next
// one); leaf: the leaf's Morton index, or -1.  64 bytes for float, 128 for
// double; built once per tree (k_line_recs).
template <class T> struct alignas(16) LineRec {
    T lo[3], hi[3];
    T c[3];
    T keep2, sr, rad2;
    T thr_lo, thr_hi;
    i32 link, leaf;
};
// The per-ray test prunes iff the line misses the box (slab test), dc2 > r2
// and p = RN(RN(RN(sqrt(dc2)) - sr)^2) > rad2.  With rr = sqrt(rad2) and the
// margin e (2^-18 for float, 2^-40 for double; every rounding above is 2^-24
// resp. 2^-53 relative):
//   dc2 >= thr_hi = (sr + rr (1 + e))^2 (1 + e): sqrt(dc2) - sr >= rr (1 + e/2),
//     so p >= rad2 (1 + e/2) > rad2 (the last term is pruned);
//   r2 < dc2 <= thr_lo = (sr + rr (1 - e))^2 (1 - e): 0 <= RN(sqrt(dc2)) - sr
//     <= rr (1 - e/2), so p < rad2 (not pruned).
// Both are formed in float64 and rounded to T (a relative 2^-24 at most, inside
// the margin).  They are NaN (no fast decision) unless r2, sr and both
// thresholds are finite and rad2 is a positive normal number.
template <class T> __device__ inline LineRec<T> line_rec_node(const NodeRec<T>& N, T Kvalid, i32 link) {
    LineRec<T> R;
    const T half = T(1) / T(2);
    T r2 = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        R.lo[d] = N.lo[d];
        R.hi[d] = N.hi[d];
        R.c[d] = (N.lo[d] + N.hi[d]) * half;
        const T hh = (N.hi[d] - N.lo[d]) * half;
        r2 += hh * hh;
    }
    R.sr = sqrt(r2);
    const T r = Kvalid * N.scale;
    R.rad2 = r * r;
    R.link = link;
    R.leaf = -1;
    const double e = std::is_same<T, float>::value ? 0x1p-18 : 0x1p-40;
    const double rr = sqrt((double)R.rad2);
    const double a = (double)R.sr + rr * (1.0 - e), b = (double)R.sr + rr * (1.0 + e);
    const T lo = T(a * a * (1.0 - e)), hi = T(b * b * (1.0 + e));
    const T tmin = std::is_same<T, float>::value ? T(__FLT_MIN__) : T(__DBL_MIN__);
    const bool ok = isfinite(r2) && isfinite(R.sr) && R.rad2 >= tmin && isfinite(R.rad2) && isfinite(lo) &&
                    isfinite(hi);
    R.thr_lo = ok ? lo : T(NAN);
    R.thr_hi = ok ? hi : T(NAN);
    // dc2 <= r2 || dc2 <= thr_lo  <=>  dc2 <= fmax(r2, thr_lo), NaNs included
    // (fmax returns the other operand of a NaN; two NaNs give NaN: never kept)
    R.keep2 = fmax(r2, R.thr_lo);
    return R;
}
template <class T>
__device__ inline LineRec<T> line_rec_leaf(const LeafRec<T>& L, T rad2, i32 k, i32 link) {
    LineRec<T> R{};
    R.lo[0] = L.x;
    R.lo[1] = L.y;
    R.lo[2] = L.z;
    R.hi[0] = L.h;
    R.hi[1] = rad2;
    R.link = link;
    R.leaf = k;
    return R;
}
// The lanes of active where d2_line_aabb_lb(l, N) > (K N.scale)^2, from the
// node's LineRec.  Called by the whole wave.  A slab hit or a line within the
// bounding sphere gives the bound 0, which no radius (a square, or NaN) is
// below.  The centre distance comes first: lanes with dc2 <= keep2 keep the
// node whatever the slab test says, and when every active lane does, the slab
// test is skipped; lanes with dc2 >= thr_hi prune unless the slab test hits;
// the rest (NaN, or within the margins) evaluate the reference's square roots.
// The decisions are lane masks: comparisons write them directly and the
// logic runs on the scalar unit.
// AX >= 0: m = +-e_AX and every coordinate finite (d2_line_point); the slab
// test then reduces to the two axes with m_d = 0 (the AX axis's interval
// [a - o, b - o] is never empty for a finite box).
template <int AX = -1, class T>
__device__ inline u64 line_node_prune(const Line<T>& l, const LineRec<T>& N, u64 active) {
    const T dc2 = d2_line_point<AX>(l, N.c);
    const u64 open = active & lanes_cmp<kCmpUGT>(dc2, N.keep2);  // !(dc2 <= keep2)
    if (open == 0) return 0;  // wave-uniform
    u64 miss = 0;
    if constexpr (AX >= 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
            if (d != AX) miss |= lanes_cmp<kCmpOLT>(l.o[d], N.lo[d]) | lanes_cmp<kCmpOGT>(l.o[d], N.hi[d]);
    } else {
        bool hit = true;
        T tmin = -T(INFINITY), tmax = T(INFINITY);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T a = N.lo[d], b = N.hi[d], o = l.o[d], m = l.m[d];
            if (m == T(0)) {
                if (o < a || o > b) hit = false;
            } else {
                T t0, t1;
                if (m == T(1)) {
                    t0 = a - o;
                    t1 = b - o;
                } else if (m == T(-1)) {
                    t0 = -(a - o);
                    t1 = -(b - o);
                } else {
                    t0 = (a - o) / m;
                    t1 = (b - o) / m;
                }
                if (t0 > t1) {
                    const T tt = t0;
                    t0 = t1;
                    t1 = tt;
                }
                tmin = fmax(tmin, t0);
                tmax = fmin(tmax, t1);
                if (tmin > tmax) hit = false;
            }
        }
        miss = wave_ballot(!hit);
    }
    u64 far = lanes_cmp<kCmpOGE>(dc2, N.thr_hi);
    const u64 band = open & ~far;
    if (band != 0) {  // wave-uniform: some lane within the margins
        // an empty asm statement (no instruction) keeps the square root inside
        // the rarely taken branch instead of being speculated above it
        T x = dc2;
        __asm__ volatile("" : "+v"(x));
        bool exact_far = false;
        if (lane_in(band)) {  // the reference's arithmetic
            const T dl = sqrt(x) - N.sr;
            exact_far = dl * dl > N.rad2;
        }
        far |= wave_ballot(exact_far);
    }
    return open & miss & far;
}

// L(u) by pointer jumping: after k_lc_init and r rounds of k_lc_jump, up[u]
// is the ancestor 2^(r+1) edges above internal node u (0 past the root) and
// lc[u] counts the left-child edges on the way there, u's own edge included.
__device__ inline i32 lc_edge(i32 c, i32 root, const i32* __restrict__ parent, const i32* __restrict__ left,
                              i32& up) {
    up = c == root ? 0 : parent[c - 1];
    return (up != 0 && left[up - 1] == c) ? 1 : 0;
}
__global__ void __launch_bounds__(kBlock) k_lc_init(i64 n, i32 root, const i32* __restrict__ parent,
                                            const i32* __restrict__ left, i32* __restrict__ up,
                                            i32* __restrict__ lc) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x + 1;
    if (i > n - 1) return;
    i32 u1 = 0, u2 = 0;
    i32 e = lc_edge((i32)i, root, parent, left, u1);
    if (u1 != 0) e += lc_edge(u1, root, parent, left, u2);
    up[i - 1] = u2;
    lc[i - 1] = e;
}
__global__ void __launch_bounds__(kBlock) k_lc_jump(i64 n, const i32* __restrict__ up_in,
                                            const i32* __restrict__ lc_in, i32* __restrict__ up_out,
                                            i32* __restrict__ lc_out) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x + 1;
    if (i > n - 1) return;
    const i32 u = up_in[i - 1];
    up_out[i - 1] = u == 0 ? 0 : up_in[u - 1];
    lc_out[i - 1] = lc_in[i - 1] + (u == 0 ? 0 : lc_in[u - 1]);
}
constexpr int kLcJumps = 2;  // pointers of 8 edges
// L(u): the left-child edges on the path from the root to internal node u
// (1-based), from the pointers of the last k_lc_jump round.
__device__ inline i64 lc_total(i64 u, i64 n, const i32* __restrict__ up, const i32* __restrict__ lc) {
    i64 Lc = lc[u - 1];
    for (i64 c = up[u - 1], steps = n; c != 0 && steps > 0; --steps) {
        Lc += lc[c - 1];
        c = up[c - 1];
    }
    return Lc;
}

// Radius rule of the strategy (A13): gather (K ha)^2, scatter (K h)^2, symmetric max.
template <class T, int STRAT> __device__ inline T radius2(T Kvalid, T r2a, T h) {
    if constexpr (STRAT == PH_GATHER) {
        return r2a;
    } else {
        const T r = Kvalid * h;
        const T rr = r * r;
        if constexpr (STRAT == PH_SCATTER) return rr;
        else return (rr > r2a) ? rr : r2a;
    }
}

// Stackless DFS: hit(k0, d2, leaf) on each accepted leaf (k0 = 0-based leaf index).
template <class T, int STRAT, class Dist, class NodeDist, class Hit>
__device__ inline void dfs(const TreeView<T>& tv, T Kvalid, T r2a, Dist leaf_d2, NodeDist node_d2, Hit hit) {
    const i64 n = tv.n;
    if (tv.root == 0) {
        for (i64 k = 0; k < n; ++k) {
            const LeafRec<T> L = tv.leaves[k];
            const T d2 = leaf_d2(L);
            if (d2 <= radius2<T, STRAT>(Kvalid, r2a, L.h)) hit(k, d2, L);
        }
        return;
    }
    i32 node = tv.root;
    // A valid tree is walked in at most 2n-1 steps; the cap only guards against
    // a corrupted tree turning into an endless loop on the device.
    for (i64 steps = 2 * n + 8; node != 0 && steps > 0; --steps) {
        if (node >= (i32)n) {
            const i64 k = (i64)node - n;
            const LeafRec<T> L = tv.leaves[k];
            const T d2 = leaf_d2(L);
            if (d2 <= radius2<T, STRAT>(Kvalid, r2a, L.h)) hit(k, d2, L);
            node = tv.leaf_escape[k];
        } else {
            const NodeRec<T> N = tv.nodes[node - 1];
            const T d2n = node_d2(N);
            node = (d2n > radius2<T, STRAT>(Kvalid, r2a, N.scale)) ? N.escape : N.left;
        }
    }
}

// LBVH_find_nearest_h (LinearBVHQueries.jl:118-140).  The reference walks the
// DFS from an infinite bound and keeps the first leaf (in DFS = Morton order) of
// minimal d2, i.e. the lexicographic minimum of (d2, leaf index).  Here a greedy
// descent (nearer child first) supplies a starting leaf and bound, then the
// escape-rope DFS prunes with `d2_node > best` and replaces the candidate when
// d2 < best or (d2 == best and the index is smaller): the same leaf, found after
// visiting far fewer nodes.
template <class T> __device__ inline T nearest_h(const TreeView<T>& tv, T px, T py, T pz) {
    const i64 n = tv.n;
    if (tv.root == 0) return tv.leaves[0].h;  // n == 1
    // greedy descent to a nearby leaf
    i32 node = tv.root;
    for (i64 steps = n; node < (i32)n && steps > 0; --steps) {
        const NodeRec<T> N = tv.nodes[node - 1];
        const i32 l = N.left;
        // the right child is the node after the left subtree: left's escape for a
        // leaf, or the escape of the left internal node
        const i32 r = (l >= (i32)n) ? tv.leaf_escape[l - n] : tv.nodes[l - 1].escape;
        const T dl = (l >= (i32)n) ? d2_point(px, py, pz, tv.leaves[l - n]) : d2_aabb(px, py, pz, tv.nodes[l - 1]);
        const T dr = (r >= (i32)n) ? d2_point(px, py, pz, tv.leaves[r - n]) : d2_aabb(px, py, pz, tv.nodes[r - 1]);
        node = (dr < dl) ? r : l;
    }
    i64 bk = (i64)node - n;
    T best = d2_point(px, py, pz, tv.leaves[bk]);
    node = tv.root;
    for (i64 steps = 2 * n + 8; node != 0 && steps > 0; --steps) {
        if (node >= (i32)n) {
            const i64 k = (i64)node - n;
            const T d2 = d2_point(px, py, pz, tv.leaves[k]);
            if (d2 < best || (d2 == best && k < bk)) {
                best = d2;
                bk = k;
            }
            node = tv.leaf_escape[k];
        } else {
            const NodeRec<T> N = tv.nodes[node - 1];
            node = (d2_aabb(px, py, pz, N) > best) ? N.escape : N.left;
        }
    }
    return tv.leaves[bk].h;
}

// ---- accumulation (general_interpolation.jl + accumulations/*.jl) ----
template <class T, int K, int STRAT, int MS, int MG, int MD, int MC> struct Acc {
    T S1;
    T sc[MS > 0 ? MS : 1];
    T gf[MG > 0 ? MG : 1][3], gb[MG > 0 ? MG : 1][3], gs[MG > 0 ? MG : 1];
    T df[MD > 0 ? MD : 1], db[MD > 0 ? MD : 1][3], ds[MD > 0 ? MD : 1][3];
    T cf[MC > 0 ? MC : 1][3], cb[MC > 0 ? MC : 1][3], cs[MC > 0 ? MC : 1][3];

    __device__ void init() {
        S1 = T(0);
#pragma unroll
        for (int j = 0; j < (MS > 0 ? MS : 1); ++j) sc[j] = T(0);
#pragma unroll
        for (int j = 0; j < (MG > 0 ? MG : 1); ++j) {
            gs[j] = T(0);
            for (int d = 0; d < 3; ++d) gf[j][d] = gb[j][d] = T(0);
        }
#pragma unroll
        for (int j = 0; j < (MD > 0 ? MD : 1); ++j) {
            df[j] = T(0);
            for (int d = 0; d < 3; ++d) db[j][d] = ds[j][d] = T(0);
        }
#pragma unroll
        for (int j = 0; j < (MC > 0 ? MC : 1); ++j)
            for (int d = 0; d < 3; ++d) cf[j][d] = cb[j][d] = cs[j][d] = T(0);
    }

    __device__ inline T Wsel(T dr, T ha, T hb) const {
        if constexpr (STRAT == PH_SYMMETRIC) return T(0.5) * (kernel_W<T, K>(dr, ha) + kernel_W<T, K>(dr, hb));
        else if constexpr (STRAT == PH_GATHER) return kernel_W<T, K>(dr, ha);
        else return kernel_W<T, K>(dr, hb);
    }
    __device__ inline void Gsel(T dx, T dy, T dz, T ha, T hb, T g[3]) const {
        if constexpr (STRAT == PH_SYMMETRIC) {
            T ax, ay, az, bx, by, bz;
            kernel_gradW<T, K>(dx, dy, dz, ha, ax, ay, az);
            kernel_gradW<T, K>(dx, dy, dz, hb, bx, by, bz);
            g[0] = T(0.5) * (ax + bx);
            g[1] = T(0.5) * (ay + by);
            g[2] = T(0.5) * (az + bz);
        } else {
            kernel_gradW<T, K>(dx, dy, dz, STRAT == PH_GATHER ? ha : hb, g[0], g[1], g[2]);
        }
    }

    // One accepted neighbour with packed record r = {m, rho, columns...}.
    __device__ inline void add(const CatChunk& c, const T* __restrict__ r, T dx, T dy, T dz, T d2, T ha, T hb) {
        const T dr = sqrt(d2);
        const T mb = r[0], rb = r[1];
        const T W = Wsel(dr, ha, hb);
        S1 += mb * W / rb;
#pragma unroll
        for (int j = 0; j < MS; ++j)
            if (j < c.ns) sc[j] += r[c.spos[j]] * (mb * W / rb);
        if constexpr (MG + MD + MC > 0) {
            if (c.ng + c.nd + c.nc == 0) return;
            T g[3];
            Gsel(dx, dy, dz, ha, hb, g);
            const T invr = T(1) / rb;
#pragma unroll
            for (int j = 0; j < MG; ++j) {
                if (j >= c.ng) break;
                const int pos = c.gpos[j];
                if (pos < 0) {  // density gradient (gradient_accumulation.jl:7-35)
                    for (int d = 0; d < 3; ++d) {
                        const T mg = mb * g[d];
                        gf[j][d] += mg;
                        gb[j][d] += mg * invr;
                    }
                    gs[j] += mb * W;
                } else {
                    const T A = r[pos];
                    for (int d = 0; d < 3; ++d) {
                        const T w = mb * invr * g[d];
                        gf[j][d] += w * A;
                        gb[j][d] += w;
                    }
                    gs[j] += A * (mb * W / rb);
                }
            }
#pragma unroll
            for (int j = 0; j < MD; ++j) {
                if (j >= c.nd) break;
                T A[3], w[3];
                for (int d = 0; d < 3; ++d) {
                    A[d] = r[c.dpos[3 * j + d]];
                    w[d] = mb * invr * g[d];
                }
                df[j] += w[0] * A[0] + w[1] * A[1] + w[2] * A[2];
                for (int d = 0; d < 3; ++d) {
                    db[j][d] += w[d];
                    ds[j][d] += A[d] * (mb * W / rb);
                }
            }
#pragma unroll
            for (int j = 0; j < MC; ++j) {
                if (j >= c.nc) break;
                T A[3], w[3];
                for (int d = 0; d < 3; ++d) {
                    A[d] = r[c.cpos[3 * j + d]];
                    w[d] = mb * invr * g[d];
                }
                cf[j][0] += A[1] * w[2] - A[2] * w[1];
                cf[j][1] += A[2] * w[0] - A[0] * w[2];
                cf[j][2] += A[0] * w[1] - A[1] * w[0];
                for (int d = 0; d < 3; ++d) {
                    cb[j][d] += w[d];
                    cs[j][d] += A[d] * (mb * W / rb);
                }
            }
        }
    }

    // add() for lanes stepping in lockstep, the record read through r(pos): a
    // lane without a pending neighbour (live = false) passes a record whose
    // mass and columns are 0 and whose density is 1, and its kernel value and
    // gradient are selected to 0 (not multiplied: a dead lane's h_a may be 0),
    // so every term it adds is +0 or -0.  The accumulators start at +0 and a
    // sum that starts at +0 never becomes -0, so adding a signed zero leaves it
    // unchanged: live lanes get add()'s bits, in add()'s order.
    template <class Rec>
    __device__ __attribute__((always_inline)) void add_live(bool live, const CatChunk& c, Rec&& r, T dx, T dy, T dz,
                                                            T d2, T ha, T hb) {
        const T dr = sqrt(d2);
        const T mb = r(0), rb = r(1);
        const T W = live ? Wsel(dr, ha, hb) : T(0);
        S1 += mb * W / rb;
#pragma unroll
        for (int j = 0; j < MS; ++j)
            if (j < c.ns) sc[j] += r(c.spos[j]) * (mb * W / rb);
        if constexpr (MG + MD + MC > 0) {
            if (c.ng + c.nd + c.nc == 0) return;
            T g[3];
            Gsel(dx, dy, dz, ha, hb, g);
#pragma unroll
            for (int d = 0; d < 3; ++d) g[d] = live ? g[d] : T(0);
            const T invr = T(1) / rb;
#pragma unroll
            for (int j = 0; j < MG; ++j) {
                if (j >= c.ng) break;
                const int pos = c.gpos[j];
                if (pos < 0) {  // density gradient (gradient_accumulation.jl:7-35)
                    for (int d = 0; d < 3; ++d) {
                        const T mg = mb * g[d];
                        gf[j][d] += mg;
                        gb[j][d] += mg * invr;
                    }
                    gs[j] += mb * W;
                } else {
                    const T A = r(pos);
                    for (int d = 0; d < 3; ++d) {
                        const T w = mb * invr * g[d];
                        gf[j][d] += w * A;
                        gb[j][d] += w;
                    }
                    gs[j] += A * (mb * W / rb);
                }
            }
#pragma unroll
            for (int j = 0; j < MD; ++j) {
                if (j >= c.nd) break;
                T A[3], w[3];
                for (int d = 0; d < 3; ++d) {
                    A[d] = r(c.dpos[3 * j + d]);
                    w[d] = mb * invr * g[d];
                }
                df[j] += w[0] * A[0] + w[1] * A[1] + w[2] * A[2];
                for (int d = 0; d < 3; ++d) {
                    db[j][d] += w[d];
                    ds[j][d] += A[d] * (mb * W / rb);
                }
            }
#pragma unroll
            for (int j = 0; j < MC; ++j) {
                if (j >= c.nc) break;
                T A[3], w[3];
                for (int d = 0; d < 3; ++d) {
                    A[d] = r(c.cpos[3 * j + d]);
                    w[d] = mb * invr * g[d];
                }
                cf[j][0] += A[1] * w[2] - A[2] * w[1];
                cf[j][1] += A[2] * w[0] - A[0] * w[2];
                cf[j][2] += A[0] * w[1] - A[1] * w[0];
                for (int d = 0; d < 3; ++d) {
                    cb[j][d] += w[d];
                    cs[j][d] += A[d] * (mb * W / rb);
                }
            }
        }
    }

    // Epilogue (general_interpolation.jl:128-214); out[k*S + s].
    __device__ inline void finish(const CatChunk& c, T* __restrict__ out, i64 S, i64 s) const {
        if (S1 == T(0)) {
            const T nan = sph_nan<T>();
#pragma unroll
            for (int j = 0; j < MS; ++j)
                if (j < c.ns) out[c.sout[j] * S + s] = nan;
#pragma unroll
            for (int j = 0; j < MG; ++j)
                if (j < c.ng)
                    for (int d = 0; d < 3; ++d) out[(c.gout[j] + d) * S + s] = nan;
#pragma unroll
            for (int j = 0; j < MD; ++j)
                if (j < c.nd) out[c.dout[j] * S + s] = nan;
#pragma unroll
            for (int j = 0; j < MC; ++j)
                if (j < c.nc)
                    for (int d = 0; d < 3; ++d) out[(c.cout[j] + d) * S + s] = nan;
            return;
        }
        const T invS1 = T(1) / S1;
#pragma unroll
        for (int j = 0; j < MS; ++j)
            if (j < c.ns) out[c.sout[j] * S + s] = c.sflag[j] ? sc[j] * invS1 : sc[j];
#pragma unroll
        for (int j = 0; j < MG; ++j)
            if (j < c.ng) {
                const T A = gs[j] * invS1;
                for (int d = 0; d < 3; ++d) out[(c.gout[j] + d) * S + s] = gf[j][d] - gb[j][d] * A;
            }
#pragma unroll
        for (int j = 0; j < MD; ++j)
            if (j < c.nd) {
                const T Ax = ds[j][0] * invS1, Ay = ds[j][1] * invS1, Az = ds[j][2] * invS1;
                const T b = Ax * db[j][0] + Ay * db[j][1] + Az * db[j][2];
                out[c.dout[j] * S + s] = df[j] - b;
            }
#pragma unroll
        for (int j = 0; j < MC; ++j)
            if (j < c.nc) {
                const T Ax = cs[j][0] * invS1, Ay = cs[j][1] * invS1, Az = cs[j][2] * invS1;
                const T bx = cb[j][0], by = cb[j][1], bz = cb[j][2];
                const T qx = Ay * bz - Az * by, qy = Az * bx - Ax * bz, qz = Ax * by - Ay * bx;
                out[(c.cout[j] + 0) * S + s] = -(cf[j][0] - qx);
                out[(c.cout[j] + 1) * S + s] = -(cf[j][1] - qy);
                out[(c.cout[j] + 2) * S + s] = -(cf[j][2] - qz);
            }
    }
};

// Sample generation.  SRC 0: list of points; SRC 1: Cartesian grid with 4x4x4 wave tiles.
template <class T, int SRC> struct Sampler;
template <class T> struct Sampler<T, 0> {
    PointList pl;
    i64 S;
    int xcd = 2;  // block order (xcd_block mode)
    __device__ inline bool get(i64& s, T& x, T& y, T& z) const {
        s = xcd_block(xcd) * kBlock + threadIdx.x;
        if (s >= S) return false;
        x = T(pl.x[s]);
        y = T(pl.y[s]);
        z = T(pl.z[s]);
        return true;
    }
};
// Position of tile t in a (tx, ty, tz) tile grid walked in bricks of kBrick^3
// tiles: z slabs of kBrick tile planes, each in y bands of kBrick rows, each in
// bricks of kBrick columns (slabs, bands and bricks at the far edges are
// clipped); Morton order inside a whole brick, x-fastest inside a clipped one.
// A bijection of [0, tx*ty*tz): aligned runs of 8^k tiles are cubes.
constexpr i64 kBrick = 8;
__device__ inline void brick_tile(i64 t, i64 tx, i64 ty, i64 tz, i64& bx, i64& by, i64& bz) {
    const i64 sz = t / (tx * ty * kBrick);  // z slab
    t -= sz * tx * ty * kBrick;
    const i64 hz = min(kBrick, tz - sz * kBrick);
    const i64 sy = t / (tx * kBrick * hz);  // y band
    t -= sy * tx * kBrick * hz;
    const i64 hy = min(kBrick, ty - sy * kBrick);
    const i64 sx = t / (kBrick * hy * hz);  // brick
    t -= sx * kBrick * hy * hz;
    const i64 wx = min(kBrick, tx - sx * kBrick);
    if (wx == kBrick && hy == kBrick && hz == kBrick) {
        bx = sx * kBrick + ((t & 1) | ((t >> 2) & 2) | ((t >> 4) & 4));
        by = sy * kBrick + (((t >> 1) & 1) | ((t >> 3) & 2) | ((t >> 5) & 4));
        bz = sz * kBrick + (((t >> 2) & 1) | ((t >> 4) & 2) | ((t >> 6) & 4));
        return;
    }
    bx = sx * kBrick + t % wx;
    by = sy * kBrick + (t / wx) % hy;
    bz = sz * kBrick + t / (wx * hy);
}

// The same for a (tx, ty) grid of image tiles: y bands of kBrick rows, bricks
// of kBrick columns, Morton order inside a whole brick (x-fastest in a clipped
// one).  A bijection of [0, tx*ty).
__device__ inline void brick_tile2(i64 t, i64 tx, i64 ty, i64& bx, i64& by) {
    const i64 sy = t / (tx * kBrick);  // y band
    t -= sy * tx * kBrick;
    const i64 hy = min(kBrick, ty - sy * kBrick);
    const i64 sx = t / (kBrick * hy);  // brick
    t -= sx * kBrick * hy;
    const i64 wx = min(kBrick, tx - sx * kBrick);
    if (wx == kBrick && hy == kBrick) {
        bx = sx * kBrick + ((t & 1) | ((t >> 1) & 2) | ((t >> 2) & 4));
        by = sy * kBrick + (((t >> 1) & 1) | ((t >> 2) & 2) | ((t >> 3) & 4));
        return;
    }
    bx = sx * kBrick + t % wx;
    by = sy * kBrick + t / wx;
}

template <class T> struct Sampler<T, 1> {
    GridAxes g;
    i64 S;
    int xcd = 2;    // block order (xcd_block mode)
    int order = 0;  // tile order: 0 x-fastest rows, 1 bricks
    __device__ inline bool get(i64& s, T& x, T& y, T& z) const {
        // one 4x4x4 tile of samples per wave (any whole number of waves per block)
        const i64 tile = xcd_block(xcd) * (i64)(blockDim.x / kWave) + threadIdx.x / kWave;
        const int lane = threadIdx.x & (kWave - 1);
        const i64 tx = (g.nx + 3) / 4, ty = (g.ny + 3) / 4, tz = (g.nz + 3) / 4;
        if (tile >= tx * ty * tz) return false;
        i64 bx, by, bz;
        if (order) brick_tile(tile, tx, ty, tz, bx, by, bz);
        else {
            bx = tile % tx;
            by = (tile / tx) % ty;
            bz = tile / (tx * ty);
        }
        const i64 ix = bx * 4 + (lane & 3), iy = by * 4 + ((lane >> 2) & 3), iz = bz * 4 + (lane >> 4);
        if (ix >= g.nx || iy >= g.ny || iz >= g.nz) return false;
        s = ix + g.nx * (iy + g.ny * iz);
        if (g.coords == PH_COORDS_CARTESIAN) {
            x = T(g.ax[ix]);
            y = T(g.ay[iy]);
            z = T(g.az[iz]);
        } else if (g.coords == PH_COORDS_CYLINDRICAL) {  // _cylin2cart (coordinate_transformations.jl:85-89)
            const T sv = T(g.ax[ix]), sp = T(g.ay[2 * iy]), cp = T(g.ay[2 * iy + 1]);
            x = sv * cp;
            y = sv * sp;
            z = T(g.az[iz]);
        } else {  // _sph2cart (coordinate_transformations.jl:198-204)
            const T r = T(g.ax[ix]), sp = T(g.ay[2 * iy]), cp = T(g.ay[2 * iy + 1]), st = T(g.az[2 * iz]),
                    ct = T(g.az[2 * iz + 1]);
            x = (r * st) * cp;
            y = (r * st) * sp;
            z = r * ct;
        }
        return true;
    }
};

// Cross-lane LDS exchange within one wave: writes of every lane are complete
// and visible to the wave's other lanes after this point.
__device__ __attribute__((always_inline)) inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lockstep iterations of a wave: the largest accepted count of its lanes.
__device__ inline int lockstep_count(u64 mask) {
    int nit = __popcll(mask);
    for (int d = kWave / 2; d > 0; d >>= 1) nit = max(nit, __shfl_xor(nit, d, kWave));
    return __builtin_amdgcn_readfirstlane(nit);
}

// Sum over the wave's 64 lanes (every lane gets the total).
template <class T> __device__ inline T wave_sum(T v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Sums of eight per-lane values over the wave, reduce-scatter style: value v
// (0..7) ends up, summed over all 64 lanes, in lanes 8v .. 8v+7 (returned by
// each lane for v = lane >> 3).  10 cross-lane moves instead of 48 for eight
// separate wave sums; the additions are float64 sums in a fixed tree order.
__device__ inline double wave_sum8(const double in[8]) {
    const int lane = threadIdx.x & (kWave - 1);
    double v[8] = {in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7]};
    {  // 8 -> 4 values: lanes with bit 5 set keep values 4..7
        const bool hi = lane & 32;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double send = hi ? v[k] : v[k + 4];
            const double keep = hi ? v[k + 4] : v[k];
            v[k] = keep + __shfl_xor(send, 32, kWave);
        }
    }
    {  // 4 -> 2 (bit 4)
        const bool hi = lane & 16;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const double send = hi ? v[k] : v[k + 2];
            const double keep = hi ? v[k + 2] : v[k];
            v[k] = keep + __shfl_xor(send, 16, kWave);
        }
    }
    {  // 2 -> 1 (bit 3)
        const bool hi = lane & 8;
        const double send = hi ? v[0] : v[1];
        const double keep = hi ? v[1] : v[0];
        v[0] = keep + __shfl_xor(send, 8, kWave);
    }
#pragma unroll
    for (int d = 4; d > 0; d >>= 1) v[0] += __shfl_xor(v[0], d, kWave);
    return v[0];
}

// The same for six values (lanes 48..63 return 0).
__device__ inline double wave_sum6(const double in[6]) {
    const double v[8] = {in[0], in[1], in[2], in[3], in[4], in[5], 0.0, 0.0};
    return wave_sum8(v);
}

__device__ inline void count_add(u64* counter, u64 v) {
    if (!counter) return;
    // wave reduction, one atomic per wave
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && v) atomicAdd((unsigned long long*)counter, (unsigned long long)v);
}

// Packs {m, rho, columns...} of every particle into records of width c.rw.
template <class T>
__global__ void __launch_bounds__(kBlock) k_pack_records(const T* __restrict__ cols, i64 n, CatChunk c, T* __restrict__ rec) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    T* r = rec + i * c.rw;
    for (int j = 0; j < c.rw; ++j) r[j] = j < c.nsrc ? cols[(i64)c.rsrc[j] * n + i] : T(0);
}

// Smoothing length of the nearest particle of every sample (gather / symmetric).
template <class T, int SRC>
__global__ void __launch_bounds__(kBlock) k_nearest_samples(TreeView<T> tv, Sampler<T, SRC> smp, T* __restrict__ ha) {
    i64 s;
    T px, py, pz;
    if (smp.get(s, px, py, pz)) ha[s] = nearest_h<T>(tv, px, py, pz);
}


}  // namespace
}  // namespace ph
