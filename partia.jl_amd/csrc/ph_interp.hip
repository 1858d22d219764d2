// ph_interp.hip -- point / grid / line interpolation in the reference's own
// precision (Float64, or Float32 for a Float32 input).
//
// One lane per sample.  Each lane walks the tree with the stackless escape-rope
// DFS of LBVH/LinearBVH/LinearBVHTraversalMacros.jl (26-108 gather, 135-236
// scatter, 268-371 symmetric; 408-796 the line variants), so accepted leaves
// arrive in DFS preorder (increasing Morton index), exactly the order in which
// the reference accumulates; the per-neighbour terms and the epilogue follow
// KI/single_point_interpolation/kernels/general_interpolation.jl and
// KI/line_integrated_interpolation/kernels/line_integrated_scalar_interpolation.jl
// operation by operation (no FMA contraction), so the float64 results are those
// of the reference algorithm, not an approximation of them.
//
// Grid samples are generated on the device (Grids/grids/StructuredGrid.jl:140-155
// column-major flattening) and each wavefront owns a 4x4x4 block of grid points,
// so the 64 lanes of a wave walk nearly the same nodes; workgroups are remapped
// so that each of the 8 XCDs (separate L2 caches) owns a contiguous run of tiles.
#include <type_traits>

#include "ph_interp.h"
#include "ph_interp_dev.h"
#include "ph_primitives.h"
#include "ph_sph.h"

namespace ph {

int split_catalog(const ph_catalog* cat, std::vector<CatChunk>& chunks) {
    chunks.clear();
    const int N = cat->n_scalar, G = cat->n_grad, D = cat->n_div, C = cat->n_curl;
    const int L = N + 3 * G + D + 3 * C;
    int si = 0, gi = 0, di = 0, ci = 0;
    while (si < N || gi < G || di < D || ci < C || chunks.empty()) {
        CatChunk c{};
        c.rsrc[0] = 3;  // m
        c.rsrc[1] = 5;  // rho
        c.nsrc = 2;
        auto pos_of = [&](int slot) {  // record position of a 1-based quantity column
            const int src = 5 + slot;
            for (int j = 2; j < c.nsrc; ++j)
                if (c.rsrc[j] == src) return j;
            c.rsrc[c.nsrc] = src;
            return c.nsrc++;
        };
        for (; c.ns < kMaxS && si < N; ++si, ++c.ns) {
            c.spos[c.ns] = pos_of(cat->scalar_slots[si]);
            c.sflag[c.ns] = cat->scalar_shepard[si] ? 1 : 0;
            c.sout[c.ns] = si;
        }
        for (; c.ng < kMaxG && gi < G; ++gi, ++c.ng) {
            c.gpos[c.ng] = cat->grad_slots[gi] == 0 ? -1 : pos_of(cat->grad_slots[gi]);
            c.gout[c.ng] = N + 3 * gi;
        }
        for (; c.nd < kMaxD && di < D; ++di, ++c.nd) {
            for (int d = 0; d < 3; ++d) c.dpos[3 * c.nd + d] = pos_of(cat->div_slots[3 * di + d]);
            c.dout[c.nd] = N + 3 * G + di;
        }
        for (; c.nc < kMaxC && ci < C; ++ci, ++c.nc) {
            for (int d = 0; d < 3; ++d) c.cpos[3 * c.nc + d] = pos_of(cat->curl_slots[3 * ci + d]);
            c.cout[c.nc] = N + 3 * G + D + 3 * ci;
        }
        c.rw = (c.nsrc + 1) & ~1;
        chunks.push_back(c);
        if (L == 0) break;
    }
    return L;
}

namespace {

constexpr int kBatch = 16;  // accepted leaves buffered per lane between accumulation rounds

// Point samples.  Each lane walks its own escape-rope DFS, buffering accepted
// leaves (in DFS order) in LDS until it holds kBatch of them or is done; then
// the whole wave accumulates its buffers in lockstep.  The expensive per-neighbour
// arithmetic thus runs without the traversal's divergence, and each sample still
// accumulates its neighbours in the reference's order.
template <class T, int K, int STRAT, int MS, int MG, int MD, int MC, int SRC>
__global__ void __launch_bounds__(kBlock) k_points(TreeView<T> tv, const T* __restrict__ rec, CatChunk c,
                                                   Sampler<T, SRC> smp, const T* __restrict__ ha_s, T* __restrict__ out,
                                                   u64* counter) {
    __shared__ i32 buf[kBatch][kBlock];
    const int tid = threadIdx.x;
    i64 s = 0;
    T px = T(0), py = T(0), pz = T(0);
    const bool valid = smp.get(s, px, py, pz);
    const i64 n = tv.n;
    const T Kvalid = T(support_of<K>());
    T ha = T(0);
    if constexpr (STRAT != PH_SCATTER) {
        if (valid) ha = ha_s[s];
    }
    const T ra = Kvalid * ha;
    const T r2a = ra * ra;
    Acc<T, K, STRAT, MS, MG, MD, MC> acc;
    acc.init();
    // n == 1 has no internal node (root 0): start at the single leaf, whose escape is 0.
    i32 node = valid ? (tv.root != 0 ? tv.root : (i32)n) : 0;
    i64 steps = 2 * n + 8;  // a valid tree is walked in at most 2n-1 steps
    u64 accepted = 0;
    while (true) {
        int cnt = 0;
        while (node != 0 && steps > 0 && cnt < kBatch) {
            --steps;
            if (node >= (i32)n) {
                const i64 k = (i64)node - n;
                const LeafRec<T> L = tv.leaves[k];
                if (d2_point(px, py, pz, L) <= radius2<T, STRAT>(Kvalid, r2a, L.h)) buf[cnt++][tid] = (i32)k;
                node = tv.leaf_escape[k];
            } else {
                const NodeRec<T> N = tv.nodes[node - 1];
                node = (d2_aabb(px, py, pz, N) > radius2<T, STRAT>(Kvalid, r2a, N.scale)) ? N.escape : N.left;
            }
        }
        accepted += cnt;
        for (int j = 0; j < cnt; ++j) {
            const i64 k = buf[j][tid];
            const LeafRec<T> L = tv.leaves[k];
            acc.add(c, rec + k * c.rw, px - L.x, py - L.y, pz - L.z, d2_point(px, py, pz, L), ha, L.h);
        }
        if (!__any(node != 0 && steps > 0)) break;
    }
    if (valid) acc.finish(c, out, smp.S, s);
    count_add(counter, accepted);
}

template <class T, int K, int STRAT, int SRC> struct LineSampler;
template <class T, int K, int STRAT> struct LineSampler<T, K, STRAT, 0> {
    LineList ll;
    i64 S;
    int xcd = 2;  // block order (xcd_block mode)
    __device__ inline bool get(i64& s, Line<T>& l) const {
        s = xcd_block(xcd) * kBlock + threadIdx.x;
        if (s >= S) return false;
        l.o[0] = T(ll.ox[s]); l.o[1] = T(ll.oy[s]); l.o[2] = T(ll.oz[s]);
        l.m[0] = T(ll.dx[s]); l.m[1] = T(ll.dy[s]); l.m[2] = T(ll.dz[s]);
        return true;
    }
};
template <class T, int K, int STRAT> struct LineSampler<T, K, STRAT, 1> {
    LineGrid lg;
    i64 S;
    int xcd = 2;    // block order (xcd_block mode)
    int order = 0;  // tile order: 0 x-fastest rows, 1 bricks
    __device__ inline bool get(i64& s, Line<T>& l) const {
        // every lane holds the direction (wave-uniform values)
        for (int d = 0; d < 3; ++d) l.m[d] = T(lg.d[d]);
        const i64 tile = xcd_block(xcd) * (kBlock / kWave) + threadIdx.x / kWave;
        const int lane = threadIdx.x & (kWave - 1);
        const i64 tx = (lg.nx + 7) / 8, ty = (lg.ny + 7) / 8;
        if (tile >= tx * ty) return false;
        i64 bx, by;
        if (order) brick_tile2(tile, tx, ty, bx, by);
        else {
            bx = tile % tx;
            by = tile / tx;
        }
        const i64 i = bx * 8 + (lane & 7), j = by * 8 + (lane >> 3);
        if (i >= lg.nx || j >= lg.ny) return false;
        s = i + lg.nx * j;
        double a, b;
        if (lg.polar) {  // _polar_plane_coordinates (coordinate.jl:189-205): xi = s cos, eta = s sin
            a = lg.ax[i] * lg.ay[2 * j + 1];
            b = lg.ax[i] * lg.ay[2 * j];
        } else {
            a = lg.ax[i];
            b = lg.ay[j];
        }
        for (int d = 0; d < 3; ++d) l.o[d] = T((lg.o[d] + lg.u[d] * a) + lg.v[d] * b);
        return true;
    }
};

// Per-ray sums of _line_integrated_quantities_interpolate_kernel, scatter
// (line_integrated_scalar_interpolation.jl:123-166): S1 = sum m_b W / rho_b and
// the catalog scalars sum A_b m_b W / rho_b, W = I(dr / h_b) / h_b from the table.
template <class T, int K> struct LineAcc {
    T S1 = T(0);
    T acc[kMaxS];
    __device__ inline LineAcc() {
#pragma unroll
        for (int j = 0; j < kMaxS; ++j) acc[j] = T(0);
    }
    __device__ inline void add(const T* __restrict__ cols, const CatChunk& c, i64 n, i64 k, T d2, T h,
                               const T* __restrict__ inorm, T dq) {
        const T dr = sqrt(d2 < T(0) ? T(0) : d2);
        const T mb = cols[3 * n + k], rb = cols[5 * n + k];
        const T W = line_W<T, K>(dr, h, inorm, dq);
        S1 += mb * W / rb;
#pragma unroll
        for (int j = 0; j < kMaxS; ++j)
            if (j < c.ns) acc[j] += cols[c.rsrc[c.spos[j]] * n + k] * (mb * W / rb);
    }
    // Same terms for a wave in lockstep: every lane evaluates them, and a lane
    // with take == false adds exact zeros (S1 and the sums start at +0 and never
    // become -0, so x + 0 == x).  mb, rb and the quantities qv are the leaf's.
    __device__ inline void add_sel(bool take, T mb, T rb, const T (&qv)[kMaxS], const CatChunk& c, T d2, T h,
                                   const T* __restrict__ inorm, T dq) {
        const T dr = sqrt(d2 < T(0) ? T(0) : d2);
        const T W = line_W_nb<T, K>(dr, h, inorm, dq);
        const T t = mb * W / rb;
        S1 += take ? t : T(0);
#pragma unroll
        for (int j = 0; j < kMaxS; ++j)
            if (j < c.ns) acc[j] += take ? qv[j] * t : T(0);
    }
    // add_sel for the first MS scalars only (c.ns <= MS), with invh = T(1) / h
    // and rdq = T(1) / dq given (line_W_rdq).
    template <int MS>
    __device__ inline void add_sel_inv(bool take, T mb, T rb, const T (&qv)[MS], const CatChunk& c, T d2, T invh,
                                       const T* __restrict__ inorm, T dq, T rdq) {
        const T dr = sqrt(d2 < T(0) ? T(0) : d2);
        const T W = line_W_rdq<T, K>(dr, invh, inorm, dq, rdq);
        const T t = mb * W / rb;
        S1 += take ? t : T(0);
#pragma unroll
        for (int j = 0; j < MS; ++j)
            if (j < c.ns) acc[j] += take ? qv[j] * t : T(0);
    }
    // NaN outputs when no particle contributes (:154-156), Shepard scalars * 1/S1.
    __device__ inline void store(const CatChunk& c, T* __restrict__ out, i64 S, i64 s) const {
        if (S1 == T(0)) {
#pragma unroll
            for (int j = 0; j < kMaxS; ++j)
                if (j < c.ns) out[c.sout[j] * S + s] = sph_nan<T>();
        } else {
            const T invS1 = T(1) / S1;
#pragma unroll
            for (int j = 0; j < kMaxS; ++j)
                if (j < c.ns) out[c.sout[j] * S + s] = c.sflag[j] ? acc[j] * invS1 : acc[j];
        }
    }
};

// One lane per ray, the reference's escape-rope DFS with the line geometry.
template <class T, int K, int STRAT, int SRC>
__global__ void __launch_bounds__(kBlock) k_lines(TreeView<T> tv, const T* __restrict__ cols, CatChunk c,
                                                  LineSampler<T, K, STRAT, SRC> smp, const T* __restrict__ inorm, T dq,
                                                  T* __restrict__ out, u64* counter) {
    i64 s;
    Line<T> l;
    const bool valid = smp.get(s, l);
    u64 accepted = 0;
    if (valid) {
        const T Kvalid = T(support_of<K>());
        LineAcc<T, K> a;
        const i64 n = tv.n;
        dfs<T, STRAT>(
            tv, Kvalid, T(0), [&](const LeafRec<T>& L) { return d2_line_leaf(l, L); },
            [&](const NodeRec<T>& N) { return d2_line_aabb_lb(l, N); },
            [&](i64 k, T d2, const LeafRec<T>& L) {
                ++accepted;
                a.add(cols, c, n, k, d2, L.h, inorm, dq);
            });
        a.store(c, out, smp.S, s);
    }
    count_add(counter, accepted);
}

// One wave per 64 rays (an 8 x 8 tile of an image plane): the wave walks the
// union of its lanes' traversals once, with every node and leaf record fetched
// once per wave at a wave-uniform address instead of once per ray.  Each lane
// replays its own traversal inside the union: a lane that prunes node u goes
// idle until the walk reaches escape(u), which the union walk always reaches
// right after u's subtree, so every lane tests exactly the nodes and leaves of
// its own DFS (LinearBVHTraversalMacros.jl:539-644), in the same order, with the
// same arithmetic as k_lines.  leaf(k, take, d2, L) is called by the whole wave
// for every leaf of the union walk; take marks the lanes that accept it.  The
// walk reads the tree's LineRec records in preorder (lrecs, 2n - 1 records and
// one of padding): each step loads a record and the one after it together,
// and when the walk goes on to the next record (a kept node's left child, or
// the record after a leaf) it tests that one without a load of its own.  The
// records come through the kernels' restrict-qualified parameters, so their
// wave-uniform loads are scalar loads.
#ifndef PH_LINE_AHEAD
#define PH_LINE_AHEAD 2
#endif
constexpr int kLineAhead = PH_LINE_AHEAD;  // records loaded per step (the array carries kLineAhead - 1 of padding)
template <int AX = -1, class T, class Leaf>
__device__ __attribute__((always_inline)) void wave_walk(const TreeView<T>& tv, const LineRec<T>* __restrict__ lrecs,
                                                         const LeafRec<T>* __restrict__ leaves, const Line<T>& l,
                                                         bool valid, T Kvalid, Leaf&& leaf) {
    const i64 n = tv.n;
    // rad2: radius2<T, PH_SCATTER>(Kvalid, 0, L.h)
    auto visit = [&](i64 k, const LeafRec<T>& L, T rad2, u64 active) {
        const T d2 = d2_line_leaf_ax<AX>(l, L);
        leaf(k, active & lanes_cmp<kCmpOLE>(d2, rad2), d2, L);
    };
    if (tv.root == 0) {  // n == 1: the reference's brute-force branch
        for (i64 k = 0; k < n; ++k)
            visit(k, leaves[k], radius2<T, PH_SCATTER>(Kvalid, T(0), leaves[k].h), wave_ballot(valid));
        return;
    }
    const i32 nrec = (i32)(2 * n - 1);
    u64 active = wave_ballot(valid);  // lanes walking (the others wait to resume)
    i32 resume = -1;                  // record at which an idle lane rejoins
    i32 p = 0;                        // the root
    auto step = [&](const LineRec<T>& R) {
        active |= wave_ballot(resume == p);
        if (R.leaf >= 0) {
            visit(R.leaf, LeafRec<T>{R.lo[0], R.lo[1], R.lo[2], R.hi[0]}, R.hi[1], active);
            p = p + 1;
        } else {
            const u64 prune = line_node_prune<AX>(l, R, active);
            resume = lane_in(prune) ? R.link : resume;
            active &= ~prune;
            p = active != 0 ? p + 1 : R.link;
        }
    };
    for (i64 steps = 2 * n + 8; p < nrec && steps > 0; --steps) {
        p = __builtin_amdgcn_readfirstlane(p);
        const i32 p0 = p;
        LineRec<T> Q[kLineAhead];
#pragma unroll
        for (int j = 0; j < kLineAhead; ++j) Q[j] = lrecs[p0 + j];
        step(Q[0]);
#pragma unroll
        for (int j = 1; j < kLineAhead; ++j)
            if (p == p0 + j && p < nrec && --steps > 0) step(Q[j]);
    }
}

// Projection with the wave walk: results identical to k_lines, bit for bit.
// The walk only tests leaves (every lane, branch-free): the leaves some lane
// accepts are listed (in walk order, i.e. increasing leaf index) with one bit
// per lane and leaf, up to 64 at a time; then the wave stages their data in
// LDS (one leaf per lane) and the lanes accumulate their own accepted leaves in
// lockstep, lowest bit first, so each ray adds its neighbours in the per-ray
// DFS's order with the same arithmetic (the distance is recomputed from the
// staged leaf with the same operations).  MS: scalar outputs staged per leaf
// (at least c.ns).  AX >= 0: the rays run along axis AX (d2_line_point).
template <int AX, int MS, class T, int K>
__device__ __attribute__((always_inline)) void lines_wave_body(
    const TreeView<T>& tv, const LineRec<T>* __restrict__ lrecs, const LeafRec<T>* __restrict__ leaves,
    const T* __restrict__ cols, const CatChunk& c, const Line<T>& l, bool valid,
    const T* tab, T dq, T rdq, T* st, LineAcc<T, K>& a, u64& accepted) {
    constexpr int kSlot = 6 + MS;  // x, y, z, 1 / h, m, rho, the scalars' columns
    const int lane = threadIdx.x & (kWave - 1);
    const i64 n = tv.n;
    const T* __restrict__ mcol = cols + 3 * n;
    const T* __restrict__ rcol = cols + 5 * n;
    const T* qcol[MS];
#pragma unroll
    for (int j = 0; j < MS; ++j) qcol[j] = cols + (i64)c.rsrc[c.spos[j < c.ns ? j : 0]] * n;
    u64 mask = 0;
    int nst = 0;    // listed leaves (wave-uniform)
    u32 kleaf = 0;  // lane j < nst: the Morton index of listed leaf j
    auto flush = [&]() {
        wave_sync();
        if (lane < nst) {
            const u32 k = kleaf;
            const LeafRec<T> L = leaves[k];
            T* d = st + lane * kSlot;
            d[0] = L.x;
            d[1] = L.y;
            d[2] = L.z;
            d[3] = T(1) / L.h;  // line_W_nb's reciprocal, once per staged leaf
            d[4] = mcol[k];
            d[5] = rcol[k];
#pragma unroll
            for (int j = 0; j < MS; ++j) d[6 + j] = qcol[j][k];
        }
        wave_sync();
        accepted += __popcll(mask);
        const int nit = lockstep_count(mask);
        for (int it = 0; it < nit; ++it) {
            const bool live = mask != 0;
            const int b = live ? __builtin_ctzll(mask) : 0;
            mask &= mask - 1;
            const T* d = st + b * kSlot;
            const LeafRec<T> L{d[0], d[1], d[2], T(0)};
            T qv[MS];
#pragma unroll
            for (int j = 0; j < MS; ++j) qv[j] = d[6 + j];
            // lanes without a neighbour left read leaf 0 and add exact zeros
            a.template add_sel_inv<MS>(live, d[4], d[5], qv, c, d2_line_leaf_ax<AX>(l, L), d[3], tab, dq, rdq);
        }
        wave_sync();  // every lane is done with the staged leaves
        nst = 0;
    };
    wave_walk<AX>(tv, lrecs, leaves, l, valid, T(support_of<K>()),
                  [&](i64 k, u64 take, T d2, const LeafRec<T>& L) {
                      (void)d2;
                      (void)L;
                      if (take == 0) return;  // wave-uniform
                      kleaf = lane == nst ? (u32)k : kleaf;
                      const u64 bit = u64(1) << nst;
                      mask = lane_in(take) ? (mask | bit) : mask;
                      if (++nst == kWave) flush();
                  });
    if (nst > 0) flush();
}

// Coordinates at most 2^(emax - 2) in magnitude: differences of two of them are finite.
template <class T> __device__ inline bool coord_ok(T x) {
    return fabs(x) <= (std::is_same<T, float>::value ? T(0x1p126f) : T(0x1p1022));
}

// The kernel.  AX >= 0 (image-plane rays along axis AX, chosen by the host)
// takes the axis-aligned body when the tree's coordinates are bounded (flag
// lflag, from k_line_nodes) and so are the wave's ray origins; otherwise, and
// for AX < 0, the general body.  The line-integrated table is staged in LDS.
template <class T, int K, int SRC, int MS, int AX>
__global__ void __launch_bounds__(kBlock) k_lines_wave(TreeView<T> tv, const LineRec<T>* __restrict__ lrecs,
                                                       const LeafRec<T>* __restrict__ leaves,
                                                       const u32* __restrict__ lflag, const T* __restrict__ cols,
                                                       CatChunk c,
                                                       LineSampler<T, K, PH_SCATTER, SRC> smp,
                                                       const T* __restrict__ inorm, T dq, T rdq,
                                                       T* __restrict__ out, u64* counter) {
    __shared__ T tab[PH_LUT_N_ENTRIES];
    __shared__ T stage[kBlock / kWave][kWave * (6 + MS)];
    for (int i = threadIdx.x; i < PH_LUT_N_ENTRIES; i += kBlock) tab[i] = inorm[i];
    __syncthreads();
    i64 s = 0;
    Line<T> l{};  // lanes past the end keep a zero line
    const bool valid = smp.get(s, l);
    if (__ballot(valid) == 0) return;  // wave-uniform; no barrier follows
    T* st = stage[threadIdx.x / kWave];
    LineAcc<T, K> a;
    u64 accepted = 0;
    bool axis = false;
    if constexpr (AX >= 0) {
        const bool ok = coord_ok(l.o[0]) && coord_ok(l.o[1]) && coord_ok(l.o[2]);
        axis = *lflag != 0 && __ballot(!ok) == 0;  // wave-uniform
    }
    if (axis)
        lines_wave_body<(AX < 0 ? 0 : AX), MS>(tv, lrecs, leaves, cols, c, l, valid, tab, dq, rdq, st, a,
                                                accepted);
    else
        lines_wave_body<-1, MS>(tv, lrecs, leaves, cols, c, l, valid, tab, dq, rdq, st, a, accepted);
    if (valid) a.store(c, out, smp.S, s);
    count_add(counter, valid ? accepted : 0);
}

// Adjoint of the projection of one catalog scalar A (no reference
// counterpart): with w_b = m_b I(q_b) / (h_b rho_b), q_b = r_b / h_b and
// out = sum A_b w_b (or, Shepard, sum A_b w_b / sum w_b), every ray adds
// cot * d out / d theta_b for theta_b = (m_b, A_b, h_b, x_b, y_b, z_b) into
// grad[6][n] at the particle's input index; rho_b is held fixed.  The
// neighbours are recomputed with the forward's wave walk (a Shepard ray walks
// twice: its sums first).  I is the table's piecewise-linear interpolant, with
// derivative (I_i - I_{i-1}) / dq on the segment of q; dr/dx_b = (D - (D.m) m) / r
// with D = x_b - o (0 at r = 0).  The terms are evaluated in T, summed over the
// wave's rays in float64 and added with one float64 atomic per wave, particle
// and parameter (so the float64 sums are order-dependent in the last bits).
template <class T, int K, int SRC, bool SHEP, int AX>
__global__ void __launch_bounds__(kBlock) k_lines_adjoint(TreeView<T> tv, const LineRec<T>* __restrict__ lrecs,
                                                          const LeafRec<T>* __restrict__ leaves,
                                                          const u32* __restrict__ lflag, const T* __restrict__ cols,
                                                          i32 qsrc,
                                                          LineSampler<T, K, PH_SCATTER, SRC> smp,
                                                          const T* __restrict__ inorm, T dq,
                                                          const T* __restrict__ cot, const u32* __restrict__ order,
                                                          double* __restrict__ grad) {
    __shared__ T tab[PH_LUT_N_ENTRIES];
    for (int i = threadIdx.x; i < PH_LUT_N_ENTRIES; i += kBlock) tab[i] = inorm[i];
    __syncthreads();
    i64 s = 0;
    Line<T> l{};  // lanes past the end keep a zero line
    const bool valid = smp.get(s, l);
    if (__ballot(valid) == 0) return;  // wave-uniform; no barrier follows
    const T Kvalid = T(support_of<K>());
    const i64 n = tv.n;
    const T* __restrict__ mcol = cols + 3 * n;
    const T* __restrict__ rcol = cols + 5 * n;
    const T* __restrict__ acol = cols + (i64)qsrc * n;
    // the walks with the axis-aligned distances (as k_lines_wave) or the general ones
    bool axis = false;
    if constexpr (AX >= 0) {
        const bool ok = coord_ok(l.o[0]) && coord_ok(l.o[1]) && coord_ok(l.o[2]);
        axis = *lflag != 0 && __ballot(!ok) == 0;  // wave-uniform
    }
    auto run = [&](auto axc) {
        constexpr int AXV = decltype(axc)::value;
        T g = valid ? cot[s] : T(0);
        T S1 = T(0), out = T(0);
        if constexpr (SHEP) {  // the forward's sums, same arithmetic
            T N = T(0);
            wave_walk<AXV>(tv, lrecs, leaves, l, valid, Kvalid,
                           [&](i64 k, u64 take_m, T d2, const LeafRec<T>& L) {
                const bool take = lane_in(take_m);
                const T mb = mcol[k], rb = rcol[k], Ab = acol[k];
                const T dr = sqrt(d2 < T(0) ? T(0) : d2);
                const T t = mb * line_W_nb<T, K>(dr, L.h, tab, dq) / rb;
                S1 += take ? t : T(0);
                N += take ? Ab * t : T(0);
            });
            if (S1 == T(0)) g = T(0);  // NaN output: no gradient
            out = S1 == T(0) ? T(0) : N * (T(1) / S1);
        }
        const T c2 = SHEP ? (S1 == T(0) ? T(0) : T(1) / S1) : T(1);
        wave_walk<AXV>(tv, lrecs, leaves, l, valid, Kvalid, [&](i64 k, u64 take_m, T d2, const LeafRec<T>& L) {
            if ((take_m & lanes_cmp<kCmpUNE>(g, T(0))) == 0) return;  // wave-uniform
            const bool take = lane_in(take_m);
            const T mb = mcol[k], rb = rcol[k], Ab = acol[k];
            const T dr = sqrt(d2 < T(0) ? T(0) : d2);
            const T invh = T(1) / L.h, q = dr * invh;
            const T idxf = q / dq + T(1);
            const T lo = T(1), hi = T(PH_LUT_N_ENTRIES - 1);
            const T cf = idxf < lo ? lo : (idxf > hi ? hi : idxf);
            i32 i = (i32)cf;
            i = i < 1 ? 1 : (i > PH_LUT_N_ENTRIES - 1 ? PH_LUT_N_ENTRIES - 1 : i);
            const T t = idxf - T(i);
            const bool cut = q >= Kvalid;
            const T I = cut ? T(0) : tab[i - 1] * (T(1) - t) + tab[i] * t;
            const T dI = cut ? T(0) : (tab[i] - tab[i - 1]) / dq;
            const T w = mb * (invh * I) / rb;
            const T c1 = SHEP ? (Ab - out) * c2 : Ab;
            const T v[3] = {g * c1 * ((invh * I) / rb), g * c2 * w, -(g * c1) * (mb / rb) * invh * invh * (I + q * dI)};
            const T f = dr > T(0) ? g * c1 * (mb / rb) * invh * invh * dI / dr : T(0);
            const T D[3] = {L.x - l.o[0], L.y - l.o[1], L.z - l.o[2]};
            const T Dm = D[0] * l.m[0] + D[1] * l.m[1] + D[2] * l.m[2];
            // lanes not taking the leaf contribute exact zeros (selected, not
            // multiplied: their terms may be inf or NaN)
            double part[6];
#pragma unroll
            for (int j = 0; j < 3; ++j) part[j] = take ? (double)v[j] : 0.0;
#pragma unroll
            for (int a = 0; a < 3; ++a) part[3 + a] = take ? (double)(f * (D[a] - Dm * l.m[a])) : 0.0;
            const double sum = wave_sum6(part);  // parameter lane >> 3
            const int lane = threadIdx.x & (kWave - 1);
            if ((lane & 7) == 0 && lane < 48 && sum != 0.0)
                unsafeAtomicAdd(&grad[(i64)(lane >> 3) * n + order[k]], sum);
        });
    };
    if (axis) run(std::integral_constant<int, (AX < 0 ? 0 : AX)>{});
    else run(std::integral_constant<int, -1>{});
}

// grad[0 .. count) = 0 (a kernel, not a memset, so the call can be recorded).
__global__ void k_zero_f64(double* __restrict__ p, i64 count) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (i64)gridDim.x * blockDim.x) p[i] = 0.0;
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_nearest(TreeView<T> tv, PointList pl, i64 S, double* ha) {
    const i64 s = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (s >= S) return;
    ha[s] = (double)nearest_h<T>(tv, T(pl.x[s]), T(pl.y[s]), T(pl.z[s]));
}

inline unsigned blocks_for_points(int src, i64 S, const GridAxes* ga) {
    if (src == 1) {
        const i64 tiles = ((ga->nx + 3) / 4) * ((ga->ny + 3) / 4) * ((ga->nz + 3) / 4);
        return (unsigned)((tiles + 3) / 4);
    }
    return (unsigned)((S + kBlock - 1) / kBlock);
}

template <class T, int K, int STRAT, int MS, int MG, int MD, int MC>
void launch_points_t(ph_lbvh* B, const CatChunk& c, const PointList* pl, const GridAxes* ga, i64 S, void* out,
                     u64* counters, hipStream_t s, hipEvent_t* ev) {
    const TreeView<T> tv = view_of<T>(B);
    B->rec.alloc(sizeof(T) * (size_t)c.rw * (size_t)B->n);
    const T* cols = B->rec.as<T>();
    k_pack_records<T><<<grid_for(B->n), kBlock, 0, s>>>(B->cols.as<T>(), B->n, c, B->rec.as<T>());
    PH_HIP(hipGetLastError());
    B->ha.alloc(sizeof(T) * (size_t)(S > 0 ? S : 1));
    T* ha = B->ha.as<T>();
    if (ga) {
        Sampler<T, 1> smp{*ga, S};
        const unsigned g = blocks_for_points(1, S, ga);
        if (STRAT != PH_SCATTER) {
            k_nearest_samples<T, 1><<<g, kBlock, 0, s>>>(tv, smp, ha);
            PH_HIP(hipGetLastError());
        }
        if (ev) PH_HIP(hipEventRecord(ev[0], s));
        k_points<T, K, STRAT, MS, MG, MD, MC, 1><<<g, kBlock, 0, s>>>(tv, cols, c, smp, ha, static_cast<T*>(out),
                                                                     counters);
        if (ev) PH_HIP(hipEventRecord(ev[1], s));
    } else {
        Sampler<T, 0> smp{*pl, S};
        const unsigned g = blocks_for_points(0, S, nullptr);
        if (STRAT != PH_SCATTER) {
            k_nearest_samples<T, 0><<<g, kBlock, 0, s>>>(tv, smp, ha);
            PH_HIP(hipGetLastError());
        }
        if (ev) PH_HIP(hipEventRecord(ev[0], s));
        k_points<T, K, STRAT, MS, MG, MD, MC, 0><<<g, kBlock, 0, s>>>(tv, cols, c, smp, ha, static_cast<T*>(out),
                                                                     counters);
        if (ev) PH_HIP(hipEventRecord(ev[1], s));
    }
    PH_HIP(hipGetLastError());
}

template <class T, int K, int STRAT>
void launch_points_k(ph_lbvh* B, const CatChunk& c, const PointList* pl, const GridAxes* ga, i64 S, void* out,
                     u64* counters, hipStream_t s, hipEvent_t* ev) {
    if (c.ng + c.nd + c.nc == 0 && c.ns <= 2)
        launch_points_t<T, K, STRAT, 2, 0, 0, 0>(B, c, pl, ga, S, out, counters, s, ev);
    else
        launch_points_t<T, K, STRAT, kMaxS, kMaxG, kMaxD, kMaxC>(B, c, pl, ga, S, out, counters, s, ev);
}

template <class T, int K>
void launch_points_s(ph_lbvh* B, const CatChunk& c, int strategy, const PointList* pl, const GridAxes* ga, i64 S,
                     void* out, u64* counters, hipStream_t s, hipEvent_t* ev) {
    switch (strategy) {
    case PH_GATHER: launch_points_k<T, K, PH_GATHER>(B, c, pl, ga, S, out, counters, s, ev); break;
    case PH_SCATTER: launch_points_k<T, K, PH_SCATTER>(B, c, pl, ga, S, out, counters, s, ev); break;
    default: launch_points_k<T, K, PH_SYMMETRIC>(B, c, pl, ga, S, out, counters, s, ev); break;
    }
}

template <class T>
void launch_points_p(ph_lbvh* B, const CatChunk& c, int strategy, const PointList* pl, const GridAxes* ga, i64 S,
                     void* out, u64* counters, hipStream_t s, hipEvent_t* ev) {
    switch (B->kernel) {
    case PH_KERNEL_M4: launch_points_s<T, PH_KERNEL_M4>(B, c, strategy, pl, ga, S, out, counters, s, ev); break;
    case PH_KERNEL_M5: launch_points_s<T, PH_KERNEL_M5>(B, c, strategy, pl, ga, S, out, counters, s, ev); break;
    case PH_KERNEL_M6: launch_points_s<T, PH_KERNEL_M6>(B, c, strategy, pl, ga, S, out, counters, s, ev); break;
    case PH_KERNEL_C2: launch_points_s<T, PH_KERNEL_C2>(B, c, strategy, pl, ga, S, out, counters, s, ev); break;
    case PH_KERNEL_C4: launch_points_s<T, PH_KERNEL_C4>(B, c, strategy, pl, ga, S, out, counters, s, ev); break;
    default: launch_points_s<T, PH_KERNEL_C6>(B, c, strategy, pl, ga, S, out, counters, s, ev); break;
    }
}

// LineRec records of a tree with kernel K (n >= 2) in DFS preorder: internal
// node u (1-based) with leaves [lo, hi] (1-based) sits at L(u) + 2 (lo - 1),
// L(u) = its left-child edges from the root (lc_total), and its subtree ends
// 2 (hi - lo) + 1 records later; a leaf child on the left follows its parent,
// one on the right (leaf hi = s + 1 of split s) sits at L(u) + 2 s.  Each node
// writes itself and its leaf children, so every record is written once.
template <class T, int K>
__global__ void __launch_bounds__(kBlock) k_line_recs(i64 n, const i32* __restrict__ left,
                                                      const i32* __restrict__ range, const i32* __restrict__ up,
                                                      const i32* __restrict__ lc, const NodeRec<T>* __restrict__ nodes,
                                                      const LeafRec<T>* __restrict__ leaves,
                                                      LineRec<T>* __restrict__ out) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x + 1;
    if (i > n - 1) return;
    const i64 lo = range[2 * (i - 1)], hi = range[2 * (i - 1) + 1];
    const i64 l = left[i - 1];
    const i64 sp = l >= n ? lo : l;  // the split: left child covers [lo, sp]
    const i64 Lc = lc_total(i, n, up, lc);
    const i64 pre = Lc + 2 * (lo - 1);
    out[pre] = line_rec_node<T>(nodes[i - 1], T(support_of<K>()), (i32)(pre + 2 * (hi - lo) + 1));
    auto leaf_rec = [&](i64 k, i32 link) {
        const LeafRec<T> L = leaves[k];
        return line_rec_leaf<T>(L, radius2<T, PH_SCATTER>(T(support_of<K>()), T(0), L.h), (i32)k, link);
    };
    if (sp == lo) out[pre + 1] = leaf_rec(lo - 1, (i32)(pre + 2));
    if (sp + 1 == hi) {
        const i64 rp = Lc + 2 * sp;
        out[rp] = leaf_rec(hi - 1, (i32)(rp + 1));
    }
}
// The flag of the axis-aligned walk: *flag (1 beforehand) is cleared when a
// leaf coordinate is not finite or not below 2^(emax - 2) in magnitude (the
// node boxes and centres lie within the leaves' range).
template <class T>
__global__ void __launch_bounds__(kBlock) k_line_flag(const LeafRec<T>* __restrict__ leaves, i64 n, u32* flag) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const LeafRec<T> L = leaves[i];
    if (!(coord_ok(L.x) && coord_ok(L.y) && coord_ok(L.z))) atomicAnd(flag, 0u);
}
__global__ void k_set_u32(u32* p, u32 v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// The tree's LineRec records (2n - 1 in preorder, then one of padding read by
// the walk's paired loads) and, after them, the axis flag; the first call
// builds them (the tree is immutable).  A one-leaf tree has none (its walk is
// the brute-force branch).
inline i64 line_rec_count(const ph_lbvh* B) { return B->n > 1 ? 2 * B->n - 1 : 0; }
template <class T, int K> const LineRec<T>* line_recs(ph_lbvh* B, hipStream_t s) {
    if (!B->lnodes_ready) {
        const i64 n = B->n, m = line_rec_count(B);
        B->lnodes.alloc(sizeof(LineRec<T>) * (size_t)(m + kLineAhead));
        LineRec<T>* lr = B->lnodes.as<LineRec<T>>();
        u32* flag = reinterpret_cast<u32*>(lr + m + kLineAhead - 1);
        k_set_u32<<<1, 1, 0, s>>>(flag, 1u);
        PH_HIP(hipGetLastError());
        k_line_flag<T><<<grid_for(n), kBlock, 0, s>>>(B->leaves.as<LeafRec<T>>(), n, flag);
        PH_HIP(hipGetLastError());
        if (n > 1) {
            DevBuf lcb;  // up, lc pointers of two rounds (released into the context's cache)
            lcb.alloc(sizeof(i32) * 4 * (size_t)(n - 1));
            i32* a = lcb.as<i32>();
            i32* b = a + 2 * (n - 1);
            k_lc_init<<<grid_for(n - 1), kBlock, 0, s>>>(n, B->root, B->parent.as<i32>(), B->left.as<i32>(), a,
                                                        a + (n - 1));
            PH_HIP(hipGetLastError());
            for (int r = 0; r < kLcJumps; ++r) {
                k_lc_jump<<<grid_for(n - 1), kBlock, 0, s>>>(n, a, a + (n - 1), b, b + (n - 1));
                PH_HIP(hipGetLastError());
                std::swap(a, b);
            }
            k_line_recs<T, K><<<grid_for(n - 1), kBlock, 0, s>>>(n, B->left.as<i32>(), B->range.as<i32>(), a,
                                                                a + (n - 1), B->nodes.as<NodeRec<T>>(),
                                                                B->leaves.as<LeafRec<T>>(), lr);
            PH_HIP(hipGetLastError());
            if (!tl_cache) PH_HIP(hipStreamSynchronize(s));
        }
        B->lnodes_ready = true;
    }
    return B->lnodes.as<LineRec<T>>();
}
template <class T> const u32* line_flag(const ph_lbvh* B) {
    return reinterpret_cast<const u32*>(B->lnodes.as<LineRec<T>>() + line_rec_count(B) + kLineAhead - 1);
}

// The axis of image-plane rays with direction exactly +-e_axis, else -1.
inline int ray_axis(const LineGrid& lg) {
    for (int a = 0; a < 3; ++a) {
        bool ok = lg.d[a] == 1.0 || lg.d[a] == -1.0;
        for (int d = 0; d < 3; ++d)
            if (d != a && lg.d[d] != 0.0) ok = false;
        if (ok) return a;
    }
    return -1;
}

template <class T, int K, int SRC, int AX, class Smp>
void launch_lines_wave_ax(const TreeView<T>& tv, const LineRec<T>* ln, const u32* flag, const T* cols,
                          const CatChunk& c, const Smp& smp, unsigned g, const T* tab, T dq, T* out, u64* counters,
                          hipStream_t s) {
    const T rdq = T(1) / dq;  // correctly rounded (host IEEE division)
    if (c.ns <= 1)
        k_lines_wave<T, K, SRC, 1, AX><<<g, kBlock, 0, s>>>(tv, ln, tv.leaves, flag, cols, c, smp, tab, dq, rdq, out,
                                                            counters);
    else
        k_lines_wave<T, K, SRC, kMaxS, AX><<<g, kBlock, 0, s>>>(tv, ln, tv.leaves, flag, cols, c, smp, tab, dq, rdq,
                                                                out, counters);
}
template <class T, int K, int SRC, class Smp>
void launch_lines_wave(const TreeView<T>& tv, const LineRec<T>* ln, const u32* flag, const T* cols,
                       const CatChunk& c, const Smp& smp, int axis, unsigned g, const T* tab, T dq, T* out,
                       u64* counters, hipStream_t s) {
    switch (SRC == 1 ? axis : -1) {
    case 0: launch_lines_wave_ax<T, K, SRC, 0>(tv, ln, flag, cols, c, smp, g, tab, dq, out, counters, s); break;
    case 1: launch_lines_wave_ax<T, K, SRC, 1>(tv, ln, flag, cols, c, smp, g, tab, dq, out, counters, s); break;
    case 2: launch_lines_wave_ax<T, K, SRC, 2>(tv, ln, flag, cols, c, smp, g, tab, dq, out, counters, s); break;
    default: launch_lines_wave_ax<T, K, SRC, -1>(tv, ln, flag, cols, c, smp, g, tab, dq, out, counters, s); break;
    }
}

template <class T, int K>
void launch_lines_k(ph_lbvh* B, const CatChunk& c, const LineList* ll, const LineGrid* lg, i64 S, void* out,
                    u64* counters, const T* inorm, bool wave, hipStream_t s, hipEvent_t* ev) {
    const TreeView<T> tv = view_of<T>(B);
    const T* cols = B->cols.as<T>();
    // Q_2 - Q_1 of the table in T (line_integrated_kernel.jl:14)
    const T dq = T((1.0 * support_of<K>()) / 4095.0) - T(0.0);
    const T* tab = inorm + (size_t)K * PH_LUT_N_ENTRIES;
    const LineRec<T>* ln = wave ? line_recs<T, K>(B, s) : nullptr;
    const u32* flag = wave ? line_flag<T>(B) : nullptr;
    if (ev) PH_HIP(hipEventRecord(ev[0], s));
    if (lg) {
        const i64 tiles = ((lg->nx + 7) / 8) * ((lg->ny + 7) / 8);
        const unsigned g = (unsigned)((tiles + 3) / 4);
        LineSampler<T, K, PH_SCATTER, 1> smp{*lg, S, line_xcd(tl_launch, g), line_order(tl_launch)};
        if (wave)
            launch_lines_wave<T, K, 1>(tv, ln, flag, cols, c, smp, ray_axis(*lg), g, tab, dq, static_cast<T*>(out),
                                       counters, s);
        else
            k_lines<T, K, PH_SCATTER, 1><<<g, kBlock, 0, s>>>(tv, cols, c, smp, tab, dq, static_cast<T*>(out),
                                                              counters);
    } else {
        const unsigned g = (unsigned)((S + kBlock - 1) / kBlock);
        LineSampler<T, K, PH_SCATTER, 0> smp{*ll, S, line_xcd(tl_launch, g)};
        if (wave)
            launch_lines_wave<T, K, 0>(tv, ln, flag, cols, c, smp, -1, g, tab, dq, static_cast<T*>(out), counters,
                                       s);
        else
            k_lines<T, K, PH_SCATTER, 0><<<g, kBlock, 0, s>>>(tv, cols, c, smp, tab, dq, static_cast<T*>(out),
                                                              counters);
    }
    PH_HIP(hipGetLastError());
    if (ev) PH_HIP(hipEventRecord(ev[1], s));
}

template <class T>
void launch_lines_p(ph_lbvh* B, const CatChunk& c, const LineList* ll, const LineGrid* lg, i64 S, void* out,
                    u64* counters, const T* inorm, bool wave, hipStream_t s, hipEvent_t* ev) {
    switch (B->kernel) {
    case PH_KERNEL_M4: launch_lines_k<T, PH_KERNEL_M4>(B, c, ll, lg, S, out, counters, inorm, wave, s, ev); break;
    case PH_KERNEL_M5: launch_lines_k<T, PH_KERNEL_M5>(B, c, ll, lg, S, out, counters, inorm, wave, s, ev); break;
    case PH_KERNEL_M6: launch_lines_k<T, PH_KERNEL_M6>(B, c, ll, lg, S, out, counters, inorm, wave, s, ev); break;
    case PH_KERNEL_C2: launch_lines_k<T, PH_KERNEL_C2>(B, c, ll, lg, S, out, counters, inorm, wave, s, ev); break;
    case PH_KERNEL_C4: launch_lines_k<T, PH_KERNEL_C4>(B, c, ll, lg, S, out, counters, inorm, wave, s, ev); break;
    default: launch_lines_k<T, PH_KERNEL_C6>(B, c, ll, lg, S, out, counters, inorm, wave, s, ev); break;
    }
}

template <class T, int K, bool SHEP>
void launch_adjoint_k(ph_lbvh* B, i32 qsrc, const LineList* ll, const LineGrid* lg, i64 S, const T* cot,
                      const T* inorm, double* grad, hipStream_t s, hipEvent_t* ev) {
    const TreeView<T> tv = view_of<T>(B);
    const T* cols = B->cols.as<T>();
    const T dq = T((1.0 * support_of<K>()) / 4095.0) - T(0.0);
    const T* tab = inorm + (size_t)K * PH_LUT_N_ENTRIES;
    const u32* order = B->order.as<u32>();
    const LineRec<T>* ln = line_recs<T, K>(B, s);
    const u32* flag = line_flag<T>(B);
    k_zero_f64<<<grid_for(6 * B->n), kBlock, 0, s>>>(grad, 6 * B->n);
    PH_HIP(hipGetLastError());
    if (ev) PH_HIP(hipEventRecord(ev[0], s));
    if (lg) {
        const i64 tiles = ((lg->nx + 7) / 8) * ((lg->ny + 7) / 8);
        const unsigned g = (unsigned)((tiles + 3) / 4);
        LineSampler<T, K, PH_SCATTER, 1> smp{*lg, S, line_xcd(tl_launch, g), line_order(tl_launch)};
        auto launch = [&](auto axc) {
            k_lines_adjoint<T, K, 1, SHEP, decltype(axc)::value><<<g, kBlock, 0, s>>>(
                tv, ln, tv.leaves, flag, cols, qsrc, smp, tab, dq, cot, order, grad);
        };
        switch (ray_axis(*lg)) {
        case 0: launch(std::integral_constant<int, 0>{}); break;
        case 1: launch(std::integral_constant<int, 1>{}); break;
        case 2: launch(std::integral_constant<int, 2>{}); break;
        default: launch(std::integral_constant<int, -1>{}); break;
        }
    } else {
        const unsigned g = (unsigned)((S + kBlock - 1) / kBlock);
        LineSampler<T, K, PH_SCATTER, 0> smp{*ll, S, line_xcd(tl_launch, g)};
        k_lines_adjoint<T, K, 0, SHEP, -1><<<g, kBlock, 0, s>>>(tv, ln, tv.leaves, flag, cols, qsrc, smp, tab, dq,
                                                                cot, order, grad);
    }
    PH_HIP(hipGetLastError());
    if (ev) PH_HIP(hipEventRecord(ev[1], s));
}

template <class T, bool SHEP>
void launch_adjoint_p(ph_lbvh* B, i32 qsrc, const LineList* ll, const LineGrid* lg, i64 S, const T* cot,
                      const T* inorm, double* grad, hipStream_t s, hipEvent_t* ev) {
    switch (B->kernel) {
    case PH_KERNEL_M4: launch_adjoint_k<T, PH_KERNEL_M4, SHEP>(B, qsrc, ll, lg, S, cot, inorm, grad, s, ev); break;
    case PH_KERNEL_M5: launch_adjoint_k<T, PH_KERNEL_M5, SHEP>(B, qsrc, ll, lg, S, cot, inorm, grad, s, ev); break;
    case PH_KERNEL_M6: launch_adjoint_k<T, PH_KERNEL_M6, SHEP>(B, qsrc, ll, lg, S, cot, inorm, grad, s, ev); break;
    case PH_KERNEL_C2: launch_adjoint_k<T, PH_KERNEL_C2, SHEP>(B, qsrc, ll, lg, S, cot, inorm, grad, s, ev); break;
    case PH_KERNEL_C4: launch_adjoint_k<T, PH_KERNEL_C4, SHEP>(B, qsrc, ll, lg, S, cot, inorm, grad, s, ev); break;
    default: launch_adjoint_k<T, PH_KERNEL_C6, SHEP>(B, qsrc, ll, lg, S, cot, inorm, grad, s, ev); break;
    }
}

}  // namespace

void launch_lines_adjoint(ph_lbvh* B, i32 qsrc, bool shepard, int precision, const LineList* ll, const LineGrid* lg,
                          i64 S, const void* cot, const void* lut, double* grad, hipStream_t s, hipEvent_t* ev) {
    if (precision == PH_PREC_FP64) {
        auto c = static_cast<const double*>(cot);
        auto t = static_cast<const double*>(lut);
        if (shepard) launch_adjoint_p<double, true>(B, qsrc, ll, lg, S, c, t, grad, s, ev);
        else launch_adjoint_p<double, false>(B, qsrc, ll, lg, S, c, t, grad, s, ev);
    } else {
        auto c = static_cast<const float*>(cot);
        auto t = static_cast<const float*>(lut);
        if (shepard) launch_adjoint_p<float, true>(B, qsrc, ll, lg, S, c, t, grad, s, ev);
        else launch_adjoint_p<float, false>(B, qsrc, ll, lg, S, c, t, grad, s, ev);
    }
}

void launch_points(ph_lbvh* B, const CatChunk& c, int strategy, int precision, const PointList* pl,
                   const GridAxes* ga, i64 S, int L, void* out, u64* counters, hipStream_t s, hipEvent_t* ev) {
    (void)L;
    if (S <= 0) return;
    if (precision == PH_PREC_FP64) launch_points_p<double>(B, c, strategy, pl, ga, S, out, counters, s, ev);
    else launch_points_p<float>(B, c, strategy, pl, ga, S, out, counters, s, ev);
}

void launch_lines(ph_lbvh* B, const CatChunk& c, int strategy, int precision, const LineList* ll, const LineGrid* lg,
                  i64 S, int L, void* out, u64* counters, const void* lut, bool wave, hipStream_t s, hipEvent_t* ev) {
    (void)L;
    (void)strategy;
    if (S <= 0) return;
    if (precision == PH_PREC_FP64)
        launch_lines_p<double>(B, c, ll, lg, S, out, counters, static_cast<const double*>(lut), wave, s, ev);
    else
        launch_lines_p<float>(B, c, ll, lg, S, out, counters, static_cast<const float*>(lut), wave, s, ev);
}

void launch_nearest_h(ph_lbvh* B, const PointList& pl, i64 S, double* ha, hipStream_t s) {
    if (S <= 0) return;
    const unsigned g = (unsigned)((S + kBlock - 1) / kBlock);
    if (B->precision == PH_PREC_FP64) k_nearest<double><<<g, kBlock, 0, s>>>(view_of<double>(B), pl, S, ha);
    else k_nearest<float><<<g, kBlock, 0, s>>>(view_of<float>(B), pl, S, ha);
    PH_HIP(hipGetLastError());
}

}  // namespace ph
