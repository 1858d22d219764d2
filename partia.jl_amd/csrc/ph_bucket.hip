// ph_bucket.hip -- bucket traversal for point and grid samples, in float64 and
// in the FP32-compensated precision: the bucket-tree build and the per-call
// kernels and launchers.  The traversal, the accumulators and the exactness
// argument are in ph_bucket_dev.h.
#include <algorithm>

#include "ph_bucket_dev.h"

namespace ph {
namespace {

// ---- bucket tree build ----

// Per internal node i: big_s[split - 1] = (leaves of i > kBucket); for a big node,
// bs[first leaf - 1] = 1 for each child that is a bucket (at most kBucket leaves).
// Up to four bucket trees are built together (kMaxTrees); tree t's flag
// segments (see build_bucket_trees) are big_s[t] and bs[t].
constexpr int kMaxTrees = 4;
struct FlagTrees {
    u32* big_s[kMaxTrees];
    u32* bs[kMaxTrees];
    int bmax[kMaxTrees];
    int ntrees;
};

// Split and bucket-start flags of every tree.
__global__ void __launch_bounds__(kBlock) k_bucket_flags(i64 n, const i32* __restrict__ left,
                                                 const i32* __restrict__ range, FlagTrees ft) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x + 1;
    if (i > n - 1) return;
    const i64 lo = range[2 * (i - 1)], hi = range[2 * (i - 1) + 1];
    const i64 l = left[i - 1];
    const i64 s = l >= n ? lo : l;  // split: the left child covers [lo, s]
    for (int t = 0; t < ft.ntrees; ++t) {
        const int bmax = ft.bmax[t];
        const bool big = hi - lo + 1 > bmax;
        ft.big_s[t][s - 1] = big ? 1u : 0u;
        if (big) {
            if (s - lo + 1 <= bmax) ft.bs[t][lo - 1] = 1u;
            if (hi - s <= bmax) ft.bs[t][s] = 1u;
        }
    }
}

// The scanned flag array's values at the segment starts of ntrees trees and
// at its closing entry (info[0 .. 2 ntrees]), and the build's refit check
// count (info[11], 0 without one).
__global__ void k_bucket_info(const u32* __restrict__ scanned, i64 n, int ntrees, const u32* __restrict__ refit_bad,
                              u32* __restrict__ info) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const i64 per_tree = 2 * n + 1;
    for (int t = 0; t < ntrees; ++t) {
        info[2 * t] = scanned ? scanned[t * per_tree] : 0u;
        info[2 * t + 1] = scanned ? scanned[t * per_tree + n] : 0u;
    }
    info[2 * ntrees] = scanned ? scanned[ntrees * per_tree] : 0u;
    info[11] = refit_bad ? *refit_bad : 0u;
}

__device__ inline void round_box(const double lo[3], const double hi[3], double scale, BucketRec& R) {
    for (int d = 0; d < 3; ++d) {
        R.lo[d] = f32_down(lo[d]);
        R.hi[d] = f32_up(hi[d]);
    }
    R.scale = f32_up(scale);
}
// Box of unified node id u (internal: its AABB and scale; leaf: its point and h).
__device__ inline void node_box(i64 u, i64 n, const NodeRec<double>* __restrict__ nodes,
                                const LeafRec<double>* __restrict__ leaves, BucketRec& R) {
    if (u >= n) {
        const LeafRec<double> L = leaves[u - n];
        const double p[3] = {L.x, L.y, L.z};
        round_box(p, p, L.h, R);
    } else {
        const NodeRec<double> N = nodes[u - 1];
        round_box(N.lo, N.hi, N.scale, R);
    }
}

// Every big internal node writes its record and those of its bucket children.
// Preorder position of a collapsed node u with first leaf lo (1-based):
//   pre(u) = L(u) + PB[lo - 1] + PS[lo - 1],
// L(u) = number of ancestors holding u in their left subtree, PB / PS = exclusive
// prefix counts of big splits / bucket starts: the nodes before u in preorder are
// its ancestors (all big) and the big nodes and buckets entirely left of lo; big
// nodes with split < lo are the latter plus the ancestors holding u on the right.
// One bucket tree's prefix arrays (segments of the scanned flag array: their
// values minus the value at the segment start), records and bucket size.
struct EmitTree {
    const u32* PB;
    const u32* PS;
    u32 pb0, ps0;
    BucketRec* recs;
    int bmax;
};

struct EmitTrees {
    EmitTree t[kMaxTrees];
    int ntrees;  // the trees with at least one split, smallest buckets first
    // the one-leaf tree (the LBVH in preorder), or null: every internal node
    // is big and every leaf a bucket, so PB(k) = PS(k) = k and its positions
    // need no flags
    BucketRec* leaf_recs;
};


// Every tree in one pass: L(u) and the boxes are properties of the LBVH,
// found once per node.
__global__ void __launch_bounds__(kBlock) k_bucket_emit(i64 n, const i32* __restrict__ left,
                                                const i32* __restrict__ range, const i32* __restrict__ up,
                                                const i32* __restrict__ lc, const NodeRec<double>* __restrict__ nodes,
                                                const LeafRec<double>* __restrict__ leaves, EmitTrees et) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x + 1;
    if (i > n - 1) return;
    const i64 lo = range[2 * (i - 1)], hi = range[2 * (i - 1) + 1];
    // a bucket (or inside one) in every tree
    if (!et.leaf_recs && (et.ntrees == 0 || hi - lo + 1 <= et.t[0].bmax)) return;
    const i64 l = left[i - 1];
    const i64 s = l >= n ? lo : l;
    BucketRec Ri, Rl, Rr;  // the node, its left and its right child
    node_box(i, n, nodes, leaves, Ri);
    node_box(s == lo ? (lo + n - 1) : s, n, nodes, leaves, Rl);
    node_box(s + 1 == hi ? (hi + n - 1) : s + 1, n, nodes, leaves, Rr);
    const i64 Lc = lc_total(i, n, up, lc);
    if (et.leaf_recs) {
        const i64 pre = Lc + 2 * (lo - 1);
        BucketRec R = Ri;
        R.link = (i32)(pre + 2 * (hi - lo) + 1);
        et.leaf_recs[pre] = R;
        if (s == lo) {  // left child a leaf, right after its parent
            R = Rl;
            encode_bucket(R, lo - 1, 1);
            et.leaf_recs[pre + 1] = R;
        }
        if (s + 1 == hi) {
            R = Rr;
            encode_bucket(R, s, 1);
            et.leaf_recs[Lc + 2 * s] = R;
        }
    }
    for (int t = 0; t < et.ntrees; ++t) {
        const EmitTree& T = et.t[t];
        if (hi - lo + 1 <= T.bmax) continue;
        auto PB = [&](i64 k) { return (i64)(T.PB[k] - T.pb0); };
        auto PS = [&](i64 k) { return (i64)(T.PS[k] - T.ps0); };
        const i64 pre = Lc + PB(lo - 1) + PS(lo - 1);
        const i64 cnt = (PB(hi - 1) - PB(lo - 1)) + (PS(hi) - PS(lo - 1));
        BucketRec R = Ri;
        R.link = (i32)(pre + cnt);
        T.recs[pre] = R;
        if (s - lo + 1 <= T.bmax) {  // left child is a bucket, right after its parent
            R = Rl;
            encode_bucket(R, lo - 1, s - lo + 1);
            T.recs[pre + 1] = R;
        }
        if (hi - s <= T.bmax) {
            const i64 rpre = Lc + PB(s) + PS(s);
            R = Rr;
            encode_bucket(R, s, hi - s);
            T.recs[rpre] = R;
        }
    }
}

// n <= kBucket: the whole tree is one bucket.
__global__ void k_bucket_root(i64 n, i32 root, const NodeRec<double>* __restrict__ nodes,
                              const LeafRec<double>* __restrict__ leaves, BucketRec* __restrict__ recs) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    BucketRec R;
    node_box(root != 0 ? root : n, n, nodes, leaves, R);
    encode_bucket(R, 0, n);
    recs[0] = R;
}

// Smoothing length of each sample's nearest particle, or NaN when no particle
// lies within Kd * max(h) (then the sample has no neighbours).
template <int SRC>
__global__ void __launch_bounds__(kBlock) k_nearest_bucket(BucketView bv, Sampler<double, SRC> smp, double Kd,
                                                           double* __restrict__ ha) {
    i64 s;
    double p[3];
    if (smp.get(s, p[0], p[1], p[2])) ha[s] = nearest_h_bucket(bv, p, neighbour_bound2(bv, Kd));
}

template <int K, int STRAT, int MS, int MG, int MD, int MC, int EXACT, int SRC, bool WIDE = false>
__global__ void __launch_bounds__(kBlock) k_bucket_c32(BucketView bv, const double* __restrict__ rec, CatChunk c,
                                                       Sampler<double, SRC> smp, const double* __restrict__ ha_s,
                                                       float* __restrict__ out, u64* counter) {
    i64 s = 0;
    double p[3] = {0.0, 0.0, 0.0};
    const bool valid = smp.get(s, p[0], p[1], p[2]);
    double ha = 0.0;
    if constexpr (STRAT != PH_SCATTER) {
        if (valid) ha = ha_s[s];
    }
    u64 accepted = 0;
    __shared__ u32 blist[kQ][kBlock];
    bucket_c32_sample<K, STRAT, MS, MG, MD, MC, EXACT, WIDE>(bv, rec, c, valid, p, ha, out, smp.S, s, accepted,
                                                             &blist[0][threadIdx.x], kBlock);
    count_add(counter, accepted);
}

template <int K, int STRAT, int MS, int MG, int MD, int MC, int SRC>
__global__ void __launch_bounds__(kBlock) k_bucket_f64(BucketView bv, const double* __restrict__ rec, CatChunk c,
                                                       Sampler<double, SRC> smp, const double* __restrict__ ha_s,
                                                       double* __restrict__ out, u64* counter) {
    i64 s = 0;
    double p[3] = {0.0, 0.0, 0.0};
    const bool valid = smp.get(s, p[0], p[1], p[2]);
    double ha = 0.0;
    if constexpr (STRAT != PH_SCATTER) {
        if (valid) ha = ha_s[s];
    }
    u64 accepted = 0;
    __shared__ u32 blist[kQ][kBlock];
    bucket_f64_sample<K, STRAT, MS, MG, MD, MC>(bv, rec, c, valid, p, ha, out, smp.S, s, accepted,
                                                &blist[0][threadIdx.x], kBlock);
    count_add(counter, accepted);
}

inline unsigned sample_blocks(const GridAxes* ga, i64 S) {
    if (ga) {
        const i64 tiles = ((ga->nx + 3) / 4) * ((ga->ny + 3) / 4) * ((ga->nz + 3) / 4);
        return (unsigned)((tiles + 3) / 4);
    }
    return (unsigned)((S + kBlock - 1) / kBlock);
}

// Dynamic LDS per workgroup that limits a kernel to `occupancy` of its default
// workgroups per CU (0 when occupancy >= 1 or the limit needs more LDS than a
// workgroup may hold).
template <class F> size_t occupancy_pad(F kern, float occupancy) {
    if (occupancy >= 1.f) return 0;
    int per_cu = 0, dev = 0, lds_cu = 0, lds_block = 0;
    PH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0));
    PH_HIP(hipGetDevice(&dev));
    PH_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    PH_HIP(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const int want = std::max(1, (int)(per_cu * occupancy + 0.5f));
    if (want >= per_cu || lds_cu <= 0) return 0;
    hipFuncAttributes fa{};
    PH_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(kern)));
    const i64 per_block = (i64)(lds_cu / want / 1024) * 1024;  // `want` fit, one more does not
    const i64 pad = per_block - (i64)fa.sharedSizeBytes;
    if (pad <= 0 || per_block > lds_block) return 0;
    return (size_t)pad;
}

template <int K, int STRAT, int MS, int MG, int MD, int MC, int EXACT>
void launch_c32_t(ph_lbvh* B, const double* rec, const CatChunk& c, const PointList* pl, const GridAxes* ga, i64 S,
                  const double* ha, void* out, u64* counters, hipStream_t s, hipEvent_t* ev) {
    const BucketView bv = bucket_view(B);
    const unsigned g = sample_blocks(ga, S);
    const LaunchCfg cfg = tl_launch;
    if (ga) {
        auto kern = k_bucket_c32<K, STRAT, MS, MG, MD, MC, EXACT, 1>;
        const size_t pad = occupancy_pad(kern, cfg.occupancy);
        if (ev) PH_HIP(hipEventRecord(ev[0], s));
        kern<<<g, kBlock, pad, s>>>(bv, rec, c, Sampler<double, 1>{*ga, S, point_xcd(cfg), point_order(cfg)}, ha,
                                    static_cast<float*>(out), counters);
    } else {
        // point lists past 32-bit offsets: the 64-bit-address variant (grids
        // then take the tile walk, launch_bucket_ks)
        auto kern = c32_wide(B->n, c.rw) ? k_bucket_c32<K, STRAT, MS, MG, MD, MC, EXACT, 0, true>
                                         : k_bucket_c32<K, STRAT, MS, MG, MD, MC, EXACT, 0>;
        const size_t pad = occupancy_pad(kern, cfg.occupancy);
        if (ev) PH_HIP(hipEventRecord(ev[0], s));
        kern<<<g, kBlock, pad, s>>>(bv, rec, c, Sampler<double, 0>{*pl, S, point_xcd(cfg)}, ha,
                                    static_cast<float*>(out), counters);
    }
    PH_HIP(hipGetLastError());
    if (ev) PH_HIP(hipEventRecord(ev[1], s));
}

template <int K, int STRAT, int MS, int MG, int MD, int MC>
void launch_f64_t(ph_lbvh* B, const double* rec, const CatChunk& c, const PointList* pl, const GridAxes* ga, i64 S,
                  const double* ha, void* out, u64* counters, hipStream_t s, hipEvent_t* ev) {
    const BucketView bv = bucket_view(B);
    const unsigned g = sample_blocks(ga, S);
    const LaunchCfg cfg = tl_launch;
    if (ga) {
        auto kern = k_bucket_f64<K, STRAT, MS, MG, MD, MC, 1>;
        const size_t pad = occupancy_pad(kern, cfg.occupancy);
        if (ev) PH_HIP(hipEventRecord(ev[0], s));
        kern<<<g, kBlock, pad, s>>>(bv, rec, c, Sampler<double, 1>{*ga, S, point_xcd(cfg), point_order(cfg)}, ha,
                                    static_cast<double*>(out), counters);
    } else {
        auto kern = k_bucket_f64<K, STRAT, MS, MG, MD, MC, 0>;
        const size_t pad = occupancy_pad(kern, cfg.occupancy);
        if (ev) PH_HIP(hipEventRecord(ev[0], s));
        kern<<<g, kBlock, pad, s>>>(bv, rec, c, Sampler<double, 0>{*pl, S, point_xcd(cfg)}, ha,
                                    static_cast<double*>(out), counters);
    }
    PH_HIP(hipGetLastError());
    if (ev) PH_HIP(hipEventRecord(ev[1], s));
}

template <int K, int STRAT>
void launch_bucket_ks(ph_lbvh* B, const CatChunk& c, int precision, const PointList* pl, const GridAxes* ga, i64 S,
                      void* out, u64* counters, hipStream_t s, hipEvent_t* ev, const CallWorkspace* ws) {
    const i64 n = B->n;
    const bool small = c.ng + c.nd + c.nc == 0 && c.ns <= 2;
    // buffers first (the pool's allocations wait for the stream), then kernels;
    // a caller's workspace replaces the tree's per-call buffers
    int ncol = 0;
    CatChunk c32{};
    double* recp = nullptr;
    double* hap = nullptr;
    if (precision == PH_PREC_FP64) {
        if (!ws) B->rec.alloc(sizeof(double) * (size_t)c.rw * (size_t)n);
        recp = ws ? ws->rec : B->rec.as<double>();
    } else {
        c32 = compact_c32(c, ncol);
        if (!ws) B->rec32.alloc(sizeof(double) * (size_t)c32.rw * (size_t)(n + 1));  // + the zero record
        recp = ws ? ws->rec : B->rec32.as<double>();
    }
    if (STRAT != PH_SCATTER) {
        if (!ws) B->ha.alloc(sizeof(double) * (size_t)S);
        hap = ws ? ws->ha : B->ha.as<double>();
    }
    // nearest-particle smoothing length of every sample (gather / symmetric)
    const double* ha = nullptr;
    if (STRAT != PH_SCATTER) {
        const BucketView bv = nearest_view(B);
        const unsigned g = sample_blocks(ga, S);
        const int xm = point_xcd(tl_launch);
        const double Kd = support_of<K>();
        if (ga && tl_launch.tile) launch_nearest_tile(B, *ga, S, Kd, hap, s);
        else if (ga)
            k_nearest_bucket<1><<<g, kBlock, 0, s>>>(bv, Sampler<double, 1>{*ga, S, xm, point_order(tl_launch)}, Kd,
                                                     hap);
        else k_nearest_bucket<0><<<g, kBlock, 0, s>>>(bv, Sampler<double, 0>{*pl, S, xm}, Kd, hap);
        PH_HIP(hipGetLastError());
        ha = hap;
    }
    if (precision == PH_PREC_FP64) {
        k_pack_records<double><<<grid_for(n), kBlock, 0, s>>>(B->cols.as<double>(), n, c, recp);
        PH_HIP(hipGetLastError());
        if (ga && tl_launch.tile) {  // one walk per wave of grid points
            launch_tile(B, recp, c, precision, STRAT, *ga, S, ha, out, counters, s, ev);
            return;
        }
        if (small) launch_f64_t<K, STRAT, 2, 0, 0, 0>(B, recp, c, pl, ga, S, ha, out, counters, s, ev);
        else launch_f64_t<K, STRAT, kMaxS, kMaxG, kMaxD, kMaxC>(B, recp, c, pl, ga, S, ha, out, counters, s, ev);
        return;
    }
    k_pack_c32<<<grid_for(n + 1), kBlock, 0, s>>>(B->cols.as<double>(), n, c32, ncol, c32.rw, recp);
    PH_HIP(hipGetLastError());
    if (ga && (tl_launch.tile || c32_wide(n, c32.rw))) {  // one walk per wave of grid points
        launch_tile(B, recp, c32, precision, STRAT, *ga, S, ha, out, counters, s, ev);
        return;
    }
    if (small) launch_c32_t<K, STRAT, 2, 0, 0, 0, 0>(B, recp, c32, pl, ga, S, ha, out, counters, s, ev);
    else if (c2_shared(c32))  // density, scalar A, grad A, div v, curl v
        launch_c32_t<K, STRAT, 2, 1, 1, 1, 2>(B, recp, c32, pl, ga, S, ha, out, counters, s, ev);
    else launch_c32_t<K, STRAT, kMaxS, kMaxG, kMaxD, kMaxC, 0>(B, recp, c32, pl, ga, S, ha, out, counters, s, ev);
}

template <int K>
void launch_bucket_k(ph_lbvh* B, const CatChunk& c, int strategy, int precision, const PointList* pl,
                     const GridAxes* ga, i64 S, void* out, u64* counters, hipStream_t s, hipEvent_t* ev,
                     const CallWorkspace* ws) {
    switch (strategy) {
    case PH_GATHER: launch_bucket_ks<K, PH_GATHER>(B, c, precision, pl, ga, S, out, counters, s, ev, ws); break;
    case PH_SCATTER: launch_bucket_ks<K, PH_SCATTER>(B, c, precision, pl, ga, S, out, counters, s, ev, ws); break;
    default: launch_bucket_ks<K, PH_SYMMETRIC>(B, c, precision, pl, ga, S, out, counters, s, ev, ws); break;
    }
}

}  // namespace

i64 bucket_record_doubles(const CatChunk& c, int precision) {
    if (precision == PH_PREC_FP64) return c.rw;
    int ncol = 0;
    return compact_c32(c, ncol).rw;
}

i64 bucket_records(i64 n, int precision) { return precision == PH_PREC_FP64 ? n : n + 1; }

bool buckets_supported(const ph_lbvh* B) {
    return B->dims == 3 && B->precision == PH_PREC_FP64 && B->n <= kMaxBucketLeaves;
}

// Builds one to four bucket trees of B (bucket sizes bmax[0] < bmax[1] < ...)
// into recs[t], nrec[t]: one flag pass and one scan for the trees of more than
// one leaf per bucket, one read-back of their record counts (with the build's
// refit check when the build left it here), one emit pass for all of them.
void build_bucket_trees(ph_lbvh* B, hipStream_t s, int ntrees, const int* bmax, DevBuf* const* recs, i64* nrec,
                        const u32* refit_bad) {
    const i64 n = B->n;
    if (n > kMaxBucketLeaves) throw Error{PH_ERR_CAPACITY, "the bucket traversal supports at most 2^30 particles"};
    bool split[kMaxTrees] = {false, false, false, false};  // split[t + 1] implies split[t]
    for (int t = 0; t < ntrees; ++t) {
        split[t] = n > bmax[t];
        nrec[t] = 1;
    }
    // the one-leaf tree has 2n - 1 records in closed form
    const int t0 = ntrees > 0 && bmax[0] == kBucketLeaf ? 1 : 0;
    if (t0 && split[0]) nrec[0] = 2 * n - 1;
    const int nf = ntrees - t0;  // trees with flags
    // buffers first.  Flags of the trees in one array, scanned as one: per
    // tree the big splits (n entries: n - 1 flags + the total) and the bucket
    // starts (n + 1 entries: n flags + the total), then one closing entry; a
    // segment's exclusive prefix is its scanned value minus that at its start.
    const bool flagged = nf > 0 && split[t0];
    const i64 per_tree = 2 * n + 1, m = nf * per_tree + 1;
    DevBuf flags, tmp, info, lcb;
    if (flagged) flags.alloc(sizeof(u32) * (size_t)m);
    if (split[0]) lcb.alloc(sizeof(i32) * 4 * (size_t)(n - 1));
    if (flagged || refit_bad) info.alloc(sizeof(u32) * 12);
    u32* f = flags.as<u32>();
    FlagTrees ft{};
    ft.ntrees = nf;
    for (int t = 0; t < nf; ++t) {
        ft.big_s[t] = f + t * per_tree;
        ft.bs[t] = f + t * per_tree + n;
        ft.bmax[t] = bmax[t0 + t];
    }
    // L(u) pointers (before the read-back, so the host waits for them too)
    i32* lcp = nullptr;  // up, then lc, of the last round
    if (split[0]) {
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
        lcp = a;
    }
    u32 h[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (flagged || refit_bad) {
        if (flagged) {
            PH_HIP(hipMemsetAsync(f, 0, sizeof(u32) * (size_t)m, s));
            k_bucket_flags<<<grid_for(n - 1), kBlock, 0, s>>>(n, B->left.as<i32>(), B->range.as<i32>(), ft);
            PH_HIP(hipGetLastError());
            exclusive_sum_u32(f, m, tmp, s);
        }
        k_bucket_info<<<1, 1, 0, s>>>(flagged ? f : nullptr, n, nf, refit_bad, info.as<u32>());
        PH_HIP(hipGetLastError());
        PH_HIP(hipMemcpyAsync(h, info.p, sizeof(h), hipMemcpyDeviceToHost, s));
        PH_HIP(hipStreamSynchronize(s));
        if (h[11] != 0)
            throw Error{PH_ERR_INTERNAL,
                        "LinearBVH refit: " + std::to_string(h[11]) + " internal nodes were not reached twice"};
    }
    for (int t = 0; t < nf; ++t)
        if (split[t0 + t]) nrec[t0 + t] = (i64)(h[2 * t + 1] - h[2 * t]) + (i64)(h[2 * t + 2] - h[2 * t + 1]);
    for (int t = 0; t < ntrees; ++t)  // record indices and escape links are int32
        if (nrec[t] > (i64)INT32_MAX - 16) throw Error{PH_ERR_CAPACITY, "bucket tree of more than 2^31 - 16 records"};
    // one record of padding: the tile walks load a record and its successor together
    for (int t = 0; t < ntrees; ++t) recs[t]->alloc(sizeof(BucketRec) * (size_t)(nrec[t] + 1));
    for (int t = 0; t < ntrees; ++t)
        if (!split[t]) {  // the whole tree is one bucket
            k_bucket_root<<<1, 1, 0, s>>>(n, B->root, B->nodes.as<NodeRec<double>>(),
                                          B->leaves.as<LeafRec<double>>(), recs[t]->as<BucketRec>());
            PH_HIP(hipGetLastError());
        }
    if (split[0]) {
        EmitTrees et{};
        et.leaf_recs = t0 ? recs[0]->as<BucketRec>() : nullptr;
        for (int t = 0; t < nf; ++t)
            if (split[t0 + t]) {
                et.t[et.ntrees++] = EmitTree{ft.big_s[t], ft.bs[t], h[2 * t], h[2 * t + 1],
                                             recs[t0 + t]->as<BucketRec>(), bmax[t0 + t]};
            }
        k_bucket_emit<<<grid_for(n - 1), kBlock, 0, s>>>(n, B->left.as<i32>(), B->range.as<i32>(), lcp,
                                                        lcp + (n - 1), B->nodes.as<NodeRec<double>>(),
                                                        B->leaves.as<LeafRec<double>>(), et);
        PH_HIP(hipGetLastError());
    }
    // flags / tmp / info are released on return: into the context's cache,
    // reusable once the stream has drained, or freed directly (wait for the kernels)
    if (!tl_cache) PH_HIP(hipStreamSynchronize(s));
}

void ensure_buckets(ph_lbvh* B, hipStream_t s, const u32* refit_bad) {
    // The bucket trees of an LBVH, built together: buckets of one leaf (the
    // nearest-h searches: the LBVH itself in preorder with float32 boxes; left
    // out when its 2n - 1 records would not fit int32 indices), of at most
    // kBucketSmall leaves (the per-lane kernels, the batch queue, adjoints and
    // queries), of at most kBucket and kBucketLarge (the grid tile walks).
    if (B->buckets_ready) return;
    const bool leaf_tree = 2 * B->n - 1 <= (i64)INT32_MAX - 16;
    const int bmax[4] = {kBucketLeaf, kBucketSmall, kBucket, kBucketLarge};
    DevBuf* recs[4] = {&B->brecs1, &B->brecs, &B->brecs8, &B->brecs16};
    i64 nrec[4] = {1, 1, 1, 1};
    const int skip = leaf_tree ? 0 : 1;
    build_bucket_trees(B, s, 4 - skip, bmax + skip, recs + skip, nrec + skip, refit_bad);
    B->nrec1 = leaf_tree ? nrec[0] : 0;
    B->nrec = nrec[1];
    B->nrec8 = nrec[2];
    B->nrec16 = nrec[3];
    B->buckets1_ready = leaf_tree;
    B->buckets_ready = true;
}

void launch_bucket(ph_lbvh* B, const CatChunk& c, int strategy, int precision, const PointList* pl, const GridAxes* ga,
                   i64 S, void* out, u64* counters, hipStream_t s, hipEvent_t* ev, const CallWorkspace* ws) {
    if (S <= 0) return;
    if (B->precision != PH_PREC_FP64) throw Error{PH_ERR_ARGUMENT, "the bucket traversal needs a float64 tree"};
    ensure_buckets(B, s);
    switch (B->kernel) {
    case PH_KERNEL_M4: launch_bucket_k<PH_KERNEL_M4>(B, c, strategy, precision, pl, ga, S, out, counters, s, ev, ws); break;
    case PH_KERNEL_M5: launch_bucket_k<PH_KERNEL_M5>(B, c, strategy, precision, pl, ga, S, out, counters, s, ev, ws); break;
    case PH_KERNEL_M6: launch_bucket_k<PH_KERNEL_M6>(B, c, strategy, precision, pl, ga, S, out, counters, s, ev, ws); break;
    case PH_KERNEL_C2: launch_bucket_k<PH_KERNEL_C2>(B, c, strategy, precision, pl, ga, S, out, counters, s, ev, ws); break;
    case PH_KERNEL_C4: launch_bucket_k<PH_KERNEL_C4>(B, c, strategy, precision, pl, ga, S, out, counters, s, ev, ws); break;
    default: launch_bucket_k<PH_KERNEL_C6>(B, c, strategy, precision, pl, ga, S, out, counters, s, ev, ws); break;
    }
}

}  // namespace ph
