// ph_tile.hip -- grid interpolation with one bucket-tree walk per wave, in the
// FP32-compensated precision (k_tile_c32) and in float64 (k_tile_f64).
//
// A wave holds a 4 x 4 x 4 block of grid points.  Instead of 64 independent
// walks (each lane loading its own records, leaves and particle records: 64
// cache lines per load instruction), the wave walks the bucket tree once and
// keeps a bucket when any lane's conservative test keeps it: records are read
// at wave-uniform addresses.  The leaves of the kept buckets are staged, 64 at
// a time, in LDS together with their particle records (one coalesced load per
// lane); every lane then applies the reference's float64 leaf test to all of
// them and accumulates its accepted leaves in lockstep from LDS.
//
// Exactness: the union of the lanes' kept buckets contains each lane's own
// kept buckets, hence all of its accepted leaves; staged leaves are in
// increasing Morton index (preorder), each lane accepts exactly the leaves the
// reference accepts (same float64 test) and accumulates them in that order with
// the same arithmetic (AccC) as the per-lane kernel k_bucket_c32.  The two
// kernels therefore give identical results, bit for bit.
#include <algorithm>

#include "ph_bucket_dev.h"

namespace ph {
namespace {

// A double in LDS (address space 3).
typedef __attribute__((address_space(3))) double LdsDouble;

// Record width of the C2 layout {vi, s0, s1, d0, d1, d2}.
constexpr int kC2Rec = 6;

// Per wave, 64 staged leaves and one zero slot of `slot` doubles each, then 64
// leaf indices per wave.  FP32C: slot = 4 + rw (x, y, z, (K h)^2, record);
// float64: slot = 5 + rw (x, y, z, (K h)^2, h, record).
inline size_t tile_lds_bytes(int slot) {
    return (size_t)(kBlock / kWave) * (kWave + 1) * (size_t)slot * sizeof(double) +
           (size_t)(kBlock / kWave) * kWave * sizeof(u32);
}

// The next batch of the wave's union walk: the leaves of the kept buckets
// (kept when any live lane's conservative test keeps them), in preorder = in
// increasing leaf index, written to lidx until one more bucket would exceed 64
// leaves.  Returns their count; 0 once the walk is done.  i and steps are
// wave-uniform; the walk visits every record at most once.
template <int STRAT>
__device__ __attribute__((always_inline)) int union_batch(const BucketView& bv, const BucketRec* __restrict__ recs,
                                                          const Probe<STRAT>& pr, u64 live_mask, i32& i, i32& steps,
                                                          u32* lidx, int lane) {
    int nl = 0;
    while (i < bv.nrec && steps > 0) {
        --steps;
        const BucketRec R = recs[i];
        const bool keep = (pr.keep_mask(R) & live_mask) != 0;
        if (R.link >= 0) {
            i = __builtin_amdgcn_readfirstlane(keep ? i + 1 : R.link);
            continue;
        }
        if (keep) {
            const int cnt = bucket_cnt(R.link);
            if (nl + cnt > kWave) {  // starts the next batch
                ++steps;
                break;
            }
            if (lane < cnt) lidx[nl + lane] = (u32)(bucket_lo(R.link) + lane);
            nl += cnt;
        }
        i = __builtin_amdgcn_readfirstlane(i + 1);
    }
    return nl;
}

// The reference's float64 leaf test of the nl staged leaves (x, y, z and
// (K h_b)^2 at slot offsets 0..3), per lane: d2 <= (K h_a)^2 or d2 <=
// (K h_b)^2, both compares evaluated (no control flow).  The leaves are tested
// from the last to the first and shifted into 32-bit halves, so leaf j lands
// at bit j.
template <int STRAT>
__device__ __attribute__((always_inline)) u64 test_staged(const LdsDouble* st, int slot, int nl, const double p[3],
                                                          double r2a) {
    auto take = [&](int j) -> u32 {
        const LdsDouble* d = st + j * slot;
        const double dx = p[0] - d[0], dy = p[1] - d[1], dz = p[2] - d[2];
        const double d2 = (dx * dx + dy * dy) + dz * dz;  // d2_point's order
        if constexpr (STRAT == PH_GATHER) return d2 <= r2a ? 1u : 0u;
        if constexpr (STRAT == PH_SCATTER) return d2 <= d[3] ? 1u : 0u;
        return (d2 <= d[3] ? 1u : 0u) | (d2 <= r2a ? 1u : 0u);
    };
    u32 mlo = 0, mhi = 0;
    const int nlo = nl < 32 ? nl : 32;
    for (int j = nl - 1; j >= 32; --j) mhi = mhi + mhi + take(j);
    for (int j = nlo - 1; j >= 0; --j) mlo = mlo + mlo + take(j);
    return ((u64)mhi << 32) | mlo;
}

// Lockstep iterations of a wave: the largest accepted count of its lanes.
__device__ inline int lockstep_count(u64 mask) {
    int nit = __popcll(mask);
    for (int d = kWave / 2; d > 0; d >>= 1) nit = max(nit, __shfl_xor(nit, d, kWave));
    return __builtin_amdgcn_readfirstlane(nit);
}

// recs is bv.recs again as a restrict-qualified argument: the records are then
// known not to be written by the kernel, and the wave-uniform record loads of
// the walk become scalar loads (one round trip per record instead of two vector
// loads).
template <int K, int STRAT, int MS, int MG, int MD, int MC, int EXACT>
__global__ void __launch_bounds__(kBlock) k_tile_c32(BucketView bv, const BucketRec* __restrict__ recs,
                                                     const double* __restrict__ rec, CatChunk c,
                                                     Sampler<double, 1> smp, const double* __restrict__ ha_s,
                                                     float* __restrict__ out, u64* counter) {
    extern __shared__ double tile_smem[];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int rw = EXACT == 2 ? kC2Rec : c.rw;  // the C2 layout: constant LDS offsets
    const int slot = 4 + rw;  // doubles per staged leaf
    // an LDS-typed pointer: 32-bit offsets in the inner loops
    LdsDouble* st = (LdsDouble*)tile_smem + wv * (kWave + 1) * slot;
    u32* lidx = reinterpret_cast<u32*>(tile_smem + (size_t)(kBlock / kWave) * (kWave + 1) * slot) + wv * kWave;

    i64 s = 0;
    double p[3] = {0.0, 0.0, 0.0};
    const bool valid = smp.get(s, p[0], p[1], p[2]);
    const double Kd = support_of<K>();
    double ha = 0.0;
    if constexpr (STRAT != PH_SCATTER) {
        if (valid) ha = ha_s[s];
    }
    // no particle within K max(h): no neighbours (NaN outputs)
    const bool none = STRAT != PH_SCATTER && !(ha == ha);
    const bool live_lane = valid && !none;
    if (!live_lane) ha = 0.0;
    const double ra = Kd * ha, r2a = ra * ra;  // float64, as the reference
    Probe<STRAT> pr;
    pr.init(p, Kd, ra);
    const double iha = STRAT == PH_SCATTER || !live_lane ? 0.0 : 1.0 / ha;
    const double iha2 = iha * iha;
    const float inv_ha = (float)iha;
    const float sig = sigma3<float, K>();
    const float Wa_norm = sig * inv_ha * inv_ha * inv_ha, Ga_norm = Wa_norm * inv_ha;
    // the zero slot read by lanes without a pending neighbour: position 0,
    // record {m / rho = 0, 1 / h = 1} and zero columns, so every term is 0
    if (lane < slot)
        st[kWave * slot + lane] = lane == 4 ? __builtin_bit_cast(double, make_float2(0.f, 1.f)) : 0.0;

    AccC<K, STRAT, MS, MG, MD, MC, EXACT> acc;
    acc.init();
    u64 accepted = 0;
    const u64 live_mask = __ballot(live_lane);
    i32 i = live_mask ? 0 : bv.nrec;  // wave-uniform from here on
    i32 steps = bv.nrec + 8;  // the union walk visits every record at most once (nrec < 2^27)
    for (;;) {
        const int nl = union_batch<STRAT>(bv, recs, pr, live_mask, i, steps, lidx, lane);
        if (nl == 0) break;
        wave_sync();
        // stage leaf lane, the square of its radius (K h)^2 and its record
        // (consecutive leaves: coalesced loads)
        if (lane < nl) {
            const u32 k = lidx[lane];
            const LeafRec<double> L = bv.leaves[k];
            LdsDouble* d = st + lane * slot;
            d[0] = L.x;
            d[1] = L.y;
            d[2] = L.z;
            const double rb = Kd * L.h;  // float64, as the reference
            d[3] = rb * rb;
            const double* r = rec + (size_t)k * rw;
            for (int q = 0; q < rw; ++q) d[4 + q] = r[q];
        }
        wave_sync();
        u64 mask = live_lane ? test_staged<STRAT>(st, slot, nl, p, r2a) : 0;
        accepted += __popcll(mask);
        // lockstep: one neighbour per lane and iteration, lowest bit first
        // (increasing leaf index); a lane with none left adds zeros
        const int nit = lockstep_count(mask);
        for (int it = 0; it < nit; ++it) {
            const bool live = mask != 0;
            const int b = __builtin_ctzll(mask | (1ull << 63));
            mask &= mask - 1;
            const LdsDouble* d = st + (live ? b : kWave) * slot;
            const double dx = p[0] - d[0], dy = p[1] - d[1], dz = p[2] - d[2];
            acc.add_r(live, c, [&](int pos) { return d[4 + pos]; }, dx, dy, dz, (dx * dx + dy * dy) + dz * dz, iha2,
                      Wa_norm, Ga_norm);
        }
        wave_sync();  // every lane is done with the staged leaves
        if (!(i < bv.nrec && steps > 0)) break;
    }
    if (valid) acc.finish(c, out, smp.S, s);
    count_add(counter, accepted);
}

// Float64 grids with the same per-wave union walk, staging and float64 leaf
// test as k_tile_c32, accumulated with the float64 arithmetic of the per-lane
// kernel k_bucket_f64 (Acc::add_live: add()'s operations in add()'s order), so
// the outputs equal k_bucket_f64's and the reference's bit for bit.  rec holds
// the float64 records {m, rho, columns} of chunk c.
template <int K, int STRAT, int MS, int MG, int MD, int MC>
__global__ void __launch_bounds__(kBlock) k_tile_f64(BucketView bv, const BucketRec* __restrict__ recs,
                                                     const double* __restrict__ rec, CatChunk c,
                                                     Sampler<double, 1> smp, const double* __restrict__ ha_s,
                                                     double* __restrict__ out, u64* counter) {
    extern __shared__ double tile_smem[];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int rw = c.rw;
    const int slot = 5 + rw;  // x, y, z, (K h)^2, h, record
    LdsDouble* st = (LdsDouble*)tile_smem + wv * (kWave + 1) * slot;
    u32* lidx = reinterpret_cast<u32*>(tile_smem + (size_t)(kBlock / kWave) * (kWave + 1) * slot) + wv * kWave;

    i64 s = 0;
    double p[3] = {0.0, 0.0, 0.0};
    const bool valid = smp.get(s, p[0], p[1], p[2]);
    const double Kd = support_of<K>();
    double ha = 0.0;
    if constexpr (STRAT != PH_SCATTER) {
        if (valid) ha = ha_s[s];
    }
    const bool none = STRAT != PH_SCATTER && !(ha == ha);
    const bool live_lane = valid && !none;
    if (!live_lane) ha = 0.0;
    const double ra = Kd * ha, r2a = ra * ra;
    Probe<STRAT> pr;
    pr.init(p, Kd, ra);
    // the zero slot: position 0, h = 1, record {m = 0, rho = 1, columns 0}
    if (lane < slot) st[kWave * slot + lane] = (lane == 4 || lane == 6) ? 1.0 : 0.0;

    Acc<double, K, STRAT, MS, MG, MD, MC> acc;
    acc.init();
    u64 accepted = 0;
    const u64 live_mask = __ballot(live_lane);
    i32 i = live_mask ? 0 : bv.nrec;
    i32 steps = bv.nrec + 8;
    for (;;) {
        const int nl = union_batch<STRAT>(bv, recs, pr, live_mask, i, steps, lidx, lane);
        if (nl == 0) break;
        wave_sync();
        if (lane < nl) {
            const u32 k = lidx[lane];
            const LeafRec<double> L = bv.leaves[k];
            LdsDouble* d = st + lane * slot;
            d[0] = L.x;
            d[1] = L.y;
            d[2] = L.z;
            const double rb = Kd * L.h;  // radius2's float64 (K h)^2
            d[3] = rb * rb;
            d[4] = L.h;
            const double* r = rec + (size_t)k * rw;
            for (int q = 0; q < rw; ++q) d[5 + q] = r[q];
        }
        wave_sync();
        u64 mask = live_lane ? test_staged<STRAT>(st, slot, nl, p, r2a) : 0;
        accepted += __popcll(mask);
        const int nit = lockstep_count(mask);
        for (int it = 0; it < nit; ++it) {
            const bool live = mask != 0;
            const int b = __builtin_ctzll(mask | (1ull << 63));
            mask &= mask - 1;
            const LdsDouble* d = st + (live ? b : kWave) * slot;
            const double dx = p[0] - d[0], dy = p[1] - d[1], dz = p[2] - d[2];
            acc.add_live(live, c, [&](int pos) { return (double)d[5 + pos]; }, dx, dy, dz,
                         (dx * dx + dy * dy) + dz * dz, ha, (double)d[4]);
        }
        wave_sync();
        if (!(i < bv.nrec && steps > 0)) break;
    }
    if (valid) acc.finish(c, out, smp.S, s);
    count_add(counter, accepted);
}

// Nearest-particle smoothing length of every grid point, one wave per 4x4x4
// tile.  The wave descends the bucket tree once (scalar loads) toward the
// centre of its points; the 64 leaves around the bucket it reaches (Morton
// neighbours) are staged in LDS and offered to every lane, which seeds each
// lane's bound with a nearby leaf; each lane then runs the pruned walk of
// nearest_h_bucket from the root.  The result is the lexicographic minimum of
// (d2, leaf index) over the leaves within the bound whatever leaves are offered
// first, so it equals the per-lane search's (k_nearest_bucket) bit for bit.
constexpr int kNnCand = 64;
__global__ void __launch_bounds__(kBlock) k_nearest_tile(BucketView bv, const BucketRec* __restrict__ recs,
                                                         Sampler<double, 1> smp, double Kd,
                                                         double* __restrict__ ha, i64* __restrict__ nn) {
    __shared__ LeafRec<double> cand[kBlock / kWave][kNnCand];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    i64 s = 0;
    double p[3] = {0.0, 0.0, 0.0};
    const bool valid = smp.get(s, p[0], p[1], p[2]);
    if (__ballot(valid) == 0) return;  // wave-uniform
    // centre of the wave's valid points (all lanes hold the same value)
    float c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float lo = valid ? (float)p[d] : INFINITY, hi = valid ? (float)p[d] : -INFINITY;
        for (int o = kWave / 2; o > 0; o >>= 1) {
            lo = fminf(lo, __shfl_xor(lo, o, kWave));
            hi = fmaxf(hi, __shfl_xor(hi, o, kWave));
        }
        c[d] = 0.5f * lo + 0.5f * hi;
    }
    // greedy descent toward the centre; decisions made wave-uniform
    i32 i = 0;
    i32 link = -1;  // bucket 0, one leaf: a valid fallback
    for (i64 steps = bv.nrec; steps > 0; --steps) {
        const BucketRec R = recs[i];
        if (R.link < 0) {
            link = R.link;
            break;
        }
        const i32 l = i + 1;
        const BucketRec RL = recs[l];
        const i32 r = RL.link >= 0 ? RL.link : l + 1;
        const BucketRec RR = recs[r];
        i = __builtin_amdgcn_readfirstlane(box_d2_32(RR, c[0], c[1], c[2]) < box_d2_32(RL, c[0], c[1], c[2]) ? r
                                                                                                             : l);
    }
    const i64 n = bv.n;
    const int m = (int)(n < kNnCand ? n : kNnCand);
    i64 first = bucket_lo(link) + bucket_cnt(link) / 2 - kNnCand / 2;
    first = first < 0 ? 0 : (first > n - m ? n - m : first);
    if (lane < m) cand[wv][lane] = bv.leaves[first + lane];
    wave_sync();
    Nearest nr{neighbour_bound2(bv, Kd)};
    for (int j = 0; j < m; ++j) nr.offer(p, cand[wv][j], first + j);
    nearest_walk(bv, p, nr);
    if (valid) {
        ha[s] = nr.h(bv);
        if (nn) nn[s] = nr.bk != Nearest::kNone ? nr.bk : -1;
    }
}

// Launches a tile kernel over the grid's 4x4x4 tiles (one wave each), with
// dyn bytes of dynamic LDS, limited to the configured occupancy.
template <class Kern, class OutT>
void launch_tile_kernel(Kern kern, size_t dyn, const BucketView& bv, const double* rec, const CatChunk& c,
                        const GridAxes& ga, i64 S, const double* ha, OutT* out, u64* counters, hipStream_t s,
                        hipEvent_t* ev) {
    const LaunchCfg cfg = tl_launch;
    size_t pad = 0;
    if (cfg.occupancy < 1.f) {  // as occupancy_pad, with this kernel's dynamic LDS
        int per_cu = 0, dev = 0, lds_cu = 0, lds_block = 0;
        PH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, dyn));
        PH_HIP(hipGetDevice(&dev));
        PH_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
        PH_HIP(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const int want = std::max(1, (int)(per_cu * cfg.occupancy + 0.5f));
        if (want < per_cu && lds_cu > 0) {
            hipFuncAttributes fa{};
            PH_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(kern)));
            const i64 per_block = (i64)(lds_cu / want / 1024) * 1024;
            const i64 extra = per_block - (i64)fa.sharedSizeBytes - (i64)dyn;
            if (extra > 0 && per_block <= lds_block) pad = (size_t)extra;
        }
    }
    const i64 tiles = ((ga.nx + 3) / 4) * ((ga.ny + 3) / 4) * ((ga.nz + 3) / 4);
    const unsigned g = (unsigned)((tiles + 3) / 4);
    if (ev) PH_HIP(hipEventRecord(ev[0], s));
    kern<<<g, kBlock, dyn + pad, s>>>(bv, bv.recs, rec, c, Sampler<double, 1>{ga, S, point_xcd(cfg), point_order(cfg)},
                                      ha, out, counters);
    PH_HIP(hipGetLastError());
    if (ev) PH_HIP(hipEventRecord(ev[1], s));
}

template <int K, int STRAT>
void launch_tile_ks(const BucketView& bv, const double* rec, const CatChunk& c, bool f64, const GridAxes& ga, i64 S,
                    const double* ha, void* out, u64* counters, hipStream_t s, hipEvent_t* ev) {
    // the per-lane kernels' selection (ph_bucket.hip), so the arithmetic is the same
    const bool small = c.ng + c.nd + c.nc == 0 && c.ns <= 2;
    if (f64) {
        double* o = static_cast<double*>(out);
        const size_t dyn = tile_lds_bytes(5 + c.rw);
        if (small) launch_tile_kernel(k_tile_f64<K, STRAT, 2, 0, 0, 0>, dyn, bv, rec, c, ga, S, ha, o, counters, s, ev);
        else
            launch_tile_kernel(k_tile_f64<K, STRAT, kMaxS, kMaxG, kMaxD, kMaxC>, dyn, bv, rec, c, ga, S, ha, o,
                               counters, s, ev);
        return;
    }
    float* o = static_cast<float*>(out);
    if (small)
        launch_tile_kernel(k_tile_c32<K, STRAT, 2, 0, 0, 0, 0>, tile_lds_bytes(4 + c.rw), bv, rec, c, ga, S, ha, o,
                           counters, s, ev);
    else if (c2_shared(c) && c.rw == kC2Rec)
        launch_tile_kernel(k_tile_c32<K, STRAT, 2, 1, 1, 1, 2>, tile_lds_bytes(4 + kC2Rec), bv, rec, c, ga, S, ha, o,
                           counters, s, ev);
    else
        launch_tile_kernel(k_tile_c32<K, STRAT, kMaxS, kMaxG, kMaxD, kMaxC, 0>, tile_lds_bytes(4 + c.rw), bv, rec, c,
                           ga, S, ha, o, counters, s, ev);
}

template <int K>
void launch_tile_k(int strategy, const BucketView& bv, const double* rec, const CatChunk& c, bool f64,
                   const GridAxes& ga, i64 S, const double* ha, void* out, u64* counters, hipStream_t s,
                   hipEvent_t* ev) {
    switch (strategy) {
    case PH_GATHER: launch_tile_ks<K, PH_GATHER>(bv, rec, c, f64, ga, S, ha, out, counters, s, ev); break;
    case PH_SCATTER: launch_tile_ks<K, PH_SCATTER>(bv, rec, c, f64, ga, S, ha, out, counters, s, ev); break;
    default: launch_tile_ks<K, PH_SYMMETRIC>(bv, rec, c, f64, ga, S, ha, out, counters, s, ev); break;
    }
}

}  // namespace

void launch_nearest_tile(ph_lbvh* B, const GridAxes& ga, i64 S, double Kd, double* ha, hipStream_t s, i64* nn) {
    ensure_buckets(B, s);
    const BucketView bv = bucket_view(B);
    const LaunchCfg cfg = tl_launch;
    const i64 tiles = ((ga.nx + 3) / 4) * ((ga.ny + 3) / 4) * ((ga.nz + 3) / 4);
    const unsigned g = (unsigned)((tiles + 3) / 4);
    hipLaunchKernelGGL(k_nearest_tile, dim3(g), dim3(kBlock), 0, s, bv, bv.recs,
                       Sampler<double, 1>{ga, S, point_xcd(cfg), point_order(cfg)}, Kd, ha, nn);
    PH_HIP(hipGetLastError());
}

void launch_tile(ph_lbvh* B, const double* rec, const CatChunk& c, int precision, int strategy, const GridAxes& ga,
                 i64 S, const double* ha, void* out, u64* counters, hipStream_t s, hipEvent_t* ev) {
    ensure_buckets(B, s);
    const BucketView bv = bucket_view(B, kBucket);
    const bool f64 = precision == PH_PREC_FP64;
    switch (B->kernel) {
    case PH_KERNEL_M4: launch_tile_k<PH_KERNEL_M4>(strategy, bv, rec, c, f64, ga, S, ha, out, counters, s, ev); break;
    case PH_KERNEL_M5: launch_tile_k<PH_KERNEL_M5>(strategy, bv, rec, c, f64, ga, S, ha, out, counters, s, ev); break;
    case PH_KERNEL_M6: launch_tile_k<PH_KERNEL_M6>(strategy, bv, rec, c, f64, ga, S, ha, out, counters, s, ev); break;
    case PH_KERNEL_C2: launch_tile_k<PH_KERNEL_C2>(strategy, bv, rec, c, f64, ga, S, ha, out, counters, s, ev); break;
    case PH_KERNEL_C4: launch_tile_k<PH_KERNEL_C4>(strategy, bv, rec, c, f64, ga, S, ha, out, counters, s, ev); break;
    default: launch_tile_k<PH_KERNEL_C6>(strategy, bv, rec, c, f64, ga, S, ha, out, counters, s, ev); break;
    }
}

}  // namespace ph
