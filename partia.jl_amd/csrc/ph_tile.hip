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

#include "ph_tile_dev.h"

namespace ph {
namespace {

// Grid kernels: one wave per 4 x 4 x 4 tile of grid points (Sampler<double, 1>).
template <int K, int STRAT, int MS, int MG, int MD, int MC, int EXACT>
__global__ void __launch_bounds__(kMaxTileBlock)
k_tile_c32(BucketView bv, const BucketRec* __restrict__ recs, const double* __restrict__ rec, CatChunk c,
           Sampler<double, 1> smp, const double* __restrict__ ha_s, float* __restrict__ out, u64* counter) {
    extern __shared__ double tile_smem[];
    const int wv = threadIdx.x / kWave;
    const int slot = kC32Head + (EXACT == 2 ? kC2Rec : c.rw);
    i64 s = 0;
    double p[3] = {0.0, 0.0, 0.0};
    const bool valid = smp.get(s, p[0], p[1], p[2]);
    double ha = 0.0;
    if constexpr (STRAT != PH_SCATTER) {
        if (valid) ha = ha_s[s];
    }
    u64 accepted = 0;
    tile_c32_sample<K, STRAT, MS, MG, MD, MC, EXACT>(bv, recs, rec, c, valid, p, ha, out, smp.S, s,
                                                     tile_smem + wv * tile_wave_doubles(slot), accepted);
    count_add(counter, accepted);
}

template <int K, int STRAT, int MS, int MG, int MD, int MC>
__global__ void __launch_bounds__(kMaxTileBlock) k_tile_f64(BucketView bv, const BucketRec* __restrict__ recs,
                                                     const double* __restrict__ rec, CatChunk c,
                                                     Sampler<double, 1> smp, const double* __restrict__ ha_s,
                                                     double* __restrict__ out, u64* counter) {
    extern __shared__ double tile_smem[];
    const int wv = threadIdx.x / kWave;
    i64 s = 0;
    double p[3] = {0.0, 0.0, 0.0};
    const bool valid = smp.get(s, p[0], p[1], p[2]);
    double ha = 0.0;
    if constexpr (STRAT != PH_SCATTER) {
        if (valid) ha = ha_s[s];
    }
    u64 accepted = 0;
    tile_f64_sample<K, STRAT, MS, MG, MD, MC>(bv, recs, rec, c, valid, p, ha, out, smp.S, s,
                                              tile_smem + wv * tile_wave_doubles(5 + c.rw), accepted);
    count_add(counter, accepted);
}

__global__ void __launch_bounds__(kBlock) k_nearest_tile(BucketView bv, const BucketRec* __restrict__ recs,
                                                         Sampler<double, 1> smp, double Kd,
                                                         double* __restrict__ ha, i64* __restrict__ nn) {
    __shared__ LeafRec<double> cand[kBlock / kWave][kNnCand];
    i64 s = 0;
    double p[3] = {0.0, 0.0, 0.0};
    const bool valid = smp.get(s, p[0], p[1], p[2]);
    i64 bk = 0;
    const double h = nearest_tile_h(bv, recs, valid, p, Kd, cand[threadIdx.x / kWave], bk);
    if (valid) {
        ha[s] = h;
        if (nn) nn[s] = bk != Nearest::kNone ? bk : -1;
    }
}

// Launches a tile kernel over the grid's 4x4x4 tiles (one wave each) in blocks
// of the configured size, with `slot` doubles per staged leaf of dynamic LDS,
// limited to the configured occupancy.
template <class Kern, class OutT>
void launch_tile_kernel(Kern kern, int slot, const BucketView& bv, const double* rec, const CatChunk& c,
                        const GridAxes& ga, i64 S, const double* ha, OutT* out, u64* counters, hipStream_t s,
                        hipEvent_t* ev) {
    const LaunchCfg cfg = tl_launch;
    const int block = cfg.block, waves = block / kWave;
    const size_t dyn = tile_lds_bytes(slot, block);
    size_t pad = 0;
    if (cfg.occupancy < 1.f) {  // as occupancy_pad, with this kernel's dynamic LDS
        int per_cu = 0, dev = 0, lds_cu = 0, lds_block = 0;
        PH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, dyn));
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
    const unsigned g = (unsigned)((tiles + waves - 1) / waves);
    if (ev) PH_HIP(hipEventRecord(ev[0], s));
    kern<<<g, block, dyn + pad, s>>>(bv, bv.recs, rec, c, Sampler<double, 1>{ga, S, point_xcd(cfg), point_order(cfg)},
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
        const int slot = 5 + c.rw;
        if (small) launch_tile_kernel(k_tile_f64<K, STRAT, 2, 0, 0, 0>, slot, bv, rec, c, ga, S, ha, o, counters, s, ev);
        else
            launch_tile_kernel(k_tile_f64<K, STRAT, kMaxS, kMaxG, kMaxD, kMaxC>, slot, bv, rec, c, ga, S, ha, o,
                               counters, s, ev);
        return;
    }
    float* o = static_cast<float*>(out);
    if (small)
        launch_tile_kernel(k_tile_c32<K, STRAT, 2, 0, 0, 0, 0>, kC32Head + c.rw, bv, rec, c, ga, S, ha, o, counters, s,
                           ev);
    else if (c2_shared(c) && c.rw == kC2Rec)
        launch_tile_kernel(k_tile_c32<K, STRAT, 2, 1, 1, 1, 2>, kC32Head + kC2Rec, bv, rec, c, ga, S, ha, o, counters,
                           s, ev);
    else
        launch_tile_kernel(k_tile_c32<K, STRAT, kMaxS, kMaxG, kMaxD, kMaxC, 0>, kC32Head + c.rw, bv, rec, c, ga, S, ha,
                           o, counters, s, ev);
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

// The grid tile walks' tree: buckets of the configured size (4, 8 or 16
// leaves).
static BucketView tile_tree(ph_lbvh* B, hipStream_t s) {
    ensure_buckets(B, s);
    const int leaves = tl_launch.leaves;
    return bucket_view(B, leaves == kBucketSmall || leaves == kBucketLarge ? leaves : kBucket);
}

void launch_nearest_tile(ph_lbvh* B, const GridAxes& ga, i64 S, double Kd, double* ha, hipStream_t s, i64* nn) {
    // the one-leaf tree whatever the accumulation walks: on 2e6 Plummer
    // particles and a 256^3 grid the search takes 2.0 ms over it, 2.2 ms over
    // two-leaf buckets, 2.5 ms over 4-leaf, 3.0 ms over 8-leaf and 3.8 ms over
    // 16-leaf ones (MI355X)
    ensure_buckets(B, s);
    const BucketView bv = nearest_view(B);
    const LaunchCfg cfg = tl_launch;
    const i64 tiles = ((ga.nx + 3) / 4) * ((ga.ny + 3) / 4) * ((ga.nz + 3) / 4);
    const unsigned g = (unsigned)((tiles + 3) / 4);
    hipLaunchKernelGGL(k_nearest_tile, dim3(g), dim3(kBlock), 0, s, bv, bv.recs,
                       Sampler<double, 1>{ga, S, point_xcd(cfg), point_order(cfg)}, Kd, ha, nn);
    PH_HIP(hipGetLastError());
}

void launch_tile(ph_lbvh* B, const double* rec, const CatChunk& c, int precision, int strategy, const GridAxes& ga,
                 i64 S, const double* ha, void* out, u64* counters, hipStream_t s, hipEvent_t* ev) {
    const BucketView bv = tile_tree(B, s);
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
