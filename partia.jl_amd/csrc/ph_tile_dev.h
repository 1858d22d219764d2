// ph_tile_dev.h -- the per-wave bodies of the grid tile kernels (ph_tile.hip),
// shared with the batched work-queue kernel (ph_batch.hip): the nearest-h
// search seeded per wave, and the union walk of the bucket tree with LDS
// staging, float64 leaf tests and lockstep accumulation in the FP32C and the
// float64 precision.  Every lane of a wave calls a body together (the bodies
// vote, shuffle and read wave-uniform values).  Included by .hip sources only.
//
// Exactness: the union of the lanes' kept buckets contains each lane's own
// kept buckets, hence all of its accepted leaves; staged leaves are in
// increasing Morton index (preorder), each lane accepts exactly the leaves the
// reference accepts (same float64 test) and accumulates them in that order with
// the per-lane kernels' arithmetic (AccC, Acc::add_live), so the bodies give
// the per-lane kernels' results bit for bit.
#pragma once

#include "ph_bucket_dev.h"

namespace ph {
namespace {

// A double in LDS (address space 3).
typedef __attribute__((address_space(3))) double LdsDouble;

// Record width of the C2 layout {vi, s0, s1, d0, d1, d2}.
constexpr int kC2Rec = 6;
// Doubles of an FP32C staged leaf before its record: x, y, z, (K h)^2, the
// float64 (1/h)^2 and the float32 norms of leaf_norms32.
constexpr int kC32Head = 6;

// LDS of one wave, in doubles: 64 staged leaves and one zero slot of `slot`
// doubles each, then 64 leaf indices, rounded up to an even count so that
// every wave's region is 16-byte aligned (it also holds the nearest-h search's
// 64 LeafRec candidates).  FP32C: slot = kC32Head + rw (x, y, z, (K h)^2,
// (1/h)^2, norms, record); float64: slot = 5 + rw (x, y, z, (K h)^2, h, record).
__host__ __device__ constexpr int tile_wave_doubles(int slot) {
    return ((kWave + 1) * slot + kWave / 2 + 1) & ~1;
}
// Dynamic LDS of a tile kernel's workgroup of `block` threads.
inline size_t tile_lds_bytes(int slot, int block = kBlock) {
    return (size_t)(block / kWave) * (size_t)tile_wave_doubles(slot) * sizeof(double);
}
// Largest workgroup of the grid tile kernels (their block size is a launch knob).
constexpr int kMaxTileBlock = 512;

// The next batch of the wave's union walk: the leaves of the kept buckets
// (kept when any live lane's conservative test keeps them), in preorder = in
// increasing leaf index, written to lidx until one more bucket would exceed 64
// leaves.  Returns their count; 0 once the walk is done.  i and steps are
// wave-uniform; the walk visits every record at most once.  Each step loads a
// record and the one after it together (the record arrays carry one record of
// padding): the walk goes on to the next record in preorder whenever it keeps
// an internal node or passes a bucket, and then that record's test needs no
// load of its own.
template <int STRAT>
__device__ __attribute__((always_inline)) int union_batch(const BucketView& bv, const BucketRec* __restrict__ recs,
                                                          const Probe<STRAT>& pr, u64 live_mask, i32& i, i32& steps,
                                                          u32* lidx, int lane) {
    int nl = 0;
    // one record: false when the batch is full (i unchanged), else i advanced
    auto visit = [&](const BucketRec& R) -> bool {
        const bool keep = (pr.keep_mask(R) & live_mask) != 0;
        if (R.link >= 0) {
            i = __builtin_amdgcn_readfirstlane(keep ? i + 1 : R.link);
            return true;
        }
        if (keep) {
            const int cnt = bucket_cnt(R);
            if (nl + cnt > kWave) return false;  // starts the next batch
            if (lane < cnt) lidx[nl + lane] = (u32)(bucket_lo(R.link) + lane);
            nl += cnt;
        }
        i = __builtin_amdgcn_readfirstlane(i + 1);
        return true;
    };
    while (i < bv.nrec && steps > 0) {
        const i32 i0 = i;
        const BucketRec R = recs[i0], R1 = recs[i0 + 1];
        --steps;
        if (!visit(R)) {
            ++steps;
            break;
        }
        if (i != i0 + 1 || !(i < bv.nrec && steps > 0)) continue;
        --steps;
        if (!visit(R1)) {
            ++steps;
            break;
        }
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

// FP32C body: sample s at p (valid) with the nearest particle's h ha;
// wave_lds is this wave's tile_wave_doubles(kC32Head + rw) doubles of LDS.  recs is
// bv.recs, restrict-qualified, so the walk's wave-uniform record loads are
// scalar loads.
template <int K, int STRAT, int MS, int MG, int MD, int MC, int EXACT>
__device__ __attribute__((always_inline)) void tile_c32_sample(const BucketView& bv,
                                                               const BucketRec* __restrict__ recs,
                                                               const double* __restrict__ rec, const CatChunk& c,
                                                               bool valid, const double p[3], double ha,
                                                               float* __restrict__ out, i64 S, i64 s,
                                                               double* wave_lds, u64& accepted) {
    const int lane = threadIdx.x & (kWave - 1);
    const int rw = EXACT == 2 ? kC2Rec : c.rw;  // the C2 layout: constant LDS offsets
    const int slot = kC32Head + rw;  // doubles per staged leaf
    // an LDS-typed pointer: 32-bit offsets in the inner loops
    LdsDouble* st = (LdsDouble*)wave_lds;
    u32* lidx = reinterpret_cast<u32*>(wave_lds + (kWave + 1) * slot);
    const double Kd = support_of<K>();
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
    float Wa_norm, Ga_norm;
    sample_norms32<K, STRAT>(inv_ha, Wa_norm, Ga_norm);
    // the zero slot read by lanes without a pending neighbour: position 0,
    // norms 0, record {m / rho = 0, 1 / h = 1} and zero columns, so every term is 0
    if (lane < slot)
        st[kWave * slot + lane] = lane == kC32Head ? __builtin_bit_cast(double, make_float2(0.f, 1.f)) : 0.0;

    AccC<K, STRAT, MS, MG, MD, MC, EXACT> acc;
    acc.init();
    const u64 live_mask = __ballot(live_lane);
    i32 i = live_mask ? 0 : bv.nrec;  // wave-uniform from here on
    i32 steps = bv.nrec + 8;  // the union walk visits every record at most once (nrec < 2^31 - 16)
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
            for (int q = 0; q < rw; ++q) d[kC32Head + q] = r[q];
            if constexpr (STRAT != PH_GATHER) {  // the neighbour's norms, once per staged leaf
                const LeafNorms nb = leaf_norms32<K, STRAT>(__builtin_bit_cast(float2, r[0]).y);
                d[4] = nb.ib2;
                d[5] = __builtin_bit_cast(double, make_float2(nb.Wn, nb.Gn));
            }
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
            const float2 nw = __builtin_bit_cast(float2, (double)d[5]);
            acc.add_rn(live, c, [&](int pos) { return d[kC32Head + pos]; }, LeafNorms{d[4], nw.x, nw.y}, dx, dy, dz,
                       (dx * dx + dy * dy) + dz * dz, iha2, Wa_norm, Ga_norm);
        }
        wave_sync();  // every lane is done with the staged leaves
        if (!(i < bv.nrec && steps > 0)) break;
    }
    if (valid) acc.finish(c, out, S, s);
}

// Float64 body: as tile_c32_sample with the float64 records {m, rho, columns}
// of chunk c and the per-lane float64 kernel's arithmetic (Acc::add_live);
// wave_lds is this wave's tile_wave_doubles(5 + rw) doubles.
template <int K, int STRAT, int MS, int MG, int MD, int MC>
__device__ __attribute__((always_inline)) void tile_f64_sample(const BucketView& bv,
                                                               const BucketRec* __restrict__ recs,
                                                               const double* __restrict__ rec, const CatChunk& c,
                                                               bool valid, const double p[3], double ha,
                                                               double* __restrict__ out, i64 S, i64 s,
                                                               double* wave_lds, u64& accepted) {
    const int lane = threadIdx.x & (kWave - 1);
    const int rw = c.rw;
    const int slot = 5 + rw;  // x, y, z, (K h)^2, h, record
    LdsDouble* st = (LdsDouble*)wave_lds;
    u32* lidx = reinterpret_cast<u32*>(wave_lds + (kWave + 1) * slot);
    const double Kd = support_of<K>();
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
    if (valid) acc.finish(c, out, S, s);
}

// Nearest-particle smoothing length of every lane's point, one search per
// wave: the wave descends the bucket tree once (scalar loads) toward the
// centre of its points; the 64 leaves around the bucket it reaches (Morton
// neighbours) are staged in LDS (cand: this wave's 64 LeafRec) and offered to
// every lane, which seeds each lane's bound with a nearby leaf; each lane then
// runs the pruned walk of nearest_h_bucket from the root.  The result is the
// lexicographic minimum of (d2, leaf index) over the leaves within the bound
// whatever leaves are offered first, so it equals the per-lane search's bit for
// bit.  Returns NaN for lanes without a particle within Kd max(h); bk: the
// nearest leaf's Morton index or Nearest::kNone.
constexpr int kNnCand = 64;
__device__ __attribute__((always_inline)) double nearest_tile_h(const BucketView& bv,
                                                               const BucketRec* __restrict__ recs, bool valid,
                                                               const double p[3], double Kd, LeafRec<double>* cand,
                                                               i64& bk) {
    const int lane = threadIdx.x & (kWave - 1);
    bk = Nearest::kNone;
    if (__ballot(valid) == 0) return (double)NAN;  // wave-uniform
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
    i64 lo = 0;  // bucket 0, one leaf: a valid fallback
    int cnt = 1;
    for (i64 steps = bv.nrec; steps > 0; --steps) {
        const BucketRec R = recs[i];
        if (R.link < 0) {
            lo = bucket_lo(R.link);
            cnt = bucket_cnt(R);
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
    i64 first = lo + cnt / 2 - kNnCand / 2;
    first = first < 0 ? 0 : (first > n - m ? n - m : first);
    if (lane < m) cand[lane] = bv.leaves[first + lane];
    wave_sync();
    Nearest nr{neighbour_bound2(bv, Kd)};
    for (int j = 0; j < m; ++j) nr.offer(p, cand[j], first + j);
    nearest_walk(bv, p, nr);
    bk = nr.bk;
    return nr.h(bv);
}

}  // namespace
}  // namespace ph
