// ph_batch.hip -- batched sampling: many PointSamples requests of uneven size
// over one tree, served by a persistent work-queue kernel recorded in a HIP graph.
//
// Work is cut into units of kBlock consecutive samples of one request.  Every
// workgroup of the persistent kernel takes the next unit from a global counter,
// finds its request by binary search over the requests' first units, generates
// the sample coordinates (explicit list, or the frame-based Cartesian plane /
// box of Grids/grids/coordinate.jl:44-143 with the reference's arithmetic),
// computes the nearest-particle smoothing length (gather / symmetric) and runs
// the per-sample body of the per-call kernels: frame-based requests (planes,
// boxes, polar planes, cylinders: each wave holds neighbouring samples) the
// grid tile kernels' per-wave union walk of the 8-leaf bucket tree or the
// per-lane bucket walk of the 4-leaf tree, chosen per wave from its samples'
// spread against their support radii (union_pays); PH_TRAVERSAL_WAVE forces the
// union walk, PH_TRAVERSAL_BUCKET / LANE the per-lane walk.  Every sample is
// computed with exactly the per-call arithmetic, so each request's outputs equal
// those of a ph_point_samples_device call on the same coordinates, bit for bit.
// Device timestamps (s_memrealtime) give each request's first-unit start and
// last-unit end.
#include <algorithm>
#include <cmath>
#include <memory>

#include "ph_tile_dev.h"

namespace ph {

// Device descriptor of one request.
// The walk of a request's waves: one bucket walk per lane (4-leaf tree), one
// union walk per wave (8-leaf tree, ph_tile_dev.h), or per wave whichever
// union_pays picks.
constexpr i32 kWalkLane = 0, kWalkUnion = 1, kWalkChoose = 2;

// The union walk pays while the box around a wave's samples, grown by the
// mean support diameter 2R, holds at most kUnionRatio times the volume of one
// sample's support cube (2R)^3: every lane tests every leaf the union holds,
// while the per-lane walk tests each lane's own leaves but waits on one chain
// of dependent loads per lane.  16 was the fastest of 0.5, 2, 4, 8, 16, 32, 64,
// 128 and 1000 on 512 mixed requests over 4e6 particles (uniform + Plummer
// clump; MI355X): 247 ms against 282 ms with the union walk for every frame
// request and 342 ms with per-lane walks only.
#ifndef PH_BATCH_UNION_RATIO
#define PH_BATCH_UNION_RATIO 16.0f
#endif
constexpr float kUnionRatio = PH_BATCH_UNION_RATIO;

struct BatchReq {
    i64 S;      // samples
    i64 unit0;  // first work unit
    i32 kind;
    i32 walk;   // kWalkLane, kWalkUnion or kWalkChoose
    const double* px;
    const double* py;
    const double* pz;
    double o[3], r[3], u[3], f[3];
    double amin[3], d[3];  // axis minimum and step (max - min) / (n - 1)
    i64 n[3];
    void* out;
    const u32* perm;    // PH_REQ_POINTS: walk order (Morton order of the points), or null
    const double* trig;  // POLAR / CYLINDER: (sin phi_j, cos phi_j) pairs, device
};

struct BatchState {
    int strategy = 0, precision = 0, nreq = 0;
    i64 units = 0, samples = 0;
    std::vector<BatchReq> host;
    std::vector<CatChunk> chunks;  // as launched (FP32C: compacted)
    std::vector<int> ncols;        // FP32C: distinct columns per chunk
    int rw_max = 2;
    int wave_doubles = 0;  // LDS per wave of the persistent kernel, in doubles
    DevBuf reqs, rec, ctl, times;  // ctl: per-chunk counters + launch timestamp
    DevBuf perm;                   // walk orders of the point-list requests
    DevBuf trig;                   // angle tables of the polar / cylindrical requests
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* cols_at_capture = nullptr;
    double clock_khz = 100000.0;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

namespace {

// _cartesian_plane_coordinates / _cartesian_box_coordinates
// (Grids/grids/coordinate.jl:44-143): xi = xmin + (i - 1) dx with
// dx = (xmax - xmin) / (nx - 1), x = x0 + xi rx + eta ux [+ zeta fx], evaluated
// left to right without contraction; sample k = i + (j - 1) nx + (k3 - 1) nx ny.
// _polar_plane_coordinates / _cylindrical_coordinates (coordinate.jl:153-276):
// s = smin + (i - 1) ds, xi = s cos(phi_j), eta = s sin(phi_j) with the angle
// table's values, then as the plane / box.
__device__ inline void batch_point(const BatchReq& R, i64 s, double p[3]) {
    if (R.kind == PH_REQ_POINTS) {
        p[0] = R.px[s];
        p[1] = R.py[s];
        p[2] = R.pz[s];
        return;
    }
    const i64 nx = R.n[0], ny = R.n[1];
    const i64 i = s % nx, j = (s / nx) % ny, k3 = s / (nx * ny);
    double xi, eta;
    if (R.kind == PH_REQ_POLAR || R.kind == PH_REQ_CYLINDER) {
        const double sv = R.amin[0] + (double)i * R.d[0];
        xi = sv * R.trig[2 * j + 1];
        eta = sv * R.trig[2 * j];
    } else {
        xi = R.amin[0] + (double)i * R.d[0];
        eta = R.amin[1] + (double)j * R.d[1];
    }
    if (R.kind == PH_REQ_PLANE || R.kind == PH_REQ_POLAR) {
#pragma unroll
        for (int a = 0; a < 3; ++a) p[a] = (R.o[a] + xi * R.r[a]) + eta * R.u[a];
    } else {
        const double zeta = R.amin[2] + (double)k3 * R.d[2];
#pragma unroll
        for (int a = 0; a < 3; ++a) p[a] = ((R.o[a] + xi * R.r[a]) + eta * R.u[a]) + zeta * R.f[a];
    }
}

// Work units of a request and the sample each thread of a unit takes: point
// lists in blocks of kBlock consecutive points (in walk order); planes in
// 8 x 8 tiles and boxes in 4 x 4 x 4 tiles of samples per wave, so that the
// lanes of a wave sample neighbouring points.  The sample index is the
// request's own (outputs land where a single call puts them); -1: none.
constexpr int kPlaneTile = 8, kBoxTile = 4;
inline i64 batch_units(const BatchReq& R) {
    if (R.kind == PH_REQ_POINTS) return (R.S + kBlock - 1) / kBlock;
    i64 tiles;
    if (R.kind == PH_REQ_PLANE || R.kind == PH_REQ_POLAR)
        tiles = ((R.n[0] + kPlaneTile - 1) / kPlaneTile) * ((R.n[1] + kPlaneTile - 1) / kPlaneTile);
    else
        tiles = ((R.n[0] + kBoxTile - 1) / kBoxTile) * ((R.n[1] + kBoxTile - 1) / kBoxTile) *
                ((R.n[2] + kBoxTile - 1) / kBoxTile);
    return (tiles + kBlock / kWave - 1) / (kBlock / kWave);
}
__device__ inline i64 batch_sample(const BatchReq& R, i64 unit, int tid) {
    if (R.kind == PH_REQ_POINTS) {
        const i64 j = unit * kBlock + tid;
        if (j >= R.S) return -1;
        return R.perm ? (i64)R.perm[j] : j;
    }
    const i64 tile = unit * (kBlock / kWave) + tid / kWave;
    const int lane = tid & (kWave - 1);
    if (R.kind == PH_REQ_PLANE || R.kind == PH_REQ_POLAR) {
        const i64 tx = (R.n[0] + kPlaneTile - 1) / kPlaneTile, ty = (R.n[1] + kPlaneTile - 1) / kPlaneTile;
        if (tile >= tx * ty) return -1;
        const i64 i = (tile % tx) * kPlaneTile + (lane % kPlaneTile), j = (tile / tx) * kPlaneTile + lane / kPlaneTile;
        if (i >= R.n[0] || j >= R.n[1]) return -1;
        return i + R.n[0] * j;
    }
    const i64 tx = (R.n[0] + kBoxTile - 1) / kBoxTile, ty = (R.n[1] + kBoxTile - 1) / kBoxTile,
              tz = (R.n[2] + kBoxTile - 1) / kBoxTile;
    if (tile >= tx * ty * tz) return -1;
    const i64 i = (tile % tx) * kBoxTile + (lane & 3), j = ((tile / tx) % ty) * kBoxTile + ((lane >> 2) & 3),
              k = (tile / (tx * ty)) * kBoxTile + (lane >> 4);
    if (i >= R.n[0] || j >= R.n[1] || k >= R.n[2]) return -1;
    return i + R.n[0] * (j + R.n[1] * k);
}

// 30-bit Morton code of each point in the list's bounding box (box6 =
// {xmin, ymin, zmin, xmax, ymax, zmax}, device) and its index, for the walk
// order of a point-list request.  Any order gives the same outputs.
__device__ inline u32 spread10(u32 v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__global__ void k_point_codes(const double* __restrict__ px, const double* __restrict__ py,
                              const double* __restrict__ pz, i64 n, const double* __restrict__ box6,
                              u32* __restrict__ keys, u32* __restrict__ vals) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v[3] = {px[i], py[i], pz[i]};
    u32 q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double ext = box6[3 + a] - box6[a];
        const double f = ext > 0.0 ? (v[a] - box6[a]) / ext : 0.5;
        q[a] = (u32)(fmin(fmax(f, 0.0), 1.0) * 1023.0);  // NaN -> 0
    }
    keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    vals[i] = (u32)i;
}

// Clears a chunk's work counter; for the first chunk also stamps the launch
// time and clears the per-request times (a kernel, not a memset: the graph is
// replayed).
__global__ void k_batch_reset(unsigned long long* __restrict__ ctr, unsigned long long* __restrict__ stamp,
                              unsigned long long* __restrict__ times, i32 nreq) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *ctr = 0ull;
        if (stamp) *stamp = (unsigned long long)wall_clock64();
    }
    if (stamp)
        for (i64 j = i; j < nreq; j += (i64)gridDim.x * blockDim.x) {
            times[j] = ~0ull;
            times[nreq + j] = 0ull;
        }
}

__device__ inline i64 uniform_i64(i64 v) {
    const u32 lo = (u32)__builtin_amdgcn_readfirstlane((int)(u32)v);
    const u32 hi = (u32)__builtin_amdgcn_readfirstlane((int)(u32)((u64)v >> 32));
    return (i64)(((u64)hi << 32) | lo);
}

// Persistent loop.  Thread 0 is the only thread with work of its own, in one
// region per iteration between barriers: it records the times of the unit the
// workgroup just finished and takes the next unit.  The unit is then read back
// into scalar registers (wave-uniform), so the loop exit is a uniform branch:
// every wave leaves in the same iteration and executes the same barriers.
// Wave-uniform: whether the union walk pays for this wave (kUnionRatio).  live
// lanes hold a sample with neighbours (support radius r = K h_a); the sums and
// the decision are taken in float, broadcast from lane 0.
__device__ inline bool union_pays(bool live, const double p[3], double r) {
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = live ? (float)p[a] : __builtin_inff();
        hi[a] = live ? (float)p[a] : -__builtin_inff();
    }
    float rs = live ? (float)r : 0.f, cnt = live ? 1.f : 0.f;
    for (int d = kWave / 2; d > 0; d >>= 1) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], __shfl_xor(lo[a], d, kWave));
            hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], d, kWave));
        }
        rs += __shfl_xor(rs, d, kWave);
        cnt += __shfl_xor(cnt, d, kWave);
    }
    const float dm = 2.f * rs / fmaxf(cnt, 1.f);
    const float v = (hi[0] - lo[0] + dm) * (hi[1] - lo[1] + dm) * (hi[2] - lo[2] + dm);
    const bool pays = cnt > 0.f && v <= kUnionRatio * dm * dm * dm;
    return __builtin_amdgcn_readfirstlane(pays ? 1 : 0) != 0;
}

template <int K, int STRAT, int MS, int MG, int MD, int MC, int EXACT, bool C32>
__global__ void __launch_bounds__(kBlock) k_batch(BucketView bv, const BucketRec* __restrict__ recs, BucketView bv8,
                                                  const BucketRec* __restrict__ recs8, BucketView nv,
                                                  const BucketRec* __restrict__ nrecs, const double* __restrict__ rec,
                                                  CatChunk c, const BatchReq* __restrict__ reqs, i32 nreq, i64 units,
                                                  unsigned long long* __restrict__ ctr,
                                                  unsigned long long* __restrict__ times, int wave_doubles) {
    extern __shared__ double batch_smem[];
    __shared__ i64 s_unit;
    __shared__ unsigned long long s_t0;
    const int lane = threadIdx.x & (kWave - 1);
    // this wave's LDS: the nearest-h candidates, then the union walk's staging
    // (or the per-lane bucket lists)
    double* wl = batch_smem + (threadIdx.x / kWave) * wave_doubles;
    i32 done_req = -1;              // thread 0: request of the unit just finished
    unsigned long long done_t0 = 0;  // thread 0: its start time
    for (;;) {
        if (threadIdx.x == 0) {
            if (done_req >= 0) {
                atomicMin(&times[done_req], done_t0);
                atomicMax(&times[nreq + done_req], (unsigned long long)wall_clock64());
            }
            s_unit = (i64)atomicAdd(ctr, 1ull);
            s_t0 = (unsigned long long)wall_clock64();
        }
        __syncthreads();
        const i64 u = uniform_i64(s_unit);
        const unsigned long long t0 = s_t0;
        __syncthreads();  // every thread has read the unit before thread 0 takes the next
        if (u >= units) break;
        // the last request whose first unit is <= u (requests without samples
        // share their first unit with the next one and are never chosen)
        i32 lo = 0, hi = nreq - 1;
        while (lo < hi) {
            const i32 mid = (lo + hi + 1) >> 1;
            if (reqs[mid].unit0 <= u) lo = mid;
            else hi = mid - 1;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        const BatchReq& R = reqs[lo];
        const i32 mode = __builtin_amdgcn_readfirstlane(R.walk);  // wave-uniform
        const i64 sv = batch_sample(R, u - R.unit0, threadIdx.x);
        const bool valid = sv >= 0;
        const i64 s = valid ? sv : 0;
        double p[3] = {0.0, 0.0, 0.0};
        if (valid) batch_point(R, s, p);
        const double Kd = support_of<K>();
        double ha = 0.0;
        bool walk = mode != kWalkLane;  // union walk (scatter: no h_a to choose by)
        if constexpr (STRAT != PH_SCATTER) {
            if (mode == kWalkUnion) {
                i64 bk = 0;
                ha = nearest_tile_h(nv, nrecs, valid, p, Kd, reinterpret_cast<LeafRec<double>*>(wl), bk);
                wave_sync();  // the candidates are read before the region is reused
            } else if (valid) {
                ha = nearest_h_bucket(nv, p, neighbour_bound2(nv, Kd));
            }
            if (!valid) ha = 0.0;
            if (mode == kWalkChoose) walk = union_pays(valid && ha == ha, p, Kd * ha);
        }
        u64 accepted = 0;
        if (walk) {
            if constexpr (C32)
                tile_c32_sample<K, STRAT, MS, MG, MD, MC, EXACT>(bv8, recs8, rec, c, valid, p, ha,
                                                                 static_cast<float*>(R.out), R.S, s, wl, accepted);
            else
                tile_f64_sample<K, STRAT, MS, MG, MD, MC>(bv8, recs8, rec, c, valid, p, ha,
                                                          static_cast<double*>(R.out), R.S, s, wl, accepted);
        } else {
            u32* list = reinterpret_cast<u32*>(wl) + lane;
            if constexpr (C32)
                bucket_c32_sample<K, STRAT, MS, MG, MD, MC, EXACT>(bv, rec, c, valid, p, ha,
                                                                   static_cast<float*>(R.out), R.S, s, accepted, list,
                                                                   kWave);
            else
                bucket_f64_sample<K, STRAT, MS, MG, MD, MC>(bv, rec, c, valid, p, ha, static_cast<double*>(R.out),
                                                            R.S, s, accepted, list, kWave);
        }
        __syncthreads();  // the unit is complete
        done_req = lo;
        done_t0 = t0;
    }
}

// LDS per wave of the persistent kernel for chunk c, in doubles: the union
// walk's staging, the per-lane bucket lists (kQ entries per lane) and the
// nearest-h candidates (64 leaves of 4 doubles) share it.
inline int batch_wave_doubles(const CatChunk& c, bool c32) {
    const int slot = c32 ? kC32Head + (c2_shared(c) && c.rw == kC2Rec ? kC2Rec : c.rw) : 5 + c.rw;
    return std::max(std::max(tile_wave_doubles(slot), kQ * kWave / 2), kNnCand * 4);
}

template <int K, int STRAT, int MS, int MG, int MD, int MC, int EXACT, bool C32>
void launch_batch_t(BatchState& st, const BucketView& bv, const BucketView& bv8, const BucketView& nv, const CatChunk& c,
                    unsigned long long* ctr, int num_cu, hipStream_t s) {
    auto kern = k_batch<K, STRAT, MS, MG, MD, MC, EXACT, C32>;
    const size_t dyn = (size_t)(kBlock / kWave) * st.wave_doubles * sizeof(double);
    int per_cu = 0;
    PH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, dyn));
    const i64 want = (i64)(per_cu > 0 ? per_cu : 1) * num_cu;
    const unsigned g = (unsigned)(want < st.units ? want : (st.units > 0 ? st.units : 1));
    kern<<<g, kBlock, dyn, s>>>(bv, bv.recs, bv8, bv8.recs, nv, nv.recs, st.rec.as<double>(), c, st.reqs.as<BatchReq>(), st.nreq,
                                st.units, ctr, st.times.as<unsigned long long>(), st.wave_doubles);
    PH_HIP(hipGetLastError());
}

template <int K, int STRAT>
void launch_batch_ks(BatchState& st, const BucketView& bv, const BucketView& bv8, const BucketView& nv, const CatChunk& c,
                     unsigned long long* ctr, int num_cu, hipStream_t s) {
    const bool small = c.ng + c.nd + c.nc == 0 && c.ns <= 2;
    if (st.precision == PH_PREC_FP64) {
        if (small) launch_batch_t<K, STRAT, 2, 0, 0, 0, 0, false>(st, bv, bv8, nv, c, ctr, num_cu, s);
        else launch_batch_t<K, STRAT, kMaxS, kMaxG, kMaxD, kMaxC, 0, false>(st, bv, bv8, nv, c, ctr, num_cu, s);
    } else {  // the per-call selection (ph_bucket.hip, ph_tile.hip), so the arithmetic is the same
        if (small) launch_batch_t<K, STRAT, 2, 0, 0, 0, 0, true>(st, bv, bv8, nv, c, ctr, num_cu, s);
        else if (c2_shared(c) && c.rw == kC2Rec)
            launch_batch_t<K, STRAT, 2, 1, 1, 1, 2, true>(st, bv, bv8, nv, c, ctr, num_cu, s);
        else launch_batch_t<K, STRAT, kMaxS, kMaxG, kMaxD, kMaxC, 0, true>(st, bv, bv8, nv, c, ctr, num_cu, s);
    }
}

template <int K>
void launch_batch_k(BatchState& st, const BucketView& bv, const BucketView& bv8, const BucketView& nv, const CatChunk& c,
                    unsigned long long* ctr, int num_cu, hipStream_t s) {
    switch (st.strategy) {
    case PH_GATHER: launch_batch_ks<K, PH_GATHER>(st, bv, bv8, nv, c, ctr, num_cu, s); break;
    case PH_SCATTER: launch_batch_ks<K, PH_SCATTER>(st, bv, bv8, nv, c, ctr, num_cu, s); break;
    default: launch_batch_ks<K, PH_SYMMETRIC>(st, bv, bv8, nv, c, ctr, num_cu, s); break;
    }
}

// Records the launch sequence of one run on stream s.  ctl[0] = launch
// timestamp, ctl[1 + q] = work counter of catalog chunk q.
void record(BatchState& st, ph_lbvh* B, int num_cu, hipStream_t s) {
    const BucketView bv = bucket_view(B), bv8 = bucket_view(B, kBucket), nv = nearest_view(B);
    unsigned long long* ctl = st.ctl.as<unsigned long long>();
    unsigned long long* times = st.times.as<unsigned long long>();
    const i64 n = B->n;
    for (size_t q = 0; q < st.chunks.size(); ++q) {
        const CatChunk& c = st.chunks[q];
        unsigned long long* ctr = ctl + 1 + q;
        k_batch_reset<<<q == 0 ? grid_for(st.nreq) : 1, kBlock, 0, s>>>(ctr, q == 0 ? ctl : nullptr, times, st.nreq);
        PH_HIP(hipGetLastError());
        if (st.precision == PH_PREC_FP64)
            k_pack_records<double><<<grid_for(n), kBlock, 0, s>>>(B->cols.as<double>(), n, c, st.rec.as<double>());
        else
            k_pack_c32<<<grid_for(n + 1), kBlock, 0, s>>>(B->cols.as<double>(), n, c, st.ncols[q], c.rw,
                                                      st.rec.as<double>());
        PH_HIP(hipGetLastError());
        if (st.units == 0) continue;
        switch (B->kernel) {
        case PH_KERNEL_M4: launch_batch_k<PH_KERNEL_M4>(st, bv, bv8, nv, c, ctr, num_cu, s); break;
        case PH_KERNEL_M5: launch_batch_k<PH_KERNEL_M5>(st, bv, bv8, nv, c, ctr, num_cu, s); break;
        case PH_KERNEL_M6: launch_batch_k<PH_KERNEL_M6>(st, bv, bv8, nv, c, ctr, num_cu, s); break;
        case PH_KERNEL_C2: launch_batch_k<PH_KERNEL_C2>(st, bv, bv8, nv, c, ctr, num_cu, s); break;
        case PH_KERNEL_C4: launch_batch_k<PH_KERNEL_C4>(st, bv, bv8, nv, c, ctr, num_cu, s); break;
        default: launch_batch_k<PH_KERNEL_C6>(st, bv, bv8, nv, c, ctr, num_cu, s); break;
        }
    }
}

void capture(BatchState& st, ph_context* ctx, ph_lbvh* B) {
    st.drop_graph();
    hipStream_t s = ctx->stream;
    PH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        record(st, B, ctx->num_cu, s);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        throw;
    }
    PH_HIP(hipStreamEndCapture(s, &st.graph));
    PH_HIP(hipGraphInstantiate(&st.exec, st.graph, nullptr, nullptr, 0));
    st.cols_at_capture = B->cols.p;
}

}  // namespace

BatchState* batch_build(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int strategy, int precision, int nreq,
                        const ph_request* reqs) {
    auto st = std::make_unique<BatchState>();
    st->strategy = strategy;
    st->precision = precision;
    st->nreq = nreq;
    st->host.resize((size_t)nreq);
    i64 unit = 0;
    for (int r = 0; r < nreq; ++r) {
        const ph_request& q = reqs[r];
        BatchReq& d = st->host[(size_t)r];
        d = BatchReq{};
        d.kind = q.kind;
        if (q.kind == PH_REQ_POINTS) {
            if (q.count < 0) throw Error{PH_ERR_ARGUMENT, "request " + std::to_string(r) + ": negative count"};
            if (q.count > 0 && !(q.px && q.py && q.pz))
                throw Error{PH_ERR_ARGUMENT, "request " + std::to_string(r) + ": null coordinate pointer"};
            d.S = q.count;
            d.px = q.px;
            d.py = q.py;
            d.pz = q.pz;
        } else if (q.kind == PH_REQ_PLANE || q.kind == PH_REQ_BOX) {
            const int axes = q.kind == PH_REQ_PLANE ? 2 : 3;
            static const char* nm[3] = {"x", "y", "z"};
            d.S = 1;
            for (int a = 0; a < axes; ++a) {
                // coordinate.jl:48-61, 94-113
                if (!(q.amax[a] > q.amin[a]))
                    throw Error{PH_ERR_ARGUMENT, std::string(nm[a]) + "max must be greater than " + nm[a] + "min."};
                if (q.n[a] < 2) throw Error{PH_ERR_ARGUMENT, std::string("n") + nm[a] + " must be at least 2."};
                d.amin[a] = q.amin[a];
                d.d[a] = (q.amax[a] - q.amin[a]) / (double)(q.n[a] - 1);
                d.n[a] = q.n[a];
                if (d.S > (i64(1) << 40) / q.n[a]) throw Error{PH_ERR_ARGUMENT, "request too large"};
                d.S *= q.n[a];
            }
            if (axes == 2) d.n[2] = 1;
            for (int a = 0; a < 3; ++a) {
                d.o[a] = q.position[a];
                d.r[a] = q.right[a];
                d.u[a] = q.up[a];
                d.f[a] = q.forward[a];
            }
        } else if (q.kind == PH_REQ_POLAR || q.kind == PH_REQ_CYLINDER) {
            // coordinate.jl:153-172, 209-234
            const double smin = q.amin[0], smax = q.amax[0], pmin = q.amin[1], pmax = q.amax[1];
            const double two_pi = 2.0 * M_PI;
            if (!(smin >= 0.0)) throw Error{PH_ERR_ARGUMENT, "smin must be nonnegative."};
            if (!(smax > smin)) throw Error{PH_ERR_ARGUMENT, "smax must be greater than smin."};
            if (q.kind == PH_REQ_CYLINDER && !(q.amax[2] > q.amin[2]))
                throw Error{PH_ERR_ARGUMENT, "zmax must be greater than zmin."};
            // isapprox(pmax, 2 pi) with Julia's default rtol = sqrt(eps)
            const bool near_2pi = std::fabs(pmax - two_pi) <= 1.4901161193847656e-8 * std::max(std::fabs(pmax), two_pi);
            if (!(pmin >= 0.0 && (pmax <= two_pi || near_2pi) && pmax > pmin))
                throw Error{PH_ERR_ARGUMENT, "angular range must satisfy 0 ≤ ϕmin < ϕmax ≤ 2π."};
            if (q.n[0] < 2) throw Error{PH_ERR_ARGUMENT, "ns must be at least 2."};
            if (q.n[1] < 1) throw Error{PH_ERR_ARGUMENT, "nϕ must be at least 1."};
            if (q.kind == PH_REQ_CYLINDER && q.n[2] < 2) throw Error{PH_ERR_ARGUMENT, "nz must be at least 2."};
            const int axes = q.kind == PH_REQ_POLAR ? 2 : 3;
            d.S = 1;
            for (int a = 0; a < axes; ++a) {
                if (d.S > (i64(1) << 40) / q.n[a]) throw Error{PH_ERR_ARGUMENT, "request too large"};
                d.S *= q.n[a];
                d.amin[a] = q.amin[a];
                d.n[a] = q.n[a];
            }
            d.d[0] = (smax - smin) / (double)(q.n[0] - 1);
            d.d[1] = (pmax - pmin) / (double)q.n[1];  // half-open in phi
            if (axes == 3) d.d[2] = (q.amax[2] - q.amin[2]) / (double)(q.n[2] - 1);
            else d.n[2] = 1;
            for (int a = 0; a < 3; ++a) {
                d.o[a] = q.position[a];
                d.r[a] = q.right[a];
                d.u[a] = q.up[a];
                d.f[a] = q.forward[a];
            }
        } else {
            throw Error{PH_ERR_ARGUMENT, "request " + std::to_string(r) + ": unknown kind"};
        }
        if (d.S > 0 && !q.out) throw Error{PH_ERR_ARGUMENT, "request " + std::to_string(r) + ": null output pointer"};
        d.out = q.out;
        d.unit0 = unit;
        unit += batch_units(d);
        st->samples += d.S;
    }
    st->units = unit;
    // angle tables of the polar / cylindrical requests: the caller's pairs, or
    // sin / cos of phi_j = phimin + (j - 1) dphi from the C library
    {
        std::vector<double> tab;
        std::vector<size_t> at((size_t)nreq, 0);
        for (int r = 0; r < nreq; ++r) {
            const BatchReq& d = st->host[(size_t)r];
            if (d.kind != PH_REQ_POLAR && d.kind != PH_REQ_CYLINDER) continue;
            at[(size_t)r] = tab.size();
            for (i64 j = 0; j < d.n[1]; ++j) {
                if (reqs[r].trig) {
                    tab.push_back(reqs[r].trig[2 * j]);
                    tab.push_back(reqs[r].trig[2 * j + 1]);
                } else {
                    const double phi = d.amin[1] + (double)j * d.d[1];
                    tab.push_back(std::sin(phi));
                    tab.push_back(std::cos(phi));
                }
            }
        }
        if (!tab.empty()) {
            st->trig.alloc(sizeof(double) * tab.size());
            PH_HIP(hipMemcpy(st->trig.p, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
            for (int r = 0; r < nreq; ++r) {
                BatchReq& d = st->host[(size_t)r];
                if (d.kind == PH_REQ_POLAR || d.kind == PH_REQ_CYLINDER) d.trig = st->trig.as<double>() + at[(size_t)r];
            }
        }
    }
    // walk orders of the point lists: Morton order within each list's box
    {
        i64 total = 0;
        for (const BatchReq& d : st->host)
            if (d.kind == PH_REQ_POINTS && d.S >= 2 * kBlock) total += d.S;
        if (total > (i64)UINT32_MAX) throw Error{PH_ERR_CAPACITY, "point-list requests exceed 2^32 points"};
        if (total > 0) {
            st->perm.alloc(sizeof(u32) * (size_t)total);
            DevBuf keys, box, tmp;
            keys.alloc(sizeof(u32) * (size_t)total);
            box.alloc(sizeof(double) * 6);
            i64 off = 0;
            for (BatchReq& d : st->host) {
                if (!(d.kind == PH_REQ_POINTS && d.S >= 2 * kBlock)) continue;
                u32* vals = st->perm.as<u32>() + off;
                minmax3<double>(d.px, d.py, d.pz, d.S, box.as<double>(), tmp, ctx->stream);
                k_point_codes<<<grid_for(d.S), kBlock, 0, ctx->stream>>>(d.px, d.py, d.pz, d.S, box.as<double>(),
                                                                          keys.as<u32>() + off, vals);
                PH_HIP(hipGetLastError());
                radix_sort_pairs(keys.as<u32>() + off, vals, d.S, 30, tmp, ctx->stream);
                d.perm = vals;
                off += d.S;
            }
            PH_HIP(hipStreamSynchronize(ctx->stream));  // the temporaries go out of scope
        }
    }
    split_catalog(cat, st->chunks);
    if (st->chunks.empty()) st->chunks.push_back(CatChunk{});
    st->ncols.assign(st->chunks.size(), 0);
    for (size_t q = 0; q < st->chunks.size(); ++q) {
        if (precision != PH_PREC_FP64) st->chunks[q] = compact_c32(st->chunks[q], st->ncols[q]);
        st->rw_max = std::max(st->rw_max, st->chunks[q].rw);
        st->wave_doubles = std::max(st->wave_doubles, batch_wave_doubles(st->chunks[q], precision != PH_PREC_FP64));
    }
    // the walk of every request's waves: chosen per wave unless the context
    // asks for the union walk (PH_TRAVERSAL_WAVE) or per-lane walks
    i32 walk = ctx->traversal == PH_TRAVERSAL_WAVE ? kWalkUnion
               : ctx->traversal == PH_TRAVERSAL_BUCKET || ctx->traversal == PH_TRAVERSAL_LANE ? kWalkLane
                                                                                           : kWalkChoose;
    // the per-lane FP32C walk reads through 32-bit buffer offsets: past them
    // every wave takes the union walk (64-bit addresses; the same results)
    if (precision != PH_PREC_FP64)
        for (const CatChunk& c : st->chunks)
            if (c32_wide(B->n, c.rw)) walk = kWalkUnion;
    for (BatchReq& d : st->host) d.walk = walk;
    const i64 n = B->n;
    ensure_buckets(B, ctx->stream);
    st->rec.alloc(sizeof(double) * (size_t)st->rw_max * (size_t)(n + 1));  // FP32C: + the zero record
    st->ctl.alloc(sizeof(unsigned long long) * (1 + st->chunks.size()));
    st->times.alloc(sizeof(unsigned long long) * 2 * (size_t)std::max(nreq, 1));
    st->reqs.alloc(sizeof(BatchReq) * (size_t)std::max(nreq, 1));
    if (nreq > 0)
        PH_HIP(hipMemcpyAsync(st->reqs.p, st->host.data(), sizeof(BatchReq) * (size_t)nreq, hipMemcpyHostToDevice,
                              ctx->stream));
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) == hipSuccess && khz > 0)
        st->clock_khz = khz;
    (void)hipGetLastError();
    PH_HIP(hipStreamSynchronize(ctx->stream));
    capture(*st, ctx, B);
    return st.release();
}

void batch_run(BatchState* st, ph_context* ctx, ph_lbvh* B) {
    if (B->cols.p != st->cols_at_capture) capture(*st, ctx, B);  // the columns were reallocated
    PH_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    PH_HIP(hipGraphLaunch(st->exec, ctx->stream));
    PH_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    PH_HIP(hipEventSynchronize(ctx->ev[1]));
    float ms = 0.f;
    PH_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    ctx->stats.interp_ms = ms;
    ctx->stats.kernel_ms = ms;
    ctx->stats.samples = st->samples;
    ctx->stats.accepted = -1;
    ctx->stats.rechecks = -1;
}

void batch_times(BatchState* st, ph_context* ctx, double* first_us, double* last_us) {
    const int nreq = st->nreq;
    std::vector<unsigned long long> t(2 * (size_t)nreq + 1);
    PH_HIP(hipStreamSynchronize(ctx->stream));
    PH_HIP(hipMemcpy(t.data(), st->ctl.p, sizeof(unsigned long long), hipMemcpyDeviceToHost));
    if (nreq > 0)
        PH_HIP(hipMemcpy(t.data() + 1, st->times.p, sizeof(unsigned long long) * 2 * (size_t)nreq,
                         hipMemcpyDeviceToHost));
    const double us_per_tick = 1000.0 / st->clock_khz;
    const unsigned long long t0 = t[0];
    for (int r = 0; r < nreq; ++r) {
        const unsigned long long a = t[1 + (size_t)r], b = t[1 + (size_t)nreq + r];
        const bool none = st->host[(size_t)r].S == 0 || a == ~0ull;
        if (first_us) first_us[r] = none ? NAN : (double)(long long)(a - t0) * us_per_tick;
        if (last_us) last_us[r] = none ? NAN : (double)(long long)(b - t0) * us_per_tick;
    }
}

void batch_free(BatchState* st) {
    if (!st) return;
    st->drop_graph();
    delete st;
}

}  // namespace ph
