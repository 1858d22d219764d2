// ph_query.hip -- the LBVH queries of LinearBVHQueries.jl on the GPU, one lane
// per query point over the bucket tree:
//   LBVH_probe_neighbors (:22-42) and LBVH_query! (:183-211): the leaves with
//     d2 <= radius^2 (d2 in float64, the reference's order of terms), counted,
//     optionally listed in DFS (= Morton) order, and the closest of them
//     (strict <, so the first in DFS order wins ties);
//   LBVH_find_nearest (:79-95): the closest leaf of the whole tree.
// Buckets are pruned with the conservative float32 box test of the
// interpolation kernels (a superset), leaves are tested exactly, so the
// accepted sets and their order are the reference's.
#include "ph_bucket_dev.h"

namespace ph {
namespace {

__global__ void __launch_bounds__(kBlock) k_query(BucketView bv, PointList pl, i64 S, const double* __restrict__ radius,
                                                  i64* __restrict__ count, i64* __restrict__ closest,
                                                  double* __restrict__ closest_d2, const i64* __restrict__ offsets,
                                                  i64* __restrict__ pool) {
    const i64 s = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (s >= S) return;
    const double p[3] = {pl.x[s], pl.y[s], pl.z[s]};
    const double r = radius[s], r2 = r * r;
    Probe<PH_GATHER> pr;
    pr.init(p, 1.0, r);  // float32 bound on radius r
    i64 cnt = 0, best = 0;
    double bd = INFINITY;  // typemax(Float64)
    i64* out = pool ? pool + offsets[s] : nullptr;
    i32 i = 0;
    for (i64 steps = 2 * (i64)bv.nrec + 8; i < bv.nrec && steps > 0; --steps) {
        const BucketRec R = bv.recs[i];
        const bool keep = pr.keep(R);
        if (R.link >= 0) {
            i = keep ? i + 1 : R.link;
            continue;
        }
        if (keep) {
            const i64 lo = bucket_lo(R.link);
            const int c = bucket_cnt(R);
            for (int j = 0; j < c; ++j) {
                const double d2 = d2_point(p[0], p[1], p[2], bv.leaves[lo + j]);
                if (d2 <= r2) {
                    if (out) out[cnt] = lo + j + 1;
                    ++cnt;
                    if (d2 < bd) {
                        bd = d2;
                        best = lo + j + 1;
                    }
                }
            }
        }
        ++i;
    }
    if (count) count[s] = cnt;
    if (closest) closest[s] = best;
    if (closest_d2) closest_d2[s] = bd;
}

__global__ void __launch_bounds__(kBlock) k_find_nearest(BucketView bv, PointList pl, i64 S, i64* __restrict__ idx,
                                                         double* __restrict__ d2) {
    const i64 s = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (s >= S) return;
    const double p[3] = {pl.x[s], pl.y[s], pl.z[s]};
    const Nearest nr = nearest_bucket(bv, p);  // unbounded: typemax
    idx[s] = nr.bk != Nearest::kNone ? nr.bk + 1 : 0;
    if (d2) d2[s] = nr.bk != Nearest::kNone ? nr.best : INFINITY;
}

}  // namespace

void launch_query(ph_lbvh* B, const PointList& pl, i64 S, const double* radius, i64* count, i64* closest,
                  double* closest_d2, const i64* offsets, i64* pool, hipStream_t s) {
    ensure_buckets(B, s);
    k_query<<<(unsigned)((S + kBlock - 1) / kBlock), kBlock, 0, s>>>(bucket_view(B), pl, S, radius, count, closest,
                                                                    closest_d2, offsets, pool);
    PH_HIP(hipGetLastError());
}

void launch_find_nearest(ph_lbvh* B, const PointList& pl, i64 S, i64* idx, double* d2, hipStream_t s) {
    ensure_buckets(B, s);
    k_find_nearest<<<(unsigned)((S + kBlock - 1) / kBlock), kBlock, 0, s>>>(nearest_view(B), pl, S, idx, d2);
    PH_HIP(hipGetLastError());
}

}  // namespace ph
