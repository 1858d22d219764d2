// ph_interp.h -- launchers of the interpolation kernels.
#pragma once

#include "ph_common.h"

namespace ph {

// One catalog chunk as the kernels see it (capacities fixed at compile time).
// The kernels read a packed per-particle record of width rw holding m, rho and
// the distinct columns the chunk uses; *pos fields index into that record.
constexpr int kMaxS = 4, kMaxG = 1, kMaxD = 1, kMaxC = 1;
constexpr int kMaxRec = 2 + kMaxS + kMaxG + 3 * kMaxD + 3 * kMaxC + 1;  // 14 (even)
struct CatChunk {
    i32 ns, ng, nd, nc;
    i32 sflag[kMaxS];
    i32 sout[kMaxS];  // output index of each entry
    i32 spos[kMaxS];  // record positions
    i32 gout[kMaxG];
    i32 gpos[kMaxG];  // -1: density gradient (reference grad slot 0)
    i32 dout[kMaxD];
    i32 dpos[3 * kMaxD];
    i32 cout[kMaxC];
    i32 cpos[3 * kMaxC];
    i32 rw;              // record width (even, >= 2)
    i32 nsrc;            // columns packed into the record
    i32 rsrc[kMaxRec];   // source column (0-based index into x,y,z,m,h,rho,quant...) of each record slot
};

// Splits a ph_catalog into chunks; returns L.
int split_catalog(const ph_catalog* cat, std::vector<CatChunk>& chunks);

// Sample sources.
struct PointList {
    const double* x;
    const double* y;
    const double* z;
};
struct GridAxes {
    const double* ax;
    const double* ay;
    const double* az;
    i64 nx, ny, nz;
    // PH_COORDS_CARTESIAN: axes x, y, z.  PH_COORDS_CYLINDRICAL: s, (sin, cos)
    // pairs of phi, z.  PH_COORDS_SPHERICAL: r, pairs of phi, pairs of theta.
    int coords = 0;
};
struct LineList {
    const double *ox, *oy, *oz, *dx, *dy, *dz;
};
struct LineGrid {
    double o[3], u[3], v[3], d[3];
    const double* ax;  // polar: the radii s_i
    const double* ay;  // polar: (sin phi_j, cos phi_j) pairs
    i64 nx, ny;
    int polar = 0;     // origins o + (s_i cos phi_j) u + (s_i sin phi_j) v
};

// Point samples (list or grid); out is L x S of T (double for FP64, float for FP32 / FP32C).
// ev (optional, 2 events) brackets the dominant traversal + accumulation kernel.
void launch_points(ph_lbvh* B, const CatChunk& c, int strategy, int precision, const PointList* pl,
                   const GridAxes* ga, i64 S, int L, void* out, u64* counters, hipStream_t s,
                   hipEvent_t* ev = nullptr);
void launch_lines(ph_lbvh* B, const CatChunk& c, int strategy, int precision, const LineList* ll, const LineGrid* lg,
                  i64 S, int L, void* out, u64* counters, const void* lut, bool wave, hipStream_t s,
                  hipEvent_t* ev = nullptr);
void launch_nearest_h(ph_lbvh* B, const PointList& pl, i64 S, double* ha, hipStream_t s);
// Per-call buffers of the bucket launchers supplied by the caller instead of
// the tree's own (which are allocated on demand): rec holds n x rw doubles of
// the chunk's packed records, ha S doubles (gather / symmetric).
struct CallWorkspace {
    double* rec;
    double* ha;
};
// Doubles per particle of the packed records the bucket launchers use for
// chunk c at the given precision (PH_PREC_FP64 or PH_PREC_FP32C).
i64 bucket_record_doubles(const CatChunk& c, int precision);
// Records the bucket launchers pack for n particles at that precision (the
// FP32C layout adds a zero record at index n).
i64 bucket_records(i64 n, int precision);
// Bucket traversal (float64 or FP32-compensated evaluation): per lane, a DFS over
// the preorder bucket tree, leaves of reached buckets tested with float64.
void launch_bucket(ph_lbvh* B, const CatChunk& c, int strategy, int precision, const PointList* pl,
                   const GridAxes* ga, i64 S, void* out, u64* counters, hipStream_t s, hipEvent_t* ev = nullptr,
                   const CallWorkspace* ws = nullptr);
// Builds B's two bucket trees if they are not built yet (buckets of at most
// 4 leaves: per-lane kernels, batches, adjoints, queries; of at most 8: the
// grid tile kernels).  ph_lbvh_build calls it for every tree that
// buckets_supported accepts (3D, float64, at most 2^30 particles).
// refit_bad: the build's refit check count on the device, read back with the
// record counts (0 expected), or null.
void ensure_buckets(ph_lbvh* B, hipStream_t s, const u32* refit_bad = nullptr);
bool buckets_supported(const ph_lbvh* B);
// Grid interpolation with one bucket-tree walk per wave (ph_tile.hip), FP32C
// (float outputs) or float64 (precision PH_PREC_FP64, double outputs); the
// caller has packed rec (the chunk's records in that precision's layout) and
// computed ha (gather / symmetric).
void launch_tile(ph_lbvh* B, const double* rec, const CatChunk& c, int precision, int strategy, const GridAxes& ga,
                 i64 S, const double* ha, void* out, u64* counters, hipStream_t s, hipEvent_t* ev);
// One output group of point / grid interpolation for the adjoint: a scalar
// (mode 0, Shepard flag shepard), the gradient of a column (1), or the
// divergence (2) or curl (3) of three; q: the columns' indices in B->cols.
struct AdjointOut {
    int mode = 0;
    i32 q[3] = {0, 0, 0};
    bool shepard = false;
    __host__ __device__ int fields() const { return mode >= 2 ? 3 : 1; }  // field columns c
    __host__ __device__ int width() const { return mode == 1 || mode == 3 ? 3 : 1; }  // values per sample
};
// Adjoint of one output group of point (pl) or grid (ga) interpolation
// (ph_adjoint.hip): grad[5 + c][n] float64, overwritten, input order; cot
// holds width() x S values of the forward's output type (double for
// PH_PREC_FP64, float for PH_PREC_FP32C), component-major like the forward's
// output rows.
void launch_points_adjoint(ph_lbvh* B, int strategy, const AdjointOut& ao, int precision, const PointList* pl,
                           const GridAxes* ga, i64 S, const void* cot, double* grad, hipStream_t s, hipEvent_t* ev);
// Nearest-particle smoothing length of every grid point (NaN: no particle
// within Kd max(h)), one wave per tile seeding its lanes' searches
// (ph_tile.hip); with nn, also the nearest leaf's Morton index (-1: none).
void launch_nearest_tile(ph_lbvh* B, const GridAxes& ga, i64 S, double Kd, double* ha, hipStream_t s,
                         i64* nn = nullptr);

// Adjoint of the line projection of one scalar (cols[qsrc]): grad[6][n] float64
// (d/dm, d/dA, d/dh, d/dx, d/dy, d/dz, input order) of sum cot[s] out[s];
// cot and lut of the precision's element type.  Zeroes grad first.
void launch_lines_adjoint(ph_lbvh* B, i32 qsrc, bool shepard, int precision, const LineList* ll, const LineGrid* lg,
                          i64 S, const void* cot, const void* lut, double* grad, hipStream_t s,
                          hipEvent_t* ev = nullptr);

// Single-quantity kernels (ph_single.hip), one lane per sample, out of the
// tree's float type: point kind PH_SINGLE_* (1 or 3 x S; ha null: nearest h);
// line quantities of chunk c with the caller's ha (L x S), or with c null the
// line density (1 x S); lut is the tables of the tree's type.
void launch_point_single(ph_lbvh* B, int kind, int strategy, const PointList& pl, const double* ha, i64 S, void* out,
                         hipStream_t s);
void launch_line_single(ph_lbvh* B, const CatChunk* c, int strategy, const LineList& ll, const double* ha, i64 S,
                        void* out, const void* lut, hipStream_t s);

// LBVH queries (ph_query.hip), float64 trees, device arrays.  launch_query:
// per point the leaves within radius[s] (count, closest 1-based Morton index or
// 0, its d2 or +Inf; with pool, their indices from pool[offsets[s]] in DFS
// order); launch_find_nearest: the closest leaf (1-based) and its d2.
void launch_query(ph_lbvh* B, const PointList& pl, i64 S, const double* radius, i64* count, i64* closest,
                  double* closest_d2, const i64* offsets, i64* pool, hipStream_t s);
void launch_find_nearest(ph_lbvh* B, const PointList& pl, i64 S, i64* idx, double* d2, hipStream_t s);

// Batched sampling (ph_batch.hip).  batch_build validates and copies the
// requests, packs nothing yet and records the launch graph; throws Error.
BatchState* batch_build(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int strategy, int precision, int nreq,
                        const ph_request* reqs);
void batch_run(BatchState* st, ph_context* ctx, ph_lbvh* B);
void batch_times(BatchState* st, ph_context* ctx, double* first_us, double* last_us);
void batch_free(BatchState* st);

}  // namespace ph
