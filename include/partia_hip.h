/* partia_hip.h -- C ABI of the MI355X (gfx950) SPH interpolation library.
 *
 * This is the drop-in boundary for Partia.jl's interpolation hot path.  A host
 * binding (Julia `ccall`, Python ctypes) calls these functions in place of the
 * reference's CPU/CUDA methods; every entry point below names the reference
 * interface it replaces (paths relative to the reference root; KI/ =
 * src/julia/KernelInterpolation/, LBVH/ = src/julia/LinearBoundingVolumeHierarchy/).
 *
 * Conventions
 *  - Plain pointers and sizes only.  `const double*` arguments named *_host are
 *    host memory read during the call; `*_dev` pointers are device memory of the
 *    context's GPU (allocate with ph_device_alloc or any HIP allocator).
 *  - Every function returns an int status: PH_OK (0) or a negative PH_ERR_* code;
 *    ph_last_error(ctx) returns the message of the last failure on that context.
 *  - One context = one GPU + one HIP stream.  Calls on one context are
 *    serialised and not thread-safe; separate contexts are independent.
 *  - Tree arrays use the reference's unified node IDs (int32): internal nodes
 *    1..n-1, leaf k (1-based) <-> n-1+k, 0 = null, root = 1 when n >= 2
 *    (LBVH/BinaryRadixTree/BinaryRadixTree.jl:37-44).
 */
#ifndef PARTIA_HIP_H
#define PARTIA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the host binding maps them to the reference's exceptions) ---- */
#define PH_OK 0
#define PH_ERR_ARGUMENT (-1)      /* ArgumentError                                   */
#define PH_ERR_DIMENSION (-2)     /* DimensionMismatch (InterpolationInput.jl:85-96)  */
#define PH_ERR_STRATEGY (-3)      /* ArgumentError: unsupported strategy (LineSamples_driver.jl:40-44) */
#define PH_ERR_LEAF_ORDER (-4)    /* ArgumentError: leaf order mismatch (PointSamples_driver.jl:100-103) */
#define PH_ERR_CAPACITY (-5)      /* ArgumentError: empty tree / > 2^30 leaves (BinaryRadixTree.jl:86-91) */
#define PH_ERR_OUT_OF_MEMORY (-6) /* device allocation failed                        */
#define PH_ERR_HIP (-7)           /* HIP runtime error                               */
#define PH_ERR_INTERNAL (-8)      /* AssertionError: refit visit count (LinearBVH.jl:89-91) */
#define PH_ERR_NO_DEVICE (-9)     /* no usable GPU                                   */

/* ---- enumerations mirroring the reference's type tags ---- */
/* kernel_function/kernels/<name>.jl */
#define PH_KERNEL_M4 0
#define PH_KERNEL_M5 1
#define PH_KERNEL_M6 2
#define PH_KERNEL_C2 3
#define PH_KERNEL_C4 4
#define PH_KERNEL_C6 5
/* interpolation_setup/InterpolationStrategy.jl:1-5 */
#define PH_GATHER 0
#define PH_SCATTER 1
#define PH_SYMMETRIC 2
/* working precision of an interpolation call */
#define PH_PREC_FP64 64  /* Float64 semantics of the reference                 */
#define PH_PREC_FP32 32  /* Float32 semantics of the reference (Float32 input)  */
#define PH_PREC_FP32C 33 /* Float64 input, FP32-compensated evaluation          */

/* ph_lbvh_get selectors */
#define PH_GET_ORDER 0      /* uint64[n]: original 1-based index (MortonEncoding.order)   */
#define PH_GET_CODES 1      /* uint64[n]: Morton codes after the duplicate rewrite        */
#define PH_GET_LEFT 2       /* int32[2n-1]                                                */
#define PH_GET_RIGHT 3      /* int32[2n-1]                                                */
#define PH_GET_ESCAPE 4     /* int32[2n-1]                                                */
#define PH_GET_PARENT 5     /* int32[2n-1]                                                */
#define PH_GET_LEAF_COOR 6  /* double[3][n]                                               */
#define PH_GET_LEAF_SCALE 7 /* double[n]                                                  */
#define PH_GET_NODE_MIN 8   /* double[3][n-1]                                             */
#define PH_GET_NODE_MAX 9   /* double[3][n-1]                                             */
#define PH_GET_NODE_SCALE 10 /* double[n-1]                                               */
#define PH_GET_COLUMNS 11   /* double[6+ncol][n]: x,y,z,m,h,rho,quant... in Morton order  */

typedef struct ph_context ph_context;
typedef struct ph_lbvh ph_lbvh;

/* Concise catalog: mirrors InterpolationCatalogConcise (KI/interpolation_setup/
 * InterpolationCatalog.jl:26-32).  Slots are the reference's 1-based column
 * indices; a gradient slot of 0 selects the density gradient.  Output order is
 * scalars, gradients (x,y,z each), divergences, curls (x,y,z each) -- L values. */
typedef struct ph_catalog {
    int32_t n_scalar;
    const int32_t* scalar_slots;
    const uint8_t* scalar_shepard; /* 1 = Shepard-normalise (all names except :rho / :ρ) */
    int32_t n_grad;
    const int32_t* grad_slots;
    int32_t n_div;
    const int32_t* div_slots; /* 3 per divergence */
    int32_t n_curl;
    const int32_t* curl_slots; /* 3 per curl */
} ph_catalog;

/* Per-call statistics and device-side timings (hipEvent on the context stream). */
typedef struct ph_stats {
    double build_ms;        /* last LBVH build                                       */
    double interp_ms;       /* last interpolation call, whole device-side work        */
    double kernel_ms;       /* last call's dominant (traversal + accumulation) kernel */
    int64_t samples;        /* samples (points or rays) of the last call             */
    int64_t accepted;       /* accepted neighbours of the last call (if counted)      */
    int64_t rechecks;       /* reserved (0): leaves are always tested in float64      */
} ph_stats;

/* ---- context ---- */
int ph_ctx_create(int device, ph_context** out);
int ph_ctx_destroy(ph_context* ctx);
const char* ph_last_error(const ph_context* ctx);
int ph_get_stats(const ph_context* ctx, ph_stats* out);
/* Count accepted neighbours during interpolation calls (adds one atomic per sample). */
int ph_set_count_neighbours(ph_context* ctx, int enable);
/* Traversal of point / grid interpolation calls on a float64 tree.  Both give the
 * reference's neighbour sets in the reference's order (DFS preorder = increasing
 * Morton index), so float64 results are identical:
 *   PH_TRAVERSAL_LANE    one lane per sample, escape-rope DFS over the LBVH leaf
 *                        by leaf (LinearBVHTraversalMacros.jl:26-371);
 *   PH_TRAVERSAL_BUCKET  one lane per sample, DFS over a preorder copy of the
 *                        LBVH whose subtrees of <= 4 leaves are buckets whose
 *                        leaves are tested together (float32 pruning rounded
 *                        outward, float64 leaf tests);
 *   PH_TRAVERSAL_AUTO    bucket.
 * PH_PREC_FP32C always uses the bucket traversal; Float32 trees the lane one.
 * The bucket traversal supports up to 2^30 particles, the LinearBVH limit
 * (BinaryRadixTree.jl:86-91; PH_ERR_CAPACITY beyond).
 * Line calls (LinearBVHTraversalMacros.jl:539-644) give identical results in
 * every mode: PH_TRAVERSAL_LANE walks one DFS per ray; otherwise the rays of an
 * image plane (ph_line_grid_device) are walked 64 at a time by one wave over
 * the union of their DFS, each ray testing exactly its own nodes, and explicit
 * ray lists use the per-ray walk unless PH_TRAVERSAL_WAVE asks for the wave
 * walk for them too (worth it when consecutive rays are close together). */
#define PH_TRAVERSAL_AUTO 0
#define PH_TRAVERSAL_LANE 1
#define PH_TRAVERSAL_BUCKET 2
#define PH_TRAVERSAL_WAVE 3
int ph_set_traversal(ph_context* ctx, int mode);

/* ---- autotuner (no reference counterpart): launch configuration of the point
 * and grid kernels of the bucket traversal.  The knobs change neither the
 * neighbour order nor any arithmetic, so every configuration gives identical
 * results; only the speed changes.  Knobs:
 *   occupancy   fraction of the default workgroups per CU, enforced with
 *               unused dynamic LDS (1, 0.75, 0.5 ...);
 *   xcd_map     block order over the 8 XCDs: 0 hardware order; 1 one
 *               contiguous range of blocks per XCD; k = 2..7 runs of 2^(k+2)
 *               consecutive blocks dealt round-robin to the XCDs (each XCD's
 *               blocks are neighbours, dense regions spread over every XCD);
 *               -1 each kernel family's measured best (2 for point and grid
 *               kernels; for line kernels 7 when the launch has at least
 *               16384 blocks, else 0);
 *   tile_order  order of the tiles (one per wavefront: 4x4x4 grid points or
 *               8x8 image-plane rays): 0 x-fastest rows, 1 bricks of 8x8x8 or
 *               8x8 tiles with Morton order inside, -1 each family's best
 *               (rows for grids, bricks for images);
 *   bucket_leaves  grid tile walks (one walk per wavefront): leaves per
 *               bucket of the walked tree, 4, 8 (default) or 16 (ph_lbvh_build
 *               builds all three trees, and a one-leaf tree for the nearest-h
 *               searches);
 *   block_size  grid tile walks: threads per workgroup, 64, 128, 256
 *               (default) or 512.
 * Line kernels follow xcd_map and tile_order; the autotuner times the point
 * and grid kernels only, and the grid tile walks' bucket_leaves and
 * block_size after their other knobs.
 * With autotuning on, the first call of each workload key (precision,
 * strategy, kernel, catalog shape, grid or list, log2 of the sample and
 * particle counts) times the candidate configurations twice on the call's own
 * data and keeps the fastest for later calls of that key; otherwise the
 * ph_set_launch configuration (default 1, -1, -1) is used. ---- */
typedef struct ph_tune_entry {
    uint64_t key;
    double occupancy;
    int32_t xcd_map;
    int32_t tile_order;
    double best_ms;    /* kernel time of the chosen configuration   */
    double default_ms; /* kernel time of the default configuration  */
    int32_t bucket_leaves;
    int32_t block_size;
} ph_tune_entry;
int ph_set_autotune(ph_context* ctx, int enable);
int ph_set_launch(ph_context* ctx, double occupancy, int32_t xcd_map, int32_t tile_order);
/* Grid tile walks: bucket_leaves 4, 8 or 16 and block_size 64, 128, 256 or
 * 512; 0 keeps the default (8, 256).  Used when autotuning is off. */
int ph_set_tile_shape(ph_context* ctx, int32_t bucket_leaves, int32_t block_size);
/* Copies up to capacity entries of the context's tuned keys; *count = all entries. */
int ph_autotune_report(ph_context* ctx, int32_t capacity, ph_tune_entry* out, int32_t* count);

/* ---- device memory helpers (for callers without their own HIP allocator) ---- */
int ph_device_alloc(ph_context* ctx, int64_t bytes, void** dptr);
int ph_device_free(ph_context* ctx, void* dptr);
int ph_memcpy_htod(ph_context* ctx, void* dst_dev, const void* src_host, int64_t bytes);
int ph_memcpy_dtoh(ph_context* ctx, void* dst_host, const void* src_dev, int64_t bytes);
int ph_synchronize(ph_context* ctx);

/* ---- LBVH build: replaces LinearBVH!(input; CodeType) (KI/interpolation_setup/
 * InterpolationInput.jl:201-221), i.e. MortonEncoding (LBVH/MortonEncoding/
 * MortonEncoding.jl:36-43), sort_by_morton! (toolbox.jl:20-33), BinaryRadixTree
 * (BinaryRadixTree.jl:82-133) and LinearBVH with BoxScale(h, true)
 * (LinearBVH.jl:56-94), all on the GPU.  code_bits is 63 (CodeType=UInt64, the
 * reference default) or 30 (CodeType=UInt32).  precision PH_PREC_FP64 keeps the
 * columns as float64; PH_PREC_FP32 converts them to float32 first and runs the
 * whole build in float32 (a Float32 InterpolationInput).  The particle columns
 * are permuted into Morton order on the device; read the permutation back with
 * PH_GET_ORDER to permute host arrays as LinearBVH! does. */
int ph_lbvh_build(ph_context* ctx, int64_t n, const double* x_host, const double* y_host, const double* z_host,
                  const double* m_host, const double* h_host, const double* rho_host, int32_t ncol,
                  const double* const* cols_host, int32_t kernel, int32_t code_bits, int32_t precision,
                  ph_lbvh** out);
/* Same, with the columns already resident on the device (float64 arrays). */
int ph_lbvh_build_device(ph_context* ctx, int64_t n, const double* x_dev, const double* y_dev, const double* z_dev,
                         const double* m_dev, const double* h_dev, const double* rho_dev, int32_t ncol,
                         const double* const* cols_dev, int32_t kernel, int32_t code_bits, int32_t precision,
                         ph_lbvh** out);
/* LinearBVH! of an InterpolationInput{2} (InterpolationInput.jl:224-242):
 * MortonEncoding(x, y; CodeType) with 16 (code_bits 30, UInt32) or 32 (code_bits
 * 63, UInt64) bits per axis (MortonEncoding.jl:80-86, toolbox.jl:77-92,
 * 166-168, 264-272), the same radix tree and refit.  The tree stores z = 0 for
 * every particle, so its boxes and distances are the 2D ones; ph_nearest_h
 * answers 2D queries (pz may be null).  The interpolation kernels are 3D only
 * (README.md:152): interpolation entry points reject a 2D tree. */
int ph_lbvh_build2d(ph_context* ctx, int64_t n, const double* x_host, const double* y_host, const double* m_host,
                    const double* h_host, const double* rho_host, int32_t ncol, const double* const* cols_host,
                    int32_t kernel, int32_t code_bits, int32_t precision, ph_lbvh** out);
int ph_lbvh_build2d_device(ph_context* ctx, int64_t n, const double* x_dev, const double* y_dev, const double* m_dev,
                           const double* h_dev, const double* rho_dev, int32_t ncol, const double* const* cols_dev,
                           int32_t kernel, int32_t code_bits, int32_t precision, ph_lbvh** out);
/* Dimension of a tree's input: 3, or 2 for ph_lbvh_build2d[_device]. */
int ph_lbvh_dims(const ph_lbvh* bvh, int32_t* dims);
int ph_lbvh_destroy(ph_lbvh* bvh);
/* Replace the m, rho and quantity columns of a built tree (host arrays already
 * in Morton order, i.e. permuted as LinearBVH! left them).  The external-LBVH
 * driver overloads (KI/grid_interpolation/drivers/PointSamples_driver.jl:98-118,
 * LineSamples_driver.jl:116-139) read these columns from the current input, so
 * the host binding refreshes them before each such call; ncol may differ from
 * the build's. */
int ph_lbvh_set_columns(ph_lbvh* bvh, const double* m_host, const double* rho_host, int32_t ncol,
                        const double* const* cols_host);
/* Tree/encoding accessors for the host mirror of LinearBVH (LBVH/LinearBVH/LinearBVH.jl:8-14). */
int ph_lbvh_info(const ph_lbvh* bvh, int64_t* n, int32_t* root, int32_t* ncol, int32_t* code_bits,
                 int32_t* precision);
int ph_lbvh_get(ph_lbvh* bvh, int32_t which, void* dst_host);

/* ---- PointSamples_interpolation: replaces the per-sample loop of
 * KI/grid_interpolation/drivers/PointSamples_driver.jl:34-48,98-118 and its CUDA
 * counterpart ext/CUDAExt/KernelInterpolation/grid_interpolation/drivers/
 * PointSamples_driver.jl:12-35,44-66: nearest-h (LBVH_find_nearest_h) then
 * _general_quantity_interpolate_kernel for every sample.  out_host is L*S
 * doubles, output k of sample i at out_host[k*S + i]. */
int ph_point_samples(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy, int32_t precision,
                     int64_t S, const double* px_host, const double* py_host, const double* pz_host,
                     double* out_host);
/* Device-resident variant: sample coordinates (float64) and output on the device.
 * Output element type is double for PH_PREC_FP64 and float for FP32 / FP32C. */
int ph_point_samples_device(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy,
                            int32_t precision, int64_t S, const double* px_dev, const double* py_dev,
                            const double* pz_dev, void* out_dev);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ph_point_samples_device with the per-call buffers (the
 * nearest-particle h of every sample and the packed particle records) in a
 * static device workspace of 64 MiB per GPU instead of allocations.  It
 * allocates nothing, needs 8 S (gather / symmetric) + 8 n rw (FP32C:
 * 8 (n + 1) rw) bytes of that workspace (rw = record width, 2..14 doubles)
 * and returns PH_ERR_CAPACITY
 * beyond it.  PH_PREC_FP64 or PH_PREC_FP32C on a float64 tree of at most 2^30
 * particles (whose bucket trees ph_lbvh_build made).  The results equal
 * ph_point_samples_device's bit for bit.  The kernels are only enqueued on the
 * context's stream (no wait; ph_synchronize before reading out_dev), and
 * ph_get_stats is not updated.  Every call on one GPU uses that GPU's one
 * workspace: calls that overlap on one GPU (other contexts' streams, other
 * threads) race. */
int ph_point_samples_device_noalloc(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy,
                                    int32_t precision, int64_t S, const double* px_dev, const double* py_dev,
                                    const double* pz_dev, void* out_dev);

/* ---- StructuredGrid_interpolation (KI/grid_interpolation/drivers/
 * StructuredGrid_driver.jl:46-64) for Cartesian grids: the flattened sample
 * coordinates (Grids/grids/StructuredGrid.jl:140-155, column-major, x fastest)
 * are generated on the device from the three axes, so no host coordinate array
 * exists.  Output layout as for ph_point_samples with S = nx*ny*nz. */
int ph_structured_grid(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy, int32_t precision,
                       const double* ax_host, int64_t nx, const double* ay_host, int64_t ny, const double* az_host,
                       int64_t nz, double* out_host);
int ph_structured_grid_device(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy,
                              int32_t precision, const double* ax_dev, int64_t nx, const double* ay_dev, int64_t ny,
                              const double* az_dev, int64_t nz, void* out_dev);
/* StructuredGrid_interpolation for Cylindrical and Spherical grids
 * (KI/grid_interpolation/drivers/StructuredGrid_driver.jl:46-64 through
 * flatten / coordinate_grid, Grids/grids/StructuredGrid.jl:176-222): the grid
 * points are generated on the device, first axis fastest, with _cylin2cart /
 * _sph2cart (Tools/coordinate_transformations.jl:85-89,198-204) in the tree's
 * precision.  Angle axes are passed as (sin, cos) pairs of the axis values,
 * as the caller computes them (Julia's sin / cos for exact parity):
 *   PH_COORDS_CYLINDRICAL  a1: s (n1); a2: pairs of phi (n2); a3: z (n3)
 *   PH_COORDS_SPHERICAL    a1: r (n1); a2: pairs of phi (n2); a3: pairs of theta (n3)
 *   PH_COORDS_CARTESIAN    as ph_structured_grid.
 * Outputs as ph_structured_grid, L x (n1 n2 n3). */
#define PH_COORDS_CARTESIAN 0
#define PH_COORDS_CYLINDRICAL 1
#define PH_COORDS_SPHERICAL 2
int ph_structured_grid_coords_device(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy,
                                     int32_t precision, int32_t coords, const double* a1_dev, int64_t n1,
                                     const double* a2_dev, int64_t n2, const double* a3_dev, int64_t n3,
                                     void* out_dev);
/* The same with host axes / tables and host outputs (float64 values). */
int ph_structured_grid_coords(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy,
                              int32_t precision, int32_t coords, const double* a1, int64_t n1, const double* a2,
                              int64_t n2, const double* a3, int64_t n3, double* out_host);

/* ---- LineSamples_interpolation: replaces KI/grid_interpolation/drivers/
 * LineSamples_driver.jl:39-57,116-139 (and the CUDA twin LineSamples_driver.jl:
 * 11-34,49-77): line-integrated scalars along infinite lines (origin, unit
 * direction).  Only PH_SCATTER is accepted (PH_ERR_STRATEGY otherwise), and the
 * catalog may hold scalars only. */
int ph_line_samples(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy, int32_t precision,
                    int64_t S, const double* ox_host, const double* oy_host, const double* oz_host,
                    const double* dx_host, const double* dy_host, const double* dz_host, double* out_host);
int ph_line_samples_device(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy, int32_t precision,
                           int64_t S, const double* ox_dev, const double* oy_dev, const double* oz_dev,
                           const double* dx_dev, const double* dy_dev, const double* dz_dev, void* out_dev);

/* Parallel-ray projection on a regular image plane (the column-density map of
 * a line-integrated StructuredGrid): ray (i, j) has origin o + u*ax[i] + v*ay[j]
 * and direction dir; the rays are generated on the device.  Output k of ray
 * (i, j) at out[k*nx*ny + i + nx*j]. */
int ph_line_grid_device(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy, int32_t precision,
                        const double* origin3, const double* u3, const double* v3, const double* dir3,
                        const double* ax_dev, int64_t nx, const double* ay_dev, int64_t ny, void* out_dev);
/* LineSamples(Polar, ParallelBeam, frame, sparams, phiparams)
 * (Grids/grids/LineSamples.jl:512-537): origins on the polar sampling of the
 * frame plane (_polar_plane_coordinates, Grids/grids/coordinate.jl:153-206),
 * origin3 + (s_i cos phi_j) u3 + (s_i sin phi_j) v3, i fastest, all along dir3.
 * s_dev: the ns radii s_i = smin + (i - 1) ds; trig_dev: nphi pairs
 * (sin phi_j, cos phi_j) of the half-open angles, as the caller computes them
 * (device float64).  Outputs and rules as ph_line_grid_device. */
int ph_line_polar_device(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy, int32_t precision,
                         const double* origin3, const double* u3, const double* v3, const double* dir3,
                         const double* s_dev, int64_t ns, const double* trig_dev, int64_t nphi, void* out_dev);

/* ---- adjoints of the line projection (no reference counterpart; the reverse
 * pass a ChainRulesCore rrule of LineSamples_interpolation would call).  For a
 * catalog of exactly one scalar A (slot s, Shepard flag as in the forward),
 * scatter strategy, with w_b = m_b I(q_b) / (h_b rho_b), q_b = r_b / h_b:
 *   out = sum_b A_b w_b            (no Shepard flag, e.g. column density :rho)
 *   out = sum_b A_b w_b / sum_b w_b (Shepard flag)
 * grad_dev (6 x n float64, overwritten, rows d/dm, d/dA, d/dh, d/dx, d/dy, d/dz,
 * columns in the input order of the particles, i.e. before LinearBVH!'s
 * permutation) receives the gradient of sum_s cot[s] * out[s]; rho_b is held
 * fixed and A_b is the value of column s.  I is the table's piecewise-linear
 * interpolant (its derivative is the segment slope).  The neighbours are
 * recomputed with the forward's traversal; terms are evaluated in the
 * precision (PH_PREC_FP64 or PH_PREC_FP32 trees) and summed in float64 with
 * atomics (sums are order-dependent in the last bits).  cot_dev holds S values
 * of the forward's output type; grad_dev must be ordinary device memory. ---- */
int ph_line_samples_adjoint_device(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy,
                                   int32_t precision, int64_t S, const double* ox_dev, const double* oy_dev,
                                   const double* oz_dev, const double* dx_dev, const double* dy_dev,
                                   const double* dz_dev, const void* cot_dev, double* grad_dev);
int ph_line_grid_adjoint_device(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy,
                                int32_t precision, const double* origin3, const double* u3, const double* v3,
                                const double* dir3, const double* ax_dev, int64_t nx, const double* ay_dev, int64_t ny,
                                const void* cot_dev, double* grad_dev);
/* The same for the polar rays of ph_line_polar_device. */
int ph_line_polar_adjoint_device(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy,
                                 int32_t precision, const double* origin3, const double* u3, const double* v3,
                                 const double* dir3, const double* s_dev, int64_t ns, const double* trig_dev,
                                 int64_t nphi, const void* cot_dev, double* grad_dev);

/* ---- adjoints of point / grid interpolation (no reference counterpart; the
 * reverse pass a ChainRulesCore rrule of PointSamples_interpolation /
 * StructuredGrid_interpolation would call).  For a catalog of exactly one
 * scalar A (slot s, Shepard flag as in the forward), any strategy, with
 * v_b = m_b / rho_b and W_ab the strategy's kernel (gather W(r, h_a), scatter
 * W(r, h_b), symmetric their mean; general_interpolation.jl:1-669):
 *   out = sum_b A_b v_b W_ab                          (no Shepard flag, e.g. :rho)
 *   out = sum_b A_b v_b W_ab / sum_b v_b W_ab         (Shepard flag)
 * grad_dev (6 x n float64, overwritten, rows d/dm, d/dA, d/dh, d/dx, d/dy,
 * d/dz, columns in the input order of the particles) receives the gradient of
 * sum_s cot[s] * out[s]; rho_b is held fixed and A_b is the value of column s.
 * For gather and symmetric, h_a is the smoothing length of the sample's
 * nearest particle (LinearBVHQueries.jl:118-140), whose h row receives
 * d out / d h_a.  Samples with a NaN output contribute nothing.  Neighbours
 * are recomputed with the forward's float64 test; terms are evaluated in
 * float64 and summed with float64 atomics (sums are order-dependent in the last
 * bits).  precision: PH_PREC_FP64 (cot_dev float64) or PH_PREC_FP32C (cot_dev
 * float32), on a tree built from Float64 columns.
 * The catalog may instead hold exactly one gradient (of a column, slot >= 1;
 * the density-gradient branch of slot 0 is not provided), one divergence or
 * one curl (general_interpolation.jl:128-214 with accumulations/{gradient,
 * divergence,curl}_accumulation.jl).  cot_dev then holds the output's 3 (or,
 * for the divergence, 1) rows of S values, laid out as the forward writes
 * them, and grad_dev has 5 + c rows: d/dm, d/d of each of the c field columns
 * (c = 1 for a gradient, 3 for a divergence or curl, in slot order), d/dh,
 * d/dx, d/dy, d/dz.  The Shepard means inside these outputs are
 * differentiated too. ---- */
int ph_point_samples_adjoint_device(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy,
                                    int32_t precision, int64_t S, const double* px_dev, const double* py_dev,
                                    const double* pz_dev, const void* cot_dev, double* grad_dev);
int ph_structured_grid_adjoint_device(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy,
                                      int32_t precision, const double* ax_dev, int64_t nx, const double* ay_dev,
                                      int64_t ny, const double* az_dev, int64_t nz, const void* cot_dev,
                                      double* grad_dev);

/* ---- the reference's single-quantity kernels (its building blocks outside
 * the drivers, used by its tests): one lane per sample walks the reference's
 * DFS and accumulates in the tree's float type, so float64 results equal the
 * reference's arithmetic bit for bit.  precision: PH_PREC_FP64 on a float64
 * tree or PH_PREC_FP32 on a Float32 tree; outputs of that type (device
 * variants) or double (host variants). ---- */
#define PH_SINGLE_DENSITY 0          /* _density_kernel: sum m_b W_ab (0 without neighbours); 1 x S */
#define PH_SINGLE_NUMBER_DENSITY 1   /* _number_density_kernel: sum W_ab; 1 x S                      */
#define PH_SINGLE_GRADIENT_DENSITY 2 /* _gradient_density_kernel: 3 x S (x, y, z rows), NaN without  */
/* _density_kernel, _number_density_kernel (KI/single_point_interpolation/
 * kernels/scalar_interpolation.jl:1-166) and _gradient_density_kernel
 * (gradient_interpolation.jl:1-170, whose mean term uses the unnormalised
 * density sum m_b W, :47-55, unlike the fused kernel's density gradient) at S
 * points.  ha (one per point, gather / symmetric) is the kernels' h_a argument;
 * null takes each point's nearest-particle h (LinearBVHQueries.jl:118-140), as
 * the drivers do. */
int ph_point_single(ph_context* ctx, ph_lbvh* bvh, int32_t kind, int32_t strategy, int32_t precision, int64_t S,
                    const double* px_host, const double* py_host, const double* pz_host, const double* ha_host,
                    double* out_host);
int ph_point_single_device(ph_context* ctx, ph_lbvh* bvh, int32_t kind, int32_t strategy, int32_t precision,
                           int64_t S, const double* px_dev, const double* py_dev, const double* pz_dev,
                           const double* ha_dev, void* out_dev);
/* _line_integrated_quantities_interpolate_kernel in every strategy
 * (line_integrated_scalar_interpolation.jl:76-214): the catalog's scalars
 * along S lines with the caller's h_a per line (ha required for gather and
 * symmetric, unused by scatter).  The drivers accept scatter only
 * (LineSamples_driver.jl:40-44); this is the kernel-level call. */
int ph_line_samples_ha(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy, int32_t precision,
                       int64_t S, const double* ox_host, const double* oy_host, const double* oz_host,
                       const double* dx_host, const double* dy_host, const double* dz_host, const double* ha_host,
                       double* out_host);
int ph_line_samples_ha_device(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy,
                              int32_t precision, int64_t S, const double* ox_dev, const double* oy_dev,
                              const double* oz_dev, const double* dx_dev, const double* dy_dev, const double* dz_dev,
                              const double* ha_dev, void* out_dev);
/* _line_integrated_density_kernel (line_integrated_scalar_interpolation.jl:1-70):
 * the column sum m_b W along each line (no rho, no Shepard; 0 without
 * neighbours), 1 x S, ha as for ph_line_samples_ha. */
int ph_line_density(ph_context* ctx, ph_lbvh* bvh, int32_t strategy, int32_t precision, int64_t S,
                    const double* ox_host, const double* oy_host, const double* oz_host, const double* dx_host,
                    const double* dy_host, const double* dz_host, const double* ha_host, double* out_host);
int ph_line_density_device(ph_context* ctx, ph_lbvh* bvh, int32_t strategy, int32_t precision, int64_t S,
                           const double* ox_dev, const double* oy_dev, const double* oz_dev, const double* dx_dev,
                           const double* dy_dev, const double* dz_dev, const double* ha_dev, void* out_dev);

/* ---- building blocks exposed for parity tests ---- */
/* LBVH_probe_neighbors and LBVH_query! (LBVH/LinearBVH/LinearBVHQueries.jl:22-42,
 * 183-211) for S points with radii radius_dev[S], on a float64 tree, device
 * arrays (any output may be null): count_dev[s] = leaves with d2 <= radius^2;
 * closest_dev[s] = 1-based Morton index of the closest of them, first in DFS
 * order on ties (0 if none); closest_d2_dev[s] = its d2 (+Inf if none).  With
 * pool_dev, the accepted leaves' 1-based indices are written in DFS (Morton)
 * order from pool_dev[offsets_dev[s]] (the caller sizes offsets from a first
 * call's counts, as LBVH_query! fills its pool).  For a 2D tree pz may be null. */
int ph_query_device(ph_context* ctx, ph_lbvh* bvh, int64_t S, const double* px_dev, const double* py_dev,
                    const double* pz_dev, const double* radius_dev, int64_t* count_dev, int64_t* closest_dev,
                    double* closest_d2_dev, const int64_t* offsets_dev, int64_t* pool_dev);
/* LBVH_find_nearest (LinearBVHQueries.jl:79-95): idx_dev[s] = 1-based Morton index of
 * the closest leaf (first in DFS order on ties), d2_dev[s] its squared distance
 * (may be null).  Float64 tree, device arrays; for a 2D tree pz may be null. */
int ph_find_nearest_device(ph_context* ctx, ph_lbvh* bvh, int64_t S, const double* px_dev, const double* py_dev,
                           const double* pz_dev, int64_t* idx_dev, double* d2_dev);
/* LBVH_find_nearest_h (LBVH/LinearBVH/LinearBVHQueries.jl:118-140) per point. */
int ph_nearest_h(ph_context* ctx, ph_lbvh* bvh, int64_t S, const double* px_host, const double* py_host,
                 const double* pz_host, double* ha_host);
/* Line-integrated kernel table (KI/kernel_function/line_integrated_kernel.jl:13-19,
 * tables of KI/table/line_integrated_kernel_tables.jl): copies the library's
 * 4096-entry float64 Inorm table of a kernel. */
int ph_line_table(int32_t kernel, double* inorm_host /* 4096 */);

/* ---- batched sampling: many PointSamples requests of uneven size over one tree
 * in one launch.  Each request is what one PointSamples_interpolation call
 * (KI/grid_interpolation/drivers/PointSamples_driver.jl:98-118, external-LBVH
 * overload) computes, and its outputs are identical to that call's
 * (ph_point_samples_device on the same coordinates), bit for bit.  The samples
 * of a request are either an explicit device list or the frame-based Cartesian
 * plane / box of PointSamples(Cartesian, frame, xparams, yparams[, zparams])
 * (Grids/grids/PointSamples.jl:328-356, coordinates by
 * Grids/grids/coordinate.jl:44-143), generated on the device with the
 * reference's arithmetic.  The frame vectors are frame_position / frame_right /
 * frame_up / frame_forward (Frames/struct/Frame.jl:145-186) evaluated by the
 * host. ---- */
#define PH_REQ_POINTS 0 /* count points at px, py, pz (device float64)                  */
#define PH_REQ_PLANE 1  /* n[0] x n[1] points: position + xi right + eta up              */
#define PH_REQ_BOX 2    /* n[0] x n[1] x n[2] points: ... + zeta forward                 */
#define PH_REQ_POLAR 3  /* n[0] x n[1] points: position + s cos(phi) right + s sin(phi) up */
#define PH_REQ_CYLINDER 4 /* n[0] x n[1] x n[2]: ... + zeta forward                       */
typedef struct ph_request {
    int32_t kind;
    int32_t reserved;
    int64_t count;                      /* PH_REQ_POINTS                                  */
    const double* px;                   /* PH_REQ_POINTS: device coordinates              */
    const double* py;
    const double* pz;
    double position[3], right[3], up[3], forward[3]; /* PLANE / BOX frame              */
    double amin[3], amax[3];            /* (xmin, ymin, zmin), (xmax, ymax, zmax)         */
    int64_t n[3];                       /* (nx, ny, nz), each >= 2; nz unused by PLANE    */
    void* out;                          /* device, L x S: double (FP64) or float (FP32C)  */
    const double* trig;                 /* POLAR / CYLINDER: host (sin phi_j, cos phi_j)  */
                                        /* pairs for j < n[1], or NULL (C library sin/cos) */
} ph_request;
typedef struct ph_batch ph_batch;
/* Validates the requests (the reference's ArgumentError rules of coordinate.jl:
 * max > min and n >= 2 per axis), copies them and records the batch as one HIP
 * graph: per catalog chunk the particle records are packed and a persistent
 * work-queue kernel serves fixed-size work units of every request (nearest-h
 * and accumulation fused) until the queue is empty.  The float64 tree must
 * stay alive; device pointers in the requests must stay valid while the batch
 * is used (their contents may change between runs). */
int ph_batch_create(ph_context* ctx, ph_lbvh* bvh, const ph_catalog* cat, int32_t strategy, int32_t precision,
                    int32_t nreq, const ph_request* reqs, ph_batch** out);
/* One launch of the recorded graph; returns when every request is done.
 * ph_get_stats: interp_ms = kernel_ms = the launch, samples = all requests. */
int ph_batch_run(ph_batch* batch);
/* Per request, from device timestamps of the last run: first_us = start of its
 * first work unit, last_us = end of its last, both in microseconds after the
 * launch began (NaN for requests without samples). */
int ph_batch_times(ph_batch* batch, double* first_us_host, double* last_us_host);
int ph_batch_destroy(ph_batch* batch);

#ifdef __cplusplus
}
#endif
#endif /* PARTIA_HIP_H */
