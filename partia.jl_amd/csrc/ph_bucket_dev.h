// ph_bucket_dev.h -- device side of the bucket traversal, shared by the
// per-call kernels (ph_bucket.hip) and the batched work-queue kernel
// (ph_batch.hip).  Included by those .hip sources only.
//
// Exactness argument.  The reference accepts leaf b for sample a iff
// d2(a, b) <= r2(a, b) in the working precision (LinearBVHTraversalMacros.jl:85,
// 209, 344) and accumulates accepted leaves in DFS preorder, i.e. in increasing
// Morton index.  Node pruning only skips leaves that fail that test (a leaf lies
// in its ancestors' boxes and its h is at most their scale), so any traversal
// that enumerates a superset of the accepted leaves in increasing index order
// and applies the same float64 leaf test gives the same neighbour sets in the
// same order.  This file uses such a traversal:
//
//  * The LBVH is copied into a "bucket tree": every maximal subtree with at most
//    kBucket leaves becomes one bucket (its leaves are consecutive in Morton
//    order), and the remaining internal nodes and the buckets are stored in DFS
//    preorder as 32-byte records with float32 boxes rounded outward and float32
//    scales rounded up.  In preorder the left child of a record is the next
//    record and a pruned subtree is skipped with one escape index, so a lane
//    walks memory forwards only.
//  * Each lane walks the bucket tree for its sample, pruning conservatively
//    (the sample is rounded to float32 and the radius inflated to cover every
//    rounding), and collects up to kQ buckets; then it tests their leaves with
//    the reference's float64 rule (independent loads, no pointer chasing) into a
//    64-bit mask; then the lanes accumulate their accepted leaves in lockstep,
//    lowest bit first, i.e. in increasing leaf index.
//
// The float64 variant accumulates with the float64 code of the per-lane path, so
// its results are bit-identical to the oracle.  The FP32-compensated variant
// evaluates kernel values and gradients in float32 from the float64 distance
// (the kernel argument q = r/h carries a single rounding; v = m/rho and 1/h are
// rounded once), and accumulates every term into float64 totals with the
// float64 field values: on MI355X float64 FMA issues at the float32 rate, so
// float64 accumulators cost less than float32 two-sum accumulators and carry the
// sums exactly to float64 rounding.  The Shepard division and the mean
// subtraction of the derivative operators (general_interpolation.jl:128-214)
// are float64; outputs are float32.
#pragma once

#include "ph_interp.h"
#include "ph_interp_dev.h"
#include "ph_primitives.h"
#include "ph_sph.h"

namespace ph {
namespace {


// Bucket geometry (compile-time; kQ * kBucket mask bits per lane and batch).
#ifndef PH_BUCKET_LEAVES
#define PH_BUCKET_LEAVES 8
#endif
#ifndef PH_BUCKET_BATCH
#define PH_BUCKET_BATCH 16
#endif
#ifndef PH_ACC_UNROLL
#define PH_ACC_UNROLL 1
#endif
constexpr int kBucket = PH_BUCKET_LEAVES;  // leaves per bucket at most (count stored in 5 bits)
// The bucket tree of the per-lane kernels, batches, adjoints and queries
// (kBucket-leaf buckets serve the grid tile kernels).
constexpr int kBucketSmall = kBucket < 4 ? kBucket : 4;
// The per-lane kernels walk the small-bucket tree: mask bit q * kBucketSmall + j.
constexpr int kLog2BucketSmall = kBucketSmall == 4 ? 2 : (kBucketSmall == 8 ? 3 : (kBucketSmall == 16 ? 4 : 5));
constexpr int kQ = PH_BUCKET_BATCH;        // buckets tested per batch
static_assert(kBucket == 4 || kBucket == 8 || kBucket == 16 || kBucket == 32, "bucket size 4, 8, 16 or 32");
static_assert(kQ * kBucketSmall <= 64, "one 64-bit mask per lane and batch");
// Leaves of a tree with bucket trees: the LinearBVH's own limit
// (LinearBoundingVolumeHierarchy/BinaryRadixTree/BinaryRadixTree.jl:86-91).
constexpr i64 kMaxBucketLeaves = i64(1) << 30;

// Preorder record: link >= 0 -> internal node, link = index after its subtree;
// link < 0 -> bucket of leaves [lo, lo + cnt), link = -(1 + (lo << 5 | (cnt - 1))).
struct alignas(16) BucketRec {
    float lo[3];
    float hi[3];
    float scale;
    i32 link;
};

// Raw buffer view (gfx9 descriptor, word 3 = 0x00020000): loads take a per-lane
// byte offset in a VGPR and a uniform one in an SGPR, so a field at a run-time
// (uniform) position costs no vector address arithmetic.  Offsets and the size
// are unsigned 32-bit (buffers below 4 GiB).
struct Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ static Buf of(const void* p, i64 bytes) {
        return Buf{__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)(u32)bytes, 0x00020000)};
    }
    __device__ double f64(u32 voff, u32 soff) const {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
    }
    __device__ LeafRec<double> leaf(u32 voff) const {
        const auto a = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, 0, 0);
        const auto b = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff + 16, 0, 0);
        const double2 xy = __builtin_bit_cast(double2, a), zh = __builtin_bit_cast(double2, b);
        return LeafRec<double>{xy.x, xy.y, zh.x, zh.y};
    }
};

// A bucket record: link = -(1 + first leaf index) (any index below 2^31 - 1),
// and the leaf count - 1 in the five low bits of the scale, which is first
// rounded up to a multiple of 32 units in the last place: the scale stays an
// upper bound of the bucket's h (only pruning tests read it; the box and the
// leaf tests are untouched).  An internal record's link is its escape index
// (>= 0) and its scale is the plain rounded-up maximum.
__device__ inline void encode_bucket(BucketRec& R, i64 lo0, i64 cnt) {
    R.link = -(i32)(1 + lo0);
    u32 b = __builtin_bit_cast(u32, R.scale);
    if (b < 0x7f800000u) b = b > 0x7f7fffe0u ? 0x7f7fffe0u : ((b + 31u) & ~31u);  // finite: round up
    R.scale = __builtin_bit_cast(float, (b & ~31u) | (u32)(cnt - 1));
}
__device__ inline i64 bucket_lo(i32 link) { return -(i64)link - 1; }
__device__ inline int bucket_cnt(const BucketRec& R) { return (int)(__builtin_bit_cast(u32, R.scale) & 31u) + 1; }

__device__ inline float next_down(float f) {
    if (f != f || f == -INFINITY) return f;
    if (f == 0.f) return -__uint_as_float(1u);
    const u32 b = __float_as_uint(f);
    return __uint_as_float(f > 0.f ? b - 1u : b + 1u);
}
__device__ inline float next_up(float f) {
    if (f != f || f == INFINITY) return f;
    if (f == 0.f) return __uint_as_float(1u);
    const u32 b = __float_as_uint(f);
    return __uint_as_float(f > 0.f ? b + 1u : b - 1u);
}
__device__ inline float f32_down(double x) {
    const float f = (float)x;
    return ((double)f > x) ? next_down(f) : f;
}
__device__ inline float f32_up(double x) {
    const float f = (float)x;
    return ((double)f < x) ? next_up(f) : f;
}



// ---- per-call records of the FP32-compensated path ----

// FP32-compensated record per particle, rw doubles: element 0 holds {m/rho, 1/h}
// as two float32 (rounded once from float64), elements 1..ncol the float64
// values of the distinct columns the catalog chunk uses (source c32.rsrc[1 + j]).
// Record n (one past the particles) is the zero record {0, 1}, 0, ...: lanes
// without a pending neighbour read it in the lockstep loops, so every term they
// add is 0 whatever the particles' columns hold (a NaN included).  Launch with
// n + 1 threads; rec holds (n + 1) rw doubles.
__global__ void __launch_bounds__(kBlock) k_pack_c32(const double* __restrict__ cols, i64 n, CatChunk c32, int ncol,
                                             int rw, double* __restrict__ rec) {
    const i64 b = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (b > n) return;
    double* r = rec + b * rw;
    if (b == n) {
        r[0] = __builtin_bit_cast(double, make_float2(0.f, 1.f));
        for (int j = 1; j < rw; ++j) r[j] = 0.0;
        return;
    }
    const float2 vi = make_float2((float)(cols[3 * n + b] / cols[5 * n + b]), (float)(1.0 / cols[4 * n + b]));
    r[0] = __builtin_bit_cast(double, vi);
    for (int j = 0; j < ncol; ++j) r[1 + j] = cols[(i64)c32.rsrc[1 + j] * n + b];
    for (int j = 1 + ncol; j < rw; ++j) r[j] = 0.0;
}

// ---- FP32-compensated accumulator ----

// The Wendland kernels: f'(q) = q g(q) with a polynomial g (C2: -5 t^3; C4:
// -14/3 t^5 (1 + 5/2 q); C6: -11/2 t^7 (1 + 7/2 q + 4 q^2); t = (1 - q/2)_+),
// so their gradient sigma f'(q) / (h^4 r) delta = sigma g(q) / h^5 delta needs
// no 1/r and is 0 at r = 0 like the reference's (kernel.jl:273-285).
template <int K> constexpr bool wendland32() {
    return K == PH_KERNEL_C2 || K == PH_KERNEL_C4 || K == PH_KERNEL_C6;
}
// Gradient norm of a kernel with 1/h = inv_h and value norm W_norm = sigma / h^3:
// sigma / h^4 (times f'(q) / r), or sigma / h^5 for the Wendland kernels (times g(q)).
template <int K> __device__ inline float grad_norm32(float W_norm, float inv_h) {
    return wendland32<K>() ? W_norm * inv_h * inv_h : W_norm * inv_h;
}

// The symmetric strategy averages the two kernels: each norm carries the
// factor 1/2 (exact, a power of two).
template <int STRAT> constexpr float mix32() { return STRAT == PH_SYMMETRIC ? 0.5f : 1.f; }
// Norms of a sample's own kernel (1/h_a = inv_ha): value and gradient.
template <int K, int STRAT> __device__ inline void sample_norms32(float inv_ha, float& Wn, float& Gn) {
    const float W = sigma3<float, K>() * inv_ha * inv_ha * inv_ha;
    Wn = mix32<STRAT>() * W;
    Gn = mix32<STRAT>() * grad_norm32<K>(W, inv_ha);
}
// Neighbour-only terms of the scatter and symmetric kernels: (1/h_b)^2 in
// float64 (exact: the square of a float32) and the float32 norms sigma / h_b^3
// and grad_norm32 (times mix32), from the record's float32 1/h_b.
struct LeafNorms {
    double ib2;
    float Wn, Gn;
};
template <int K, int STRAT> __device__ inline LeafNorms leaf_norms32(float ib) {
    const double ibd = ib;
    const float W = sigma3<float, K>() * (ib * ib * ib);
    return LeafNorms{ibd * ibd, mix32<STRAT>() * W, mix32<STRAT>() * grad_norm32<K>(W, ib)};
}

// Shape f(q) and df = f'(q) (the Wendland kernels: g(q) = f'(q) / q) in
// float32 without branches: every piece is written with clamped factors
// (x)_+ = max(x, 0), which reproduces the piecewise definitions of
// KI/kernel_function/kernels/*.jl (zero beyond the support); q >= 0 here.
template <int K> __device__ __attribute__((always_inline)) void shape32(float q, float& f, float& df) {
    if constexpr (K == PH_KERNEL_M4) {
        const float a = fmaxf(2.f - q, 0.f), b = fmaxf(1.f - q, 0.f);
        const float a2 = a * a, b2 = b * b;
        f = fmaf(0.25f * a2, a, -(b2 * b));
        df = fmaf(-0.75f, a2, 3.f * b2);
    } else if constexpr (K == PH_KERNEL_M5) {
        const float a = fmaxf(2.5f - q, 0.f), b = fmaxf(1.5f - q, 0.f), c = fmaxf(0.5f - q, 0.f);
        const float a3 = a * a * a, b3 = b * b * b, c3 = c * c * c;
        f = fmaf(a3, a, fmaf(-5.f * b3, b, 10.f * c3 * c));
        df = fmaf(-4.f, a3, fmaf(20.f, b3, -40.f * c3));
    } else if constexpr (K == PH_KERNEL_M6) {
        const float a = fmaxf(3.f - q, 0.f), b = fmaxf(2.f - q, 0.f), c = fmaxf(1.f - q, 0.f);
        const float a2 = a * a, b2 = b * b, c2 = c * c;
        const float a4 = a2 * a2, b4 = b2 * b2, c4 = c2 * c2;
        f = fmaf(a4, a, fmaf(-6.f * b4, b, 15.f * c4 * c));
        df = fmaf(-5.f, a4, fmaf(30.f, b4, -75.f * c4));
    } else {
        const float t = fmaxf(fmaf(-0.5f, q, 1.f), 0.f);
        const float t2 = t * t, t3 = t2 * t, t4 = t2 * t2;
        if constexpr (K == PH_KERNEL_C2) {
            f = t4 * fmaf(2.f, q, 1.f);
            df = -5.f * t3;
        } else if constexpr (K == PH_KERNEL_C4) {
            const float t5 = t4 * t, t6 = t4 * t2;
            f = t6 * fmaf(fmaf(35.f / 12.f, q, 3.f), q, 1.f);
            df = (-14.f / 3.f) * t5 * fmaf(2.5f, q, 1.f);
        } else {
            const float t7 = t4 * t3, t8 = t4 * t4;
            f = t8 * fmaf(fmaf(fmaf(4.f, q, 6.25f), q, 4.f), q, 1.f);
            df = -5.5f * t7 * fmaf(fmaf(4.f, q, 3.5f), q, 1.f);
        }
    }
}

template <int STRAT> __device__ inline float radius2_32(float Kv, float r2a, float h) {
    if constexpr (STRAT == PH_GATHER) {
        return r2a;
    } else {
        const float r = Kv * h;
        const float rr = r * r;
        if constexpr (STRAT == PH_SCATTER) return rr;
        else return fmaxf(rr, r2a);
    }
}

// FP32-compensated accumulator.  Kernel values and gradients are evaluated in
// float32 from the float64 distance (so the kernel argument q carries a single
// rounding) with v = m/rho and 1/h rounded once to float32; every term goes into
// float64 totals with the float64 field values, where the products are exact to
// float64 rounding.  Every output keeps its own totals (static register
// indices), and a lane without a pending neighbour adds exact zeros, so the
// lockstep loop has no divergent branch.  Record positions index the
// FP32-compensated record (position 0: {m/rho, 1/h}; >= 1: float64 columns).  EXACT: the catalog chunk has
// exactly MS scalars, MG gradients, MD divergences and MC curls; otherwise the
// unused output slots read position 0 and add zeros (no branch either).
// EXACT = 2 additionally shares sums the C2-shaped catalog computes twice (the
// gradient field is the second scalar's column, the curl the divergence's
// columns): the same fma sequence over the same operands, so identical values;
// its record positions are the fixed compact layout c2_shared checks, so the
// record reads have constant offsets.
template <int K, int STRAT, int MS, int MG, int MD, int MC, int EXACT> struct AccC {
    static constexpr bool SHARE = EXACT == 2;
    // record positions: the chunk's, or the C2 layout {vi, s0, s1 = g0, d0..d2 = c0..c2}
    __device__ static int spos(const CatChunk& c, int j) { return SHARE ? 1 + j : c.spos[j]; }
    __device__ static int gpos(const CatChunk& c, int j) { return SHARE ? 2 : c.gpos[j]; }
    __device__ static int dpos(const CatChunk& c, int k) { return SHARE ? 3 + k : c.dpos[k]; }
    __device__ static int cpos(const CatChunk& c, int k) { return SHARE ? 3 + k : c.cpos[k]; }
    static constexpr int NS = MS > 0 ? MS : 1, G = MG > 0 ? MG : 1, D = MD > 0 ? MD : 1, CC = MC > 0 ? MC : 1;
    double S1, Q[3];
    double ts[NS];
    double tgs[G], tg[G][3];
    double tds[D][3], td[D];
    double tcs[CC][3], tc[CC][3];

    __device__ __attribute__((always_inline)) void init() {
        S1 = 0.0;
        Q[0] = Q[1] = Q[2] = 0.0;
#pragma unroll
        for (int j = 0; j < NS; ++j) ts[j] = 0.0;
#pragma unroll
        for (int j = 0; j < G; ++j) tgs[j] = tg[j][0] = tg[j][1] = tg[j][2] = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) td[j] = tds[j][0] = tds[j][1] = tds[j][2] = 0.0;
#pragma unroll
        for (int j = 0; j < CC; ++j) tcs[j][0] = tcs[j][1] = tcs[j][2] = tc[j][0] = tc[j][1] = tc[j][2] = 0.0;
    }

    // One neighbour (live = false: adds zeros): d = sample - particle and d2 in
    // float64, r = record; iha2 = (1/ha)^2 and the float32 norms of the sample's
    // kernel.
    __device__ __attribute__((always_inline)) void add(bool live, const CatChunk& c, const Buf& rb, u32 voff,
                                                       double dx, double dy, double dz, double d2, double iha2,
                                                       float Wa_norm, float Ga_norm) {
        add_r(live, c, [&](int pos) { return rb.f64(voff, (u32)pos * 8u); }, dx, dy, dz, d2, iha2, Wa_norm, Ga_norm);
    }
    // The same with the record read through r(pos).
    template <class Rec>
    __device__ __attribute__((always_inline)) void add_r(bool live, const CatChunk& c, Rec&& r, double dx, double dy,
                                                         double dz, double d2, double iha2, float Wa_norm,
                                                         float Ga_norm) {
        LeafNorms nb{0.0, 0.f, 0.f};
        if constexpr (STRAT != PH_GATHER) nb = leaf_norms32<K, STRAT>(__builtin_bit_cast(float2, r(0)).y);
        add_rn(live, c, r, nb, dx, dy, dz, d2, iha2, Wa_norm, Ga_norm);
    }
    // The same with the neighbour's norms nb given (leaf_norms32 of its 1/h).
    // A lane without a pending neighbour (live false) reads a zero record
    // (m / rho = 0, zero columns) at a finite position, so every term it adds
    // is an exact zero; no select is needed.
    template <class Rec>
    __device__ __attribute__((always_inline)) void add_rn(bool live, const CatChunk& c, Rec&& r, const LeafNorms& nb,
                                                          double dx, double dy, double dz, double d2, double iha2,
                                                          float Wa_norm, float Ga_norm) {
        const float2 vi = __builtin_bit_cast(float2, r(0));
        const float d2f = (float)d2;
        float W, coef;
        if constexpr (STRAT == PH_GATHER) {
            (void)nb;
            float f, df;
            shape32<K>(__builtin_amdgcn_sqrtf((float)(d2 * iha2)), f, df);
            W = Wa_norm * f;
            coef = Ga_norm * df;
        } else {
            float fb, dfb;
            shape32<K>(__builtin_amdgcn_sqrtf((float)(d2 * nb.ib2)), fb, dfb);
            const float Wb = nb.Wn * fb, Gb = nb.Gn * dfb;
            if constexpr (STRAT == PH_SCATTER) {
                W = Wb;
                coef = Gb;
            } else {  // the norms carry the mean's 1/2
                float fa, dfa;
                shape32<K>(__builtin_amdgcn_sqrtf((float)(d2 * iha2)), fa, dfa);
                W = Wa_norm * fa + Wb;
                coef = Ga_norm * dfa + Gb;
            }
        }
        (void)live;
        // float64 accumulation with fused products (explicit fma: the build
        // disables contraction so that the float64 path stays reference-exact)
        const double w = (double)(vi.x * W);
        S1 += w;
#pragma unroll
        for (int j = 0; j < MS; ++j) {
            const double wj = (EXACT || j < c.ns) ? w : 0.0;
            ts[j] = fma(r(spos(c, j)), wj, ts[j]);
        }
        if constexpr (MG + MD + MC > 0) {
            // grad W = delta f'(q) sigma / (r h^4); zero at r = 0 (kernel.jl:273-285)
            double s;
            if constexpr (wendland32<K>()) s = (double)(vi.x * coef);
            else s = d2 > 0.0 ? (double)(vi.x * coef * __builtin_amdgcn_rsqf(d2f)) : 0.0;
            const double g0 = dx * s, g1 = dy * s, g2 = dz * s;
            Q[0] += g0;
            Q[1] += g1;
            Q[2] += g2;
#pragma unroll
            for (int j = 0; j < MG; ++j) {
                const bool on = EXACT || j < c.ng;
                const double A = on ? r(gpos(c, j)) : 0.0;
                if constexpr (!SHARE) tgs[j] = fma(A, w, tgs[j]);
                tg[j][0] = fma(A, g0, tg[j][0]);
                tg[j][1] = fma(A, g1, tg[j][1]);
                tg[j][2] = fma(A, g2, tg[j][2]);
            }
#pragma unroll
            for (int j = 0; j < MD; ++j) {
                const bool on = EXACT || j < c.nd;
                const double ax = on ? r(dpos(c, 3 * j)) : 0.0, ay = on ? r(dpos(c, 3 * j + 1)) : 0.0,
                             az = on ? r(dpos(c, 3 * j + 2)) : 0.0;
                tds[j][0] = fma(ax, w, tds[j][0]);
                tds[j][1] = fma(ay, w, tds[j][1]);
                tds[j][2] = fma(az, w, tds[j][2]);
                td[j] = fma(az, g2, fma(ay, g1, fma(ax, g0, td[j])));
            }
#pragma unroll
            for (int j = 0; j < MC; ++j) {
                const bool on = EXACT || j < c.nc;
                const double ax = on ? r(cpos(c, 3 * j)) : 0.0, ay = on ? r(cpos(c, 3 * j + 1)) : 0.0,
                             az = on ? r(cpos(c, 3 * j + 2)) : 0.0;
                if constexpr (!SHARE) {
                    tcs[j][0] = fma(ax, w, tcs[j][0]);
                    tcs[j][1] = fma(ay, w, tcs[j][1]);
                    tcs[j][2] = fma(az, w, tcs[j][2]);
                }
                tc[j][0] = fma(-az, g1, fma(ay, g2, tc[j][0]));
                tc[j][1] = fma(-ax, g2, fma(az, g0, tc[j][1]));
                tc[j][2] = fma(-ay, g0, fma(ax, g1, tc[j][2]));
            }
        }
    }

    // Epilogue in float64 (general_interpolation.jl:128-214); outputs float32.
    __device__ __attribute__((always_inline)) void finish(const CatChunk& c, float* __restrict__ out, i64 S,
                                                          i64 s) const {
        if (S1 == 0.0) {
            const float nan = __builtin_nanf("");
#pragma unroll
            for (int j = 0; j < MS; ++j)
                if (j < c.ns) out[c.sout[j] * S + s] = nan;
#pragma unroll
            for (int j = 0; j < MG; ++j)
                if (j < c.ng) {
                    out[(c.gout[j] + 0) * S + s] = nan;
                    out[(c.gout[j] + 1) * S + s] = nan;
                    out[(c.gout[j] + 2) * S + s] = nan;
                }
#pragma unroll
            for (int j = 0; j < MD; ++j)
                if (j < c.nd) out[c.dout[j] * S + s] = nan;
#pragma unroll
            for (int j = 0; j < MC; ++j)
                if (j < c.nc) {
                    out[(c.cout[j] + 0) * S + s] = nan;
                    out[(c.cout[j] + 1) * S + s] = nan;
                    out[(c.cout[j] + 2) * S + s] = nan;
                }
            return;
        }
        const double invS1 = 1.0 / S1;
#pragma unroll
        for (int j = 0; j < MS; ++j)
            if (j < c.ns) out[c.sout[j] * S + s] = (float)(c.sflag[j] ? ts[j] * invS1 : ts[j]);
#pragma unroll
        for (int j = 0; j < MG; ++j)
            if (j < c.ng) {
                const double A = (SHARE ? ts[1] : tgs[j]) * invS1;
                out[(c.gout[j] + 0) * S + s] = (float)(tg[j][0] - A * Q[0]);
                out[(c.gout[j] + 1) * S + s] = (float)(tg[j][1] - A * Q[1]);
                out[(c.gout[j] + 2) * S + s] = (float)(tg[j][2] - A * Q[2]);
            }
#pragma unroll
        for (int j = 0; j < MD; ++j)
            if (j < c.nd) {
                const double ax = tds[j][0] * invS1, ay = tds[j][1] * invS1, az = tds[j][2] * invS1;
                out[c.dout[j] * S + s] = (float)(td[j] - ((ax * Q[0] + ay * Q[1]) + az * Q[2]));
            }
#pragma unroll
        for (int j = 0; j < MC; ++j)
            if (j < c.nc) {
                const double(&cs)[3] = SHARE ? tds[j] : tcs[j];
                const double ax = cs[0] * invS1, ay = cs[1] * invS1, az = cs[2] * invS1;
                out[(c.cout[j] + 0) * S + s] = (float)(-(tc[j][0] - (ay * Q[2] - az * Q[1])));
                out[(c.cout[j] + 1) * S + s] = (float)(-(tc[j][1] - (az * Q[0] - ax * Q[2])));
                out[(c.cout[j] + 2) * S + s] = (float)(-(tc[j][2] - (ax * Q[1] - ay * Q[0])));
            }
    }
};

struct BucketView {
    i64 n;
    i32 nrec;
    const BucketRec* __restrict__ recs;
    const LeafRec<double>* __restrict__ leaves;
};

constexpr float kInflate = 1.0f + 9.5367432e-7f;  // 1 + 2^-20

// Conservative float32 pruning data of one sample.
template <int STRAT> struct Probe {
    float px, py, pz;
    float Kf, Rg2, eps;
    __device__ inline void init(const double p[3], double Kd, double ra) {
        px = (float)p[0];
        py = (float)p[1];
        pz = (float)p[2];
        // |p32 - p| <= 2^-24 |p| per axis; eps covers it in every direction
        eps = 2.3841858e-7f * fmaxf(fabsf(px), fmaxf(fabsf(py), fabsf(pz))) + 1e-30f;
        Kf = (float)Kd;
        Rg2 = 0.f;
        if constexpr (STRAT != PH_SCATTER) {
            const float Rg = f32_up(ra) * kInflate + eps;
            Rg2 = Rg * Rg * kInflate;
        }
    }
    // true when the record's subtree may hold an accepted leaf
    __device__ inline bool keep(const BucketRec& R) const {
        const float p[3] = {px, py, pz};
        float d2 = 0.f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float g = fmaxf(fmaxf(R.lo[d] - p[d], p[d] - R.hi[d]), 0.f);
            d2 += g * g;
        }
        if constexpr (STRAT == PH_GATHER) return !(d2 > Rg2);
        const float rs = Kf * R.scale * kInflate + eps;
        const float rs2 = rs * rs * kInflate;
        if constexpr (STRAT == PH_SCATTER) return !(d2 > rs2);
        return !(d2 > rs2) || !(d2 > Rg2);  // d2 <= max(rs2, Rg2)
    }
    // The same test as a wave mask (every lane active): the compares give the
    // lane mask directly, with no boolean round trip through a vector register.
    __device__ inline u64 keep_mask(const BucketRec& R) const {
        constexpr int kUle = 13;  // unordered or <=, i.e. !(d2 > R2)
        const float p[3] = {px, py, pz};
        float d2 = 0.f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float g = fmaxf(fmaxf(R.lo[d] - p[d], p[d] - R.hi[d]), 0.f);
            d2 += g * g;
        }
        if constexpr (STRAT == PH_GATHER) return __builtin_amdgcn_fcmpf(d2, Rg2, kUle);
        const float rs = Kf * R.scale * kInflate + eps;
        const float rs2 = rs * rs * kInflate;
        if constexpr (STRAT == PH_SCATTER) return __builtin_amdgcn_fcmpf(d2, rs2, kUle);
        return __builtin_amdgcn_fcmpf(d2, rs2, kUle) | __builtin_amdgcn_fcmpf(d2, Rg2, kUle);
    }
};

// Per-lane list of up to kQ buckets in LDS (entry q at slot[q * stride]): the
// bucket code -link - 1 = first leaf << 5 | (count - 1).
// A lane's listed buckets of the kBucketSmall-leaf tree: (first leaf <<
// kLog2BucketSmall) | (count - 1) per entry (first leaf < 2^30).
struct BucketList {
    u32* slot;
    int n;
    int stride = kBlock;
    __device__ __attribute__((always_inline)) void push(const BucketRec& R) {
        slot[n++ * stride] = ((u32)bucket_lo(R.link) << kLog2BucketSmall) | (u32)(bucket_cnt(R) - 1);
    }
    __device__ __attribute__((always_inline)) u32 code(int q) const { return slot[q * stride]; }
    __device__ __attribute__((always_inline)) u32 lo(int q) const { return code(q) >> kLog2BucketSmall; }
    __device__ __attribute__((always_inline)) int cnt(int q) const {
        return (int)(code(q) & (u32)(kBucketSmall - 1)) + 1;
    }
};

// Walks from record i until kQ buckets are collected or the tree is done.
template <int STRAT>
__device__ __attribute__((always_inline)) void collect(const BucketView& bv, const Probe<STRAT>& pr, i32& i,
                                                       i64& steps, BucketList& bl) {
    bl.n = 0;
    while (i < bv.nrec && steps > 0 && bl.n < kQ) {
        --steps;
        const BucketRec R = bv.recs[i];
        const bool k = pr.keep(R);
        if (R.link >= 0) {
            i = k ? i + 1 : R.link;
        } else {
            if (k) bl.push(R);
            ++i;
        }
    }
}

// Float64 leaf tests of the listed buckets: bit q * kBucket + j for leaf j of bucket q.
template <int STRAT>
__device__ __attribute__((always_inline)) u64 test_buckets(const BucketView& bv, const BucketList& bl,
                                                           const double p[3], double Kd, double r2a) {
    u64 mask = 0;
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        if (q < bl.n) {
            const i64 lo = bl.lo(q);
            int cnt = bl.cnt(q);
            // the count's range (1..4) is hidden from the optimiser: the loop
            // below keeps its partial unrolling instead of four loads in flight
            // per lane (registers: occupancy of the per-lane and batch kernels)
            __asm__ volatile("" : "+v"(cnt));
#pragma unroll 4
            for (int j = 0; j < cnt; ++j) {
                const LeafRec<double> L = bv.leaves[lo + j];
                if (d2_point(p[0], p[1], p[2], L) <= radius2<double, STRAT>(Kd, r2a, L.h))
                    mask |= 1ull << (q * kBucketSmall + j);
            }
        }
    }
    return mask;
}

// LBVH_find_nearest_h (LinearBVHQueries.jl:118-140) on the bucket tree: the
// lexicographic minimum of (float64 d2, leaf index), i.e. the first nearest leaf
// in DFS order.  A greedy descent seeds the bound; pruning compares float32 box
// distances with the bound inflated to cover rounding, so no candidate is lost.
// Only leaves with d2 <= bound2 are candidates: NaN when there is none (with
// bound2 = infinity, only for an empty tree).
// Nearest-particle search state: the lexicographic minimum of (d2, leaf
// index) over the leaves seen so far with d2 <= the initial bound.  The result
// does not depend on the order in which leaves are offered, so any superset of
// the leaves the pruned walk must see gives the same answer.
struct Nearest {
    static constexpr i64 kNone = INT64_MAX;
    double best;  // a leaf at exactly the bound is a candidate (k < kNone)
    i64 bk = kNone;
    __device__ inline void offer(const double p[3], const LeafRec<double>& L, i64 k) {
        const double d2 = d2_point(p[0], p[1], p[2], L);
        if (d2 < best || (d2 == best && k < bk)) {
            best = d2;
            bk = k;
        }
    }
    __device__ inline double h(const BucketView& bv) const { return bk != kNone ? bv.leaves[bk].h : (double)NAN; }
};

__device__ inline float box_d2_32(const BucketRec& R, float px, float py, float pz) {
    const float q[3] = {px, py, pz};
    float d2 = 0.f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float g = fmaxf(fmaxf(R.lo[d] - q[d], q[d] - R.hi[d]), 0.f);
        d2 += g * g;
    }
    return d2;
}

__device__ inline void nearest_scan(const BucketView& bv, const double p[3], const BucketRec& R, Nearest& nr) {
    const i64 lo = bucket_lo(R.link);
    const int cnt = bucket_cnt(R);
    for (int j = 0; j < cnt; ++j) nr.offer(p, bv.leaves[lo + j], lo + j);
}

// The pruned walk from the root: every bucket whose box may hold a leaf within
// the current best distance (float32 test inflated past rounding) is scanned.
__device__ inline void nearest_walk(const BucketView& bv, const double p[3], Nearest& nr) {
    const float px = (float)p[0], py = (float)p[1], pz = (float)p[2];
    const float eps = 2.3841858e-7f * fmaxf(fabsf(px), fmaxf(fabsf(py), fabsf(pz))) + 1e-30f;
    i32 i = 0;
    for (i64 steps = 2 * (i64)bv.nrec + 8; i < bv.nrec && steps > 0; --steps) {
        const BucketRec R = bv.recs[i];
        const float bound = (float)nr.best;
        const float lim = (sqrtf(bound) * kInflate + eps);
        const bool keep = !(box_d2_32(R, px, py, pz) > lim * lim * kInflate);
        if (R.link >= 0) {
            i = keep ? i + 1 : R.link;
        } else {
            if (keep) nearest_scan(bv, p, R, nr);
            ++i;
        }
    }
}

__device__ inline Nearest nearest_bucket(const BucketView& bv, const double p[3], double bound2 = INFINITY) {
    const float px = (float)p[0], py = (float)p[1], pz = (float)p[2];
    Nearest nr{bound2};
    // greedy descent: nearer child first (left child = next record; right child
    // = the record after the left subtree); its bucket seeds the bound
    i32 i = 0;
    for (i64 steps = bv.nrec; steps > 0; --steps) {
        const BucketRec R = bv.recs[i];
        if (R.link < 0) {
            nearest_scan(bv, p, R, nr);
            break;
        }
        const i32 l = i + 1;
        const BucketRec RL = bv.recs[l];
        const i32 r = RL.link >= 0 ? RL.link : l + 1;
        i = box_d2_32(bv.recs[r], px, py, pz) < box_d2_32(RL, px, py, pz) ? r : l;
    }
    nearest_walk(bv, p, nr);
    return nr;
}
__device__ inline double nearest_h_bucket(const BucketView& bv, const double p[3], double bound2 = INFINITY) {
    return nearest_bucket(bv, p, bound2).h(bv);
}

// Bound of the nearest-h search of a sample that is interpolated with support
// K: a particle farther than K * max(h) from the sample is no neighbour in any
// strategy (every radius is K h_a or K h_b, h_a being some particle's h), so
// when no particle lies within it the outputs are NaN whatever h_a is, and the
// search may stop there.  The root record's scale is max(h) rounded up.
__device__ inline double neighbour_bound2(const BucketView& bv, double Kd) {
    const double r = Kd * (double)bv.recs[0].scale;
    return r * r;
}


// ---- per-sample bodies (one lane = one sample; every lane of a wave calls
// them together, valid or not, since they use wave-wide votes) ----

// FP32-compensated interpolation of sample s at p (ha: smoothing length of its
// nearest particle, gather / symmetric).  out: L x S float32.
// list: this lane's kQ bucket-list entries in LDS, entry q at list[q * stride].
template <int K, int STRAT, int MS, int MG, int MD, int MC, int EXACT, bool WIDE = false>
__device__ __attribute__((always_inline)) void bucket_c32_sample(const BucketView& bv, const double* __restrict__ rec,
                                                                 const CatChunk& c, bool valid, const double p[3],
                                                                 double ha, float* __restrict__ out, i64 S, i64 s,
                                                                 u64& accepted, u32* list, int stride) {
    const double Kd = support_of<K>();
    // no particle within K max(h): no neighbours (NaN outputs), no walk
    const bool none = STRAT != PH_SCATTER && !(ha == ha);
    if (!valid || none) ha = 0.0;
    const double ra = Kd * ha, r2a = ra * ra;  // float64, as the reference
    Probe<STRAT> pr;
    pr.init(p, Kd, ra);
    const double iha = STRAT == PH_SCATTER || !valid || none ? 0.0 : 1.0 / ha;
    const double iha2 = iha * iha;
    const float inv_ha = (float)iha;
    float Wa_norm, Ga_norm;
    sample_norms32<K, STRAT>(inv_ha, Wa_norm, Ga_norm);
    const int rw = c.rw;

    // per-lane bucket list in LDS: read back with a per-lane index in the
    // accumulation loop, which thus has no divergent branch
    const u32 nlast = (u32)(bv.n - 1);
    const Buf rb = Buf::of(rec, ((i64)bv.n + 1) * rw * 8);  // with the zero record n
    const Buf lb = Buf::of(bv.leaves, (i64)bv.n * (i64)sizeof(LeafRec<double>));
    // WIDE: the leaves or the records pass 32-bit byte offsets (more than 2^27
    // leaves, or records of 4 GiB): the same loads through 64-bit addresses

    AccC<K, STRAT, MS, MG, MD, MC, EXACT> acc;
    acc.init();
    i32 i = valid && !none ? 0 : bv.nrec;
    i64 steps = 2 * (i64)bv.nrec + 8;  // a valid tree is walked in at most nrec steps
    BucketList bl{list, 0, stride};
    while (true) {
        collect<STRAT>(bv, pr, i, steps, bl);
        u64 mask = test_buckets<STRAT>(bv, bl, p, Kd, r2a);
        accepted += __popcll(mask);
        // lockstep: one neighbour per lane and iteration, lowest bit first
        // (increasing leaf index); a lane with none left adds zeros
        int nit = __popcll(mask);
        for (int d = kWave / 2; d > 0; d >>= 1) nit = max(nit, __shfl_xor(nit, d, kWave));
        nit = __builtin_amdgcn_readfirstlane((nit + PH_ACC_UNROLL - 1) / PH_ACC_UNROLL);
        for (int it = 0; it < nit; ++it) {
            bool live[PH_ACC_UNROLL];
            u32 k[PH_ACC_UNROLL];
            LeafRec<double> L[PH_ACC_UNROLL];
#pragma unroll
            for (int u = 0; u < PH_ACC_UNROLL; ++u) {  // loads of all unrolled neighbours first
                live[u] = mask != 0;
                const int b = __builtin_ctzll(mask | (1ull << 63));
                mask &= mask - 1;
                const u32 kk = bl.lo(b >> kLog2BucketSmall) + (u32)(b & (kBucketSmall - 1));
                k[u] = kk < nlast ? kk : nlast;
                if constexpr (WIDE) L[u] = bv.leaves[k[u]];
                else L[u] = lb.leaf(k[u] * (u32)sizeof(LeafRec<double>));
            }
#pragma unroll
            for (int u = 0; u < PH_ACC_UNROLL; ++u) {
                const double dx = p[0] - L[u].x, dy = p[1] - L[u].y, dz = p[2] - L[u].z;
                const double d2 = (dx * dx + dy * dy) + dz * dz;
                // a lane without a pending neighbour reads the zero record
                const u32 kr = live[u] ? k[u] : (u32)bv.n;
                if constexpr (WIDE) {
                    const double* r = rec + (i64)kr * rw;
                    acc.add_r(live[u], c, [&](int pos) { return r[pos]; }, dx, dy, dz, d2, iha2, Wa_norm, Ga_norm);
                } else {
                    acc.add(live[u], c, rb, kr * (u32)rw * 8u, dx, dy, dz, d2, iha2, Wa_norm, Ga_norm);
                }
            }
        }
        if (!__any(i < bv.nrec && steps > 0)) break;
    }
    if (valid) acc.finish(c, out, S, s);
}

// Float64 interpolation of sample s with the accumulator of the per-lane path.
template <int K, int STRAT, int MS, int MG, int MD, int MC>
__device__ __attribute__((always_inline)) void bucket_f64_sample(const BucketView& bv, const double* __restrict__ rec,
                                                                 const CatChunk& c, bool valid, const double p[3],
                                                                 double ha, double* __restrict__ out, i64 S, i64 s,
                                                                 u64& accepted, u32* list, int stride) {
    const double Kd = support_of<K>();
    const bool none = STRAT != PH_SCATTER && !(ha == ha);  // as in bucket_c32_sample
    if (!valid || none) ha = 0.0;
    const double ra = Kd * ha, r2a = ra * ra;
    Probe<STRAT> pr;
    pr.init(p, Kd, ra);
    Acc<double, K, STRAT, MS, MG, MD, MC> acc;
    acc.init();
    i32 i = valid && !none ? 0 : bv.nrec;
    i64 steps = 2 * (i64)bv.nrec + 8;
    BucketList bl{list, 0, stride};
    while (true) {
        collect<STRAT>(bv, pr, i, steps, bl);
        u64 mask = test_buckets<STRAT>(bv, bl, p, Kd, r2a);
        accepted += __popcll(mask);
        while (__any(mask != 0)) {
            if (mask != 0) {
                const int b = __builtin_ctzll(mask);
                mask &= mask - 1;
                const i64 k = (i64)bl.lo(b >> kLog2BucketSmall) + (b & (kBucketSmall - 1));
                const LeafRec<double> L = bv.leaves[k];
                acc.add(c, rec + k * c.rw, p[0] - L.x, p[1] - L.y, p[2] - L.z, d2_point(p[0], p[1], p[2], L), ha, L.h);
            }
        }
        if (!__any(i < bv.nrec && steps > 0)) break;
    }
    if (valid) acc.finish(c, out, S, s);
}

constexpr int kBucketLarge = 16;  // the largest buckets, for grid tile walks only
constexpr int kBucketLeaf = 1;    // buckets of one leaf: the nearest-h searches' tree

// Whether the per-lane FP32C walk needs 64-bit addresses: the leaves or the
// records (rw doubles each, plus the zero record) pass 4 GiB.
inline bool c32_wide(i64 n, int rw) {
    return (n + 1) * (i64)rw * 8 > (i64)0xffffffff || n * (i64)sizeof(LeafRec<double>) > (i64)0xffffffff;
}

BucketView bucket_view(const ph_lbvh* B, int leaves = kBucketSmall);

// The tree of the nearest-h searches: the one-leaf bucket tree when it was
// built, else the kBucketSmall one (the same results: the search is exact over
// any bucket tree of the LBVH).
BucketView nearest_view(const ph_lbvh* B) { return bucket_view(B, B->buckets1_ready ? kBucketLeaf : kBucketSmall); }

BucketView bucket_view(const ph_lbvh* B, int leaves) {
    if (leaves == kBucketLeaf)
        return BucketView{B->n, (i32)B->nrec1, B->brecs1.as<BucketRec>(), B->leaves.as<LeafRec<double>>()};
    if (leaves == kBucketLarge)
        return BucketView{B->n, (i32)B->nrec16, B->brecs16.as<BucketRec>(), B->leaves.as<LeafRec<double>>()};
    if (leaves == kBucket)
        return BucketView{B->n, (i32)B->nrec8, B->brecs8.as<BucketRec>(), B->leaves.as<LeafRec<double>>()};
    return BucketView{B->n, (i32)B->nrec, B->brecs.as<BucketRec>(), B->leaves.as<LeafRec<double>>()};
}

// A compacted FP32C chunk of the C2 shape (2 scalars, 1 gradient, 1 divergence,
// 1 curl) whose gradient field is the second scalar's column and whose curl
// reads the divergence's columns: the shared-sum accumulator applies.
inline bool c2_shared(const CatChunk& c) {
    // two scalars, the second one's gradient, a divergence and a curl of the
    // same three columns, in the compact layout {vi, s0, s1, d0, d1, d2}
    if (!(c.ns == 2 && c.ng == 1 && c.nd == 1 && c.nc == 1)) return false;
    if (c.spos[0] != 1 || c.spos[1] != 2 || c.gpos[0] != 2) return false;
    for (int k = 0; k < 3; ++k)
        if (c.dpos[k] != 3 + k || c.cpos[k] != 3 + k) return false;
    return true;
}

// Compact FP32-compensated record layout of a catalog chunk: positions 1..ncol
// are the distinct source columns the chunk reads (rsrc[1 + j]); the density
// gradient (reference grad slot 0) reads the rho column.
CatChunk compact_c32(const CatChunk& c, int& ncol) {
    CatChunk o = c;
    ncol = 0;
    auto map = [&](int src) {
        for (int j = 0; j < ncol; ++j)
            if (o.rsrc[1 + j] == src) return 1 + j;
        o.rsrc[1 + ncol] = src;
        return 1 + ncol++;
    };
    for (int j = 0; j < c.ns; ++j) o.spos[j] = map(c.rsrc[c.spos[j]]);
    for (int j = 0; j < c.ng; ++j) o.gpos[j] = map(c.gpos[j] < 0 ? 5 : c.rsrc[c.gpos[j]]);
    for (int j = 0; j < 3 * c.nd; ++j) o.dpos[j] = map(c.rsrc[c.dpos[j]]);
    for (int j = 0; j < 3 * c.nc; ++j) o.cpos[j] = map(c.rsrc[c.cpos[j]]);
    for (int j = c.ns; j < kMaxS; ++j) o.spos[j] = 0;
    o.nsrc = 1 + ncol;
    o.rw = (1 + ncol + 1) & ~1;
    return o;
}

}  // namespace
}  // namespace ph
