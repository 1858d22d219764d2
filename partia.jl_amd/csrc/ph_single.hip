// ph_single.hip -- the reference's single-quantity kernels on the GPU:
// _density_kernel, _number_density_kernel and _gradient_density_kernel per
// point (KI/single_point_interpolation/kernels/scalar_interpolation.jl:1-166,
// gradient_interpolation.jl:1-170), and per line
// _line_integrated_quantities_interpolate_kernel with the caller's h_a in every
// strategy and _line_integrated_density_kernel
// (KI/line_integrated_interpolation/kernels/line_integrated_scalar_interpolation.jl:1-214).
// The drivers never call them (the fused kernel serves the point drivers, the
// scatter projection the line drivers); they are the reference's building
// blocks for other callers and its tests.  One lane per sample walks the
// reference's escape-rope DFS over the LBVH (dfs) and accumulates its accepted
// leaves in DFS order with the reference's arithmetic in the tree's float type,
// so float64 results equal the reference's restatement bit for bit.
#include "ph_interp.h"
#include "ph_interp_dev.h"
#include "ph_primitives.h"
#include "ph_sph.h"

namespace ph {
namespace {

// The strategy's kernel value and gradient (accumulations/*.jl): gather h_a,
// scatter h_b, symmetric the mean of both.
template <class T, int K, int STRAT> __device__ inline T w_sel(T dr, T ha, T hb) {
    if constexpr (STRAT == PH_SYMMETRIC) return T(0.5) * (kernel_W<T, K>(dr, ha) + kernel_W<T, K>(dr, hb));
    else return kernel_W<T, K>(dr, STRAT == PH_GATHER ? ha : hb);
}
template <class T, int K, int STRAT> __device__ inline void g_sel(T dx, T dy, T dz, T ha, T hb, T g[3]) {
    if constexpr (STRAT == PH_SYMMETRIC) {
        T ax, ay, az, bx, by, bz;
        kernel_gradW<T, K>(dx, dy, dz, ha, ax, ay, az);
        kernel_gradW<T, K>(dx, dy, dz, hb, bx, by, bz);
        g[0] = T(0.5) * (ax + bx);
        g[1] = T(0.5) * (ay + by);
        g[2] = T(0.5) * (az + bz);
    } else {
        kernel_gradW<T, K>(dx, dy, dz, STRAT == PH_GATHER ? ha : hb, g[0], g[1], g[2]);
    }
}

// KIND 0: sum m_b W (0 without neighbours); 1: sum W; 2: the density gradient
// f - rho b with f = sum m_b grad W, b = sum m_b grad W inv(rho_b) and the
// unnormalised rho = sum m_b W (gradient_interpolation.jl:47-55), NaN x 3 when
// rho == 0.  grad W is taken of ra - rb (kernel.jl:470-477); the distance of
// kernel.jl:183-190 sums the squares from zero, which is d2_point's value.
// ha_s: the caller's h_a per point, or null for the nearest particle's h.
template <class T, int K, int STRAT, int KIND>
__global__ void __launch_bounds__(kBlock) k_point_single(TreeView<T> tv, const T* __restrict__ cols, PointList pl,
                                                         const double* __restrict__ ha_s, i64 S, T* __restrict__ out) {
    const i64 s = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (s >= S) return;
    const T px = T(pl.x[s]), py = T(pl.y[s]), pz = T(pl.z[s]);
    const i64 n = tv.n;
    const T Kvalid = T(support_of<K>());
    T ha = T(0);
    if constexpr (STRAT != PH_SCATTER) ha = ha_s ? T(ha_s[s]) : nearest_h<T>(tv, px, py, pz);
    const T ra = Kvalid * ha, r2a = ra * ra;
    T rho = T(0);
    T f[3] = {T(0), T(0), T(0)}, b[3] = {T(0), T(0), T(0)};
    dfs<T, STRAT>(
        tv, Kvalid, r2a, [&](const LeafRec<T>& L) { return d2_point(px, py, pz, L); },
        [&](const NodeRec<T>& N) { return d2_aabb(px, py, pz, N); },
        [&](i64 k, T d2, const LeafRec<T>& L) {
            const T dr = sqrt(d2);
            const T mb = cols[3 * n + k];
            const T W = w_sel<T, K, STRAT>(dr, ha, L.h);
            if constexpr (KIND == 0) {
                rho += mb * W;
            } else if constexpr (KIND == 1) {
                rho += W;
            } else {
                T g[3];
                g_sel<T, K, STRAT>(px - L.x, py - L.y, pz - L.z, ha, L.h, g);
                const T invrb = T(1) / cols[5 * n + k];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const T mg = mb * g[d];
                    f[d] += mg;
                    b[d] += mg * invrb;
                }
                rho += mb * W;
            }
        });
    if constexpr (KIND != 2) {
        out[s] = rho;
    } else {
        if (rho == T(0)) {
#pragma unroll
            for (int d = 0; d < 3; ++d) out[(i64)d * S + s] = sph_nan<T>();
            return;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            b[d] *= rho;
            out[(i64)d * S + s] = f[d] - b[d];
        }
    }
}

// Per line with the caller's h_a (gather / symmetric; ignored by scatter):
// DENSITY 0, the catalog scalars of _line_integrated_quantities_interpolate_kernel
// (S1 = sum m_b W / rho_b, sums A_b m_b W / rho_b, NaN when S1 == 0, Shepard
// flags); DENSITY 1, _line_integrated_density_kernel: sum m_b W (0 without
// neighbours).  W = I(r / h) / h from the table; r = sqrt(d2) with a negative
// d2 clamped to 0.
template <class T, int K, int STRAT, int DENSITY>
__global__ void __launch_bounds__(kBlock) k_line_single(TreeView<T> tv, const T* __restrict__ cols, CatChunk c,
                                                        LineList ll, const double* __restrict__ ha_s, i64 S,
                                                        const T* __restrict__ inorm, T dq, T* __restrict__ out) {
    const i64 s = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (s >= S) return;
    Line<T> l;
    l.o[0] = T(ll.ox[s]);
    l.o[1] = T(ll.oy[s]);
    l.o[2] = T(ll.oz[s]);
    l.m[0] = T(ll.dx[s]);
    l.m[1] = T(ll.dy[s]);
    l.m[2] = T(ll.dz[s]);
    const i64 n = tv.n;
    const T Kvalid = T(support_of<K>());
    const T ha = STRAT != PH_SCATTER ? T(ha_s[s]) : T(0);
    const T ra = Kvalid * ha, r2a = ra * ra;
    T S1 = T(0);
    T acc[kMaxS];
#pragma unroll
    for (int j = 0; j < kMaxS; ++j) acc[j] = T(0);
    dfs<T, STRAT>(
        tv, Kvalid, r2a, [&](const LeafRec<T>& L) { return d2_line_leaf(l, L); },
        [&](const NodeRec<T>& N) { return d2_line_aabb_lb(l, N); },
        [&](i64 k, T d2, const LeafRec<T>& L) {
            const T dr = sqrt(d2 < T(0) ? T(0) : d2);
            T W;
            if constexpr (STRAT == PH_SYMMETRIC)
                W = T(0.5) * (line_W<T, K>(dr, ha, inorm, dq) + line_W<T, K>(dr, L.h, inorm, dq));
            else
                W = line_W<T, K>(dr, STRAT == PH_GATHER ? ha : L.h, inorm, dq);
            const T mb = cols[3 * n + k];
            if constexpr (DENSITY) {
                S1 += mb * W;
            } else {
                const T rb = cols[5 * n + k];
                S1 += mb * W / rb;
#pragma unroll
                for (int j = 0; j < kMaxS; ++j)
                    if (j < c.ns) acc[j] += cols[c.rsrc[c.spos[j]] * n + k] * (mb * W / rb);
            }
        });
    if constexpr (DENSITY) {
        out[s] = S1;
    } else if (S1 == T(0)) {
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

template <class T, int K, int STRAT>
void point_single_ks(ph_lbvh* B, int kind, const PointList& pl, const double* ha, i64 S, T* out, hipStream_t s) {
    const TreeView<T> tv = view_of<T>(B);
    const T* cols = B->cols.as<T>();
    const unsigned g = (unsigned)((S + kBlock - 1) / kBlock);
    if (kind == PH_SINGLE_DENSITY)
        k_point_single<T, K, STRAT, 0><<<g, kBlock, 0, s>>>(tv, cols, pl, ha, S, out);
    else if (kind == PH_SINGLE_NUMBER_DENSITY)
        k_point_single<T, K, STRAT, 1><<<g, kBlock, 0, s>>>(tv, cols, pl, ha, S, out);
    else
        k_point_single<T, K, STRAT, 2><<<g, kBlock, 0, s>>>(tv, cols, pl, ha, S, out);
    PH_HIP(hipGetLastError());
}

template <class T, int K>
void point_single_k(ph_lbvh* B, int kind, int strategy, const PointList& pl, const double* ha, i64 S, T* out,
                    hipStream_t s) {
    switch (strategy) {
    case PH_GATHER: point_single_ks<T, K, PH_GATHER>(B, kind, pl, ha, S, out, s); break;
    case PH_SCATTER: point_single_ks<T, K, PH_SCATTER>(B, kind, pl, ha, S, out, s); break;
    default: point_single_ks<T, K, PH_SYMMETRIC>(B, kind, pl, ha, S, out, s); break;
    }
}

template <class T>
void point_single_t(ph_lbvh* B, int kind, int strategy, const PointList& pl, const double* ha, i64 S, T* out,
                    hipStream_t s) {
    switch (B->kernel) {
    case PH_KERNEL_M4: point_single_k<T, PH_KERNEL_M4>(B, kind, strategy, pl, ha, S, out, s); break;
    case PH_KERNEL_M5: point_single_k<T, PH_KERNEL_M5>(B, kind, strategy, pl, ha, S, out, s); break;
    case PH_KERNEL_M6: point_single_k<T, PH_KERNEL_M6>(B, kind, strategy, pl, ha, S, out, s); break;
    case PH_KERNEL_C2: point_single_k<T, PH_KERNEL_C2>(B, kind, strategy, pl, ha, S, out, s); break;
    case PH_KERNEL_C4: point_single_k<T, PH_KERNEL_C4>(B, kind, strategy, pl, ha, S, out, s); break;
    default: point_single_k<T, PH_KERNEL_C6>(B, kind, strategy, pl, ha, S, out, s); break;
    }
}

template <class T, int K, int STRAT>
void line_single_ks(ph_lbvh* B, const CatChunk* c, const LineList& ll, const double* ha, i64 S, T* out,
                    const T* inorm, hipStream_t s) {
    const TreeView<T> tv = view_of<T>(B);
    const T* cols = B->cols.as<T>();
    // Q_2 - Q_1 of the table in T (line_integrated_kernel.jl:14)
    const T dq = T((1.0 * support_of<K>()) / 4095.0) - T(0.0);
    const T* tab = inorm + (size_t)K * PH_LUT_N_ENTRIES;
    const unsigned g = (unsigned)((S + kBlock - 1) / kBlock);
    if (c) k_line_single<T, K, STRAT, 0><<<g, kBlock, 0, s>>>(tv, cols, *c, ll, ha, S, tab, dq, out);
    else k_line_single<T, K, STRAT, 1><<<g, kBlock, 0, s>>>(tv, cols, CatChunk{}, ll, ha, S, tab, dq, out);
    PH_HIP(hipGetLastError());
}

template <class T, int K>
void line_single_k(ph_lbvh* B, const CatChunk* c, int strategy, const LineList& ll, const double* ha, i64 S, T* out,
                   const T* inorm, hipStream_t s) {
    switch (strategy) {
    case PH_GATHER: line_single_ks<T, K, PH_GATHER>(B, c, ll, ha, S, out, inorm, s); break;
    case PH_SCATTER: line_single_ks<T, K, PH_SCATTER>(B, c, ll, ha, S, out, inorm, s); break;
    default: line_single_ks<T, K, PH_SYMMETRIC>(B, c, ll, ha, S, out, inorm, s); break;
    }
}

template <class T>
void line_single_t(ph_lbvh* B, const CatChunk* c, int strategy, const LineList& ll, const double* ha, i64 S, T* out,
                   const T* inorm, hipStream_t s) {
    switch (B->kernel) {
    case PH_KERNEL_M4: line_single_k<T, PH_KERNEL_M4>(B, c, strategy, ll, ha, S, out, inorm, s); break;
    case PH_KERNEL_M5: line_single_k<T, PH_KERNEL_M5>(B, c, strategy, ll, ha, S, out, inorm, s); break;
    case PH_KERNEL_M6: line_single_k<T, PH_KERNEL_M6>(B, c, strategy, ll, ha, S, out, inorm, s); break;
    case PH_KERNEL_C2: line_single_k<T, PH_KERNEL_C2>(B, c, strategy, ll, ha, S, out, inorm, s); break;
    case PH_KERNEL_C4: line_single_k<T, PH_KERNEL_C4>(B, c, strategy, ll, ha, S, out, inorm, s); break;
    default: line_single_k<T, PH_KERNEL_C6>(B, c, strategy, ll, ha, S, out, inorm, s); break;
    }
}

}  // namespace

void launch_point_single(ph_lbvh* B, int kind, int strategy, const PointList& pl, const double* ha, i64 S, void* out,
                         hipStream_t s) {
    if (S <= 0) return;
    if (B->precision == PH_PREC_FP64) point_single_t<double>(B, kind, strategy, pl, ha, S, static_cast<double*>(out), s);
    else point_single_t<float>(B, kind, strategy, pl, ha, S, static_cast<float*>(out), s);
}

void launch_line_single(ph_lbvh* B, const CatChunk* c, int strategy, const LineList& ll, const double* ha, i64 S,
                        void* out, const void* lut, hipStream_t s) {
    if (S <= 0) return;
    if (B->precision == PH_PREC_FP64)
        line_single_t<double>(B, c, strategy, ll, ha, S, static_cast<double*>(out), static_cast<const double*>(lut), s);
    else
        line_single_t<float>(B, c, strategy, ll, ha, S, static_cast<float*>(out), static_cast<const float*>(lut), s);
}

}  // namespace ph
