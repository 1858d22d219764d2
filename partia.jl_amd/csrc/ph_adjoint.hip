// ph_adjoint.hip -- reverse pass of one scalar output of point / grid
// interpolation (no reference counterpart: what a ChainRulesCore rrule of
// PointSamples_interpolation / StructuredGrid_interpolation would call).
//
// For a catalog of one scalar A (Shepard flag s) the forward of
// general_interpolation.jl:1-669 computes, per sample a,
//   N = sum_b A_b v_b W_ab, S1 = sum_b v_b W_ab, out = s ? N / S1 : N,
// v_b = m_b / rho_b and W_ab = W(r, h_a) (gather), W(r, h_b) (scatter) or their
// mean (symmetric), h_a the smoothing length of a's nearest particle.  Given
// the cotangent g_a of every output, the kernel adds to each particle
//   d/dm_b  g c_b W_ab / rho_b            d/dA_b  g v_b W_ab (/ S1)
//   d/dh_b  g c_b v_b dW_ab/dh_b          d/dx_b  g c_b v_b dW/dr (-(x_a - x_b) / r)
// with c_b = A_b, or (A_b - out) / S1 under Shepard, and to a's nearest
// particle's h the term g d out / d h_a.  W = s3 h^-3 f(q), q = r / h:
// dW/dr = s3 h^-4 f'(q), dW/dh = -s3 h^-4 (3 f(q) + q f'(q)).  rho_b is held
// fixed.  The oracle's point_adjoint states the same sums.
//
// One wave walks the bucket tree once for its 64 samples (as k_tile_c32), with
// the forward's float64 leaf test, so every lane sees exactly its forward
// neighbours.  Pass 1 recomputes N, S1 and the h_a sums; pass 2 forms each
// staged leaf's six partials on every lane, sums them over the wave in float64
// and adds them with one float64 atomic per wave, particle and parameter
// (so the float64 totals are order-dependent in the last bits).
//
// Gradient, divergence and curl outputs (the epilogue of
// general_interpolation.jl:128-214 with accumulations/{gradient,divergence,
// curl}_accumulation.jl) are all of the form sum_b v_b P(A_b - At) . G_ab, with
// G_ab = grad_a W_ab = (x_a - x_b) F(r), F = s3 h^-4 f'(q) / r, and the Shepard
// means At = sum A v W / S1.  With b = sum v G and the output cotangent g:
//   gradient   P(G) = g . G,  q = -(g . b) / S1,   omega = (A - At) g
//   divergence P(G) = g G,    q = -g b / S1,       omega = g (A - At)
//   curl       P(G) = g x G,  q = (b x g) / S1,    omega = (A - At) x g
// every particle's change enters through u = P(G) + q W:
//   d/dA_b = v u,  d/dm_b = (A - At) . u / rho_b,
//   d/dh_b = v (omega . D dF/dh + kappa dW/dh),  kappa = (A - At) . q,
//   d/dx_b = v (-omega F - (omega . D) F_r D / r - kappa W_r D / r),
// D = x_a - x_b, F_r = s3 h^-4 (q f'' - f') / r^2, dF/dh = -s3 h^-5 (4 f' + q f'') / r,
// and the h_a terms to the nearest particle.  k_field_adjoint forms the 5 + c
// partials of a staged leaf (c = 1 field column for the gradient, 3 for
// divergence and curl) and scatters them with wave_sum8.  The oracle's
// field_adjoint states the same sums.
#include <algorithm>
#include <type_traits>

#include "ph_bucket_dev.h"

namespace ph {
namespace {

// Nearest particle of every sample: smoothing length and Morton index (-1 and
// NaN when no particle lies within K max(h)).
template <int SRC>
__global__ void __launch_bounds__(kBlock) k_nearest_index(BucketView bv, Sampler<double, SRC> smp, double Kd,
                                                          double* __restrict__ ha, i64* __restrict__ nn) {
    i64 s;
    double p[3];
    if (!smp.get(s, p[0], p[1], p[2])) return;
    const Nearest nr = nearest_bucket(bv, p, neighbour_bound2(bv, Kd));
    ha[s] = nr.h(bv);
    nn[s] = nr.bk != Nearest::kNone ? nr.bk : -1;
}

__global__ void k_zero_grad(double* __restrict__ p, i64 count) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (i64)gridDim.x * blockDim.x) p[i] = 0.0;
}

// The union walk of k_tile_c32: calls batch(nl) with lidx[0, nl) holding the
// Morton indices of up to 64 leaves of buckets some live lane keeps, in
// increasing order over the whole walk.  Wave-uniform throughout.
template <int STRAT, class Batch>
__device__ __attribute__((always_inline)) void union_walk(const BucketView& bv, const BucketRec* __restrict__ recs,
                                                          const Probe<STRAT>& pr, bool live, u32* lidx,
                                                          Batch&& batch) {
    const int lane = threadIdx.x & (kWave - 1);
    const u64 live_mask = __ballot(live);
    i32 i = live_mask ? 0 : bv.nrec;
    i32 steps = bv.nrec + 8;  // every record at most once (nrec < 2^27)
    for (;;) {
        int nl = 0;
        while (i < bv.nrec && steps > 0) {
            --steps;
            const BucketRec R = recs[i];  // wave-uniform: a scalar load
            const bool keep = (pr.keep_mask(R) & live_mask) != 0;
            if (R.link >= 0) {
                i = __builtin_amdgcn_readfirstlane(keep ? i + 1 : R.link);
                continue;
            }
            if (keep) {
                const int cnt = bucket_cnt(R);
                if (nl + cnt > kWave) {  // starts the next batch
                    ++steps;
                    break;
                }
                if (lane < cnt) lidx[nl + lane] = (u32)(bucket_lo(R.link) + lane);
                nl += cnt;
            }
            i = __builtin_amdgcn_readfirstlane(i + 1);
        }
        if (nl == 0) break;
        wave_sync();
        batch(nl);
        wave_sync();  // lidx is rewritten by the next batch
        if (!(i < bv.nrec && steps > 0)) break;
    }
}

// W, dW/dr and dW/dh at (r, h) from one evaluation of f and f'.
struct KernelD {
    double W, dr, dh;
};
template <int K> __device__ inline KernelD kernel_d(double r, double h) {
    const double ih = 1.0 / h, q = r * ih;
    const double f = shape_f<double, K>(q), df = shape_df<double, K>(q);
    const double s3 = sigma3<double, K>() * (ih * ih * ih), s4 = s3 * ih;
    return KernelD{s3 * f, s4 * df, -s4 * (3.0 * f + q * df)};
}

// W, dW/dr, dW/dh and F = (dW/dr) / r, dF/dr, dF/dh at (r, h); the F terms are
// zero at r = 0.
struct KernelD2 {
    double W, Wr, Wh, F, Fr, Fh;
};
template <int K> __device__ inline KernelD2 kernel_d2(double r, double h) {
    const double ih = 1.0 / h, q = r * ih;
    const double f = shape_f<double, K>(q), f1 = shape_df<double, K>(q), f2 = shape_d2f<double, K>(q);
    const double s3 = sigma3<double, K>() * (ih * ih * ih), s4 = s3 * ih, s5 = s4 * ih;
    const double ir = r > 0.0 ? 1.0 / r : 0.0;
    return KernelD2{s3 * f, s4 * f1, -s4 * (3.0 * f + q * f1), s4 * f1 * ir, s4 * (q * f2 - f1) * (ir * ir),
                    -s5 * (4.0 * f1 + q * f2) * ir};
}

// W and F at (r, h) (pass 1 of k_field_adjoint).
template <int K> __device__ inline void kernel_wf(double r, double h, double& W, double& F) {
    const double ih = 1.0 / h, q = r * ih;
    const double s3 = sigma3<double, K>() * (ih * ih * ih);
    W = s3 * shape_f<double, K>(q);
    F = r > 0.0 ? s3 * ih * shape_df<double, K>(q) / r : 0.0;
}

// Adjoint of a gradient (mode 1), divergence (2) or curl (3) output.  Rows of
// the per-leaf partials: m, A_0, A_1, A_2, h, x, y, z (the gradient's A_1 and
// A_2 are zero and not written); grad rows m, the c columns, h, x, y, z.
template <int K, int STRAT, int SRC>
__global__ void __launch_bounds__(kBlock) k_field_adjoint(BucketView bv, const BucketRec* __restrict__ recs,
                                                          const LeafRec<double>* __restrict__ leaves,
                                                          const double* __restrict__ cols, AdjointOut ao,
                                                          Sampler<double, SRC> smp, const double* __restrict__ ha_s,
                                                          const i64* __restrict__ nn_s, const void* __restrict__ cot,
                                                          int cot32, const u32* __restrict__ order,
                                                          double* __restrict__ grad) {
    __shared__ u32 lidx_all[kBlock];
    u32* lidx = lidx_all + (threadIdx.x / kWave) * kWave;
    i64 s = 0;
    double p[3] = {0.0, 0.0, 0.0};
    const bool valid = smp.get(s, p[0], p[1], p[2]);
    if (__ballot(valid) == 0) return;  // wave-uniform; no block barrier follows
    const double Kd = support_of<K>();
    const i64 n = bv.n, S = smp.S;
    const int mode = ao.mode, c = ao.fields(), w = ao.width();
    double ha = 0.0;
    i64 nn = -1;
    if constexpr (STRAT != PH_SCATTER) {
        if (valid) {
            ha = ha_s[s];
            nn = nn_s[s];
        }
    }
    const bool live = valid && (STRAT == PH_SCATTER || nn >= 0);
    if (!live) ha = 0.0;
    const double ra = Kd * ha, r2a = ra * ra;
    Probe<STRAT> pr;
    pr.init(p, Kd, ra);
    constexpr double wa = STRAT == PH_SYMMETRIC ? 0.5 : (STRAT == PH_GATHER ? 1.0 : 0.0);
    constexpr double wb = STRAT == PH_SYMMETRIC ? 0.5 : (STRAT == PH_SCATTER ? 1.0 : 0.0);
    const double* __restrict__ mcol = cols + 3 * n;
    const double* __restrict__ rcol = cols + 5 * n;
    const double* __restrict__ a0 = cols + (i64)ao.q[0] * n;
    const double* __restrict__ a1 = cols + (i64)ao.q[c == 3 ? 1 : 0] * n;
    const double* __restrict__ a2 = cols + (i64)ao.q[c == 3 ? 2 : 0] * n;
    auto field = [&](u32 k, double A[3]) {
        A[0] = a0[k];
        A[1] = c == 3 ? a1[k] : 0.0;
        A[2] = c == 3 ? a2[k] : 0.0;
    };

    // pass 1: S1, the field sums and b = sum v G
    double S1 = 0.0, SA[3] = {0.0, 0.0, 0.0}, bs[3] = {0.0, 0.0, 0.0};
    union_walk<STRAT>(bv, recs, pr, live, lidx, [&](int nl) {
        for (int j = 0; j < nl; ++j) {
            const u32 k = __builtin_amdgcn_readfirstlane(lidx[j]);
            const LeafRec<double> L = leaves[k];
            const double d2 = d2_point(p[0], p[1], p[2], L);
            const bool take = live && d2 <= radius2<double, STRAT>(Kd, r2a, L.h);
            const double r = sqrt(d2), v = mcol[k] / rcol[k];
            double W = 0.0, F = 0.0;
            if constexpr (wa != 0.0) {
                double Wa, Fa;
                kernel_wf<K>(r, take ? ha : 1.0, Wa, Fa);
                W += wa * Wa;
                F += wa * Fa;
            }
            if constexpr (wb != 0.0) {
                double Wb, Fb;
                kernel_wf<K>(r, L.h, Wb, Fb);
                W += wb * Wb;
                F += wb * Fb;
            }
            double A[3];
            field(k, A);
            const double vW = take ? v * W : 0.0, vF = take ? v * F : 0.0;
            S1 += vW;
            SA[0] += A[0] * vW;
            SA[1] += A[1] * vW;
            SA[2] += A[2] * vW;
            bs[0] += (p[0] - L.x) * vF;
            bs[1] += (p[1] - L.y) * vF;
            bs[2] += (p[2] - L.z) * vF;
        }
    });
    const bool has = live && S1 != 0.0;  // otherwise the output is NaN: no gradient
    double g[3] = {0.0, 0.0, 0.0};
    if (has) {
        for (int j = 0; j < 3; ++j) {
            if (j < w) g[j] = cot32 ? (double)static_cast<const float*>(cot)[j * S + s]
                                    : static_cast<const double*>(cot)[j * S + s];
        }
    }
    const bool nz = has && (g[0] != 0.0 || g[1] != 0.0 || g[2] != 0.0);
    const double iS1 = has ? 1.0 / S1 : 0.0;
    const double At[3] = {SA[0] * iS1, SA[1] * iS1, SA[2] * iS1};
    double qv[3] = {0.0, 0.0, 0.0};
    if (mode == 1) {
        qv[0] = -(g[0] * bs[0] + g[1] * bs[1] + g[2] * bs[2]) * iS1;
    } else if (mode == 2) {
        for (int d = 0; d < 3; ++d) qv[d] = -g[0] * bs[d] * iS1;
    } else {
        qv[0] = (bs[1] * g[2] - bs[2] * g[1]) * iS1;
        qv[1] = (bs[2] * g[0] - bs[0] * g[2]) * iS1;
        qv[2] = (bs[0] * g[1] - bs[1] * g[0]) * iS1;
    }

    // pass 2: per staged leaf, eight partials summed over the wave
    const int lane = threadIdx.x & (kWave - 1);
    const int vrow = lane >> 3;  // the partial this lane receives
    const bool emit = (lane & 7) == 0 && (c == 3 || vrow < 2 || vrow >= 4);
    const int orow = c == 3 || vrow < 2 ? vrow : vrow - 2;
    double dha = 0.0;
    union_walk<STRAT>(bv, recs, pr, live, lidx, [&](int nl) {
        for (int j = 0; j < nl; ++j) {
            const u32 k = __builtin_amdgcn_readfirstlane(lidx[j]);
            const LeafRec<double> L = leaves[k];
            const double d2 = d2_point(p[0], p[1], p[2], L);
            const bool take = nz && d2 <= radius2<double, STRAT>(Kd, r2a, L.h);
            if (__ballot(take) == 0) continue;  // wave-uniform
            const double r = sqrt(d2), rb = rcol[k], v = mcol[k] / rb;
            KernelD2 ka{}, kb{};
            if constexpr (wa != 0.0) ka = kernel_d2<K>(r, take ? ha : 1.0);
            if constexpr (wb != 0.0) kb = kernel_d2<K>(r, L.h);
            const double W = wa * ka.W + wb * kb.W, F = wa * ka.F + wb * kb.F;
            const double Wr = wa * ka.Wr + wb * kb.Wr, Fr = wa * ka.Fr + wb * kb.Fr;
            const double D[3] = {p[0] - L.x, p[1] - L.y, p[2] - L.z};
            const double G[3] = {D[0] * F, D[1] * F, D[2] * F};
            double a[3];
            field(k, a);
            a[0] -= At[0];
            a[1] -= At[1];
            a[2] -= At[2];
            double u[3] = {0.0, 0.0, 0.0}, om[3];
            if (mode == 1) {
                u[0] = g[0] * G[0] + g[1] * G[1] + g[2] * G[2] + qv[0] * W;
                for (int d = 0; d < 3; ++d) om[d] = a[0] * g[d];
            } else if (mode == 2) {
                for (int d = 0; d < 3; ++d) u[d] = g[0] * G[d] + qv[d] * W;
                for (int d = 0; d < 3; ++d) om[d] = g[0] * a[d];
            } else {
                u[0] = g[1] * G[2] - g[2] * G[1] + qv[0] * W;
                u[1] = g[2] * G[0] - g[0] * G[2] + qv[1] * W;
                u[2] = g[0] * G[1] - g[1] * G[0] + qv[2] * W;
                om[0] = a[1] * g[2] - a[2] * g[1];
                om[1] = a[2] * g[0] - a[0] * g[2];
                om[2] = a[0] * g[1] - a[1] * g[0];
            }
            const double au = a[0] * u[0] + a[1] * u[1] + a[2] * u[2];
            const double kap = a[0] * qv[0] + a[1] * qv[1] + a[2] * qv[2];
            const double omD = om[0] * D[0] + om[1] * D[1] + om[2] * D[2];
            const double ir = r > 0.0 ? 1.0 / r : 0.0;
            const double radial = omD * Fr * ir + kap * Wr * ir;
            // lanes not taking the leaf contribute exact zeros (selected, not
            // multiplied: their terms may be inf or NaN)
            const double t[8] = {take ? au / rb : 0.0,
                                 take ? v * u[0] : 0.0,
                                 take ? v * u[1] : 0.0,
                                 take ? v * u[2] : 0.0,
                                 take && wb != 0.0 ? v * wb * (omD * kb.Fh + kap * kb.Wh) : 0.0,
                                 take ? -v * (om[0] * F + radial * D[0]) : 0.0,
                                 take ? -v * (om[1] * F + radial * D[1]) : 0.0,
                                 take ? -v * (om[2] * F + radial * D[2]) : 0.0};
            if constexpr (wa != 0.0) dha += take ? v * wa * (omD * ka.Fh + kap * ka.Wh) : 0.0;
            const double sum = wave_sum8(t);
            if (emit && sum != 0.0) unsafeAtomicAdd(&grad[(i64)orow * n + order[k]], sum);
        }
    });
    if constexpr (wa != 0.0) {
        if (nz && dha != 0.0) unsafeAtomicAdd(&grad[(i64)(1 + c) * n + order[nn]], dha);
    }
}

template <int K, int STRAT, int SRC, bool SHEP, class CT>
__global__ void __launch_bounds__(kBlock) k_points_adjoint(BucketView bv, const BucketRec* __restrict__ recs,
                                                           const LeafRec<double>* __restrict__ leaves,
                                                           const double* __restrict__ cols, i32 qsrc,
                                                           Sampler<double, SRC> smp, const double* __restrict__ ha_s,
                                                           const i64* __restrict__ nn_s, const CT* __restrict__ cot,
                                                           const u32* __restrict__ order, double* __restrict__ grad) {
    __shared__ u32 lidx_all[kBlock];
    u32* lidx = lidx_all + (threadIdx.x / kWave) * kWave;
    i64 s = 0;
    double p[3] = {0.0, 0.0, 0.0};
    const bool valid = smp.get(s, p[0], p[1], p[2]);
    if (__ballot(valid) == 0) return;  // wave-uniform; no block barrier follows
    const double Kd = support_of<K>();
    const i64 n = bv.n;
    double ha = 0.0;
    i64 nn = -1;
    if constexpr (STRAT != PH_SCATTER) {
        if (valid) {
            ha = ha_s[s];
            nn = nn_s[s];
        }
    }
    const bool live = valid && (STRAT == PH_SCATTER || nn >= 0);
    if (!live) ha = 0.0;
    const double ra = Kd * ha, r2a = ra * ra;  // the forward's radius
    Probe<STRAT> pr;
    pr.init(p, Kd, ra);
    constexpr double wa = STRAT == PH_SYMMETRIC ? 0.5 : (STRAT == PH_GATHER ? 1.0 : 0.0);  // h_a kernel weight
    constexpr double wb = STRAT == PH_SYMMETRIC ? 0.5 : (STRAT == PH_SCATTER ? 1.0 : 0.0);  // h_b kernel weight
    const double* __restrict__ mcol = cols + 3 * n;
    const double* __restrict__ rcol = cols + 5 * n;
    const double* __restrict__ acol = cols + (i64)qsrc * n;

    // pass 1: the forward's sums and the h_a sums
    double S1 = 0.0, N = 0.0, HN = 0.0, HS = 0.0;
    union_walk<STRAT>(bv, recs, pr, live, lidx, [&](int nl) {
        for (int j = 0; j < nl; ++j) {
            const u32 k = __builtin_amdgcn_readfirstlane(lidx[j]);
            const LeafRec<double> L = leaves[k];
            const double d2 = d2_point(p[0], p[1], p[2], L);
            const bool take = live && d2 <= radius2<double, STRAT>(Kd, r2a, L.h);
            const double r = sqrt(d2), v = mcol[k] / rcol[k], A = acol[k];
            double W = 0.0;
            if constexpr (wa != 0.0) {
                const KernelD a = kernel_d<K>(r, take ? ha : 1.0);
                W += wa * a.W;
                HN += take ? A * v * (wa * a.dh) : 0.0;
                HS += take ? v * (wa * a.dh) : 0.0;
            }
            if constexpr (wb != 0.0) W += wb * kernel_W<double, K>(r, L.h);
            S1 += take ? v * W : 0.0;
            N += take ? A * v * W : 0.0;
        }
    });
    const bool has = live && S1 != 0.0;  // otherwise the output is NaN: no gradient
    const double g = has ? (double)cot[s] : 0.0;
    const double out = has ? (SHEP ? N / S1 : N) : 0.0;
    const double iS1 = has && SHEP ? 1.0 / S1 : 1.0;

    // pass 2: per staged leaf, six partials summed over the wave
    union_walk<STRAT>(bv, recs, pr, live, lidx, [&](int nl) {
        for (int j = 0; j < nl; ++j) {
            const u32 k = __builtin_amdgcn_readfirstlane(lidx[j]);
            const LeafRec<double> L = leaves[k];
            const double d2 = d2_point(p[0], p[1], p[2], L);
            const bool take = has && g != 0.0 && d2 <= radius2<double, STRAT>(Kd, r2a, L.h);
            if (__ballot(take) == 0) continue;  // wave-uniform
            const double r = sqrt(d2), mb = mcol[k], rb = rcol[k], v = mb / rb, A = acol[k];
            double W = 0.0, dr = 0.0, dhb = 0.0;
            if constexpr (wa != 0.0) {
                const KernelD a = kernel_d<K>(r, take ? ha : 1.0);
                W += wa * a.W;
                dr += wa * a.dr;
            }
            if constexpr (wb != 0.0) {
                const KernelD b = kernel_d<K>(r, L.h);
                W += wb * b.W;
                dr += wb * b.dr;
                dhb = wb * b.dh;
            }
            const double c = SHEP ? (A - out) * iS1 : A;
            const double gc = g * c;
            const double fx = r > 0.0 ? -gc * v * dr / r : 0.0;
            // lanes not taking the leaf contribute exact zeros (selected, not
            // multiplied: their terms may be inf or NaN)
            const double t[6] = {take ? gc * W / rb : 0.0,         take ? g * v * W * iS1 : 0.0,
                                 take ? gc * v * dhb : 0.0,        take ? fx * (p[0] - L.x) : 0.0,
                                 take ? fx * (p[1] - L.y) : 0.0,   take ? fx * (p[2] - L.z) : 0.0};
            const double sum = wave_sum6(t);  // parameter lane >> 3
            const int lane = threadIdx.x & (kWave - 1);
            if ((lane & 7) == 0 && lane < 48 && sum != 0.0)
                unsafeAtomicAdd(&grad[(i64)(lane >> 3) * n + order[k]], sum);
        }
    });
    // d out / d h_a to the nearest particle's h
    if constexpr (wa != 0.0) {
        if (has && g != 0.0) {
            const double dha = g * (SHEP ? (HN - out * HS) * iS1 : HN);
            if (dha != 0.0) unsafeAtomicAdd(&grad[2 * n + order[nn]], dha);
        }
    }
}

// Nearest particle (smoothing length and Morton index) of every sample, for
// gather and symmetric adjoints.
template <int K>
void adjoint_nearest(ph_lbvh* B, const BucketView& bv, const PointList* pl, const GridAxes* ga, i64 S, unsigned g,
                     int xm, double*& ha, i64*& nn, hipStream_t s) {
    const double Kd = support_of<K>();
    B->ha.alloc(sizeof(double) * (size_t)S);
    B->nn.alloc(sizeof(i64) * (size_t)S);
    ha = B->ha.as<double>();
    nn = B->nn.as<i64>();
    if (ga) launch_nearest_tile(B, *ga, S, Kd, ha, s, nn);
    else k_nearest_index<0><<<g, kBlock, 0, s>>>(nearest_view(B), Sampler<double, 0>{*pl, S, xm}, Kd, ha, nn);
    PH_HIP(hipGetLastError());
}

unsigned adjoint_blocks(const GridAxes* ga, i64 S) {
    if (ga) {
        const i64 tiles = ((ga->nx + 3) / 4) * ((ga->ny + 3) / 4) * ((ga->nz + 3) / 4);
        return (unsigned)((tiles + 3) / 4);
    }
    return (unsigned)((S + kBlock - 1) / kBlock);
}

template <int K, int STRAT, bool SHEP, class CT>
void launch_adjoint_t(ph_lbvh* B, const AdjointOut& ao, const PointList* pl, const GridAxes* ga, i64 S,
                      const CT* cot, double* grad, hipStream_t s, hipEvent_t* ev) {
    const BucketView bv = bucket_view(B);
    const LaunchCfg cfg = tl_launch;
    const int xm = point_xcd(cfg);
    const unsigned g = adjoint_blocks(ga, S);
    double* ha = nullptr;
    i64* nn = nullptr;
    if (STRAT != PH_SCATTER) adjoint_nearest<K>(B, bv, pl, ga, S, g, xm, ha, nn, s);
    const double* cols = B->cols.as<double>();
    const u32* order = B->order.as<u32>();
    if (ev) PH_HIP(hipEventRecord(ev[0], s));
    if (ao.mode != 0) {
        const int c32 = std::is_same<CT, float>::value ? 1 : 0;
        if (ga)
            k_field_adjoint<K, STRAT, 1><<<g, kBlock, 0, s>>>(bv, bv.recs, bv.leaves, cols, ao,
                                                              Sampler<double, 1>{*ga, S, xm, point_order(cfg)}, ha, nn,
                                                              cot, c32, order, grad);
        else
            k_field_adjoint<K, STRAT, 0><<<g, kBlock, 0, s>>>(bv, bv.recs, bv.leaves, cols, ao,
                                                              Sampler<double, 0>{*pl, S, xm}, ha, nn, cot, c32, order,
                                                              grad);
    } else if (ga) {
        k_points_adjoint<K, STRAT, 1, SHEP, CT><<<g, kBlock, 0, s>>>(bv, bv.recs, bv.leaves, cols, ao.q[0],
                                                                     Sampler<double, 1>{*ga, S, xm, point_order(cfg)}, ha,
                                                                     nn, cot, order, grad);
    } else {
        k_points_adjoint<K, STRAT, 0, SHEP, CT><<<g, kBlock, 0, s>>>(bv, bv.recs, bv.leaves, cols, ao.q[0],
                                                                     Sampler<double, 0>{*pl, S, xm}, ha, nn, cot,
                                                                     order, grad);
    }
    PH_HIP(hipGetLastError());
    if (ev) PH_HIP(hipEventRecord(ev[1], s));
}

template <int K, bool SHEP, class CT>
void launch_adjoint_s(ph_lbvh* B, int strategy, const AdjointOut& ao, const PointList* pl, const GridAxes* ga, i64 S,
                      const CT* cot, double* grad, hipStream_t s, hipEvent_t* ev) {
    switch (strategy) {
    case PH_GATHER: launch_adjoint_t<K, PH_GATHER, SHEP, CT>(B, ao, pl, ga, S, cot, grad, s, ev); break;
    case PH_SCATTER: launch_adjoint_t<K, PH_SCATTER, SHEP, CT>(B, ao, pl, ga, S, cot, grad, s, ev); break;
    default: launch_adjoint_t<K, PH_SYMMETRIC, SHEP, CT>(B, ao, pl, ga, S, cot, grad, s, ev); break;
    }
}

template <bool SHEP, class CT>
void launch_adjoint_k(ph_lbvh* B, int strategy, const AdjointOut& ao, const PointList* pl, const GridAxes* ga, i64 S,
                      const CT* cot, double* grad, hipStream_t s, hipEvent_t* ev) {
    switch (B->kernel) {
    case PH_KERNEL_M4: launch_adjoint_s<PH_KERNEL_M4, SHEP, CT>(B, strategy, ao, pl, ga, S, cot, grad, s, ev); break;
    case PH_KERNEL_M5: launch_adjoint_s<PH_KERNEL_M5, SHEP, CT>(B, strategy, ao, pl, ga, S, cot, grad, s, ev); break;
    case PH_KERNEL_M6: launch_adjoint_s<PH_KERNEL_M6, SHEP, CT>(B, strategy, ao, pl, ga, S, cot, grad, s, ev); break;
    case PH_KERNEL_C2: launch_adjoint_s<PH_KERNEL_C2, SHEP, CT>(B, strategy, ao, pl, ga, S, cot, grad, s, ev); break;
    case PH_KERNEL_C4: launch_adjoint_s<PH_KERNEL_C4, SHEP, CT>(B, strategy, ao, pl, ga, S, cot, grad, s, ev); break;
    default: launch_adjoint_s<PH_KERNEL_C6, SHEP, CT>(B, strategy, ao, pl, ga, S, cot, grad, s, ev); break;
    }
}

}  // namespace

void launch_points_adjoint(ph_lbvh* B, int strategy, const AdjointOut& ao, int precision, const PointList* pl,
                           const GridAxes* ga, i64 S, const void* cot, double* grad, hipStream_t s, hipEvent_t* ev) {
    const i64 count = (i64)(5 + ao.fields()) * B->n;
    k_zero_grad<<<(unsigned)std::min<i64>((count + kBlock - 1) / kBlock, 4096), kBlock, 0, s>>>(grad, count);
    PH_HIP(hipGetLastError());
    if (S <= 0) return;
    ensure_buckets(B, s);
    // the field kernels take the Shepard flag as unused (always true) and the
    // cotangent type at run time
    const bool shepard = ao.mode == 0 ? ao.shepard : true;
    if (precision == PH_PREC_FP64) {
        auto c = static_cast<const double*>(cot);
        if (shepard) launch_adjoint_k<true, double>(B, strategy, ao, pl, ga, S, c, grad, s, ev);
        else launch_adjoint_k<false, double>(B, strategy, ao, pl, ga, S, c, grad, s, ev);
    } else {
        auto c = static_cast<const float*>(cot);
        if (shepard) launch_adjoint_k<true, float>(B, strategy, ao, pl, ga, S, c, grad, s, ev);
        else launch_adjoint_k<false, float>(B, strategy, ao, pl, ga, S, c, grad, s, ev);
    }
}

}  // namespace ph
