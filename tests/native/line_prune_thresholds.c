/* Randomised check of the line walk's node test with centre-distance thresholds
 * (line_node / line_node_prune in partia.jl_amd/csrc/ph_interp_dev.h) against
 * the per-ray test (d2_line_aabb_lb > (K scale)^2, the reference's
 * _line_intersects_aabb and bounding-sphere bound), in float and in double, on
 * lines placed at distances within a relative 2^-30 .. 2^-2 of the pruning
 * threshold, through or beside the box, and on degenerate nodes (point boxes,
 * zero and huge scales, infinities, NaN).  Same operations as the device code
 * (built with -ffp-contract=off; sqrt and / correctly rounded).  Prints the
 * number of cases and mismatches; exit status 1 on any mismatch. */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static uint64_t rs = 0x9E3779B97F4A7C15ull;
static uint64_t next64(void) {
    uint64_t z = (rs += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double unif(void) { return (double)(next64() >> 11) * 0x1p-53; }

#define DEFINE(T, SUF, EPS, TMIN)                                                                        \
    typedef struct {                                                                                     \
        T lo[3], hi[3], c[3], r2, sr, rad2, thr_lo, thr_hi;                                              \
    } Node##SUF;                                                                                         \
    static T d2_ref_##SUF(const T o[3], const T m[3], const T lo[3], const T hi[3]) {                    \
        T tmin = -(T)INFINITY, tmax = (T)INFINITY;                                                       \
        int hit = 1;                                                                                     \
        for (int d = 0; d < 3; ++d) {                                                                    \
            const T a = lo[d], b = hi[d];                                                                \
            if (m[d] == (T)0) {                                                                          \
                if (o[d] < a || o[d] > b) hit = 0;                                                       \
            } else {                                                                                     \
                T t0, t1;                                                                                \
                if (m[d] == (T)1) {                                                                      \
                    t0 = a - o[d];                                                                       \
                    t1 = b - o[d];                                                                       \
                } else if (m[d] == (T)-1) {                                                              \
                    t0 = -(a - o[d]);                                                                    \
                    t1 = -(b - o[d]);                                                                    \
                } else {                                                                                 \
                    t0 = (a - o[d]) / m[d];                                                              \
                    t1 = (b - o[d]) / m[d];                                                              \
                }                                                                                        \
                if (t0 > t1) {                                                                           \
                    const T tt = t0;                                                                     \
                    t0 = t1;                                                                             \
                    t1 = tt;                                                                             \
                }                                                                                        \
                tmin = fmax##SUF(tmin, t0);                                                              \
                tmax = fmin##SUF(tmax, t1);                                                              \
                if (tmin > tmax) hit = 0;                                                                \
            }                                                                                            \
        }                                                                                                \
        if (hit) return (T)0;                                                                            \
        const T half = (T)1 / (T)2;                                                                      \
        T c[3], r2 = (T)0;                                                                               \
        for (int d = 0; d < 3; ++d) c[d] = (lo[d] + hi[d]) * half;                                       \
        for (int d = 0; d < 3; ++d) {                                                                    \
            const T hh = (hi[d] - lo[d]) * half;                                                         \
            r2 += hh * hh;                                                                               \
        }                                                                                                \
        T D2 = (T)0, Dm = (T)0;                                                                          \
        for (int d = 0; d < 3; ++d) {                                                                    \
            const T D = c[d] - o[d];                                                                     \
            D2 += D * D;                                                                                 \
            Dm += D * m[d];                                                                              \
        }                                                                                                \
        const T dc2 = D2 - Dm * Dm;                                                                      \
        if (dc2 <= r2) return (T)0;                                                                      \
        const T dl = sqrt##SUF(dc2) - sqrt##SUF(r2);                                                     \
        return dl * dl;                                                                                  \
    }                                                                                                    \
    static Node##SUF node_##SUF(const T lo[3], const T hi[3], T scale, T K) {                            \
        Node##SUF R;                                                                                     \
        const T half = (T)1 / (T)2;                                                                      \
        T r2 = (T)0;                                                                                     \
        for (int d = 0; d < 3; ++d) {                                                                    \
            R.lo[d] = lo[d];                                                                             \
            R.hi[d] = hi[d];                                                                             \
            R.c[d] = (lo[d] + hi[d]) * half;                                                             \
            const T hh = (hi[d] - lo[d]) * half;                                                         \
            r2 += hh * hh;                                                                               \
        }                                                                                                \
        R.r2 = r2;                                                                                       \
        R.sr = sqrt##SUF(r2);                                                                            \
        const T r = K * scale;                                                                           \
        R.rad2 = r * r;                                                                                  \
        const double e = EPS;                                                                            \
        const double rr = sqrt((double)R.rad2);                                                          \
        const double a = (double)R.sr + rr * (1.0 - e), b = (double)R.sr + rr * (1.0 + e);               \
        const T tl = (T)(a * a * (1.0 - e)), th = (T)(b * b * (1.0 + e));                                \
        const int ok = isfinite(R.r2) && isfinite(R.sr) && R.rad2 >= (T)TMIN && isfinite(R.rad2) &&      \
                       isfinite(tl) && isfinite(th);                                                     \
        R.thr_lo = ok ? tl : (T)NAN;                                                                     \
        R.thr_hi = ok ? th : (T)NAN;                                                                     \
        return R;                                                                                        \
    }                                                                                                    \
    static int prune_new_##SUF(const T o[3], const T m[3], const Node##SUF* N) {                         \
        T D2 = (T)0, Dm = (T)0;                                                                          \
        for (int d = 0; d < 3; ++d) {                                                                    \
            const T D = N->c[d] - o[d];                                                                  \
            D2 += D * D;                                                                                 \
            Dm += D * m[d];                                                                              \
        }                                                                                                \
        const T dc2 = D2 - Dm * Dm;                                                                      \
        const int keep = dc2 <= N->r2 || dc2 <= N->thr_lo;                                               \
        if (keep) return 0;                                                                              \
        int hit = 1;                                                                                     \
        {                                                                                                \
            T tmin = -(T)INFINITY, tmax = (T)INFINITY;                                                   \
            for (int d = 0; d < 3; ++d) {                                                                \
                const T a = N->lo[d], b = N->hi[d];                                                      \
                if (m[d] == (T)0) {                                                                      \
                    if (o[d] < a || o[d] > b) hit = 0;                                                   \
                } else {                                                                                 \
                    T t0, t1;                                                                            \
                    if (m[d] == (T)1) {                                                                  \
                        t0 = a - o[d];                                                                   \
                        t1 = b - o[d];                                                                   \
                    } else if (m[d] == (T)-1) {                                                          \
                        t0 = -(a - o[d]);                                                                \
                        t1 = -(b - o[d]);                                                                \
                    } else {                                                                             \
                        t0 = (a - o[d]) / m[d];                                                          \
                        t1 = (b - o[d]) / m[d];                                                          \
                    }                                                                                    \
                    if (t0 > t1) {                                                                       \
                        const T tt = t0;                                                                 \
                        t0 = t1;                                                                         \
                        t1 = tt;                                                                         \
                    }                                                                                    \
                    tmin = fmax##SUF(tmin, t0);                                                          \
                    tmax = fmin##SUF(tmax, t1);                                                          \
                    if (tmin > tmax) hit = 0;                                                            \
                }                                                                                        \
            }                                                                                            \
        }                                                                                                \
        int far = dc2 >= N->thr_hi;                                                                      \
        if (!far) {                                                                                      \
            const T dl = sqrt##SUF(dc2) - N->sr;                                                         \
            far = dl * dl > N->rad2;                                                                     \
        }                                                                                                \
        return !hit && far;                                                                              \
    }                                                                                                    \
    static uint64_t run_##SUF(long cases, uint64_t* fast) {                                              \
        uint64_t bad = 0;                                                                                \
        for (long it = 0; it < cases; ++it) {                                                            \
            T lo[3], hi[3], o[3], m[3];                                                                  \
            const int kind = (int)(next64() % 16);                                                       \
            const double ext = kind == 0 ? 0.0 : pow(10.0, -6.0 * unif());                               \
            for (int d = 0; d < 3; ++d) {                                                                \
                const double c = unif();                                                                 \
                const double w = ext * unif();                                                           \
                lo[d] = (T)(c - w);                                                                      \
                hi[d] = (T)(c + w);                                                                      \
            }                                                                                            \
            T K = (T)(2.0 + 0.5 * (double)(next64() % 3));                                               \
            T scale = (T)(pow(10.0, -4.0 * unif()) * 0.05);                                              \
            if (kind == 1) scale = (T)0;                                                                 \
            if (kind == 2) scale = (T)INFINITY;                                                          \
            if (kind == 3) scale = (T)NAN;                                                               \
            if (kind == 4) scale = (T)1e-30;                                                             \
            if (kind == 5) lo[0] = -(T)INFINITY;                                                         \
            const int axis = (int)(next64() % 4);                                                        \
            if (axis < 3) {                                                                              \
                for (int d = 0; d < 3; ++d) m[d] = (T)(d == axis ? ((next64() & 1) ? 1.0 : -1.0) : 0.0); \
            } else {                                                                                     \
                double v[3], n2 = 0;                                                                     \
                for (int d = 0; d < 3; ++d) {                                                            \
                    v[d] = unif() - 0.5;                                                                 \
                    n2 += v[d] * v[d];                                                                   \
                }                                                                                        \
                for (int d = 0; d < 3; ++d) m[d] = (T)(v[d] / sqrt(n2));                                 \
            }                                                                                            \
            const Node##SUF N = node_##SUF(lo, hi, scale, K);                                            \
            /* a unit vector orthogonal to m, then a distance near the threshold */                     \
            double q[3] = {1.0, 0.0, 0.0};                                                               \
            if (fabs((double)m[0]) > 0.5) {                                                              \
                q[0] = 0.0;                                                                              \
                q[1] = 1.0;                                                                              \
            }                                                                                            \
            double qm = q[0] * m[0] + q[1] * m[1] + q[2] * m[2], qn = 0;                                 \
            for (int d = 0; d < 3; ++d) {                                                                \
                q[d] -= qm * m[d];                                                                       \
                qn += q[d] * q[d];                                                                       \
            }                                                                                            \
            const double thr = sqrt((double)N.r2) + (double)K * (double)scale;                           \
            double dist = thr * (1.0 + (unif() * 2.0 - 1.0) * pow(2.0, -2.0 - 28.0 * unif()));           \
            if (kind == 6) dist = thr * unif();                                                          \
            if (kind == 7) dist = 0.0;                                                                   \
            const double along = (unif() - 0.5) * 4.0;                                                   \
            for (int d = 0; d < 3; ++d)                                                                  \
                o[d] = (T)((double)N.c[d] + dist * q[d] / sqrt(qn) + along * (double)m[d]);              \
            const T rad = K * scale;                                                                     \
            const int ref = d2_ref_##SUF(o, m, lo, hi) > rad * rad;                                      \
            const int got = prune_new_##SUF(o, m, &N);                                                   \
            if (ref != got) {                                                                            \
                if (bad < 5) printf("mismatch (" #T ") kind %d ref %d got %d\n", kind, ref, got);        \
                ++bad;                                                                                   \
            }                                                                                            \
            {                                                                                            \
                T D2 = (T)0, Dm = (T)0;                                                                  \
                for (int d = 0; d < 3; ++d) {                                                            \
                    const T D = N.c[d] - o[d];                                                           \
                    D2 += D * D;                                                                         \
                    Dm += D * m[d];                                                                      \
                }                                                                                        \
                const T dc2 = D2 - Dm * Dm;                                                              \
                if (dc2 <= N.thr_lo || dc2 >= N.thr_hi) ++*fast;                                         \
            }                                                                                            \
        }                                                                                                \
        return bad;                                                                                      \
    }

DEFINE(float, f, 0x1p-18, FLT_MIN)
DEFINE(double, , 0x1p-40, DBL_MIN)

int main(int argc, char** argv) {
    const long cases = argc > 1 ? atol(argv[1]) : 4000000;
    uint64_t ff = 0, fd = 0;
    const uint64_t bf = run_f(cases, &ff);
    const uint64_t bd = run_(cases, &fd);
    printf("float: %ld cases (%llu decided by the thresholds), %llu mismatches\n", cases, (unsigned long long)ff,
           (unsigned long long)bf);
    printf("double: %ld cases (%llu decided by the thresholds), %llu mismatches\n", cases, (unsigned long long)fd,
           (unsigned long long)bd);
    return (bf || bd) ? 1 : 0;
}
