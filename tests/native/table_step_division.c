/* Exhaustive check of the float32 line-integrated table argument of the
 * per-wave line kernel (line_W_rdq): for the table steps dq = K / 4095 of the
 * supports K = 2, 2.5 and 3 (float32, as line_integrated_kernel.jl:14 forms
 * them) and every float32 q with 2^-40 <= |q| <= 16,
 *     fl(fl(q / dq) + 1) == fl(fma(fma(-q0, dq, q), y, q0) + 1),
 *     y = RN(1 / dq), q0 = RN(q y).
 * Below 2^-40 both sides are exactly 1 (|q / dq| < 2^-28).  Prints the count
 * of values checked and of mismatches; exit status 1 on any mismatch. */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#define NTHREADS 8
static float steps[3];
typedef struct {
    int t;
    uint64_t bad, n;
} Part;

static float from_bits(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static uint32_t to_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

static void* run(void* p) {
    Part* a = (Part*)p;
    const uint32_t lo = to_bits(ldexpf(1.0f, -40)), hi = to_bits(16.0f);
    for (int k = 0; k < 3; ++k) {
        volatile float dq = steps[k];
        const float y = 1.0f / dq;
        for (uint32_t u = lo + (uint32_t)a->t; u <= hi; u += NTHREADS) {
            for (int sgn = 0; sgn < 2; ++sgn) {
                const float q = from_bits(u | (sgn ? 0x80000000u : 0u));
                const float ref = q / dq + 1.0f;
                const float q0 = q * y;
                const float e = fmaf(-q0, dq, q);
                const float got = fmaf(e, y, q0) + 1.0f;
                a->n++;
                if (to_bits(ref) != to_bits(got)) {
                    if (a->bad < 5) printf("mismatch: dq=%a q=%a ref=%a got=%a\n", (double)dq, (double)q, (double)ref, (double)got);
                    a->bad++;
                }
            }
        }
    }
    return 0;
}

int main(void) {
    const double K[3] = {2.0, 2.5, 3.0};
    for (int k = 0; k < 3; ++k) steps[k] = (float)((1.0 * K[k]) / 4095.0) - 0.0f;
    pthread_t th[NTHREADS];
    Part parts[NTHREADS];
    for (int t = 0; t < NTHREADS; ++t) {
        parts[t].t = t;
        parts[t].bad = parts[t].n = 0;
        pthread_create(&th[t], 0, run, &parts[t]);
    }
    uint64_t bad = 0, n = 0;
    for (int t = 0; t < NTHREADS; ++t) {
        pthread_join(th[t], 0);
        bad += parts[t].bad;
        n += parts[t].n;
    }
    printf("checked %llu table arguments, %llu mismatches\n", (unsigned long long)n, (unsigned long long)bad);
    return bad != 0;
}
