/* kmx_demo.c -- a plain C caller of the operator-apply C ABI (include/kmx.h),
 * the way a MATLAB MEX gateway would drive it: host arrays in MATLAB layout
 * (1-based indices, column-major 1D matrices), kmx_plan_create once,
 * kmx_apply_host per apply.
 *
 * It builds a 2D full-grid operator (lev 2, deg 2, two terms) from a fixed
 * linear-congruential sequence, applies it through the library in FP64 and
 * FP32, checks against a direct sum over element pairs and Kronecker blocks
 * computed here, and prints "kmx_demo ok <rel err fp64> <rel err fp32>".
 * Exit status 0 on success. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../include/kmx.h"

static uint64_t state = 88172645463325252ull;
static double rnd(void) { /* xorshift64 -> [-1, 1) */
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return (double)(state >> 11) / 4503599627370496.0 - 1.0;
}

int main(void) {
    enum { D = 2, DEG = 2, LEV = 2, T = 2 };
    const int nb = 1 << LEV, n = DEG * nb, B = DEG * DEG, N = nb * nb;
    double* mats[T * D];
    int rows[D] = {n, n};
    for (int i = 0; i < T * D; ++i) {
        mats[i] = malloc(sizeof(double) * n * n);
        for (int e = 0; e < n * n; ++e) mats[i][e] = rnd();
    }
    /* full grid: element e = (I1, I2) with I2 fastest, 1-based 1D indices */
    int idx[N * D];
    for (int e = 0; e < N; ++e) {
        idx[e * D + 0] = e / nb + 1;
        idx[e * D + 1] = e % nb + 1;
    }
    double x[N * B], y64[N * B], y32[N * B], ref[N * B];
    for (int i = 0; i < N * B; ++i) x[i] = rnd();
    /* direct reference: y_i[a1,a2] = sum_j sum_t A1[I1 a1, J1 b1] A2[I2 a2, J2 b2] x_j[b1,b2] */
    for (int i = 0; i < N; ++i)
        for (int a1 = 0; a1 < DEG; ++a1)
            for (int a2 = 0; a2 < DEG; ++a2) {
                double s = 0.0;
                for (int j = 0; j < N; ++j)
                    for (int t = 0; t < T; ++t)
                        for (int b1 = 0; b1 < DEG; ++b1)
                            for (int b2 = 0; b2 < DEG; ++b2) {
                                const int r1 = (idx[i * D] - 1) * DEG + a1, c1 = (idx[j * D] - 1) * DEG + b1;
                                const int r2 = (idx[i * D + 1] - 1) * DEG + a2, c2 = (idx[j * D + 1] - 1) * DEG + b2;
                                s += mats[t * D][c1 * n + r1] * mats[t * D + 1][c2 * n + r2] * x[j * B + b1 * DEG + b2];
                            }
                ref[i * B + a1 * DEG + a2] = s;
            }
    kmx_plan_desc desc = {D, DEG, N, T, rows, (const double* const*)mats, idx, NULL, NULL, 1, KMX_PLAN_FP32, NULL};
    kmx_plan* plan = NULL;
    if (kmx_plan_create(&desc, &plan) != KMX_OK) {
        fprintf(stderr, "kmx_plan_create: %s\n", kmx_last_error());
        return 1;
    }
    if (kmx_apply_host(plan, x, y64, KMX_PREC_FP64) != KMX_OK || kmx_apply_host(plan, x, y32, KMX_PREC_FP32) != KMX_OK) {
        fprintf(stderr, "kmx_apply_host: %s\n", kmx_last_error());
        return 1;
    }
    double e64 = 0, e32 = 0, nr = 0;
    for (int i = 0; i < N * B; ++i) {
        e64 += (y64[i] - ref[i]) * (y64[i] - ref[i]);
        e32 += (y32[i] - ref[i]) * (y32[i] - ref[i]);
        nr += ref[i] * ref[i];
    }
    e64 = sqrt(e64 / nr);
    e32 = sqrt(e32 / nr);
    /* error path: a bad index is rejected with a code, not a crash */
    idx[0] = 99;
    kmx_plan* bad = NULL;
    const int rc = kmx_plan_create(&desc, &bad);
    kmx_plan_destroy(plan);
    for (int i = 0; i < T * D; ++i) free(mats[i]);
    printf("kmx_demo ok %.3e %.3e rc_bad=%d\n", e64, e32, rc);
    return (e64 <= 1e-13 && e32 <= 5e-6 && rc == KMX_E_ARG) ? 0 : 2;
}
