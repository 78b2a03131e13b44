/* Sanitizer harness for the oracle's CVR-64 and CSR5-64 emulations (built
 * with -fsanitize=address,undefined by tests/test_sanitizers.py). Random CSR
 * matrices with empty rows, long rows and ragged sizes; every layout is built
 * and multiplied, and y is checked against the serial FMA sum within the
 * blocked-sum bound. Exit status 0 = all good. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "oracle.h"

static unsigned long long st = 88172645463325252ull;
static unsigned long long rnd(void) {
    st ^= st << 13, st ^= st >> 7, st ^= st << 17;
    return st;
}

static int check(int m, const int *rp, const double *va, const int *ci, const double *x, const double *y,
                 const double *yr, int extra) {
    for (int r = 0; r < m; ++r) {
        double s = 0;
        for (int j = rp[r]; j < rp[r + 1]; ++j) s += fabs(va[j] * x[ci[j]]);
        const double bound = 2.0 * (rp[r + 1] - rp[r] + extra) * ldexp(1.0, -53) * s;
        if (fabs(y[r] - yr[r]) > bound) {
            fprintf(stderr, "row %d: %g vs %g\n", r, y[r], yr[r]);
            return 1;
        }
    }
    return 0;
}

int main(void) {
    int fails = 0;
    for (int it = 0; it < 60; ++it) {
        const int m = 1 + (int)(rnd() % 900), n = 1 + (int)(rnd() % 700);
        int *rp = (int *)malloc(sizeof(int) * (size_t)(m + 1));
        rp[0] = 0;
        for (int r = 0; r < m; ++r) {
            const unsigned long long k = rnd() % 100;
            int len = k < 25 ? 0 : k < 90 ? (int)(rnd() % 12) : (int)(rnd() % 3000);  /* empty, short, long */
            rp[r + 1] = rp[r] + len;
        }
        const int nnz = rp[m];
        int *ci = (int *)malloc(sizeof(int) * (size_t)(nnz + 1));
        double *va = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
        for (int j = 0; j < nnz; ++j) ci[j] = (int)(rnd() % (unsigned)n), va[j] = (double)(rnd() % 2001) / 1000.0 - 1.0;
        double *x = (double *)malloc(sizeof(double) * (size_t)n), *y = (double *)malloc(sizeof(double) * (size_t)m),
               *yr = (double *)malloc(sizeof(double) * (size_t)m);
        for (int c = 0; c < n; ++c) x[c] = (double)(rnd() % 2001) / 1000.0 - 1.0;
        orc_csr_spmv_fma(m, rp, ci, va, x, yr);
        const int Ss[3] = {8, 64, 256};
        for (int q = 0; q < 3; ++q) {  /* CVR-64, FP64 */
            const int S = Ss[q];
            int nc;
            long long L, sc;
            orc_cvr64_sizes(m, nnz, S, &nc, &L, &sc);
            double *ov = (double *)malloc(sizeof(double) * (size_t)L);
            int *oc = (int *)malloc(sizeof(int) * (size_t)L);
            int *chunk_rows = (int *)malloc(sizeof(int) * (size_t)(nc + 1));
            int *seg = (int *)malloc(sizeof(int) * (size_t)sc);
            const size_t nc1 = (size_t)(nc > 1 ? nc : 1);
            int *segc = (int *)malloc(sizeof(int) * nc1), *spn = (int *)malloc(sizeof(int) * nc1);
            int *spt = (int *)malloc(sizeof(int) * 144 * nc1);
            int *fix = (int *)malloc(sizeof(int) * 2 * nc1);
            int info[3];
            if (orc_cvr64_build(m, n, nnz, rp, ci, va, S, 0, ov, oc, chunk_rows, seg, segc, spn, spt, fix, info)) {
                fprintf(stderr, "cvr build failed\n");
                ++fails;
            } else {
                orc_cvr64_spmv(m, nnz, S, 0, ov, oc, chunk_rows, seg, spn, spt, fix, rp, x, y);
                fails += check(m, rp, va, ci, x, y, yr, 64);
            }
            free(ov), free(oc), free(chunk_rows), free(seg), free(segc), free(spn), free(spt), free(fix);
        }
        const int sgs[3] = {1, 7, 32};
        for (int q = 0; q < 3; ++q) {  /* CSR5-64 */
            const int sg = sgs[q], nt = orc_csr5_ntiles(nnz, sg);
            const size_t L = (size_t)nt * 64 * (size_t)sg;
            /* exact sizes (no slack), so an off-by-one in the emulation is an ASan error */
            double *tv = (double *)malloc(sizeof(double) * L);
            int *tc = (int *)malloc(sizeof(int) * L);
            uint32_t *fl = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nt * 64);
            int *yo = (int *)malloc(sizeof(int) * (size_t)nt * 64), *tb = (int *)malloc(sizeof(int) * (size_t)nt);
            int *ne = (int *)malloc(sizeof(int) * (size_t)m), *fx = (int *)malloc(sizeof(int) * 2 * (size_t)nt);
            unsigned char *co = (unsigned char *)malloc((size_t)nt);
            if (orc_csr5_build(m, nnz, rp, ci, va, sg, tv, tc, fl, yo, tb, ne, fx, co)) {
                fprintf(stderr, "csr5 build failed\n");
                ++fails;
            } else {
                orc_csr5_spmv(m, nnz, sg, tv, tc, fl, yo, tb, ne, fx, co, x, y);
                fails += check(m, rp, va, ci, x, y, yr, 64 + nnz / (64 * sg) + 2);
            }
            free(tv), free(tc), free(fl), free(yo), free(tb), free(ne), free(fx), free(co);
        }
        free(rp), free(ci), free(va), free(x), free(y), free(yr);
    }
    printf("oracle harness: %d failures\n", fails);
    return fails != 0;
}
