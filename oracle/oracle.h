/*
 * oracle.h -- CPU oracle for the CVR SpMV path. TEST INFRASTRUCTURE ONLY:
 * used by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py as the checker; never linked into or called by the product.
 */
#ifndef PAVA_ORACLE_H
#define PAVA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* y_i = sum_j a_ij x_j, sequential in j, one rounding per term via fma
 * (utils/basicMatrix.cpp:538-556 referenceSpMV, with the contraction fixed). */
void orc_csr_spmv_fma(int m, const int *rp, const int *ci, const double *va, const double *x, double *y);
/* Same order, separate multiply and add (two roundings per term). */
void orc_csr_spmv_muladd(int m, const int *rp, const int *ci, const double *va, const double *x, double *y);
/* Double-double accumulation (error-free TwoProduct by fma + TwoSum, Ogita,
 * Rump, Oishi, SIAM J. Sci. Comput. 26(6), 2005, Algorithm Dot2): result as
 * if computed in ~twice the working precision, then rounded. */
void orc_csr_spmv_dd(int m, const int *rp, const int *ci, const double *va, const double *x, double *y);
/* s_i = sum_j |a_ij x_j| (the scale of row i's rounding error). */
void orc_row_abs(int m, const int *rp, const int *ci, const double *va, const double *x, double *s);
/* Same as orc_csr_spmv_fma, rows split over `threads` OpenMP threads. */
void orc_csr_spmv_omp(int m, const int *rp, const int *ci, const double *va, const double *x, double *y,
                      int threads);
/* Rows with |y - y_ref| > 0.01 |y_ref| (utils/basicMatrix.cpp:559-587). */
int orc_check_results(int n, const double *y, const double *y_ref);

/* ---------------------------------------------------------------- CVR-64
 * CPU emulation of the CVR-64 layout built by the GPU conversion and of the
 * GPU SpMV's summation order. prec: 0 FP64, 1 FP32, 2 BF16 value storage. */
void orc_cvr64_sizes(int m, int nnz, int S, int *nchunks, long long *layout_elems, long long *seg_cap);
int orc_cvr64_build(int m, int n, int nnz, const int *rp, const int *ci, const double *va, int S, int prec,
                    void *ovals, int *ocols, int *chunk_rows, int *seg, int *seg_count, int *split_n,
                    int *split_tab, int *fix, int *info /* [0] max slots, [1] n fix rows, [2] has empty */);
/* y = A x over a built layout. prec 0: x, y double*; prec 1/2: x, y float*. */
void orc_cvr64_spmv(int m, int nnz, int S, int prec, const void *vals, const int *cols, const int *chunk_rows,
                    const int *seg, const int *split_n, const int *split_tab, const int *fix, const int *rp,
                    const void *x, void *y);

#ifdef __cplusplus
}
#endif
#endif
