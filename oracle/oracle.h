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

/* The same with the FP64 serial-parity pass of the GPU (prec 0): split rows
 * whose running error bound cannot show |y_cvr - y_serial| <= parity_tol *
 * |y_serial| (rigorous for rows <= 4096 elements, also probabilistic with
 * lambda = 10 beyond) are recomputed as the serial FMA sum from the CSR (ci,
 * va). bound (may be NULL, m entries): u * A of each split row (the rigorous
 * bound on |y_cvr - y_serial|), 0 on unsplit rows. parity_tol = 0 with bound
 * != NULL computes the bounds of the plain CVR order. */
void orc_cvr64_spmv_par(int m, int nnz, int S, int prec, const void *vals, const int *cols, const int *chunk_rows,
                        const int *seg, const int *split_n, const int *split_tab, const int *fix, const int *rp,
                        const int *ci, const double *va, const void *x, void *y, double parity_tol, double *bound,
                        int *n_recomputed);

/* ---------------------------------------------------------------- CSR5-64
 * CPU emulation of the GPU csr5_hip format (CSR5 re-derived for 64 lanes:
 * csr5/detail/avx512/format_avx512.h, csr5_spmv_avx512.h) and of its
 * summation order. Tile t holds elements [t*64*sigma, (t+1)*64*sigma); lane l
 * owns elements t*64*sigma + l*sigma + i (i < sigma), stored transposed at
 * t*64*sigma + i*64 + l; elements past nnz are padding (value 0, column 0).
 *   flags[t*64+l]  bit i: element i of the lane starts a non-empty row
 *   yoff[t*64+l]   row starts in lanes < l of the tile
 *   tile_base[t]   non-empty rows starting before the tile (index into nerows)
 *   nerows[]       the non-empty rows in order
 *   fix[2t]        row whose last element is in tile t and which began in an
 *                  earlier tile (else -1); fix[2t+1] the tile it began in
 *   cont[t]        1 if the row open at the end of tile t continues past it */
int orc_csr5_ntiles(int nnz, int sigma);
int orc_csr5_build(int m, int nnz, const int *rp, const int *ci, const double *va, int sigma, double *vals_t,
                   int *cols_t, uint32_t *flags, int *yoff, int *tile_base, int *nerows, int *fix,
                   unsigned char *cont);
/* y = A x over a built CSR5-64 layout, in the GPU kernel's summation order:
 * segments summed by fma from 0 in element order, a row crossing lanes as
 * (its owner lane's open segment) + (whole following lanes) + (leading
 * segment of the lane where it ends), in lane order, a row crossing tiles as
 * tail + whole tiles + head in tile order; rows with no element get 0. */
void orc_csr5_spmv(int m, int nnz, int sigma, const double *vals_t, const int *cols_t, const uint32_t *flags,
                   const int *yoff, const int *tile_base, const int *nerows, const int *fix,
                   const unsigned char *cont, const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
