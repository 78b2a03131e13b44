/*
 * pava_hip.h -- C ABI of the MI355X CVR SpMV path ("cvr_hip" format).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns
 * PAVA_HIP_OK (0) or a negative PAVA_HIP_E* code.  pava_hip_last_error() gives a
 * message for the most recent failure of the calling host thread.
 *
 * Streams are passed as `void*` holding a hipStream_t (NULL = the null stream),
 * so this header needs no HIP include.
 *
 * Reference interfaces replaced (paths relative to the pava source tree):
 *   pava_coo_matrix / pava_csr_matrix  <- struct _PAVA_COOMatrix / _PAVA_CSRMatrix
 *                                         utils/basicMatrix.h:38-60
 *   pava_read_coo_matrix               <- readCOOMatrix       utils/basicMatrix.cpp:63-253
 *   pava_convert_coo_to_csr            <- convertCOO2CSR      utils/basicMatrix.cpp:350-415
 *   pava_delete_{coo,csr}_matrix       <- deleteCOOMatrix / deleteCSRMatrix basicMatrix.h:124-125
 *   pava_init_vectors                  <- initVectors         utils/basicMatrix.cpp:499-520
 *   pava_reference_spmv                <- referenceSpMV       utils/basicMatrix.cpp:538-556
 *   pava_check_results                 <- checkResults        utils/basicMatrix.cpp:559-587
 *   pava_print_performance             <- printPerformance    utils/basicMatrix.cpp:491-495
 *   pava_microtime                     <- microtime           utils/microtime.h:9-17
 *   conduct_cvr_hip                    <- conduct_cvr         cvr/cvr.h:1953-2186
 *   cvr_hip_create / cvr_hip_create_device
 *                                      <- pre_processing      cvr/cvr.h:583-1060
 *   cvr_hip_spmv / cvr_hip_spmv_host   <- spmv_compute_kernel cvr/cvr.h:1062-1884
 *   pava_bench_file                    <- benchFile + conduct_benchmark (CVR block)
 *                                         main.cpp:61-114, 203-207, 554-571
 */
#ifndef PAVA_HIP_H
#define PAVA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ errors */
#define PAVA_HIP_OK 0
#define PAVA_HIP_EINVAL (-1)   /* bad argument or malformed matrix */
#define PAVA_HIP_ENOMEM (-2)   /* host or device allocation failed */
#define PAVA_HIP_EDEVICE (-3)  /* HIP runtime error / no device */
#define PAVA_HIP_EIO (-4)      /* file could not be opened / read / written */
#define PAVA_HIP_EFORMAT (-5)  /* Matrix Market banner or body not supported */
#define PAVA_HIP_ERANGE (-6)   /* index out of range or size beyond int32 */
#define PAVA_HIP_ESTATE (-7)   /* handle used with the wrong precision / API */

/* Largest numRows, numCols and nnz the GPU formats accept: 2^31 - 2^16, so that
 * element positions plus a chunk's padding or a 512-element stride stay in
 * int32 (larger matrices: PAVA_HIP_ERANGE). The reader and the CSR cache take
 * up to INT32_MAX. */
#define PAVA_HIP_MAX_SIZE 2147418112

/* Reader return codes follow readCOOMatrix (utils/basicMatrix.cpp:84-132):
 * -1 bad banner, -2 not a matrix, -3 not sparse (coordinate), -4 bad size line,
 * -5 allocation failure; this library adds -6 (entry out of range / truncated
 * body) and -7 (file cannot be opened; the reference calls exit(1) there). */
#define PAVA_MM_EBANNER (-1)
#define PAVA_MM_ENOTMATRIX (-2)
#define PAVA_MM_ENOTSPARSE (-3)
#define PAVA_MM_ESIZE (-4)
#define PAVA_MM_ENOMEM (-5)
#define PAVA_MM_EBODY (-6)
#define PAVA_MM_EOPEN (-7)

/* --------------------------------------------------------- host matrices */
/* Same fields, order and meaning as _PAVA_COOMatrix / _PAVA_CSRMatrix
 * (utils/basicMatrix.h:38-60): 0-based int32 indices, FP64 values. Arrays
 * returned by this library are 64-byte aligned and released by the matching
 * pava_delete_* call. */
typedef struct pava_coo_matrix {
    int nnz;
    int numRows;
    int numCols;
    double *vals;
    int *cols;
    int *rows;
} pava_coo_matrix;

typedef struct pava_csr_matrix {
    int nnz;
    int numRows;
    int numCols;
    double *vals;
    int *cols;
    int *rowOffsets;
} pava_csr_matrix;

/* Matrix Market (coordinate; real/integer/pattern/complex-real-part;
 * general/symmetric/skew-symmetric/hermitian) -> COO with 0-based indices,
 * symmetric halves mirrored. Unlike the reference no padding entries are
 * added and indices are decremented to 0-based. */
int pava_read_coo_matrix(const char *path, pava_coo_matrix *out);
/* COO -> CSR sorted by (row, col); duplicate entries are summed in file order. */
int pava_convert_coo_to_csr(const pava_coo_matrix *coo, pava_csr_matrix *out);
/* Convenience: read + convert. A "<path>.csrbin" cache newer than the file is
 * read instead of parsing (the fast input path); the body of a large file is
 * parsed by all OpenMP threads. */
int pava_read_csr_matrix(const char *path, pava_csr_matrix *out);
/* Binary CSR cache: "PAVACSR1", int32 numRows, numCols, int64 nnz, then
 * rowOffsets, cols (int32) and vals (f64), little-endian. */
int pava_write_csr_binary(const char *path, const pava_csr_matrix *A);
int pava_read_csr_binary(const char *path, pava_csr_matrix *out);
/* Write a CSR as "%%MatrixMarket matrix coordinate real general", 1-based,
 * values printed with 17 significant digits (round-trips FP64 exactly). */
int pava_write_matrix_market(const char *path, const pava_csr_matrix *A);
void pava_delete_coo_matrix(pava_coo_matrix *m);
void pava_delete_csr_matrix(pava_csr_matrix *m);
/* Validate a CSR (monotone offsets, indices in range). */
int pava_csr_validate(const pava_csr_matrix *A);

/* initVectors: y = 0, y_ref = 0, x[i] = i % 3 (any pointer may be NULL). */
void pava_init_vectors(int numRows, int numCols, double *x, double *y, double *y_ref);
/* x[i] ~ U[lo, hi) from a counter-based generator keyed by (seed, i). */
void pava_fill_uniform(double *x, long long n, unsigned long long seed, double lo, double hi);
/* The driver's own verification SpMV (main.cpp:318 calls referenceSpMV):
 * sequential FP64 sum per row with fused multiply-add, OpenMP over rows. */
void pava_reference_spmv(const pava_csr_matrix *A, const double *x, double *y_ref);
/* Number of rows with |y - y_ref| > 0.01 |y_ref|; prints "Checking Pass" or
 * the warning line exactly like checkResults. */
int pava_check_results(int size, const double *y, const double *y_ref);
/* " Performance <matrix> <format> <threads> <convert_s> <exec_s> " */
int pava_print_performance(const char *matrixName, const char *format, int numThreads,
                           double convertTime, double executionTime);
double pava_microtime(void);

/* ------------------------------------------------- synthetic generators */
/* 2D 5-point Laplacian on nx*ny grid: diagonal 4, neighbours -1. */
int pava_gen_laplacian2d(int nx, int ny, pava_csr_matrix *out);
/* 3D 27-point stencil on n^3 grid: diagonal 26, off-diagonal -1 + U(-jitter, jitter). */
int pava_gen_stencil27(int n, double jitter, unsigned long long seed, pava_csr_matrix *out);
/* R-MAT: 2^scale square, edge_factor*2^scale edges, quadrant probabilities
 * a,b,c (d = 1-a-b-c), values U(0,1], duplicates summed. No vertex permutation. */
int pava_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                  unsigned long long seed, pava_csr_matrix *out);

/* Row-length distributions and column structures of the synthetic families. */
#define PAVA_RL_CONSTANT 0
#define PAVA_RL_POISSON 1
#define PAVA_RL_LOGNORMAL 2
#define PAVA_RL_POWERLAW 3
#define PAVA_COL_BANDED 0
#define PAVA_COL_BLOCKDIAG 1
#define PAVA_COL_UNIFORM 2
#define PAVA_COL_MIXED 3 /* band_frac of the entries within +-half_band, rest uniform */
typedef struct pava_gen_params {
    int rows;            /* square matrix */
    int row_dist;        /* PAVA_RL_* */
    double len_param;    /* mean length (constant / Poisson / lognormal median) or x_min (power law) */
    double shape;        /* lognormal sigma, or power-law alpha */
    int col_struct;      /* PAVA_COL_* */
    int half_band;       /* banded / mixed: columns within row +- half_band */
    int block;           /* block-diagonal block size */
    double band_frac;    /* mixed: fraction of banded entries */
    int value_dist;      /* 0: N(0,1); 1: U(0,1] */
    int max_len;         /* cap on a row length (<= rows) */
    unsigned long long seed;
} pava_gen_params;
int pava_gen_synthetic(const pava_gen_params *p, pava_csr_matrix *out);

/* ------------------------------------------------------------- CVR-HIP */
#define CVR_HIP_LANES 64
#define CVR_HIP_PREC_F64 0   /* FP64 values, x, y; FP64 accumulation */
#define CVR_HIP_PREC_F32 1   /* FP32 values, x, y; FP32 compensated accumulation */
#define CVR_HIP_PREC_BF16 2  /* BF16 values, FP32 x, y; FP32 compensated accumulation */

typedef struct cvr_hip_opts {
    int steps_per_chunk;     /* S: 64-lane steps per chunk; multiple of 8, 8..4096; 0 = automatic
                                (default): about four chunks per resident wave slot of the device,
                                at most 256 (cvr_hip_get_info reports the value used) */
    int precision;           /* CVR_HIP_PREC_* */
    int col_index_bits;      /* 32; 16: chunk-relative offsets plus a 64-bit end mask per step
                                (2.125 B/nnz instead of 4; PAVA_HIP_ERANGE if a chunk spans
                                65,536 columns or more; no hot-x staging); 0 = automatic
                                (default): 16 when every chunk fits and x_hot_kb <= 0, else 32 */
    int recompute;           /* mixed precision: recompute rows with kappa > kappa_threshold in FP64 */
    double kappa_threshold;  /* sum|a*x| / |sum a*x| above which a row is recomputed (default 64:
                                FP32 storage then stays within ~1e-5 relative of FP64) */
    int keep_fp64;           /* keep the FP64 CSR resident (required by recompute) */
    int device;              /* HIP device ordinal */
    int x_hot_kb;            /* hot-x staging: the most used columns (each used >= 2x, at most
                                x_hot_kb KiB of x) are renumbered, most used first, into a dense
                                copy of x refreshed before each SpMV; 0 = off; -1 = automatic
                                (default): 2048 KiB, staged only when the most used eighth of
                                the used columns is used at least twice as often as an average
                                used column (32-bit layouts only) */
    int x_lds_kb;            /* of the hot-x copy, the first x_lds_kb KiB are held in LDS by each
                                workgroup of the SpMV; -1 = as much as fits (default), 0 = none */
    double parity_tol;       /* FP64 handles: a row that the CVR order splits is recomputed as one
                                serial FMA chain (the order of the reference's referenceSpMV,
                                utils/basicMatrix.cpp:538-556) when its running error bound cannot
                                show |y_cvr - y_serial| <= parity_tol * |y_serial|. Rows of at most
                                4,096 elements use the rigorous bound; longer rows also accept a
                                probabilistic bound (lambda = 10). Unsplit rows equal the serial sum
                                bit for bit. Default 1e-12; 0 = off (pure CVR order). Keeps the FP64
                                CSR resident (12 B per nonzero + 4 B per row) */
} cvr_hip_opts;

typedef struct cvr_hip_info {
    int numRows, numCols, nnz;
    int nchunks, steps_per_chunk, precision, col_index_bits;
    long long layout_elems;  /* 64 * total steps (nnz + padding) */
    long long seg_capacity;  /* entries allocated for the segment list */
    int max_slots;           /* largest number of split-segment slots in one chunk */
    int has_empty_rows;
    int n_fixup_rows;        /* rows spanning more than one chunk */
    long long device_bytes;  /* device memory held by the handle */
    double convert_ms;       /* GPU time of the CSR->CVR conversion (hipEvent) */
    int n_hot_cols;          /* columns staged in the hot-x copy */
    int n_lds_cols;          /* of those, columns the SpMV reads from LDS */
    double hot_setup_ms;     /* wall time of the hot-column staging (counts, selection, remap) */
    double parity_tol;       /* in effect (0: off) */
    int n_spec_rows;         /* FP64 parity without the LDS table: rows spanning chunks (at least
                                4,096 elements, the 1,024 longest) whose serial sum is started at
                                the beginning of each SpMV, before their bound is known; used when
                                the bound fails, abandoned when it passes (results unchanged) */
} cvr_hip_info;

typedef struct cvr_hip_handle cvr_hip_handle;

void cvr_hip_default_opts(cvr_hip_opts *o);
/* Host CSR -> device upload -> GPU CSR->CVR conversion. Synchronous. */
int cvr_hip_create(const pava_csr_matrix *A, const cvr_hip_opts *opts, cvr_hip_handle **out);
/* Same from a CSR already in device memory (the arrays are not retained). The
 * call first waits for all work on the device (the CSR may still be written
 * by the caller's streams), then checks the CSR on the device: offsets from 0
 * to nnz, non-decreasing (PAVA_HIP_EINVAL otherwise) and every column in
 * [0, numCols) (PAVA_HIP_ERANGE). 32-bit column layouts (col_index_bits 32, or
 * automatic when a chunk spans 65,536 columns or more) take at most 2^30
 * columns (PAVA_HIP_ERANGE); 16-bit layouts take up to PAVA_HIP_MAX_SIZE. */
int cvr_hip_create_device(int numRows, int numCols, int nnz, const int *d_rowOffsets,
                          const int *d_cols, const double *d_vals, const cvr_hip_opts *opts,
                          cvr_hip_handle **out);
int cvr_hip_destroy(cvr_hip_handle *h);
int cvr_hip_get_info(const cvr_hip_handle *h, cvr_hip_info *info);

/* y = A*x, FP64 handle, device pointers, enqueued on `stream`, asynchronous.
 * y is fully written (empty rows get 0); no pre-zeroing needed. A handle
 * holds per-SpMV device state (the hot-x copy, carries of rows spanning
 * chunks, their arrival counters and the shared chunk-queue counters, each
 * left re-armed by the kernel for the next call): SpMVs and SpMMs on one
 * handle must not overlap (one stream, or ordered streams). */
int cvr_hip_spmv(cvr_hip_handle *h, const double *d_x, double *d_y, void *stream);
/* Host-pointer variant (copies in and out, synchronous). */
int cvr_hip_spmv_host(cvr_hip_handle *h, const double *x, double *y);
/* y = A*x for FP32/BF16 handles: x and y FP32. d_x64 (may be NULL) is the FP64
 * x used by the FP64 recompute of high-cancellation rows. */
int cvr_hip_spmv_mixed(cvr_hip_handle *h, const float *d_x32, const double *d_x64, float *d_y,
                       void *stream);
/* Y = A*X with k right-hand sides, X (numCols x k) and Y (numRows x k) row-major
 * FP32, device pointers; 1 <= k <= 32. Any precision of handle. The first call
 * on a handle with nonzeros allocates its cross-chunk buffer (nchunks * 512 B):
 * make one call before capturing the call in a graph. cvr_hip_spmv and
 * cvr_hip_spmv_mixed enqueue kernels only and can be captured at any time. */
int cvr_hip_spmm(cvr_hip_handle *h, int k, const float *d_X, float *d_Y, void *stream);
/* SpMM engines (no counterpart in pava, which is single-RHS): per-lane FP64
 * FMA on the vector ALUs (PAVA_HIP_SPMM_FMA, the default), or the matrix cores
 * (PAVA_HIP_SPMM_MFMA): v_mfma_f64_4x4x4f64 over windows of 4 steps of 4 CVR
 * lanes, each nonzero scaling a tile of its X row into its row segment's
 * accumulator. Both accumulate exact products in FP64 and round once to FP32;
 * they may differ only in the order of the additions inside a window. The
 * engine applies to cvr_hip_spmm and cvr_hip_spmm_noalloc on this handle;
 * EINVAL for another value. */
#define PAVA_HIP_SPMM_FMA 0
#define PAVA_HIP_SPMM_MFMA 1
int cvr_hip_set_spmm_engine(cvr_hip_handle *h, int engine);
/* cvr_hip_spmm without any allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). Same kernels,
 * arguments and results (bit for bit) as cvr_hip_spmm, but the cross-chunk
 * partial sums go to a 16 MiB scratch array that the library holds on every
 * GPU it runs on until the process ends, instead of a buffer the handle
 * allocates on its first SpMM; so it may be captured in a graph at any time.
 * Needs 2 * nchunks * k * 8 bytes <= 16 MiB (PAVA_HIP_ERANGE otherwise, before
 * any device work). All calls on one GPU share the array: calls that overlap
 * on one GPU (several streams, host threads or graphs in flight) race. */
int cvr_hip_spmm_noalloc(cvr_hip_handle *h, int k, const float *d_X, float *d_Y, void *stream);
/* Batched multi-RHS SpMM: the chunks of many matrices (same device and
 * precision, 16- or 32-bit columns, nnz > 0) form one work queue consumed by a
 * persistent grid sized to the GPU's resident capacity; a fixup kernel closes
 * rows spanning chunks and rearms the queue, so the launch pair can be captured
 * once in a hipGraph and replayed. The handles must outlive the batch; a
 * handle may appear once per batch (EINVAL otherwise), and must not be used by
 * cvr_hip_spmm or another batch while a batch holding it is in flight (they
 * share its cross-chunk buffer). Accumulation is FP64 over FP32 inputs (one
 * rounding per output). */
typedef struct cvr_hip_batch cvr_hip_batch;
int cvr_hip_batch_create(int nmat, cvr_hip_handle *const *handles, int k, cvr_hip_batch **out);
/* Device pointers of each matrix's X (numCols x k) and Y (numRows x k), FP32
 * row-major, 16-byte aligned when k is 8, 16 or 32. Synchronous upload. */
int cvr_hip_batch_set_operands(cvr_hip_batch *b, const float *const *d_X, float *const *d_Y);
int cvr_hip_batch_spmm(cvr_hip_batch *b, void *stream);
/* X staging for k <= 8 (no counterpart in pava: its batching is this
 * library's own): with wgs_per_cu = 1 or 2, workgroups of 8 waves take groups
 * of 4 consecutive chunks of one matrix and first stage in LDS the rows of
 * that matrix's X around the group's rows (as many as fit with wgs_per_cu
 * workgroups per CU); nonzeros whose column falls there read X from LDS. Y is
 * unchanged bit for bit. 0 turns it off (the default). EINVAL for k > 8,
 * ERANGE if no window of 256 rows fits. Drops a captured graph: call
 * cvr_hip_batch_graph again after it. */
int cvr_hip_batch_set_x_window(cvr_hip_batch *b, int wgs_per_cu);
/* The batch's SpMM engine (PAVA_HIP_SPMM_FMA, the default, or
 * PAVA_HIP_SPMM_MFMA; see cvr_hip_set_spmm_engine). The MFMA engine does not
 * take the X window (EINVAL with a window on, and the window EINVAL with the
 * MFMA engine). Drops a captured graph: call cvr_hip_batch_graph again after
 * it. */
int cvr_hip_batch_set_engine(cvr_hip_batch *b, int engine);
/* Captures one batched SpMM on `stream` (non-null) into a hipGraph. */
int cvr_hip_batch_graph(cvr_hip_batch *b, void *stream);
int cvr_hip_batch_replay(cvr_hip_batch *b, void *stream);
/* *items = work items of the queue (chunks; groups of 4 chunks with the X
 * window), *grid = workgroups of the persistent launch. */
int cvr_hip_batch_info(const cvr_hip_batch *b, int *items, int *grid, long long *nnz);
/* Queue check of the last completed batched SpMM (synchronises the device).
 * The queue is cut into *nranges = min(grid, 8) ranges; range q holds work
 * items [total*q/n, total*(q+1)/n). counts[q] (8 entries) = pops from range q:
 * each workgroup pops until it draws an index past the range's end, so every
 * item handed out once gives counts[q] = range size + grid (0 beyond nranges). */
int cvr_hip_batch_queue_counts(const cvr_hip_batch *b, int *counts, int *nranges);
int cvr_hip_batch_destroy(cvr_hip_batch *b);

/* "csr_hip": plain CSR SpMV with lanes_per_row (power of two, 1..64) lanes per
 * row, FP64 or FP32 values (x, y of the same type). The GPU stand-in for the
 * MKL CSR block benchmark_CSR_SpMV (utils/pava_formats.cpp:39-47) and the
 * selector's second format. */
typedef struct csr_hip_handle csr_hip_handle;
int csr_hip_create(const pava_csr_matrix *A, int lanes_per_row, int precision, int device, csr_hip_handle **out);
int csr_hip_spmv(csr_hip_handle *h, const void *d_x, void *d_y, void *stream);
int csr_hip_destroy(csr_hip_handle *h);

/* "coo_hip": COO SpMV, FP64. The CSR's nonzeros in row order with a row index
 * per nonzero (the conversion expands the row pointers on the GPU, as
 * convertCSR2COO does on the CPU, utils/pava_formats.cpp:147+). The GPU
 * stand-in for the MKL COO block benchmark_COO_SpMV (utils/pava_formats.cpp:
 * 50-61). One wave per tile of 512 nonzeros, segmented sums in a fixed order:
 * deterministic; y is written completely (rows without nonzeros get 0). */
typedef struct coo_hip_handle coo_hip_handle;
int coo_hip_create(const pava_csr_matrix *A, int device, coo_hip_handle **out);
int coo_hip_spmv(coo_hip_handle *h, const double *d_x, double *d_y, void *stream);
int coo_hip_get_info(coo_hip_handle *h, int *ntiles, double *convert_ms);
int coo_hip_destroy(coo_hip_handle *h);

/* "csr5_hip": CSR5 (csr5/detail/avx512/format_avx512.h, csr5_spmv_avx512.h;
 * anonymouslibHandle::asCSR5 / spmv, csr5/anonymouslib_avx512.h:107-221)
 * re-derived for 64-lane wavefronts: tiles of 64 x sigma nonzeros stored
 * transposed, per-lane row-start bit flags, lane-order segmented sums and a
 * tile-order calibrate step. FP64. sigma in {4, 8, 16, 32}, 0 -> 16.
 * csr5_hip_export_layout copies the device arrays out (vals/cols: ntiles *
 * 64 * sigma; flags/yoff: ntiles * 64; tile_base: ntiles; nerows: non-empty
 * rows; fix: 2 * ntiles; cont: ntiles). */
typedef struct csr5_hip_handle csr5_hip_handle;
int csr5_hip_create(const pava_csr_matrix *A, int sigma, int device, csr5_hip_handle **out);
int csr5_hip_spmv(csr5_hip_handle *h, const double *d_x, double *d_y, void *stream);
int csr5_hip_spmv_host(csr5_hip_handle *h, const double *x, double *y);
int csr5_hip_get_info(const csr5_hip_handle *h, int *ntiles, int *sigma, double *convert_ms);
int csr5_hip_export_layout(const csr5_hip_handle *h, double *vals, int *cols, unsigned *flags, int *yoff,
                           int *tile_base, int *nerows, int *fix, unsigned char *cont);
int csr5_hip_destroy(csr5_hip_handle *h);

/* ------------------------------------------------------------- selector
 * No counterpart in pava (it benchmarks every format instead of choosing).
 * 16 cheap CSR features (log sizes, row-length statistics, bandwidth, gather
 * locality), then a 3-layer MLP (n_in -> n_hidden -> n_hidden -> n_out, ReLU,
 * FP32) trained and run on the GPU by this library's own kernels. Parameter
 * layout: W1[n_hidden x n_in], b1, W2[n_hidden x n_hidden], b2,
 * W3[n_out x n_hidden], b3 (row-major). Loss: masked mean squared error. */
#define PAVA_N_FEATURES 16
int pava_csr_features(const pava_csr_matrix *A, float *out /* PAVA_N_FEATURES */);
typedef struct pava_mlp pava_mlp;
int pava_mlp_create(int n_in, int n_hidden, int n_out, unsigned long long seed, int device, pava_mlp **out);
int pava_mlp_num_params(const pava_mlp *m);
int pava_mlp_get_params(const pava_mlp *m, float *host);
int pava_mlp_set_params(pava_mlp *m, const float *host);
/* Y[batch x n_out] = MLP(X[batch x n_in]); host pointers. */
int pava_mlp_predict(pava_mlp *m, int batch, const float *X, float *Y);
/* Loss and flat gradient of one batch (host pointers; T targets, M mask). */
int pava_mlp_loss_grad(pava_mlp *m, int batch, const float *X, const float *T, const float *M, float *loss,
                       float *grad);
/* Adam on shuffled minibatches; loss_hist[epoch] = mean minibatch loss. */
int pava_mlp_train(pava_mlp *m, int n, const float *X, const float *T, const float *M, int epochs, int batch,
                   float lr, unsigned long long seed, float *loss_hist);
int pava_mlp_destroy(pava_mlp *m);

/* Number of rows recomputed in FP64 by the last mixed SpMV, or in serial order
 * by the last FP64 SpMV with parity_tol > 0 (synchronises). */
int cvr_hip_last_recompute_count(cvr_hip_handle *h, int *count);

/* Timing of `iters` back-to-back SpMVs after `warmup`, with hipEvents on the
 * handle's stream: ms_total = average wall per SpMV (all kernels), ms_main =
 * average duration of the main SpMV kernel alone (events around it). */
int cvr_hip_time_spmv(cvr_hip_handle *h, const void *d_x, void *d_y, int warmup, int iters,
                      double *ms_total, double *ms_main);
/* Same for a mixed-precision handle built with recompute (d_x64 = FP64 x). */
int cvr_hip_time_spmv_x64(cvr_hip_handle *h, const void *d_x, const double *d_x64, void *d_y, int warmup,
                          int iters, double *ms_total, double *ms_main);
/* The same measurement with the spread of the main kernel: stats[0] = mean ms
 * per SpMV (all kernels), stats[1..4] = mean, median, min and max ms of the
 * main SpMV kernel over the iters launches (d_x64 may be null unless the
 * handle recomputes). */
int cvr_hip_time_spmv_stats(cvr_hip_handle *h, const void *d_x, const double *d_x64, void *d_y, int warmup,
                            int iters, double *stats);

/* Copies of the layout for checking (host buffers sized from cvr_hip_info):
 * vals (layout_elems * value size), cols (layout_elems int32, or uint16 when
 * col_index_bits == 16), chunk_rows (nchunks+1), seg (seg_capacity),
 * seg_count (nchunks), split_n (nchunks), split_tab (2*nchunks*CVR_HIP_SPLIT_CAP),
 * fix (2*nchunks), col_base (nchunks, 16-bit layouts only), end_mask
 * (layout_elems/64 uint64, 16-bit layouts only). Any pointer may be NULL. */
/* hot_cols (n_hot_cols entries): the column of each hot index; a layout column
 * with bit 30 set refers to hot_cols[col & 0x3fffffff]. */
int cvr_hip_export_hot(const cvr_hip_handle *h, int *hot_cols);
#define CVR_HIP_SPLIT_CAP 72
#define CVR_HIP_SEG_PER_CHUNK_EXTRA 66
int cvr_hip_export_layout(const cvr_hip_handle *h, void *vals, void *cols, int *chunk_rows,
                          int *seg, int *seg_count, int *split_n, int *split_tab, int *fix,
                          int *col_base, uint64_t *end_mask);

/* Drop-in for conduct_cvr (cvr/cvr.h:1953): same arguments; converts (timed),
 * runs one verification SpMV into refOut, numIterations timed SpMVs, prints the
 * printPerformance line with format "CVR_HIP". Owns refOut completely (no
 * pre-zeroing needed, empty rows written). nItems_raw needs no padding; alpha
 * is accepted and ignored as in the reference. Returns 0 or PAVA_HIP_E*. */
int conduct_cvr_hip(int numRows, int numCols, int nItems_raw, int *h_rowDelimiters, int *h_cols,
                    double *h_val, double *h_vec, double *refOut, int alpha, char *matrixName,
                    int numIterations);

/* The csr_hip block (same protocol as conduct_cvr_hip, report format
 * "CSR_HIP", thread column = lanes per row); stands in for the MKL CSR block
 * (main.cpp CSR block, utils/pava_formats.cpp:39-47). */
int conduct_csr_hip(int numRows, int numCols, int nItems, int *h_rowDelimiters, int *h_cols, double *h_val,
                    double *h_vec, double *refOut, int lanes_per_row, char *matrixName, int numIterations);

/* The coo_hip block (same protocol, report format "COO_HIP", thread column =
 * tiles); stands in for the MKL COO block (main.cpp COO block,
 * utils/pava_formats.cpp:50-61). */
int conduct_coo_hip(int numRows, int numCols, int nItems, int *h_rowDelimiters, int *h_cols, double *h_val,
                    double *h_vec, double *refOut, char *matrixName, int numIterations);

/* conduct_csr5 (csr5/csr5.h:72-179) on the GPU, same arguments (alpha is a
 * double there and ignored here as there): conversion (timed, sigma 16), one
 * verification SpMV into y (written completely), numIterations timed SpMVs,
 * the report line with format "CSR5_HIP" and thread column = tiles. */
int conduct_csr5_hip(int m, int n, int nnz, int *csrRowPtr, int *csrColIdx, double *csrVal, double *x,
                     double *y, double alpha, char *matrixName, int numIterations);

/* benchFile with the GPU format registry: read, size filters (main.cpp:82-99),
 * x = i%3, referenceSpMV, then each format block in turn (cvr_hip, csr5_hip, csr_hip, coo_hip)
 * with checkResults. Returns the total error count (>= 0), -1 if the matrix
 * was skipped, or a PAVA_HIP_E* code. */
int pava_bench_file(const char *path, int numIterations);

/* Test hooks (no counterpart in the reference; for tests of the failure
 * paths). pava_hip_test_fail_alloc(h, d): counted from this call, the h-th
 * host allocation of the matrix readers and converters and the d-th device
 * allocation of the CVR-HIP handle and batch code fail as if memory had run
 * out (0 = no failure). pava_hip_test_live_device_allocs: device allocations
 * of that code currently held. */
void pava_hip_test_fail_alloc(int host_nth, int device_nth);
long long pava_hip_test_live_device_allocs(void);

const char *pava_hip_last_error(void);
/* Library version string. */
const char *pava_hip_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PAVA_HIP_H */
