// cvr_kernels.h -- layout constants and kernel launch declarations shared by
// the CVR-64 device code and the host handle code.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace cvr {

constexpr int kLanes = 64;          // one wavefront = one CVR "SIMD unit"
constexpr int kWavesPerBlock = 4;   // 256-thread workgroups, one chunk per wave
constexpr int kWavesLds = 8;        // LDS hot-x variant: 512-thread workgroups, one per CU (the
                                    // smaller static LDS leaves room for a larger table)
constexpr int kStepGroup = 8;       // steps per chunk are a multiple of this
constexpr int kSplitCap = 72;       // split-table entries per chunk (groups + sentinel)
constexpr int kSegExtra = 66;       // segment-list slack per chunk (steals + padding piece)
constexpr int kSlotCap = 160;       // split-segment slots per chunk (bound: 64 thieves + 67 owners)
// Queue ranges of the batched SpMM (one per XCD). Its counter array holds the
// live pop counters in [0, kBatchQueues) and, in [kBatchQueues, 2 kBatchQueues),
// their final values in the last completed launch (saved by the fixup kernel).
constexpr int kBatchQueues = 8;
constexpr int kEndFlag = (int)0x80000000u;  // sign bit of a column: last element of a segment
constexpr int kHotFlag = 0x40000000;        // bit 30: column renumbered into the hot-x buffer
constexpr int kColMask = 0x3fffffff;

// Piece kinds (bit set)
constexpr int kHead = 1;  // row started in an earlier chunk
constexpr int kTail = 2;  // row continues in a later chunk
constexpr int kPad = 4;   // padding piece of the last chunk (discarded)

// Split-table codes (row ids are >= 0)
constexpr int kCodeHead = -1;     // partial of the chunk's first row -> carry[2c]
constexpr int kCodeTail = -2;     // partial of the chunk's last row  -> carry[2c+1]
constexpr int kCodeDiscard = -3;  // padding

// s_waitcnt vmcnt(0) with the other counters unconstrained: issued after a
// loop's prologue so the loop is entered with no loads in flight (the
// compiler's wait counts inside the loop then follow the loop's own loads).
constexpr int kWaitVm0 = 0x0f70;

// Global-memory access through pointers whose address space the compiler
// cannot see (integer addresses, pointers read from memory): without the cast
// they become flat accesses, which are also waited for as possible LDS
// accesses.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T *p) {
    return reinterpret_cast<gptr<T>>(reinterpret_cast<uintptr_t>(p));
}
template <typename T>
__device__ __forceinline__ T load_global(uint64_t addr) {
    return *reinterpret_cast<gptr<const T>>(addr);
}

struct ConvertArgs {
    const int *rp;
    const int *ci;
    const void *va;      // const double*
    int M, N, nnz;
    int S;               // steps per full chunk
    int nchunks;
    int val_kind;        // 0: f64, 1: f32, 2: bf16
    int col16;           // 1: 16-bit chunk-relative columns + end mask
    void *ovals;         // layout values
    void *ocols;         // int32 (flagged) or uint16
    uint64_t *end_mask;  // 16-bit layouts: one 64-bit mask per step
    int *col_base;       // 16-bit layouts: per-chunk column base
    int *chunk_rows;     // nchunks + 1
    int *seg;            // segment destinations
    int *seg_count;      // per chunk
    int *split_n;        // per chunk
    int2 *split_tab;     // per chunk kSplitCap x (code, first slot)
    int *fix;            // per chunk (row or -1, first chunk of that row)
    int *status;         // [0] max slots, [1] error flags, [2] col16 overflow, [3] n fixup rows
};

struct SpmvArgs {
    const void *vals;
    const void *cols;
    const uint64_t *end_mask;
    const int *col_base;
    const int *chunk_rows;
    const int *seg;
    const int *split_n;
    const int2 *split_tab;
    const unsigned *empty_bits;
    const void *x;
    const void *xhot;    // dense copy of the hot columns of x (or null)
    void *y;
    double *carry;       // 2 per chunk: head / tail partial sums
    double *carry_abs;   // 2 per chunk: head / tail partial sum|a x| (recompute only)
    unsigned char *rflag;  // per row: 1 = recompute in FP64 (recompute only)
    float kappa;         // recompute threshold on sum|a x| / |sum a x|
    long long layout_elems;
    int nchunks, S, M, has_empty, recompute;
    int n_lds;       // hot columns [0, n_lds) read from an LDS copy (0: no LDS table)
    int lds_blocks;  // grid of the LDS variant
    // LDS variant: [0] = items of the shared chunk queue handed out, [1] =
    // workgroups finished; both 0 between SpMVs (the last workgroup re-arms them)
    int *wg_next;
    // FP64 serial parity (parity_tol > 0): the resident FP64 CSR, the bound
    // terms of each head / tail carry, and the rows recomputed per chunk
    const int *rp;
    const int *ci;  // with ci_hot: hot columns as kHotFlag | hot index (the layout's renumbering)
    const double *va;
    int ci_hot;
    double2 *carry_bnd;  // 2 per chunk: (A, packed Q|N) of the head / tail carry
    int *rc_count;       // [c]: rows of chunk c recomputed in serial order by the last SpMV;
                         // [nchunks + c]: 1 if the row closed by chunk c was
    const int *fix;      // per chunk (row closed by the chunk or -1, its first chunk)
    // per chunk c, what its end needs, read when the chunk starts:
    //   pinfo[2c]     = (row closed by c or -1, that row's first chunk,
    //                    chunk k closing the row c's tail carry belongs to or -1, flags)
    //   pinfo[2c + 1] = (row closed by k, that row's first chunk, speculative chain + 1 of
    //                    the row closed by k, of the row closed by c; 0: none)
    // flags: kPiHeadIntra / kPiTailIntra -- every chunk of the row closed by c /
    // by k runs in one workgroup
    const int4 *pinfo;
    int *close_cnt;      // per chunk: contributing chunks arrived (0 between SpMVs)
    // FP64 parity, plain kernel: serial chains of the longest rows spanning
    // chunks, started at kernel start by the first spec_blocks workgroups
    // (one row per workgroup) instead of by the row's last arriving chunk
    const int *spec_rows;    // [i]: row of speculative chain i
    int *spec_state;         // [i]: kSpecDone | kSpecOk | kSpecNeed; 0 between SpMVs
    double *spec_val;        // [i]: chain i's serial sum
    int n_spec, spec_blocks;
    const int *wg_list;      // LDS variant: the chunks of each workgroup, then the shared queue
    const int *wg_start;     // LDS variant: workgroup w runs wg_list[wg_start[w] .. wg_start[w + 1]), then
                             // takes items of the queue wg_list[wg_start[G] .. nchunks) (G workgroups)
    double parity_tol;
};

// ---------------------------------------------------------- FP64 parity
// A row that the CVR-64 order splits into segments (steals, chunk
// boundaries) is summed as per-segment FMA chains added in element order;
// the serial reference sums it as one FMA chain. Both are within a running
// error bound of the exact sum (Wilkinson's running error analysis: an FMA
// chain's error is at most u * sum_j |s_j| over its computed partial sums
// s_j). For a row cut into segments k = 0..K-1 with per-segment partial sums
// t_j, q_k = sum t_j^2 and lengths n_k, merged as S_k = S_(k-1) + seg_k:
//   CVR order:    u * (sum_k sum|t_j| + sum_(k>=1) |S_k|)
//   serial order: u * sum_k sum_j |S_(k-1) + t_j| <= u * sum_k (n_k |S_(k-1)| + sum|t_j|)
// with sum|t_j| <= sqrt(n_k q_k) (Cauchy-Schwarz), so
//   |y_cvr - y_serial| <= u * A,  A = sum_k 2 sqrt(n_k q_k) + sum_(k>=1) (n_k |S_(k-1)| + |S_k|).
// For rows longer than kParRigorousLen the probabilistic bound of Higham and
// Mary (SIAM J. Sci. Comput. 41(5), 2019: with independent mean-zero
// rounding errors, |error| <= lambda u sqrt(sum_j s_j^2) with probability at
// least 1 - 2 exp(-lambda^2 / 2)) is also accepted, with lambda = 10 and
//   Q = sum_k 3 q_k + sum_(k>=1) (2 n_k S_(k-1)^2 + S_k^2)
// covering both orders ((S + t)^2 <= 2 S^2 + 2 t^2). A row whose bound can
// exceed parity_tol * |y| is recomputed as one serial FMA chain, which is
// bit-identical to the reference. Underflow is covered by an absolute term of
// 2^-1074 per element.
constexpr int kPiHeadIntra = 1, kPiTailIntra = 2;  // pinfo flags
// Speculative serial chains (plain kernel): rows spanning chunks with at least
// kSpecLen elements, the kSpecMax longest. State bits: the chain is done (its
// sum stored), the row's bound passed (the CVR-order sum is written), or the
// row needs the serial sum. Each side sets its bit with one atomic OR; the
// side that finds the other's bit writes y where needed and clears the word.
constexpr int kSpecLen = 4096, kSpecMax = 1024;
constexpr int kSpecDone = 1, kSpecOk = 2, kSpecNeed = 4;
constexpr int kParRigorousLen = 4096;
constexpr double kParLambda = 10.0;
constexpr long long kParNMask = 0xFFFFF;  // low 20 bits of a packed Q hold N (N <= 64 * 4096)

__host__ __device__ inline double par_pack(double q, int n) {
    long long b;
    __builtin_memcpy(&b, &q, 8);
    b = (b & ~kParNMask) | (long long)n;
    double r;
    __builtin_memcpy(&r, &b, 8);
    return r;
}
__host__ __device__ inline void par_unpack(double p, double *q, int *n) {
    long long b;
    __builtin_memcpy(&b, &p, 8);
    *n = (int)(b & kParNMask);
    b &= ~kParNMask;
    __builtin_memcpy(q, &b, 8);
}
// true when |y_cvr - y_serial| <= tol * |y_serial| is certain (rigorous
// bound) or, for rows longer than kParRigorousLen, holds under the
// probabilistic bound; false (including NaN) -> recompute the row serially
__host__ __device__ inline bool par_ok(double s, double A, double Q, int N, double tol) {
    const double u = 0x1p-53;
    const double lim = tol * __builtin_fabs(s);
    const double Ar = A + (double)N * 0x1p-1021;  // underflow: 2^-1074 = u * 2^-1021 per element
    if (u * Ar * (1.0 + 2e-6) <= lim) return true;
    if (N > kParRigorousLen && Q > 0x1p-900) {
        const double p = kParLambda * u;
        if (p * p * Q * (1.0 + 4e-6) <= lim * lim) return true;
    }
    return false;
}


// FP64 recompute after a mixed-precision SpMV: one wave per 64 chunks for the
// cross-chunk rows, then one wave per block of kFinishRows rows for the rows
// flagged by the SpMV.
constexpr int kFinishRows = 1024;
__host__ __device__ inline int finish_fix_waves(int nchunks) { return (nchunks + 63) / 64; }
__host__ __device__ inline int finish_row_blocks(int M) { return (M + kFinishRows - 1) / kFinishRows; }
// flag buffer bytes: whole blocks (read 16 flags per lane) plus slack
inline size_t rflag_bytes(int M) { return (size_t)finish_row_blocks(M) * kFinishRows + 64; }
// recompute counters: one per wave of the recompute pass
inline size_t rc_count_len(int M, int nchunks) {
    return (size_t)finish_fix_waves(nchunks) + (size_t)finish_row_blocks(M);
}

// Flagged rows longer than kLongRow are recomputed apart from the others, in
// pieces of kPiece elements (one wave each), the pieces then added in order.
constexpr int kLongRow = 128;
constexpr int kPiece = 512;

struct FinishArgs {
    const int *fix;
    const double *carry;
    const double *carry_abs;
    const int *chunk_rows;
    const int *rp;
    const int *ci;
    const double *va;
    const double *x64;
    unsigned char *rflag;
    float *y;
    int *rc_count;       // rows recomputed by the last SpMV, per wave (rc_count_len entries)
    float kappa;
    int nchunks, M;
};

// One matrix of a (batched) multi-RHS SpMM. 32-bit flagged columns only.
struct SpmmDesc {
    const void *vals;
    const int *cols;
    const int *chunk_rows;
    const int *seg;
    const int *split_n;
    const int2 *split_tab;
    const unsigned *empty_bits;
    const int *fix;
    const int *hot_cols;  // column of each hot index (layouts with hot-x renumbering)
    const int *col_base;          // 16-bit layouts: smallest column of each chunk
    const uint64_t *end_mask;     // 16-bit layouts: segment-end mask of each step
    const float *X;   // numCols x k, row-major
    float *Y;         // numRows x k, row-major
    double *carry;    // 2 * nchunks * k
    long long layout_elems;
    int nchunks, S, M, N, k, has_empty, col16;
};

int spmm_kp(int k);
hipError_t launch_zero_f32(float *y, long long n, hipStream_t st);
size_t spmm_lds_bytes(int max_slots, int k);
// SpMM engines: per-lane FP64 FMA (the default) or the matrix cores
// (v_mfma_f64_4x4x4f64); same results within FP64 rounding.
constexpr int kSpmmFma = 0, kSpmmMfma = 1;
hipError_t launch_spmm(const SpmmDesc &a, int prec, int max_slots, int engine, hipStream_t st);
hipError_t spmm_batch_grid(int prec, int k, size_t lds, int engine, int *grid);
hipError_t launch_spmm_batch(const SpmmDesc *d_descs, const int2 *d_queue, int total, int *d_counter, int prec, int k,
                             size_t lds, int grid, int engine, hipStream_t st);
// The batched SpMM with an LDS window of X (k <= 8): workgroups of 8 waves
// take groups of kWinChunks consecutive chunks of one matrix.
constexpr int kWinChunks = 4;
// For wgs_per_cu workgroups per CU: the window rows that fit beside the
// slot arrays (0 if fewer than kWinMinRows), the dynamic LDS and the grid.
constexpr int kWinMinRows = 256, kWinMaxRows = 8192;
hipError_t spmm_batch_win_plan(int prec, int max_slots, int wgs_per_cu, int *win_rows, size_t *lds, int *grid);
// d_gqueue: (matrix, first chunk) of each group; d_queue: every (matrix,
// chunk), for the fixup that closes rows spanning chunks
hipError_t launch_spmm_batch_win(const SpmmDesc *d_descs, const int2 *d_gqueue, int total_g, const int2 *d_queue,
                                 int total, int *d_counter, int prec, int max_slots, int win_rows, size_t lds, int grid,
                                 hipStream_t st);

hipError_t launch_empty_rows(const int *rp, int M, unsigned *bits, int *status, hipStream_t st);
// *bad |= 1 if the device CSR's offsets are not 0 .. nnz and monotone, |= 2 if a column is outside [0, N)
hipError_t launch_validate_csr(const int *rp, const int *ci, int M, int N, long long nnz, int *bad, hipStream_t st);
// *flag |= 1 if a chunk of C elements of ci spans 65,536 columns or more (flag cleared by the caller)
hipError_t launch_chunk_span(const int *ci, long long nnz, long long C, int nchunks, int *flag, hipStream_t st);
hipError_t launch_convert(const ConvertArgs &a, hipStream_t st);
hipError_t launch_spmv(const SpmvArgs &a, int prec, int col16, hipStream_t st);
// LDS table size for a handle: n_lds (0 = no LDS variant) and its grid.
// max_bytes < 0: as much as the workgroup can hold. bound: the variant that
// carries per-segment bound terms (mixed-precision recompute, FP64 parity).
hipError_t plan_spmv_lds(int prec, int bound, int n_hot, long long max_bytes, int nchunks, int *n_lds,
                         int *lds_blocks);
hipError_t launch_fixup(const int *fix, const double *carry, void *y, int nchunks, int prec, hipStream_t st);
// len[c] = length of the row chunk c closes, or -1
hipError_t launch_closer_len(const int *fix, const int *rp, int *len, int nchunks, hipStream_t st);
hipError_t launch_finish(const FinishArgs &a, hipStream_t st);

// Long flagged rows (static list of the rows longer than kLongRow).
struct LongArgs {
    const int *rp;
    const int *ci;
    const double *va;
    const double *x64;
    unsigned char *rflag;
    const int *long_rows;  // n_long rows, ascending
    const int *poff;       // n_long + 1: first piece of each long row
    const int *piece_row;  // n_pieces: index into long_rows
    double *partial;       // n_pieces
    float *y;
    int *rc_long;          // n_long: 1 if the row was recomputed by the last SpMV
    int n_long, n_pieces;
};
hipError_t launch_long_rows(const LongArgs &a, hipStream_t st);
// Hot-x staging. min_skew > 0: stage nothing unless the most used eighth of
// the used columns is used at least min_skew times as often as an average
// used column.
hipError_t launch_hot_columns(const int *ci, long long nnz, int N, int max_hot, double min_skew, int *cnt, int *colmap,
                              int *hot_cols, int *cols, long long L, int *n_hot, hipStream_t st);
// The hot renumbering of launch_hot_columns (its colmap, still allocated)
// applied to more column indices: the CSR copy kept for the parity recompute.
hipError_t launch_remap_hot(int *cols, long long n, const int *colmap, hipStream_t st);
hipError_t launch_hot_gather(const void *x, const int *hot_cols, void *xhot, int n_hot, int xsize, hipStream_t st);

}  // namespace cvr
