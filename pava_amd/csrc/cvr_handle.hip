// cvr_handle.hip -- the extern "C" CVR-HIP handle API: device upload, GPU
// CSR->CVR conversion (pava's pre_processing, cvr/cvr.h:583-1060) and the SpMV
// launches (pava's spmv_compute_kernel, cvr/cvr.h:1062-1884).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <new>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>

#include "cvr_kernels.h"
#include "pava_hip.h"
#include "pava_internal.h"


using namespace pava;

struct cvr_hip_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    cvr_hip_opts opts{};
    int M = 0, N = 0, nnz = 0, nchunks = 0, S = 256, prec = 0, col16 = 0;
    int auto_s = 0;  // steps per chunk chosen by the library (auto_steps)
    long long layout_elems = 0, seg_cap = 0;
    int has_empty = 0, n_fix = 0, max_slots = 0;
    void *d_vals = nullptr, *d_cols = nullptr;
    uint64_t *d_end_mask = nullptr;
    int *d_col_base = nullptr;
    int *d_chunk_rows = nullptr, *d_seg = nullptr, *d_seg_count = nullptr, *d_split_n = nullptr;
    int2 *d_split_tab = nullptr;
    int *d_fix = nullptr;
    double *d_carry = nullptr, *d_carry_abs = nullptr;
    double *d_carry_mm = nullptr;  // SpMM carries: 2 * nchunks * 32
    int spmm_engine = 0;           // cvr::kSpmmFma or cvr::kSpmmMfma (cvr_hip_set_spmm_engine)
    // FP64 serial parity (opts.parity_tol > 0): bound terms of the carries
    int par = 0;
    double2 *d_carry_bnd = nullptr;
    int *d_close_cnt = nullptr;    // rows spanning chunks closed inside the SpMV: arrival counters
    int4 *d_pinfo = nullptr;       // per chunk: closing and carry-scope information (SpmvArgs::pinfo)
    int *d_wg_list = nullptr, *d_wg_start = nullptr;  // LDS variant: chunks of each workgroup
    unsigned char *d_rflag = nullptr;
    int *d_rc_count = nullptr;
    // recompute of long rows: the rows longer than kLongRow, cut in pieces
    int n_long = 0, n_pieces = 0;
    int *d_long_rows = nullptr, *d_long_poff = nullptr, *d_piece_row = nullptr, *d_rc_long = nullptr;
    double *d_partial = nullptr;
    unsigned *d_empty = nullptr;
    int *d_status = nullptr;
    // FP64 CSR kept resident (recompute); owned by the handle when set
    int *d_rp = nullptr, *d_ci = nullptr;
    double *d_va = nullptr;
    // staging for the host-pointer API
    void *d_hx = nullptr, *d_hy = nullptr, *d_hx64 = nullptr;
    long long device_bytes = 0;
    double convert_ms = 0.0;
    double hot_setup_ms = 0.0;  // wall time of the hot-column staging
    // hot-x staging
    int *d_hot_cols = nullptr;
    void *d_xhot = nullptr;
    int n_hot = 0;
    int n_lds = 0, lds_blocks = 0;  // hot columns held in LDS by the SpMV, and its grid
    int *d_wg_next = nullptr;       // LDS variant: per-workgroup pop counters + finished count
    // FP64 parity, plain kernel: speculative serial chains of the longest rows spanning chunks
    int n_spec = 0, spec_blocks = 0;
    int ci_hot = 0;  // d_ci carries the hot renumbering (FP64 parity with hot-x staging)
    int *d_spec_rows = nullptr, *d_spec_state = nullptr;
    double *d_spec_val = nullptr;
    std::vector<void *> allocs;
    // one device allocation holding the arrays whose sizes follow from the
    // layout's shape, carved in order by dev_alloc (fewer hipMalloc/hipFree
    // calls per handle); anything beyond it is allocated on its own
    char *arena = nullptr;
    size_t arena_size = 0, arena_used = 0;
};

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess)                                                                           \
            return set_error(PAVA_HIP_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                             __FILE__, __LINE__);                                                       \
    } while (0)

// Every device allocation of the CVR handle code goes through dmalloc/dfree:
// they honour the failure injection of pava_hip_test_fail_alloc and count the
// live allocations (pava_hip_test_live_device_allocs), so tests can check that
// a failed creation releases everything it took.
static std::atomic<long long> g_live_device_allocs{0};
template <typename T>
static hipError_t dmalloc(T **p, size_t bytes) {
    *p = nullptr;
    if (inject_device_failure()) return hipErrorOutOfMemory;
    void *q = nullptr;
    const hipError_t e = hipMalloc(&q, bytes);
    if (e == hipSuccess) {
        *p = (T *)q;
        ++g_live_device_allocs;
    }
    return e;
}
static void dfree(void *p) {
    if (!p) return;
    (void)hipFree(p);
    --g_live_device_allocs;
}

extern "C" long long pava_hip_test_live_device_allocs(void) { return g_live_device_allocs.load(); }

// a failed HIP call: PAVA_HIP_ENOMEM when it ran out of memory, else EDEVICE
static int hip_fail(hipError_t e, const char *what) {
    return set_error(e == hipErrorOutOfMemory ? PAVA_HIP_ENOMEM : PAVA_HIP_EDEVICE, "%s: %s", what,
                     hipGetErrorString(e));
}

constexpr size_t kArenaAlign = 256;
static size_t arena_round(size_t b) { return (std::max<size_t>(b, 1) + kArenaAlign - 1) / kArenaAlign * kArenaAlign; }

static int dev_alloc(cvr_hip_handle *h, void **p, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (h->arena && h->arena_used + arena_round(bytes) <= h->arena_size) {
        *p = h->arena + h->arena_used;
        h->arena_used += arena_round(bytes);
        h->device_bytes += (long long)bytes;
        return 0;
    }
    hipError_t e = dmalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(PAVA_HIP_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    h->allocs.push_back(*p);
    h->device_bytes += (long long)bytes;
    return 0;
}

// 32-bit CVR layouts store a column with flags in bits 31 (segment end) and 30
// (staged hot column): at most 2^30 columns.
constexpr long long kCol32Limit = 1LL << 30;
static int check_col_limit(const cvr_hip_opts *opts, long long numCols) {
    if (opts && opts->col_index_bits == 32 && numCols > kCol32Limit)
        return set_error(PAVA_HIP_ERANGE, "32-bit CVR column layouts take at most 2^30 columns, got %lld", numCols);
    return 0;
}

extern "C" void cvr_hip_default_opts(cvr_hip_opts *o) {
    memset(o, 0, sizeof *o);
    o->steps_per_chunk = 0;  // automatic
    o->precision = CVR_HIP_PREC_F64;
    o->col_index_bits = 0;  // automatic
    o->recompute = 0;
    o->kappa_threshold = 64.0;
    o->keep_fp64 = 0;
    o->device = 0;
    o->x_hot_kb = -1;  // automatic
    o->x_lds_kb = -1;
    o->parity_tol = 1e-12;
}

static int value_size(int prec) { return prec == CVR_HIP_PREC_F64 ? 8 : prec == CVR_HIP_PREC_F32 ? 4 : 2; }

// Automatic hot-x staging (x_hot_kb = -1): a copy of at most 2 MiB (half an
// XCD's L2), made only when the most used eighth of the used columns is used
// at least twice as often as an average used column (power-law column use;
// not stencils or FEM meshes).
constexpr int kAutoRounds = 4;  // automatic chunk size: chunks per resident wave slot
// ... and with FP64 parity on a 16-bit layout: each chunk end then also counts
// in and may close the rows spanning chunks, and fewer, longer chunks pay
// that less often (27-point stencil 128^3: 0.140 ms at four rounds, S = 56;
// 0.132 ms at two, S = 112)
constexpr int kAutoRoundsParity16 = 2;
static int auto_steps(long long nnz, int device, int rounds = kAutoRounds);
constexpr int kAutoHotKb = 2048;
constexpr double kAutoHotSkew = 2.0;

extern "C" int cvr_hip_destroy(cvr_hip_handle *h) {
    if (!h) return 0;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void *p : h->allocs) dfree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

static int validate_opts(const cvr_hip_opts *o) {
    if (o->steps_per_chunk != 0 &&
        (o->steps_per_chunk < 8 || o->steps_per_chunk > 4096 || o->steps_per_chunk % cvr::kStepGroup))
        return set_error(PAVA_HIP_EINVAL, "steps_per_chunk must be a multiple of 8 in [8, 4096], got %d",
                         o->steps_per_chunk);
    if (o->precision < 0 || o->precision > 2) return set_error(PAVA_HIP_EINVAL, "bad precision %d", o->precision);
    if (o->col_index_bits != 0 && o->col_index_bits != 32 && o->col_index_bits != 16)
        return set_error(PAVA_HIP_EINVAL, "col_index_bits must be 0 (automatic), 16 or 32, got %d", o->col_index_bits);
    if (o->x_hot_kb < -1) return set_error(PAVA_HIP_EINVAL, "x_hot_kb must be >= -1, got %d", o->x_hot_kb);
    if (o->x_lds_kb < -1) return set_error(PAVA_HIP_EINVAL, "x_lds_kb must be >= -1, got %d", o->x_lds_kb);
    if (!(o->parity_tol >= 0.0)) return set_error(PAVA_HIP_EINVAL, "parity_tol must be >= 0, got %g", o->parity_tol);
    return 0;
}

static int plan_parity(cvr_hip_handle *h);
static int upload_lds_lists(cvr_hip_handle *h, const std::vector<std::vector<int>> &lst);

// Core: conversion from a CSR in device memory.
static int create_core_body(int M, int N, int nnz, const int *d_rp, const int *d_ci, const double *d_va,
                            const cvr_hip_opts *o, cvr_hip_handle *h);
// host containers may throw std::bad_alloc: it becomes PAVA_HIP_ENOMEM, and the
// caller destroys the partly built handle as for any other failure
static int create_core(int M, int N, int nnz, const int *d_rp, const int *d_ci, const double *d_va,
                       const cvr_hip_opts *o, cvr_hip_handle *h) {
    try {
        return create_core_body(M, N, nnz, d_rp, d_ci, d_va, o, h);
    } catch (const std::bad_alloc &) {
        return set_error(PAVA_HIP_ENOMEM, "host allocation failed while building the CVR layout");
    }
}

// The arena: the sizes of the layout arrays, the per-chunk tables, the
// parity tables and the largest hot-x staging the options allow, each
// rounded up as dev_alloc carves them.
static int reserve_arena(cvr_hip_handle *h, const cvr_hip_opts *o) {
    const size_t L = (size_t)h->layout_elems, nc = (size_t)h->nchunks, M = (size_t)h->M, N = (size_t)h->N;
    auto r = arena_round;
    size_t t = r(L * value_size(h->prec)) + r(L * (h->col16 ? 2 : 4));
    if (h->col16) t += r(L / 64 * 8) + r(nc * 4);
    t += r((nc + 1) * 4) + r((size_t)h->seg_cap * 4) + 2 * r(nc * 4) + r(nc * cvr::kSplitCap * 8) + r(nc * 8) +
         r(nc * 16 + 16);
    if (o->recompute && h->prec != CVR_HIP_PREC_F64)
        t += r(nc * 16 + 16) + r(cvr::rflag_bytes((int)M)) + r(cvr::rc_count_len((int)M, (int)nc) * 4 + 4);
    if (h->par)
        t += r(nc * 2 * sizeof(double2) + 32) + r(nc * 8 + 4) + r(nc * 4) + r(nc * 2 * sizeof(int4)) + r(nc * 4) +
             r((size_t)4096 * 4);
    t += r((M + 31) / 32 * 4) + r(64);
    const long long max_hot = (long long)(o->x_hot_kb < 0 ? kAutoHotKb : o->x_hot_kb) * 1024 /
                              (h->prec == CVR_HIP_PREC_F64 ? 8 : 4);
    if (max_hot > 0 && !h->col16) {
        const size_t nh = (size_t)std::min<long long>(max_hot, (long long)N);
        t += r(nh * 4 + 4) + r(nh * (h->prec == CVR_HIP_PREC_F64 ? 8 : 4));
    }
    void *p = nullptr;
    if (dmalloc(&p, t) != hipSuccess)
        return set_error(PAVA_HIP_ENOMEM, "hipMalloc(%zu) of the handle's arrays failed", t);
    h->allocs.push_back(p);
    h->arena = static_cast<char *>(p);
    h->arena_size = t;
    h->arena_used = 0;
    return 0;
}

static int create_core_body(int M, int N, int nnz, const int *d_rp, const int *d_ci, const double *d_va,
                            const cvr_hip_opts *o, cvr_hip_handle *h) {
    h->opts = *o;
    h->M = M;
    h->N = N;
    h->nnz = nnz;
    h->S = o->steps_per_chunk;
    h->prec = o->precision;
    long long C = 64LL * h->S;
    h->nchunks = nnz > 0 ? (int)((nnz + C - 1) / C) : 0;
    h->col16 = (o->col_index_bits == 16);
    h->par = (h->prec == CVR_HIP_PREC_F64 && o->parity_tol > 0.0 && h->d_rp && nnz > 0) ? 1 : 0;
    // every chunk of Cs elements spans fewer than 65,536 columns?
    auto spans_fit = [&](long long Cs, int ncs, int *fit) -> int {
        if (N <= 65536) {
            *fit = 1;
            return 0;
        }
        int *d_flag = nullptr, flag = 1;
        hipError_t e = dmalloc(&d_flag, 4);
        if (e == hipSuccess) e = hipMemsetAsync(d_flag, 0, 4, h->stream);
        if (e == hipSuccess) e = cvr::launch_chunk_span(d_ci, nnz, Cs, ncs, d_flag, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&flag, d_flag, 4, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        dfree(d_flag);
        if (e != hipSuccess) return hip_fail(e, "column span check");
        *fit = flag == 0;
        return 0;
    };
    // FP64 parity on a 16-bit layout: the longer automatic chunks, if 16-bit
    // columns still fit them
    auto try_parity_size = [&](int *fit) -> int {
        *fit = 0;
        if (!(h->par && h->auto_s && h->nchunks > 0)) return 0;
        const int S2 = auto_steps(nnz, h->device, kAutoRoundsParity16);
        if (S2 == h->S) return 0;
        const long long C2 = 64LL * S2;
        const int nc2 = (int)((nnz + C2 - 1) / C2);
        int rc0 = spans_fit(C2, nc2, fit);
        if (rc0) return rc0;
        if (*fit) {
            h->S = S2;
            h->opts.steps_per_chunk = S2;
            h->nchunks = nc2;
        }
        return 0;
    };
    int rc0;
    if (o->col_index_bits == 0 && h->nchunks > 0 && o->x_hot_kb <= 0) {
        // automatic: 16-bit chunk-relative columns when every chunk's nonzeros
        // span fewer than 65,536 columns (an explicit hot-x request keeps 32)
        int fit = 0;
        if ((rc0 = try_parity_size(&fit))) return rc0;
        if (!fit && (rc0 = spans_fit(64LL * h->S, h->nchunks, &fit))) return rc0;
        h->col16 = fit;
    } else if (h->col16) {
        int fit = 0;
        if ((rc0 = try_parity_size(&fit))) return rc0;
    }
    C = 64LL * h->S;
    if (!h->col16 && h->nchunks > 0 && (long long)N > kCol32Limit)
        return set_error(PAVA_HIP_ERANGE, "a chunk spans 65,536 columns or more and there are more than 2^30 columns: "
                                          "no CVR column layout fits");
    if (h->nchunks > 0) {
        long long nlast = nnz - (long long)(h->nchunks - 1) * C;
        long long steps_last = (nlast + 63) / 64;
        steps_last = (steps_last + cvr::kStepGroup - 1) / cvr::kStepGroup * cvr::kStepGroup;
        h->layout_elems = (long long)(h->nchunks - 1) * C + 64 * steps_last;
    }
    // layout positions (padding included) and segment-list offsets are int32
    if (h->layout_elems > PAVA_HIP_MAX_SIZE)
        return set_error(PAVA_HIP_ERANGE, "layout of %lld elements beyond %d", h->layout_elems, PAVA_HIP_MAX_SIZE);
    h->seg_cap = (long long)M + (long long)cvr::kSegExtra * h->nchunks + 512;
    if (h->seg_cap > INT32_MAX)
        return set_error(PAVA_HIP_ERANGE, "segment list of %lld entries beyond int32 (use more steps per chunk)",
                         h->seg_cap);
    int rc;
    const int vs = value_size(h->prec);
    if ((rc = reserve_arena(h, o))) return rc;
    if ((rc = dev_alloc(h, &h->d_vals, (size_t)h->layout_elems * vs))) return rc;
    if ((rc = dev_alloc(h, &h->d_cols, (size_t)h->layout_elems * (h->col16 ? 2 : 4)))) return rc;
    if (h->col16) {
        if ((rc = dev_alloc(h, (void **)&h->d_end_mask, (size_t)(h->layout_elems / 64) * 8))) return rc;
        if ((rc = dev_alloc(h, (void **)&h->d_col_base, (size_t)h->nchunks * 4))) return rc;
    }
    if ((rc = dev_alloc(h, (void **)&h->d_chunk_rows, (size_t)(h->nchunks + 1) * 4))) return rc;
    if ((rc = dev_alloc(h, (void **)&h->d_seg, (size_t)h->seg_cap * 4))) return rc;
    if ((rc = dev_alloc(h, (void **)&h->d_seg_count, (size_t)h->nchunks * 4))) return rc;
    if ((rc = dev_alloc(h, (void **)&h->d_split_n, (size_t)h->nchunks * 4))) return rc;
    if ((rc = dev_alloc(h, (void **)&h->d_split_tab, (size_t)h->nchunks * cvr::kSplitCap * 8))) return rc;
    if ((rc = dev_alloc(h, (void **)&h->d_fix, (size_t)h->nchunks * 8))) return rc;
    if ((rc = dev_alloc(h, (void **)&h->d_carry, (size_t)h->nchunks * 16 + 16))) return rc;
    if (o->recompute && h->prec != CVR_HIP_PREC_F64) {
        if ((rc = dev_alloc(h, (void **)&h->d_carry_abs, (size_t)h->nchunks * 16 + 16))) return rc;
        if ((rc = dev_alloc(h, (void **)&h->d_rflag, cvr::rflag_bytes(M)))) return rc;
        if ((rc = dev_alloc(h, (void **)&h->d_rc_count, cvr::rc_count_len(M, h->nchunks) * 4 + 4))) return rc;
    }
    if (h->par) {
        if ((rc = dev_alloc(h, (void **)&h->d_carry_bnd, (size_t)h->nchunks * 2 * sizeof(double2) + 32))) return rc;
        if ((rc = dev_alloc(h, (void **)&h->d_rc_count, (size_t)h->nchunks * 8 + 4)))
            return rc;
        HIP_TRY(hipMemsetAsync(h->d_carry_bnd, 0, (size_t)h->nchunks * 2 * sizeof(double2) + 32, h->stream));
        HIP_TRY(hipMemsetAsync(h->d_rc_count, 0, (size_t)h->nchunks * 8 + 4,
                               h->stream));
        if ((rc = dev_alloc(h, (void **)&h->d_close_cnt, (size_t)h->nchunks * 4))) return rc;
        HIP_TRY(hipMemsetAsync(h->d_close_cnt, 0, (size_t)h->nchunks * 4, h->stream));
    }
    if ((rc = dev_alloc(h, (void **)&h->d_empty, (size_t)((M + 31) / 32) * 4))) return rc;
    if ((rc = dev_alloc(h, (void **)&h->d_status, 64))) return rc;

    HIP_TRY(hipMemsetAsync(h->d_status, 0, 64, h->stream));
    HIP_TRY(hipMemsetAsync(h->d_chunk_rows, 0, (size_t)(h->nchunks + 1) * 4, h->stream));
    // the segment list beyond the written entries is read by the SpMV window
    // (values unused); clear it once so it is defined
    HIP_TRY(hipMemsetAsync(h->d_seg, 0, (size_t)h->seg_cap * 4, h->stream));
    HIP_TRY(hipMemsetAsync(h->d_carry, 0, (size_t)h->nchunks * 16 + 16, h->stream));
    if (h->d_rflag) {
        HIP_TRY(hipMemsetAsync(h->d_rflag, 0, cvr::rflag_bytes(M), h->stream));
        HIP_TRY(hipMemsetAsync(h->d_carry_abs, 0, (size_t)h->nchunks * 16 + 16, h->stream));
        HIP_TRY(hipMemsetAsync(h->d_rc_count, 0, cvr::rc_count_len(M, h->nchunks) * 4 + 4, h->stream));
    }

    if (h->d_rflag) {
        // the rows longer than kLongRow and their pieces (static per matrix)
        std::vector<int> rp((size_t)M + 1);
        HIP_TRY(hipMemcpy(rp.data(), d_rp, ((size_t)M + 1) * 4, hipMemcpyDeviceToHost));
        std::vector<int> rows, poff, prow;
        for (int r = 0; r < M; ++r) {
            const int len = rp[(size_t)r + 1] - rp[(size_t)r];
            if (len <= cvr::kLongRow) continue;
            poff.push_back((int)prow.size());
            for (int k = 0; k < (len + cvr::kPiece - 1) / cvr::kPiece; ++k) prow.push_back((int)rows.size());
            rows.push_back(r);
        }
        poff.push_back((int)prow.size());
        h->n_long = (int)rows.size();
        h->n_pieces = (int)prow.size();
        if (h->n_long > 0) {
            if ((rc = dev_alloc(h, (void **)&h->d_long_rows, rows.size() * 4)) ||
                (rc = dev_alloc(h, (void **)&h->d_long_poff, poff.size() * 4)) ||
                (rc = dev_alloc(h, (void **)&h->d_piece_row, prow.size() * 4)) ||
                (rc = dev_alloc(h, (void **)&h->d_partial, prow.size() * 8)) ||
                (rc = dev_alloc(h, (void **)&h->d_rc_long, rows.size() * 4)))
                return rc;
            HIP_TRY(hipMemcpy(h->d_long_rows, rows.data(), rows.size() * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(h->d_long_poff, poff.data(), poff.size() * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(h->d_piece_row, prow.data(), prow.size() * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemsetAsync(h->d_rc_long, 0, rows.size() * 4, h->stream));
        }
    }

    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, h->stream));
    HIP_TRY(cvr::launch_empty_rows(d_rp, M, h->d_empty, h->d_status, h->stream));
    cvr::ConvertArgs a{};
    a.rp = d_rp;
    a.ci = d_ci;
    a.va = d_va;
    a.M = M;
    a.N = N;
    a.nnz = nnz;
    a.S = h->S;
    a.nchunks = h->nchunks;
    a.val_kind = h->prec;
    a.col16 = h->col16;
    a.ovals = h->d_vals;
    a.ocols = h->d_cols;
    a.end_mask = h->d_end_mask;
    a.col_base = h->d_col_base;
    a.chunk_rows = h->d_chunk_rows;
    a.seg = h->d_seg;
    a.seg_count = h->d_seg_count;
    a.split_n = h->d_split_n;
    a.split_tab = h->d_split_tab;
    a.fix = h->d_fix;
    a.status = h->d_status;
    HIP_TRY(cvr::launch_convert(a, h->stream));
    HIP_TRY(hipEventRecord(e1, h->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    h->convert_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int status[16];
    HIP_TRY(hipMemcpy(status, h->d_status, sizeof status, hipMemcpyDeviceToHost));
    h->max_slots = status[0];
    if (status[1]) return set_error(PAVA_HIP_EDEVICE, "CVR conversion overflowed its per-chunk tables (flags %d)", status[1]);
    if (status[2])
        return set_error(PAVA_HIP_ERANGE, "a chunk spans 65536 columns or more: 16-bit column layout not possible");
    h->n_fix = status[3];
    h->has_empty = status[4] ? 1 : 0;

    // hot-x staging: renumber the most used columns into a dense copy of x
    const int xsize = h->prec == CVR_HIP_PREC_F64 ? 8 : 4;
    // x_hot_kb < 0: automatic (kAutoHotKb budget, staged only on skewed column use)
    const bool auto_hot = o->x_hot_kb < 0;
    const long long max_hot = (long long)(auto_hot ? kAutoHotKb : o->x_hot_kb) * 1024 / xsize;
    if (max_hot > 0 && !h->col16 && nnz > 0 && N < (1 << 30)) {
        const auto t0 = std::chrono::steady_clock::now();
        int *cnt = nullptr, *colmap = nullptr;
        hipError_t e = hipSuccess;
        int nh = 0;
        // the column counts and the column map: one temporary allocation
        const size_t half = arena_round((size_t)N * 4 + 4);
        if (dmalloc(&cnt, 2 * half) == hipSuccess) colmap = reinterpret_cast<int *>(reinterpret_cast<char *>(cnt) + half);
        if (!cnt || dev_alloc(h, (void **)&h->d_hot_cols, (size_t)std::min<long long>(max_hot, N) * 4 + 4) != 0) {
            e = hipErrorOutOfMemory;
        } else {
            e = cvr::launch_hot_columns(d_ci, nnz, N, (int)std::min<long long>(max_hot, N), auto_hot ? kAutoHotSkew : 0.0,
                                        cnt, colmap, h->d_hot_cols, (int *)h->d_cols, h->layout_elems, &nh, h->stream);
            // FP64 parity: the serial recompute reads the handle's own copy of
            // the CSR; its columns get the same renumbering, so that its
            // gathers of hot columns read the LDS table or the hot-x copy, as
            // the SpMV's do (R-MAT scale 22: 1.9M recomputed elements, ~85% on
            // hot columns)
            if (e == hipSuccess && nh > 0 && h->par && h->d_ci) {
                e = cvr::launch_remap_hot(h->d_ci, nnz, colmap, h->stream);
                if (e == hipSuccess) h->ci_hot = 1;
            }
        }
        (void)hipStreamSynchronize(h->stream);
        dfree(cnt);
        if (e != hipSuccess) return hip_fail(e, "hot-column staging");
        h->n_hot = nh;
        if (nh > 0) {
            int rc2 = dev_alloc(h, &h->d_xhot, (size_t)nh * xsize);
            if (rc2) return rc2;
            const long long lds_bytes = o->x_lds_kb < 0 ? -1 : (long long)o->x_lds_kb * 1024;
            HIP_TRY(cvr::plan_spmv_lds(h->prec, h->d_rflag != nullptr || h->par, nh, lds_bytes, h->nchunks, &h->n_lds,
                                       &h->lds_blocks));
            if (h->n_lds > 0) {
                rc2 = dev_alloc(h, (void **)&h->d_wg_next, 2 * sizeof(int));
                if (rc2) return rc2;
                HIP_TRY(hipMemsetAsync(h->d_wg_next, 0, 2 * sizeof(int), h->stream));
                HIP_TRY(hipStreamSynchronize(h->stream));
            }
        }
        h->hot_setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (h->par) {
        if ((rc = plan_parity(h))) return rc;
    } else if (h->n_lds > 0) {
        // contiguous ranges of nchunks / G chunks
        const int G = h->lds_blocks, nc = h->nchunks;
        std::vector<std::vector<int>> lst((size_t)G);
        for (int g = 0; g < G; ++g)
            for (int c = (int)((long long)nc * g / G); c < (int)((long long)nc * (g + 1) / G); ++c)
                lst[(size_t)g].push_back(c);
        if ((rc = upload_lds_lists(h, lst))) return rc;
    }
    return 0;
}

// The LDS variant's chunk lists: workgroup w keeps the first half of its list
// lst[w] (wg_list[wg_start[w] .. wg_start[w + 1])); the second halves of all
// lists, interleaved (the first tail chunk of every workgroup, then the
// second, ...), form the queue wg_list[wg_start[G] .. nchunks) that every
// workgroup pops from once its own part is done. Shared fraction measured on
// R-MAT scale 22, FP64 with parity: 1/8 0.438 ms, 1/4 0.434, 1/2 0.431 (only
// private ranges: 0.446); without parity 0.380-0.382 (0.392).
constexpr int kShareDiv = 2;
static int upload_lds_lists(cvr_hip_handle *h, const std::vector<std::vector<int>> &lst) {
    const int G = (int)lst.size(), nc = h->nchunks;
    std::vector<int> list, start((size_t)G + 1, 0);
    std::vector<size_t> keep((size_t)G);
    list.reserve((size_t)nc);
    size_t max_tail = 0;
    for (int g = 0; g < G; ++g) {
        const size_t n = lst[(size_t)g].size();
        keep[(size_t)g] = n - n / kShareDiv;
        max_tail = std::max(max_tail, n - keep[(size_t)g]);
        start[(size_t)g] = (int)list.size();
        list.insert(list.end(), lst[(size_t)g].begin(), lst[(size_t)g].begin() + (long)keep[(size_t)g]);
    }
    start[(size_t)G] = (int)list.size();
    for (size_t j = 0; j < max_tail; ++j)
        for (int g = 0; g < G; ++g)
            if (keep[(size_t)g] + j < lst[(size_t)g].size()) list.push_back(lst[(size_t)g][keep[(size_t)g] + j]);
    if ((int)list.size() != nc) return set_error(PAVA_HIP_EDEVICE, "chunk lists hold %zu of %d chunks", list.size(), nc);
    int rc;
    if ((rc = dev_alloc(h, (void **)&h->d_wg_list, (size_t)nc * 4))) return rc;
    if ((rc = dev_alloc(h, (void **)&h->d_wg_start, ((size_t)G + 1) * 4))) return rc;
    HIP_TRY(hipMemcpy(h->d_wg_list, list.data(), (size_t)nc * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->d_wg_start, start.data(), ((size_t)G + 1) * 4, hipMemcpyHostToDevice));
    return 0;
}

// FP64 parity: which workgroup runs each chunk, and for each row spanning
// chunks whether all its chunks run in one workgroup (its carries then pass
// through ordinary memory with workgroup-scope ordering, else through
// agent-scope atomics). The plain kernel runs chunks 4g..4g+3 in one
// workgroup. The LDS variant's workgroup w starts from a list: first the
// chunks of the long rows (longer than kSpreadLen) assigned to it -- long rows
// are dealt to the least loaded workgroups, longest first, so that their
// serial recomputes, when the bound requires them, start early and on
// different CUs -- then a contiguous run of the other chunks, filled up to its
// share nchunks/G. A chunk holding the end of one long row and the start of
// another goes with the first; the second then spans two workgroups. The last
// half of each list goes to a queue shared by all workgroups
// (upload_lds_lists), so a chunk may run anywhere: in the LDS variant every
// row spanning chunks passes its carries through agent-scope atomics.
constexpr int kSpreadLen = 2048;
static int plan_parity(cvr_hip_handle *h) {
    const int nc = h->nchunks;
    std::vector<int> fix((size_t)nc * 2), len((size_t)nc);
    int *d_len = nullptr;
    if (dmalloc(&d_len, (size_t)nc * 4) != hipSuccess) return set_error(PAVA_HIP_ENOMEM, "parity plan: hipMalloc failed");
    hipError_t e = cvr::launch_closer_len(h->d_fix, h->d_rp, d_len, nc, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(len.data(), d_len, (size_t)nc * 4, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(fix.data(), h->d_fix, (size_t)nc * 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    dfree(d_len);
    if (e != hipSuccess) return hip_fail(e, "parity plan");
    std::vector<int> wg_of((size_t)nc);
    int rc;
    if (h->n_lds > 0) {
        const int G = h->lds_blocks;
        std::vector<int> owner((size_t)nc, -1);
        std::vector<std::vector<int>> lst((size_t)G);
        std::vector<long long> load((size_t)G, 0);
        std::vector<int> longk;
        for (int k = 0; k < nc; ++k)
            if (fix[2 * (size_t)k] >= 0 && len[(size_t)k] > kSpreadLen) longk.push_back(k);
        std::stable_sort(longk.begin(), longk.end(), [&](int p, int q) { return len[(size_t)p] > len[(size_t)q]; });
        std::vector<int> row_wg(longk.size());
        for (size_t i = 0; i < longk.size(); ++i) {  // chunks c0+1 .. k belong to the long row closed by k
            const int k = longk[i];
            const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            row_wg[i] = w;
            for (int c = fix[2 * (size_t)k + 1] + 1; c <= k; ++c) {
                owner[(size_t)c] = w;
                lst[(size_t)w].push_back(c);
            }
            load[(size_t)w] += k - fix[2 * (size_t)k + 1];
        }
        // the first chunk too, unless it closes another long row (then it runs
        // early in that row's workgroup)
        for (size_t i = 0; i < longk.size(); ++i) {
            const int c0 = fix[2 * (size_t)longk[i] + 1], w = row_wg[i];
            if (owner[(size_t)c0] >= 0) continue;
            owner[(size_t)c0] = w;
            lst[(size_t)w].insert(lst[(size_t)w].begin(), c0);
            ++load[(size_t)w];
        }
        for (int c = 0; c < nc; ++c)
            if (owner[(size_t)c] >= 0) wg_of[(size_t)c] = owner[(size_t)c];
        int w = 0;
        long long filled = load[0];
        for (int c = 0; c < nc; ++c) {
            if (owner[(size_t)c] >= 0) continue;
            while (w < G - 1 && filled >= (long long)nc * (w + 1) / G - (long long)nc * w / G) {
                ++w;
                filled = load[(size_t)w];
            }
            lst[(size_t)w].push_back(c);
            wg_of[(size_t)c] = w;
            ++filled;
        }
        if ((rc = upload_lds_lists(h, lst))) return rc;
    } else {
        for (int c = 0; c < nc; ++c) wg_of[(size_t)c] = c / cvr::kWavesPerBlock;
    }
    // per chunk: the row it closes, the chunk closing the row its tail carry
    // belongs to, and whether each of those rows runs in one workgroup
    std::vector<int4> pinfo((size_t)nc * 2);
    for (int c = 0; c < nc; ++c) {
        pinfo[2 * (size_t)c] = make_int4(fix[2 * (size_t)c], fix[2 * (size_t)c + 1], -1, 0);
        pinfo[2 * (size_t)c + 1] = make_int4(-1, 0, 0, 0);
    }
    for (int k = 0; k < nc; ++k) {
        const int row = fix[2 * (size_t)k], c0 = fix[2 * (size_t)k + 1];
        if (row < 0) continue;
        bool same = h->n_lds == 0;  // LDS variant: chunks move between workgroups
        for (int c = c0; c < k && same; ++c) same = wg_of[(size_t)c] == wg_of[(size_t)k];
        if (same) pinfo[2 * (size_t)k].w |= cvr::kPiHeadIntra;
        for (int c = c0; c < k; ++c) {
            int4 &p = pinfo[2 * (size_t)c];
            p.z = k;
            if (same) p.w |= cvr::kPiTailIntra;
            pinfo[2 * (size_t)c + 1] = make_int4(row, c0, 0, 0);
        }
    }
    // Plain kernel: the kSpecMax longest rows spanning chunks with at least
    // kSpecLen elements get a speculative serial chain, started at kernel
    // start by a wave of their own instead of after the row's last chunk
    // arrives (the LDS variant runs the chunks of long rows first instead).
    // Such a chain stops once the row's bound has passed; otherwise its sum
    // is the row's y, as the closer's own chain would have been.
    if (h->n_lds == 0) {
        std::vector<int> spec_k;
        for (int k = 0; k < nc; ++k)
            if (fix[2 * (size_t)k] >= 0 && len[(size_t)k] >= cvr::kSpecLen) spec_k.push_back(k);
        std::stable_sort(spec_k.begin(), spec_k.end(), [&](int a, int b) { return len[(size_t)a] > len[(size_t)b]; });
        if ((int)spec_k.size() > cvr::kSpecMax) spec_k.resize(cvr::kSpecMax);
        std::vector<int> rows(spec_k.size());
        for (size_t i = 0; i < spec_k.size(); ++i) {
            const int k = spec_k[i], c0 = fix[2 * (size_t)k + 1];
            rows[i] = fix[2 * (size_t)k];
            pinfo[2 * (size_t)k + 1].w = (int)i + 1;  // the row chunk k closes
            for (int c = c0; c < k; ++c) pinfo[2 * (size_t)c + 1].z = (int)i + 1;  // their tail carries' row
        }
        h->n_spec = (int)spec_k.size();
        if (h->n_spec > 0) {
            h->spec_blocks = (h->n_spec + 7) / 8 * 8;  // whole XCD rounds: the chunk blocks keep their XCDs
            if ((rc = dev_alloc(h, (void **)&h->d_spec_rows, (size_t)h->n_spec * 4))) return rc;
            if ((rc = dev_alloc(h, (void **)&h->d_spec_state, (size_t)h->n_spec * 4))) return rc;
            if ((rc = dev_alloc(h, (void **)&h->d_spec_val, (size_t)h->n_spec * 8))) return rc;
            HIP_TRY(hipMemcpy(h->d_spec_rows, rows.data(), (size_t)h->n_spec * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemset(h->d_spec_state, 0, (size_t)h->n_spec * 4));
        }
    }
    if ((rc = dev_alloc(h, (void **)&h->d_pinfo, (size_t)nc * 2 * sizeof(int4)))) return rc;
    HIP_TRY(hipMemcpy(h->d_pinfo, pinfo.data(), (size_t)nc * 2 * sizeof(int4), hipMemcpyHostToDevice));
    return 0;
}

// Automatic chunk size: about four chunks per resident wave slot of the
// device (16 waves per CU), at most 256 steps. One chunk per wave runs start
// to end, so the last round leaves a tail as long as a chunk; four rounds
// keep it short while the per-chunk work (split rows, the row spanning chunks)
// stays small. Measured on the 27-point stencil 128^3: 0.149 ms at two rounds
// (112 steps), 0.140 ms at four (56); R-MAT scale 22: 0.395 -> 0.389 ms.
static int auto_steps(long long nnz, int device, int rounds) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu <= 0)
        ncu = 256;
    const long long per = 64LL * rounds * 16 * ncu;  // elements per chunk-step over the rounds
    long long s = (nnz + per - 1) / per;
    s = (s + cvr::kStepGroup - 1) / cvr::kStepGroup * cvr::kStepGroup;
    return (int)std::min<long long>(std::max<long long>(s, cvr::kStepGroup), 256);
}

static int begin_handle(const cvr_hip_opts *opts, long long nnz, cvr_hip_opts *o, cvr_hip_handle **out,
                        cvr_hip_handle **hp) {
    if (!out) return set_error(PAVA_HIP_EINVAL, "null output handle");
    *out = nullptr;
    if (opts)
        *o = *opts;
    else
        cvr_hip_default_opts(o);
    if (o->kappa_threshold <= 0) o->kappa_threshold = 64.0;
    int rc = validate_opts(o);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(PAVA_HIP_EDEVICE, "no HIP device");
    if (o->device < 0 || o->device >= ndev) return set_error(PAVA_HIP_EINVAL, "device %d out of range", o->device);
    HIP_TRY(hipSetDevice(o->device));
    const int auto_s = o->steps_per_chunk == 0;
    if (auto_s) o->steps_per_chunk = auto_steps(nnz, o->device);
    cvr_hip_handle *h = new (std::nothrow) cvr_hip_handle();
    if (!h) return set_error(PAVA_HIP_ENOMEM, "host allocation of the handle failed");
    h->device = o->device;
    h->auto_s = auto_s;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return set_error(PAVA_HIP_EDEVICE, "hipStreamCreate failed");
    }
    *hp = h;
    return 0;
}

// Upload of a caller's pageable host array: through two pinned staging
// pieces, each filled by the host's threads while the previous piece's DMA
// runs (a pageable hipMemcpy moves ~27 GB/s). The pinned memory is allocated
// once per process and shared under a mutex; small arrays, or a failed
// pinned allocation, take the plain hipMemcpy.
constexpr size_t kStagePiece = size_t(32) << 20;
static std::mutex g_stage_mu;
static char *g_stage = nullptr;  // 2 * kStagePiece pinned bytes

static hipError_t upload_host(void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes < 2 * kStagePiece) return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    std::lock_guard<std::mutex> lk(g_stage_mu);
    if (!g_stage && hipHostMalloc((void **)&g_stage, 2 * kStagePiece, hipHostMallocPortable) != hipSuccess) {
        g_stage = nullptr;
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipSuccess;
    for (hipEvent_t &x : ev)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
    bool used[2] = {false, false};
    for (size_t off = 0, i = 0; e == hipSuccess && off < bytes; off += kStagePiece, ++i) {
        const int b = (int)(i & 1);
        const size_t n = std::min(kStagePiece, bytes - off);
        if (used[b] && (e = hipEventSynchronize(ev[b])) != hipSuccess) break;
        char *buf = g_stage + (size_t)b * kStagePiece;
        pava::parallel_memcpy(buf, static_cast<const char *>(src) + off, n);
        if ((e = hipMemcpyAsync(static_cast<char *>(dst) + off, buf, n, hipMemcpyHostToDevice, st)) != hipSuccess)
            break;
        if ((e = hipEventRecord(ev[b], st)) != hipSuccess) break;
        used[b] = true;
    }
    const hipError_t e2 = hipStreamSynchronize(st);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    return e != hipSuccess ? e : e2;
}

extern "C" int cvr_hip_create(const pava_csr_matrix *A, const cvr_hip_opts *opts, cvr_hip_handle **out) {
    int rc = 0;
    if (A && (rc = check_gpu_sizes(A->numRows, A->numCols, A->nnz))) return rc;
    if (A && (rc = check_col_limit(opts, A->numCols))) return rc;
    if ((rc = pava_csr_validate(A))) return rc;
    cvr_hip_opts o;
    cvr_hip_handle *h = nullptr;
    if ((rc = begin_handle(opts, A->nnz, &o, out, &h))) return rc;
    const int M = A->numRows, nnz = A->nnz;
    int *d_rp = nullptr, *d_ci = nullptr;
    double *d_va = nullptr;
    const bool keep = o.keep_fp64 || o.recompute || (o.precision == CVR_HIP_PREC_F64 && o.parity_tol > 0.0);
    auto fail = [&](int code) {
        if (!keep) {
            dfree(d_rp);
            dfree(d_ci);
            dfree(d_va);
        }
        cvr_hip_destroy(h);
        return code;
    };
    if (keep) {
        if ((rc = dev_alloc(h, (void **)&d_rp, (size_t)(M + 1) * 4)) || (rc = dev_alloc(h, (void **)&d_ci, (size_t)nnz * 4)) ||
            (rc = dev_alloc(h, (void **)&d_va, (size_t)nnz * 8)))
            return fail(rc);
    } else {
        if (dmalloc(&d_rp, (size_t)(M + 1) * 4 + 4) != hipSuccess || dmalloc(&d_ci, (size_t)nnz * 4 + 4) != hipSuccess ||
            dmalloc(&d_va, (size_t)nnz * 8 + 8) != hipSuccess)
            return fail(set_error(PAVA_HIP_ENOMEM, "hipMalloc of the CSR upload failed"));
    }
    if (upload_host(d_rp, A->rowOffsets, (size_t)(M + 1) * 4, h->stream) != hipSuccess ||
        (nnz && upload_host(d_ci, A->cols, (size_t)nnz * 4, h->stream) != hipSuccess) ||
        (nnz && upload_host(d_va, A->vals, (size_t)nnz * 8, h->stream) != hipSuccess))
        return fail(set_error(PAVA_HIP_EDEVICE, "CSR upload failed"));
    if (keep) {
        h->d_rp = d_rp;
        h->d_ci = d_ci;
        h->d_va = d_va;
    }
    rc = create_core(M, A->numCols, nnz, d_rp, d_ci, d_va, &o, h);
    if (rc) return fail(rc);
    if (!keep) {
        dfree(d_rp);
        dfree(d_ci);
        dfree(d_va);
    }
    *out = h;
    return 0;
}

extern "C" int cvr_hip_create_device(int numRows, int numCols, int nnz, const int *d_rowOffsets, const int *d_cols,
                                     const double *d_vals, const cvr_hip_opts *opts, cvr_hip_handle **out) {
    if (numRows < 0 || numCols < 0 || nnz < 0) return set_error(PAVA_HIP_EINVAL, "negative size");
    if (!d_rowOffsets || (nnz > 0 && (!d_cols || !d_vals))) return set_error(PAVA_HIP_EINVAL, "null device CSR");
    int rc = check_gpu_sizes(numRows, numCols, nnz);
    if (rc) return rc;
    if ((rc = check_col_limit(opts, numCols))) return rc;
    cvr_hip_opts o;
    cvr_hip_handle *h = nullptr;
    if ((rc = begin_handle(opts, nnz, &o, out, &h))) return rc;
    auto fail = [&](int code) {
        cvr_hip_destroy(h);
        return code;
    };
    // the CSR may still be written by work the caller enqueued on any stream:
    // the conversion, on the handle's own stream, starts after all of it
    if (hipDeviceSynchronize() != hipSuccess)
        return fail(set_error(PAVA_HIP_EDEVICE, "work pending on the device failed: %s",
                              hipGetErrorString(hipGetLastError())));
    {
        int *d_bad = nullptr, bad = 0;
        hipError_t e = dmalloc(&d_bad, 4);
        if (e == hipSuccess) e = hipMemsetAsync(d_bad, 0, 4, h->stream);
        if (e == hipSuccess) e = cvr::launch_validate_csr(d_rowOffsets, d_cols, numRows, numCols, nnz, d_bad, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        dfree(d_bad);
        if (e != hipSuccess) return fail(hip_fail(e, "device CSR check"));
        if (bad & 1) return fail(set_error(PAVA_HIP_EINVAL, "device CSR: row offsets not 0..nnz and non-decreasing"));
        if (bad & 2) return fail(set_error(PAVA_HIP_ERANGE, "device CSR: a column is outside [0, %d)", numCols));
    }
    if (o.recompute || o.keep_fp64 || (o.precision == CVR_HIP_PREC_F64 && o.parity_tol > 0.0)) {
        // keep a private copy of the FP64 CSR for the recompute pass
        if ((rc = dev_alloc(h, (void **)&h->d_rp, (size_t)(numRows + 1) * 4)) ||
            (rc = dev_alloc(h, (void **)&h->d_ci, (size_t)nnz * 4)) || (rc = dev_alloc(h, (void **)&h->d_va, (size_t)nnz * 8)))
            return fail(rc);
        hipError_t e = hipMemcpyAsync(h->d_rp, d_rowOffsets, (size_t)(numRows + 1) * 4, hipMemcpyDeviceToDevice, h->stream);
        if (e == hipSuccess && nnz) e = hipMemcpyAsync(h->d_ci, d_cols, (size_t)nnz * 4, hipMemcpyDeviceToDevice, h->stream);
        if (e == hipSuccess && nnz) e = hipMemcpyAsync(h->d_va, d_vals, (size_t)nnz * 8, hipMemcpyDeviceToDevice, h->stream);
        if (e != hipSuccess) return fail(set_error(PAVA_HIP_EDEVICE, "copy of the device CSR failed: %s", hipGetErrorString(e)));
    }
    rc = create_core(numRows, numCols, nnz, d_rowOffsets, d_cols, d_vals, &o, h);
    if (rc) {
        cvr_hip_destroy(h);
        return rc;
    }
    *out = h;
    return 0;
}

extern "C" int cvr_hip_get_info(const cvr_hip_handle *h, cvr_hip_info *info) {
    if (!h || !info) return set_error(PAVA_HIP_EINVAL, "null argument");
    info->numRows = h->M;
    info->numCols = h->N;
    info->nnz = h->nnz;
    info->nchunks = h->nchunks;
    info->steps_per_chunk = h->S;
    info->precision = h->prec;
    info->col_index_bits = h->col16 ? 16 : 32;
    info->layout_elems = h->layout_elems;
    info->seg_capacity = h->seg_cap;
    info->max_slots = h->max_slots;
    info->has_empty_rows = h->has_empty;
    info->n_fixup_rows = h->n_fix;
    info->device_bytes = h->device_bytes;
    info->convert_ms = h->convert_ms;
    info->n_hot_cols = h->n_hot;
    info->n_lds_cols = h->n_lds;
    info->hot_setup_ms = h->hot_setup_ms;
    info->parity_tol = h->par ? h->opts.parity_tol : 0.0;
    info->n_spec_rows = h->n_spec;
    return 0;
}

static cvr::SpmvArgs spmv_args(const cvr_hip_handle *h, const void *d_x, void *d_y) {
    cvr::SpmvArgs a{};
    a.vals = h->d_vals;
    a.cols = h->d_cols;
    a.end_mask = h->d_end_mask;
    a.col_base = h->d_col_base;
    a.chunk_rows = h->d_chunk_rows;
    a.seg = h->d_seg;
    a.split_n = h->d_split_n;
    a.split_tab = h->d_split_tab;
    a.empty_bits = h->d_empty;
    a.x = d_x;
    a.xhot = h->d_xhot;
    a.y = d_y;
    a.carry = h->d_carry;
    a.carry_abs = h->d_carry_abs;
    a.rflag = h->d_rflag;
    a.kappa = (float)h->opts.kappa_threshold;
    a.layout_elems = h->layout_elems;
    a.nchunks = h->nchunks;
    a.S = h->S;
    a.M = h->M;
    a.has_empty = h->has_empty;
    a.recompute = h->d_rflag != nullptr;
    a.n_lds = h->n_lds;
    a.lds_blocks = h->lds_blocks;
    a.wg_next = h->d_wg_next;
    if (h->par) {
        a.rp = h->d_rp;
        a.ci = h->d_ci;
        a.va = h->d_va;
        a.ci_hot = h->ci_hot;
        a.carry_bnd = h->d_carry_bnd;
        a.rc_count = h->d_rc_count;
        a.fix = h->d_fix;
        a.pinfo = h->d_pinfo;
        a.close_cnt = h->d_close_cnt;
        a.parity_tol = h->opts.parity_tol;
        a.spec_rows = h->d_spec_rows;
        a.spec_state = h->d_spec_state;
        a.spec_val = h->d_spec_val;
        a.n_spec = h->n_spec;
        a.spec_blocks = h->spec_blocks;
    }
    a.wg_list = h->d_wg_list;
    a.wg_start = h->d_wg_start;
    return a;
}

// Kernels after the main SpMV: cross-chunk rows, and for mixed precision with
// recompute the FP64 pass over rows flagged for cancellation.
static int enqueue_tail(cvr_hip_handle *h, const void *d_x, void *d_y, const double *d_x64, hipStream_t st) {
    if (h->d_rflag) {
        cvr::FinishArgs f{};
        f.fix = h->d_fix;
        f.carry = h->d_carry;
        f.carry_abs = h->d_carry_abs;
        f.chunk_rows = h->d_chunk_rows;
        f.rp = h->d_rp;
        f.ci = h->d_ci;
        f.va = h->d_va;
        f.x64 = d_x64;
        f.rflag = h->d_rflag;
        f.y = (float *)d_y;
        f.rc_count = h->d_rc_count;
        f.kappa = (float)h->opts.kappa_threshold;
        f.nchunks = h->nchunks;
        f.M = h->M;
        HIP_TRY(cvr::launch_finish(f, st));
        if (h->n_long > 0) {
            cvr::LongArgs l{};
            l.rp = h->d_rp;
            l.ci = h->d_ci;
            l.va = h->d_va;
            l.x64 = d_x64;
            l.rflag = h->d_rflag;
            l.long_rows = h->d_long_rows;
            l.poff = h->d_long_poff;
            l.piece_row = h->d_piece_row;
            l.partial = h->d_partial;
            l.y = (float *)d_y;
            l.rc_long = h->d_rc_long;
            l.n_long = h->n_long;
            l.n_pieces = h->n_pieces;
            if (d_x64) HIP_TRY(cvr::launch_long_rows(l, st));
        }
    } else if (h->n_fix > 0 && !h->par) {  // FP64 parity closes these rows inside the SpMV
        HIP_TRY(cvr::launch_fixup(h->d_fix, h->d_carry, d_y, h->nchunks, h->prec, st));
    }
    return 0;
}

static int enqueue_spmv(cvr_hip_handle *h, const void *d_x, void *d_y, hipStream_t st, const double *d_x64 = nullptr) {
    cvr::SpmvArgs a = spmv_args(h, d_x, d_y);
    if (h->n_hot > 0)
        HIP_TRY(cvr::launch_hot_gather(d_x, h->d_hot_cols, h->d_xhot, h->n_hot,
                                       h->prec == CVR_HIP_PREC_F64 ? 8 : 4, st));
    HIP_TRY(cvr::launch_spmv(a, h->prec, h->col16, st));
    return enqueue_tail(h, d_x, d_y, d_x64, st);
}

extern "C" int cvr_hip_spmv(cvr_hip_handle *h, const double *d_x, double *d_y, void *stream) {
    if (!h) return set_error(PAVA_HIP_EINVAL, "null handle");
    if (h->prec != CVR_HIP_PREC_F64) return set_error(PAVA_HIP_ESTATE, "cvr_hip_spmv needs an FP64 handle");
    if ((!d_x && h->N > 0) || (!d_y && h->M > 0)) return set_error(PAVA_HIP_EINVAL, "null vector");
    HIP_TRY(hipSetDevice(h->device));
    return enqueue_spmv(h, d_x, d_y, (hipStream_t)stream);
}

extern "C" int cvr_hip_spmv_mixed(cvr_hip_handle *h, const float *d_x32, const double *d_x64, float *d_y, void *stream) {
    if (!h) return set_error(PAVA_HIP_EINVAL, "null handle");
    if (h->prec == CVR_HIP_PREC_F64) return set_error(PAVA_HIP_ESTATE, "cvr_hip_spmv_mixed needs an FP32/BF16 handle");
    if ((!d_x32 && h->N > 0) || (!d_y && h->M > 0)) return set_error(PAVA_HIP_EINVAL, "null vector");
    if (h->d_rflag && !d_x64 && h->N > 0)
        return set_error(PAVA_HIP_EINVAL, "a handle built with recompute needs the FP64 x");
    HIP_TRY(hipSetDevice(h->device));
    return enqueue_spmv(h, d_x32, d_y, (hipStream_t)stream, d_x64);
}

extern "C" int cvr_hip_spmv_host(cvr_hip_handle *h, const double *x, double *y) {
    if (!h || (!x && h->N > 0) || (!y && h->M > 0)) return set_error(PAVA_HIP_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(h->device));
    int rc;
    if (!h->d_hx) {
        if ((rc = dev_alloc(h, &h->d_hx, (size_t)std::max(h->N, 1) * 8))) return rc;
        if ((rc = dev_alloc(h, &h->d_hy, (size_t)std::max(h->M, 1) * 8))) return rc;
    }
    if (h->prec == CVR_HIP_PREC_F64) {
        HIP_TRY(hipMemcpyAsync(h->d_hx, x, (size_t)h->N * 8, hipMemcpyHostToDevice, h->stream));
        if ((rc = enqueue_spmv(h, h->d_hx, h->d_hy, h->stream))) return rc;
        HIP_TRY(hipMemcpyAsync(y, h->d_hy, (size_t)h->M * 8, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        return 0;
    }
    std::vector<float> xf((size_t)h->N), yf((size_t)h->M);
    for (int i = 0; i < h->N; ++i) xf[(size_t)i] = (float)x[i];
    HIP_TRY(hipMemcpyAsync(h->d_hx, xf.data(), (size_t)h->N * 4, hipMemcpyHostToDevice, h->stream));
    if (h->d_rflag) {
        if (!h->d_hx64 && (rc = dev_alloc(h, &h->d_hx64, (size_t)std::max(h->N, 1) * 8))) return rc;
        HIP_TRY(hipMemcpyAsync(h->d_hx64, x, (size_t)h->N * 8, hipMemcpyHostToDevice, h->stream));
    }
    if ((rc = enqueue_spmv(h, h->d_hx, h->d_hy, h->stream, (const double *)h->d_hx64))) return rc;
    HIP_TRY(hipMemcpyAsync(yf.data(), h->d_hy, (size_t)h->M * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int i = 0; i < h->M; ++i) y[i] = yf[(size_t)i];
    return 0;
}

// The SpMM descriptor of a handle, its cross-chunk buffer (2 * nchunks * k
// doubles) taken from `carry`; allocates nothing.
static int spmm_desc_fill(cvr_hip_handle *h, int k, const float *d_X, float *d_Y, double *carry, cvr::SpmmDesc *a) {
    if (k < 1 || k > 32) return set_error(PAVA_HIP_EINVAL, "k must be in [1, 32], got %d", k);
    if ((!d_X && h->N > 0) || (!d_Y && h->M > 0)) return set_error(PAVA_HIP_EINVAL, "null X or Y");
    if (k == cvr::spmm_kp(k) && (((uintptr_t)d_X | (uintptr_t)d_Y) & 15))
        return set_error(PAVA_HIP_EINVAL, "X and Y must be 16-byte aligned");
    *a = cvr::SpmmDesc{};
    a->vals = h->d_vals;
    a->cols = (const int *)h->d_cols;
    a->chunk_rows = h->d_chunk_rows;
    a->seg = h->d_seg;
    a->split_n = h->d_split_n;
    a->split_tab = h->d_split_tab;
    a->empty_bits = h->d_empty;
    a->fix = h->d_fix;
    a->hot_cols = h->n_hot > 0 ? h->d_hot_cols : nullptr;
    a->col_base = h->d_col_base;
    a->end_mask = h->d_end_mask;
    a->col16 = h->col16;
    a->X = d_X;
    a->Y = d_Y;
    a->carry = carry;
    a->layout_elems = h->layout_elems;
    a->nchunks = h->nchunks;
    a->S = h->S;
    a->M = h->M;
    a->N = h->N;
    a->k = k;
    a->has_empty = h->has_empty;
    return 0;
}

// As spmm_desc_fill, with the handle's own cross-chunk buffer (allocated by
// the first call; none for an empty matrix).
static int spmm_desc(cvr_hip_handle *h, int k, const float *d_X, float *d_Y, cvr::SpmmDesc *a) {
    if (k < 1 || k > 32) return set_error(PAVA_HIP_EINVAL, "k must be in [1, 32], got %d", k);
    const size_t need = (size_t)h->nchunks * 2 * 32 * 8 + 64;
    if (!h->d_carry_mm && h->nchunks > 0) {
        int rc = dev_alloc(h, (void **)&h->d_carry_mm, need);
        if (rc) return rc;
    }
    return spmm_desc_fill(h, k, d_X, d_Y, h->d_carry_mm, a);
}


extern "C" int cvr_hip_spmm(cvr_hip_handle *h, int k, const float *d_X, float *d_Y, void *stream) {
    if (!h) return set_error(PAVA_HIP_EINVAL, "null handle");
    HIP_TRY(hipSetDevice(h->device));
    cvr::SpmmDesc a;
    int rc = spmm_desc(h, k, d_X, d_Y, &a);
    if (rc) return rc;
    if (h->nchunks == 0) {
        HIP_TRY(cvr::launch_zero_f32(d_Y, (long long)h->M * k, (hipStream_t)stream));
        return 0;
    }
    HIP_TRY(cvr::launch_spmm(a, h->prec, h->max_slots, h->spmm_engine, (hipStream_t)stream));
    return 0;
}

static_assert(PAVA_HIP_SPMM_FMA == cvr::kSpmmFma && PAVA_HIP_SPMM_MFMA == cvr::kSpmmMfma, "engine codes");

extern "C" int cvr_hip_set_spmm_engine(cvr_hip_handle *h, int engine) {
    if (!h) return set_error(PAVA_HIP_EINVAL, "null handle");
    if (engine != PAVA_HIP_SPMM_FMA && engine != PAVA_HIP_SPMM_MFMA)
        return set_error(PAVA_HIP_EINVAL, "unknown SpMM engine %d", engine);
    h->spmm_engine = engine;
    return 0;
}

// ------------------------------------------------ SpMM without allocation
// cvr_hip_spmm_noalloc: the same kernels as cvr_hip_spmm, with the cross-chunk
// buffer in `scratch`, a file-scope device array each GPU holds from module
// load until the process ends. Its device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and kept in
// g_scratch_ptr; host code passes that address, never the array's name. Every
// call on a GPU shares the array: calls that overlap on one GPU race. Each
// entry read by a call is written earlier in the same call (the head and tail
// partials of every chunk of a row spanning chunks), so nothing needs
// clearing.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;  // 16 MiB: 2 * nchunks * k doubles
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices];

// the scratch address on device `dev` (current), looked up on first use
static int scratch_for(int dev, double **out) {
    if (dev < 0 || dev >= kMaxDevices) return set_error(PAVA_HIP_EINVAL, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(PAVA_HIP_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                             hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" int cvr_hip_spmm_noalloc(cvr_hip_handle *h, int k, const float *d_X, float *d_Y, void *stream) {
    if (!h) return set_error(PAVA_HIP_EINVAL, "null handle");
    if (k < 1 || k > 32) return set_error(PAVA_HIP_EINVAL, "k must be in [1, 32], got %d", k);
    const size_t need = (size_t)h->nchunks * 2 * (size_t)k * sizeof(double);
    if (need > kNoAllocScratchBytes)
        return set_error(PAVA_HIP_ERANGE,
                         "cvr_hip_spmm_noalloc: %d chunks x %d right-hand sides need %zu bytes of scratch, more than "
                         "its %zu (use cvr_hip_spmm, or more steps per chunk)",
                         h->nchunks, k, need, kNoAllocScratchBytes);
    HIP_TRY(hipSetDevice(h->device));
    double *scr = nullptr;
    int rc = scratch_for(h->device, &scr);
    if (rc) return rc;
    cvr::SpmmDesc a;
    if ((rc = spmm_desc_fill(h, k, d_X, d_Y, scr, &a))) return rc;
    if (h->nchunks == 0) {
        HIP_TRY(cvr::launch_zero_f32(d_Y, (long long)h->M * k, (hipStream_t)stream));
        return 0;
    }
    HIP_TRY(cvr::launch_spmm(a, h->prec, h->max_slots, h->spmm_engine, (hipStream_t)stream));
    return 0;
}

// ------------------------------------------------------------ batched SpMM
struct cvr_hip_batch {
    int device = 0, nmat = 0, k = 0, prec = 0, total = 0, grid = 0, max_slots = 1;
    size_t lds = 0;
    std::vector<cvr_hip_handle *> mats;
    cvr::SpmmDesc *d_descs = nullptr;
    int2 *d_queue = nullptr;
    // X window (k <= 8): groups of kWinChunks chunks, win workgroups per CU (0: off)
    int2 *d_gqueue = nullptr;
    int total_g = 0, win = 0, win_rows = 0, win_grid = 0;
    size_t win_lds = 0;
    int engine = 0;  // cvr::kSpmmFma or cvr::kSpmmMfma
    int *d_counter = nullptr;
    std::vector<cvr::SpmmDesc> descs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long long nnz = 0;
};

static int batch_create_body(int nmat, cvr_hip_handle *const *handles, int k, cvr_hip_batch **out);
extern "C" int cvr_hip_batch_create(int nmat, cvr_hip_handle *const *handles, int k, cvr_hip_batch **out) {
    try {
        return batch_create_body(nmat, handles, k, out);
    } catch (const std::bad_alloc &) {
        return set_error(PAVA_HIP_ENOMEM, "host allocation failed while building the batch");
    }
}

static int batch_create_body(int nmat, cvr_hip_handle *const *handles, int k, cvr_hip_batch **out) {
    if (!out || nmat <= 0 || !handles) return set_error(PAVA_HIP_EINVAL, "bad batch arguments");
    *out = nullptr;
    if (k < 1 || k > 32) return set_error(PAVA_HIP_EINVAL, "k must be in [1, 32]");
    // every entry's cross-chunk carries live in its handle: a handle may appear
    // once per batch
    {
        std::vector<const cvr_hip_handle *> seen(handles, handles + nmat);
        for (const cvr_hip_handle *x : seen)
            if (!x) return set_error(PAVA_HIP_EINVAL, "null handle in a batch");
        std::sort(seen.begin(), seen.end());
        if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
            return set_error(PAVA_HIP_EINVAL, "a handle appears more than once in a batch");
    }
    cvr_hip_batch *b = new (std::nothrow) cvr_hip_batch();
    if (!b) return set_error(PAVA_HIP_ENOMEM, "host allocation of the batch failed");
    std::unique_ptr<cvr_hip_batch> guard(b);  // released once the batch is complete
    b->nmat = nmat;
    b->k = k;
    b->mats.assign(handles, handles + nmat);
    b->device = handles[0]->device;
    b->prec = handles[0]->prec;
    std::vector<int2> queue;
    for (int i = 0; i < nmat; ++i) {
        cvr_hip_handle *h = handles[i];
        if (h->device != b->device || h->prec != b->prec || h->nchunks == 0)
            return set_error(PAVA_HIP_EINVAL, "batch matrices must be non-empty and share device and precision");
        b->max_slots = std::max(b->max_slots, h->max_slots);
        b->nnz += h->nnz;
        for (int c = 0; c < h->nchunks; ++c) queue.push_back(make_int2(i, c));
    }
    b->total = (int)queue.size();
    std::vector<int2> gqueue;
    for (int i = 0; i < nmat; ++i)
        for (int c = 0; c < handles[i]->nchunks; c += cvr::kWinChunks) gqueue.push_back(make_int2(i, c));
    b->total_g = (int)gqueue.size();
    b->descs.resize(nmat);
    auto fail = [&](int code) {  // the guard deletes b
        dfree(b->d_descs);
        dfree(b->d_queue);
        dfree(b->d_gqueue);
        dfree(b->d_counter);
        return code;
    };
    if (hipSetDevice(b->device) != hipSuccess) return fail(set_error(PAVA_HIP_EDEVICE, "hipSetDevice failed"));
    if (dmalloc(&b->d_descs, sizeof(cvr::SpmmDesc) * nmat) != hipSuccess ||
        dmalloc(&b->d_queue, sizeof(int2) * std::max(b->total, 1)) != hipSuccess ||
        dmalloc(&b->d_gqueue, sizeof(int2) * std::max(b->total_g, 1)) != hipSuccess ||
        dmalloc(&b->d_counter, 256) != hipSuccess)
        return fail(set_error(PAVA_HIP_ENOMEM, "batch allocation failed"));
    if (hipMemcpy(b->d_queue, queue.data(), sizeof(int2) * queue.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_gqueue, gqueue.data(), sizeof(int2) * gqueue.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b->d_counter, 0, 256) != hipSuccess)
        return fail(set_error(PAVA_HIP_EDEVICE, "batch upload failed"));
    b->lds = cvr::spmm_lds_bytes(b->max_slots, k);
    if (cvr::spmm_batch_grid(b->prec, k, b->lds, b->engine, &b->grid) != hipSuccess)
        return fail(set_error(PAVA_HIP_EDEVICE, "occupancy query failed"));
    *out = guard.release();
    return 0;
}

extern "C" int cvr_hip_batch_set_operands(cvr_hip_batch *b, const float *const *d_X, float *const *d_Y) {
    if (!b || !d_X || !d_Y) return set_error(PAVA_HIP_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(b->device));
    for (int i = 0; i < b->nmat; ++i) {
        int rc = spmm_desc(b->mats[i], b->k, d_X[i], d_Y[i], &b->descs[i]);
        if (rc) return rc;
    }
    HIP_TRY(hipMemcpy(b->d_descs, b->descs.data(), sizeof(cvr::SpmmDesc) * b->nmat, hipMemcpyHostToDevice));
    return 0;
}

static int batch_enqueue(cvr_hip_batch *b, hipStream_t st) {
    if (b->win > 0)
        HIP_TRY(cvr::launch_spmm_batch_win(b->d_descs, b->d_gqueue, b->total_g, b->d_queue, b->total, b->d_counter,
                                           b->prec, b->max_slots, b->win_rows, b->win_lds, b->win_grid, st));
    else
        HIP_TRY(cvr::launch_spmm_batch(b->d_descs, b->d_queue, b->total, b->d_counter, b->prec, b->k, b->lds, b->grid,
                                       b->engine, st));
    return 0;
}

extern "C" int cvr_hip_batch_set_x_window(cvr_hip_batch *b, int wgs_per_cu) {
    if (!b) return set_error(PAVA_HIP_EINVAL, "null batch");
    if (wgs_per_cu < 0 || wgs_per_cu > 4) return set_error(PAVA_HIP_EINVAL, "wgs_per_cu must be in [0, 4]");
    if (wgs_per_cu > 0 && b->k > 8) return set_error(PAVA_HIP_EINVAL, "the X window needs k <= 8, got %d", b->k);
    if (wgs_per_cu > 0 && b->engine == cvr::kSpmmMfma)
        return set_error(PAVA_HIP_EINVAL, "the X window runs on the FMA engine only");
    HIP_TRY(hipSetDevice(b->device));
    // a captured launch no longer matches: it is dropped
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    if (wgs_per_cu == 0) {
        b->win = 0;
        return 0;
    }
    int rows = 0, grid = 0;
    size_t lds = 0;
    HIP_TRY(cvr::spmm_batch_win_plan(b->prec, b->max_slots, wgs_per_cu, &rows, &lds, &grid));
    if (rows == 0)
        return set_error(PAVA_HIP_ERANGE, "no X window of %d rows or more fits %d workgroups per CU",
                         cvr::kWinMinRows, wgs_per_cu);
    b->win = wgs_per_cu;
    b->win_rows = rows;
    b->win_lds = lds;
    b->win_grid = grid;
    return 0;
}

extern "C" int cvr_hip_batch_set_engine(cvr_hip_batch *b, int engine) {
    if (!b) return set_error(PAVA_HIP_EINVAL, "null batch");
    if (engine != PAVA_HIP_SPMM_FMA && engine != PAVA_HIP_SPMM_MFMA)
        return set_error(PAVA_HIP_EINVAL, "unknown SpMM engine %d", engine);
    if (engine == PAVA_HIP_SPMM_MFMA && b->win > 0)
        return set_error(PAVA_HIP_EINVAL, "the X window runs on the FMA engine only");
    HIP_TRY(hipSetDevice(b->device));
    int grid = 0;
    HIP_TRY(cvr::spmm_batch_grid(b->prec, b->k, b->lds, engine, &grid));
    // a captured launch no longer matches: it is dropped
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->engine = engine;
    b->grid = grid;
    return 0;
}

extern "C" int cvr_hip_batch_spmm(cvr_hip_batch *b, void *stream) {
    if (!b) return set_error(PAVA_HIP_EINVAL, "null batch");
    HIP_TRY(hipSetDevice(b->device));
    return batch_enqueue(b, (hipStream_t)stream);
}

extern "C" int cvr_hip_batch_graph(cvr_hip_batch *b, void *stream) {
    if (!b || !stream) return set_error(PAVA_HIP_EINVAL, "graph capture needs a non-null stream");
    HIP_TRY(hipSetDevice(b->device));
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(b, st);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) return rc;
    if (e != hipSuccess) return set_error(PAVA_HIP_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    return 0;
}

extern "C" int cvr_hip_batch_replay(cvr_hip_batch *b, void *stream) {
    if (!b || !b->exec) return set_error(PAVA_HIP_ESTATE, "no captured graph");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return 0;
}

extern "C" int cvr_hip_batch_info(const cvr_hip_batch *b, int *items, int *grid, long long *nnz) {
    if (!b) return set_error(PAVA_HIP_EINVAL, "null batch");
    if (items) *items = b->win > 0 ? b->total_g : b->total;
    if (grid) *grid = b->win > 0 ? b->win_grid : b->grid;
    if (nnz) *nnz = b->nnz;
    return 0;
}

extern "C" int cvr_hip_batch_queue_counts(const cvr_hip_batch *b, int *counts, int *nranges) {
    if (!b || !counts) return set_error(PAVA_HIP_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(counts, b->d_counter + cvr::kBatchQueues, cvr::kBatchQueues * sizeof(int), hipMemcpyDeviceToHost));
    if (nranges) *nranges = std::min(b->win > 0 ? b->win_grid : b->grid, cvr::kBatchQueues);
    return 0;
}

extern "C" int cvr_hip_batch_destroy(cvr_hip_batch *b) {
    if (!b) return 0;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    dfree(b->d_descs);
    dfree(b->d_queue);
    dfree(b->d_gqueue);
    dfree(b->d_counter);
    delete b;
    return 0;
}

extern "C" int cvr_hip_last_recompute_count(cvr_hip_handle *h, int *count) {
    if (!h || !count) return set_error(PAVA_HIP_EINVAL, "null argument");
    *count = 0;
    if (!h->d_rc_count || h->nchunks == 0) return 0;
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipDeviceSynchronize());
    if (h->par) {
        std::vector<int> v((size_t)h->nchunks * 2);
        HIP_TRY(hipMemcpy(v.data(), h->d_rc_count, v.size() * 4, hipMemcpyDeviceToHost));
        long long t = 0;
        for (int c : v) t += c;
        *count = (int)t;
        return 0;
    }
    std::vector<int> v(cvr::rc_count_len(h->M, h->nchunks) + (size_t)h->n_long);
    HIP_TRY(hipMemcpy(v.data(), h->d_rc_count, cvr::rc_count_len(h->M, h->nchunks) * 4, hipMemcpyDeviceToHost));
    if (h->n_long > 0)
        HIP_TRY(hipMemcpy(v.data() + cvr::rc_count_len(h->M, h->nchunks), h->d_rc_long, (size_t)h->n_long * 4,
                          hipMemcpyDeviceToHost));
    long long t = 0;
    for (int c : v) t += c;
    *count = (int)t;
    return 0;
}

extern "C" int cvr_hip_time_spmv(cvr_hip_handle *h, const void *d_x, void *d_y, int warmup, int iters, double *ms_total,
                                 double *ms_main) {
    return cvr_hip_time_spmv_x64(h, d_x, nullptr, d_y, warmup, iters, ms_total, ms_main);
}

// Timing on the handle's stream: warmup SpMVs, then `iters` SpMVs bracketed
// by two events (stats[0]: mean ms per SpMV, every kernel included), then
// `iters` SpMVs with events around each launch of the main SpMV kernel
// (stats[1..4]: its mean, median, min and max in ms).
extern "C" int cvr_hip_time_spmv_stats(cvr_hip_handle *h, const void *d_x, const double *d_x64, void *d_y, int warmup,
                                       int iters, double *stats) {
    if (!h || iters <= 0 || !stats) return set_error(PAVA_HIP_EINVAL, "bad timing arguments");
    if (h->d_rflag && !d_x64) return set_error(PAVA_HIP_EINVAL, "a handle built with recompute needs the FP64 x");
    HIP_TRY(hipSetDevice(h->device));
    std::vector<hipEvent_t> ev;  // [0, 1]: whole SpMVs; then a pair per main-kernel launch
    auto finish = [&](int code) {
        (void)hipStreamSynchronize(h->stream);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        return code;
    };
    for (size_t i = 0; i < 2 + 2 * (size_t)iters; ++i) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return finish(set_error(PAVA_HIP_EDEVICE, "hipEventCreate failed"));
        ev.push_back(e);
    }
    int rc;
    for (int i = 0; i < warmup; ++i)
        if ((rc = enqueue_spmv(h, d_x, d_y, h->stream, d_x64))) return finish(rc);
    if (hipEventRecord(ev[0], h->stream) != hipSuccess) return finish(set_error(PAVA_HIP_EDEVICE, "event record"));
    for (int i = 0; i < iters; ++i)
        if ((rc = enqueue_spmv(h, d_x, d_y, h->stream, d_x64))) return finish(rc);
    if (hipEventRecord(ev[1], h->stream) != hipSuccess) return finish(set_error(PAVA_HIP_EDEVICE, "event record"));
    cvr::SpmvArgs sa = spmv_args(h, d_x, d_y);
    for (int i = 0; i < iters; ++i) {
        hipError_t e = hipSuccess;
        if (h->n_hot > 0)
            e = cvr::launch_hot_gather(d_x, h->d_hot_cols, h->d_xhot, h->n_hot, h->prec == CVR_HIP_PREC_F64 ? 8 : 4,
                                       h->stream);
        if (e == hipSuccess) e = hipEventRecord(ev[2 + 2 * (size_t)i], h->stream);
        if (e == hipSuccess) e = cvr::launch_spmv(sa, h->prec, h->col16, h->stream);
        if (e == hipSuccess) e = hipEventRecord(ev[3 + 2 * (size_t)i], h->stream);
        if (e != hipSuccess) return finish(set_error(PAVA_HIP_EDEVICE, "timed launch: %s", hipGetErrorString(e)));
        if ((rc = enqueue_tail(h, d_x, d_y, d_x64, h->stream))) return finish(rc);
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess) return finish(set_error(PAVA_HIP_EDEVICE, "timed SpMVs failed"));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) return finish(set_error(PAVA_HIP_EDEVICE, "event time"));
    stats[0] = ms / iters;
    std::vector<double> t((size_t)iters);
    for (int i = 0; i < iters; ++i) {
        float x = 0.f;
        if (hipEventElapsedTime(&x, ev[2 + 2 * (size_t)i], ev[3 + 2 * (size_t)i]) != hipSuccess)
            return finish(set_error(PAVA_HIP_EDEVICE, "event time"));
        t[(size_t)i] = x;
    }
    double tot = 0;
    for (double x : t) tot += x;
    std::sort(t.begin(), t.end());
    stats[1] = tot / iters;
    stats[2] = iters % 2 ? t[(size_t)iters / 2] : 0.5 * (t[(size_t)iters / 2 - 1] + t[(size_t)iters / 2]);
    stats[3] = t.front();
    stats[4] = t.back();
    return finish(0);
}

extern "C" int cvr_hip_time_spmv_x64(cvr_hip_handle *h, const void *d_x, const double *d_x64, void *d_y, int warmup,
                                     int iters, double *ms_total, double *ms_main) {
    double st[5];
    const int rc = cvr_hip_time_spmv_stats(h, d_x, d_x64, d_y, warmup, iters, st);
    if (rc) return rc;
    if (ms_total) *ms_total = st[0];
    if (ms_main) *ms_main = st[1];
    return 0;
}

extern "C" int cvr_hip_export_hot(const cvr_hip_handle *h, int *hot_cols) {
    if (!h || !hot_cols) return set_error(PAVA_HIP_EINVAL, "null argument");
    if (h->n_hot == 0) return 0;
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipMemcpy(hot_cols, h->d_hot_cols, (size_t)h->n_hot * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int cvr_hip_export_layout(const cvr_hip_handle *h, void *vals, void *cols, int *chunk_rows, int *seg,
                                     int *seg_count, int *split_n, int *split_tab, int *fix, int *col_base,
                                     uint64_t *end_mask) {
    if (!h) return set_error(PAVA_HIP_EINVAL, "null handle");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const size_t L = (size_t)h->layout_elems;
    if (vals && L) HIP_TRY(hipMemcpy(vals, h->d_vals, L * value_size(h->prec), hipMemcpyDeviceToHost));
    if (cols && L) HIP_TRY(hipMemcpy(cols, h->d_cols, L * (h->col16 ? 2 : 4), hipMemcpyDeviceToHost));
    if (chunk_rows) HIP_TRY(hipMemcpy(chunk_rows, h->d_chunk_rows, (size_t)(h->nchunks + 1) * 4, hipMemcpyDeviceToHost));
    if (seg) HIP_TRY(hipMemcpy(seg, h->d_seg, (size_t)h->seg_cap * 4, hipMemcpyDeviceToHost));
    if (h->nchunks) {
        if (seg_count) HIP_TRY(hipMemcpy(seg_count, h->d_seg_count, (size_t)h->nchunks * 4, hipMemcpyDeviceToHost));
        if (split_n) HIP_TRY(hipMemcpy(split_n, h->d_split_n, (size_t)h->nchunks * 4, hipMemcpyDeviceToHost));
        if (split_tab)
            HIP_TRY(hipMemcpy(split_tab, h->d_split_tab, (size_t)h->nchunks * cvr::kSplitCap * 8, hipMemcpyDeviceToHost));
        if (fix) HIP_TRY(hipMemcpy(fix, h->d_fix, (size_t)h->nchunks * 8, hipMemcpyDeviceToHost));
        if (col_base && h->col16) HIP_TRY(hipMemcpy(col_base, h->d_col_base, (size_t)h->nchunks * 4, hipMemcpyDeviceToHost));
        if (end_mask && h->col16 && L) HIP_TRY(hipMemcpy(end_mask, h->d_end_mask, L / 64 * 8, hipMemcpyDeviceToHost));
    }
    return 0;
}
