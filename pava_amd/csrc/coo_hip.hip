// coo_hip.hip -- "coo_hip": a COO SpMV on gfx950, the GPU stand-in for
// pava's MKL COO block (utils/pava_formats.cpp:50-61, benchmark_COO_SpMV, fed
// by convertCSR2COO from the CSR, utils/pava_formats.cpp:147+). The COO is the
// CSR's nonzeros in row order with an explicit row index per nonzero; the
// conversion expands the row pointers on the GPU.
//
// SpMV: one wave per tile of kTile consecutive nonzeros, in kSteps steps of 64
// (one element per lane, coalesced). Per step every lane forms its product,
// the wave runs an inclusive segmented scan over the lanes (rows are
// contiguous runs), the partial carried from the previous step is added to
// the step's first run, and the lane that ends a run stores it. A run still
// open at the tile's end becomes the tile's tail partial; the partial of a row
// begun in an earlier tile becomes its head partial; a fix-up kernel adds
// them in tile order. Every sum has a fixed order (deterministic). y is zeroed
// first so that rows without nonzeros read 0.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "pava_hip.h"
#include "pava_internal.h"

namespace coo {

constexpr int kSteps = 8;
constexpr int kTile = 64 * kSteps;  // nonzeros per wave
constexpr int kWaves = 4;           // waves per workgroup

// rows[j] = the row of nonzero j (one thread per row writes its range)
__global__ void rows_kernel(const int *__restrict__ rp, int M, int *__restrict__ rows) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M) return;
    for (int j = rp[r]; j < rp[r + 1]; ++j) rows[j] = r;
}

__global__ void zero_kernel(double *__restrict__ y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = 0.0;
}

struct Args {
    const int *rows;
    const int *cols;
    const double *vals;
    const double *x;
    double *y;
    double *head;  // per tile: partial of the row begun before the tile
    double *tail;  // per tile: partial of the row continuing after the tile
    long long nnz;
    int ntiles;
};

__global__ __launch_bounds__(64 * kWaves) void spmv_kernel(Args a) {
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (t >= a.ntiles) return;
    const long long e0 = (long long)t * kTile;
    const long long e1 = e0 + kTile < a.nnz ? e0 + kTile : a.nnz;
    const int r_prev = t > 0 ? a.rows[e0 - 1] : -1;    // row of the nonzero before the tile
    const int r_next = e1 < a.nnz ? a.rows[e1] : -1;   // row of the nonzero after it
    auto store = [&](int row, double v) {
        if (row == r_prev)
            a.head[t] = v;  // begun before the tile (and perhaps continuing after it)
        else if (row == r_next)
            a.tail[t] = v;
        else
            a.y[row] = v;
    };
    // the step's loads for all steps first, then the gathers
    int r[kSteps], c[kSteps];
    double v[kSteps], xv[kSteps];
#pragma unroll
    for (int s = 0; s < kSteps; ++s) {
        const long long e = e0 + 64 * s + lane;
        const bool ok = e < e1;
        r[s] = ok ? a.rows[e] : -1;
        c[s] = ok ? a.cols[e] : 0;
        v[s] = ok ? a.vals[e] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < kSteps; ++s) xv[s] = a.x[c[s]];
    int carry_row = -1;
    double carry = 0.0;
#pragma unroll
    for (int s = 0; s < kSteps; ++s) {
        const int row = r[s];
        double p = v[s] * xv[s];
        // inclusive segmented scan over the lanes (Hillis-Steele, fixed order)
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const double q = __shfl_up(p, d);
            const int rq = __shfl_up(row, d);
            if (lane >= d && rq == row) p += q;
        }
        // the run open at the previous step's last lane either continues into
        // this step's first run or ended with that step
        if (lane == 0 && carry_row >= 0 && row != carry_row) store(carry_row, carry);
        if (row == carry_row) p = carry + p;
        const int rn = __shfl_down(row, 1);
        if (row >= 0 && lane < 63 && rn != row) store(row, p);  // a run ending inside the step
        carry_row = __shfl(row, 63);
        carry = __shfl(p, 63);
    }
    // the run open at the tile's last nonzero
    if (lane == 0 && carry_row >= 0) store(carry_row, carry);
}

// Rows spanning tiles: the tile where such a row ends adds, in tile order, the
// tail partial of the tile where it begins, the partials of the tiles it
// covers whole and its own head partial.
__global__ void fix_kernel(const int *__restrict__ rows, long long nnz, const double *__restrict__ head,
                           const double *__restrict__ tail, double *__restrict__ y, int ntiles) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= 0 || t >= ntiles) return;
    const long long e0 = (long long)t * kTile;
    const int r0 = rows[e0];
    if (rows[e0 - 1] != r0) return;  // the tile's first row begins inside it
    const long long e1 = e0 + kTile < nnz ? e0 + kTile : nnz;
    if (e1 < nnz && rows[e1] == r0) return;  // the row continues: a later tile closes it
    int u0 = t - 1;  // the tile where the row begins
    while (u0 > 0 && rows[(long long)u0 * kTile] == r0 && rows[(long long)u0 * kTile - 1] == r0) --u0;
    double s = tail[u0];
    for (int u = u0 + 1; u < t; ++u) s += head[u];
    s += head[t];
    y[r0] = s;
}

}  // namespace coo

using namespace pava;

struct coo_hip_handle {
    int device = 0, M = 0, N = 0, nnz = 0, ntiles = 0;
    int *d_rows = nullptr, *d_cols = nullptr;
    double *d_vals = nullptr, *d_head = nullptr, *d_tail = nullptr;
    double convert_ms = 0.0;
};

extern "C" int coo_hip_destroy(coo_hip_handle *h) {
    if (!h) return 0;
    (void)hipSetDevice(h->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(h->d_rows);
    (void)hipFree(h->d_cols);
    (void)hipFree(h->d_vals);
    (void)hipFree(h->d_head);
    (void)hipFree(h->d_tail);
    delete h;
    return 0;
}

extern "C" int coo_hip_create(const pava_csr_matrix *A, int device, coo_hip_handle **out) {
    if (!out) return set_error(PAVA_HIP_EINVAL, "null output");
    *out = nullptr;
    int rc = 0;
    if (A && (rc = check_gpu_sizes(A->numRows, A->numCols, A->nnz))) return rc;
    if ((rc = pava_csr_validate(A))) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(PAVA_HIP_EDEVICE, "no HIP device");
    if (hipSetDevice(device) != hipSuccess) return set_error(PAVA_HIP_EDEVICE, "hipSetDevice failed");
    coo_hip_handle *h = new coo_hip_handle();
    h->device = device;
    h->M = A->numRows;
    h->N = A->numCols;
    h->nnz = A->nnz;
    h->ntiles = (int)(((long long)h->nnz + coo::kTile - 1) / coo::kTile);
    int *d_rp = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const size_t nz = (size_t)std::max(h->nnz, 1);
    bool ok = hipMalloc(&d_rp, ((size_t)h->M + 1) * 4) == hipSuccess &&
              hipMalloc(&h->d_rows, nz * 4) == hipSuccess && hipMalloc(&h->d_cols, nz * 4) == hipSuccess &&
              hipMalloc(&h->d_vals, nz * 8) == hipSuccess &&
              hipMalloc(&h->d_head, (size_t)std::max(h->ntiles, 1) * 8) == hipSuccess &&
              hipMalloc(&h->d_tail, (size_t)std::max(h->ntiles, 1) * 8) == hipSuccess;
    ok = ok && hipMemcpy(d_rp, A->rowOffsets, ((size_t)h->M + 1) * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && (h->nnz == 0 || (hipMemcpy(h->d_cols, A->cols, (size_t)h->nnz * 4, hipMemcpyHostToDevice) == hipSuccess &&
                                hipMemcpy(h->d_vals, A->vals, (size_t)h->nnz * 8, hipMemcpyHostToDevice) == hipSuccess));
    ok = ok && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
         hipEventRecord(e0, nullptr) == hipSuccess;
    if (ok && h->M > 0) {
        hipLaunchKernelGGL(coo::rows_kernel, dim3((unsigned)((h->M + 255) / 256)), dim3(256), 0, nullptr, d_rp, h->M,
                           h->d_rows);
        ok = hipGetLastError() == hipSuccess;
    }
    float ms = 0.f;
    ok = ok && hipEventRecord(e1, nullptr) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
         hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    h->convert_ms = ms;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_rp);
    if (!ok) {
        coo_hip_destroy(h);
        return set_error(PAVA_HIP_EDEVICE, "coo_hip upload or conversion failed");
    }
    *out = h;
    return 0;
}

extern "C" int coo_hip_spmv(coo_hip_handle *h, const double *d_x, double *d_y, void *stream) {
    if (!h) return set_error(PAVA_HIP_EINVAL, "null handle");
    if ((!d_x && h->N > 0) || (!d_y && h->M > 0)) return set_error(PAVA_HIP_EINVAL, "null vector");
    if (hipSetDevice(h->device) != hipSuccess) return set_error(PAVA_HIP_EDEVICE, "hipSetDevice failed");
    hipStream_t st = (hipStream_t)stream;
    if (h->M > 0)
        hipLaunchKernelGGL(coo::zero_kernel, dim3((unsigned)((h->M + 255) / 256)), dim3(256), 0, st, d_y, h->M);
    if (h->ntiles > 0) {
        coo::Args a{h->d_rows, h->d_cols, h->d_vals, d_x, d_y, h->d_head, h->d_tail, (long long)h->nnz, h->ntiles};
        hipLaunchKernelGGL(coo::spmv_kernel, dim3((unsigned)((h->ntiles + coo::kWaves - 1) / coo::kWaves)),
                           dim3(64 * coo::kWaves), 0, st, a);
        hipLaunchKernelGGL(coo::fix_kernel, dim3((unsigned)((h->ntiles + 255) / 256)), dim3(256), 0, st, h->d_rows,
                           (long long)h->nnz, h->d_head, h->d_tail, d_y, h->ntiles);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(PAVA_HIP_EDEVICE, "coo_hip launch: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int coo_hip_get_info(coo_hip_handle *h, int *ntiles, double *convert_ms) {
    if (!h) return set_error(PAVA_HIP_EINVAL, "null handle");
    if (ntiles) *ntiles = h->ntiles;
    if (convert_ms) *convert_ms = h->convert_ms;
    return 0;
}
