// csr_vector.hip -- "csr_hip": a plain CSR SpMV on gfx950 with L lanes per
// row (L = 1, 2, 4, ..., 64). The lanes of a row segment stride over the row
// and combine their partial sums with a fixed xor butterfly (deterministic).
// It is the GPU stand-in for pava's MKL CSR block (utils/pava_formats.cpp:39-47,
// benchmark_CSR_SpMV) in the format registry, and the second format the
// selector can choose: short regular rows favour small L, long rows large L.

#include <hip/hip_runtime.h>

#include "pava_hip.h"

namespace csrv {

template <typename VT, typename XT, int L>
__global__ __launch_bounds__(256) void csr_vector_kernel(const int *__restrict__ rp, const int *__restrict__ ci,
                                                         const VT *__restrict__ va, const XT *__restrict__ x,
                                                         XT *__restrict__ y, int M) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = tid / L;
    const int sub = tid & (L - 1);
    double s = 0.0;
    if (row < M) {
        const int b = rp[row], e = rp[row + 1];
        for (int j = b + sub; j < e; j += L) s = __builtin_fma((double)va[j], (double)x[ci[j]], s);
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
    if (row < M && sub == 0) y[row] = (XT)s;
}

template <typename VT, typename XT>
hipError_t launch_vt(int L, const int *rp, const int *ci, const void *va, const void *x, void *y, int M,
                     hipStream_t st) {
    const long long threads = (long long)M * L;
    dim3 grid((unsigned)((threads + 255) / 256)), block(256);
    const VT *v = (const VT *)va;
    const XT *xx = (const XT *)x;
    XT *yy = (XT *)y;
    switch (L) {
    case 1: hipLaunchKernelGGL((csr_vector_kernel<VT, XT, 1>), grid, block, 0, st, rp, ci, v, xx, yy, M); break;
    case 2: hipLaunchKernelGGL((csr_vector_kernel<VT, XT, 2>), grid, block, 0, st, rp, ci, v, xx, yy, M); break;
    case 4: hipLaunchKernelGGL((csr_vector_kernel<VT, XT, 4>), grid, block, 0, st, rp, ci, v, xx, yy, M); break;
    case 8: hipLaunchKernelGGL((csr_vector_kernel<VT, XT, 8>), grid, block, 0, st, rp, ci, v, xx, yy, M); break;
    case 16: hipLaunchKernelGGL((csr_vector_kernel<VT, XT, 16>), grid, block, 0, st, rp, ci, v, xx, yy, M); break;
    case 32: hipLaunchKernelGGL((csr_vector_kernel<VT, XT, 32>), grid, block, 0, st, rp, ci, v, xx, yy, M); break;
    default: hipLaunchKernelGGL((csr_vector_kernel<VT, XT, 64>), grid, block, 0, st, rp, ci, v, xx, yy, M); break;
    }
    return hipGetLastError();
}

__global__ void to_f32_kernel(const double *__restrict__ a, float *__restrict__ b, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

}  // namespace csrv

#include <vector>

#include "pava_internal.h"

using namespace pava;

struct csr_hip_handle {
    int device = 0, M = 0, N = 0, nnz = 0, L = 8, prec = 0;
    int *d_rp = nullptr, *d_ci = nullptr;
    void *d_va = nullptr;
    double convert_ms = 0;
};

extern "C" int csr_hip_create(const pava_csr_matrix *A, int lanes_per_row, int precision, int device,
                              csr_hip_handle **out) {
    if (!out) return set_error(PAVA_HIP_EINVAL, "null output");
    *out = nullptr;
    int rc = 0;
    if (A && (rc = check_gpu_sizes(A->numRows, A->numCols, A->nnz))) return rc;
    if ((rc = pava_csr_validate(A))) return rc;
    if (lanes_per_row < 1 || lanes_per_row > 64 || (lanes_per_row & (lanes_per_row - 1)))
        return set_error(PAVA_HIP_EINVAL, "lanes_per_row must be a power of two in [1, 64]");
    if (precision != CVR_HIP_PREC_F64 && precision != CVR_HIP_PREC_F32)
        return set_error(PAVA_HIP_EINVAL, "csr_hip supports FP64 and FP32 values");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(PAVA_HIP_EDEVICE, "no HIP device");
    if (hipSetDevice(device) != hipSuccess) return set_error(PAVA_HIP_EDEVICE, "hipSetDevice failed");
    csr_hip_handle *h = new csr_hip_handle();
    h->device = device;
    h->M = A->numRows;
    h->N = A->numCols;
    h->nnz = A->nnz;
    h->L = lanes_per_row;
    h->prec = precision;
    const size_t vs = precision == CVR_HIP_PREC_F64 ? 8 : 4;
    double *tmp = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = hipMalloc(&h->d_rp, (size_t)(h->M + 1) * 4) == hipSuccess &&
              hipMalloc(&h->d_ci, (size_t)h->nnz * 4 + 4) == hipSuccess &&
              hipMalloc(&h->d_va, (size_t)h->nnz * vs + 8) == hipSuccess;
    ok = ok && hipMemcpy(h->d_rp, A->rowOffsets, (size_t)(h->M + 1) * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && (h->nnz == 0 || hipMemcpy(h->d_ci, A->cols, (size_t)h->nnz * 4, hipMemcpyHostToDevice) == hipSuccess);
    if (ok && h->nnz) {
        if (precision == CVR_HIP_PREC_F64) {
            ok = hipMemcpy(h->d_va, A->vals, (size_t)h->nnz * 8, hipMemcpyHostToDevice) == hipSuccess;
        } else {
            ok = hipMalloc(&tmp, (size_t)h->nnz * 8) == hipSuccess &&
                 hipMemcpy(tmp, A->vals, (size_t)h->nnz * 8, hipMemcpyHostToDevice) == hipSuccess &&
                 hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
                 hipEventRecord(e0, nullptr) == hipSuccess;
            if (ok) {
                hipLaunchKernelGGL(csrv::to_f32_kernel, dim3((unsigned)((h->nnz + 255) / 256)), dim3(256), 0, nullptr,
                                   tmp, (float *)h->d_va, (long long)h->nnz);
                float ms = 0.f;
                ok = hipGetLastError() == hipSuccess && hipEventRecord(e1, nullptr) == hipSuccess &&
                     hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
                h->convert_ms = ms;
            }
        }
    }
    if (tmp) (void)hipFree(tmp);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (!ok) {
        (void)hipFree(h->d_rp);
        (void)hipFree(h->d_ci);
        (void)hipFree(h->d_va);
        delete h;
        return set_error(PAVA_HIP_EDEVICE, "csr_hip upload failed");
    }
    *out = h;
    return 0;
}

extern "C" int csr_hip_spmv(csr_hip_handle *h, const void *d_x, void *d_y, void *stream) {
    if (!h) return set_error(PAVA_HIP_EINVAL, "null handle");
    if ((!d_x && h->N > 0) || (!d_y && h->M > 0)) return set_error(PAVA_HIP_EINVAL, "null vector");
    if (hipSetDevice(h->device) != hipSuccess) return set_error(PAVA_HIP_EDEVICE, "hipSetDevice failed");
    if (h->M == 0) return 0;
    hipError_t e = h->prec == CVR_HIP_PREC_F64
                       ? csrv::launch_vt<double, double>(h->L, h->d_rp, h->d_ci, h->d_va, d_x, d_y, h->M, (hipStream_t)stream)
                       : csrv::launch_vt<float, float>(h->L, h->d_rp, h->d_ci, h->d_va, d_x, d_y, h->M, (hipStream_t)stream);
    if (e != hipSuccess) return set_error(PAVA_HIP_EDEVICE, "csr_vector launch: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int csr_hip_destroy(csr_hip_handle *h) {
    if (!h) return 0;
    (void)hipSetDevice(h->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(h->d_rp);
    (void)hipFree(h->d_ci);
    (void)hipFree(h->d_va);
    delete h;
    return 0;
}
