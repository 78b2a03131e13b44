// csr5_hip.hip -- the csr5_hip format: CSR5 re-derived for 64-lane wavefronts
// on gfx950 (a second GPU format beside cvr_hip).
//
// pava's CSR5 (csr5/detail/avx512/format_avx512.h, csr5_spmv_avx512.h) cuts
// the nonzeros into omega x sigma tiles (omega = 8 AVX-512 lanes), stores each
// tile transposed, and describes it with per-lane bit flags marking the
// elements that start a row; a segmented sum over the lanes and a calibrate
// step for rows crossing tiles finish y. Here omega = 64 (one wavefront per
// tile), sigma in {4, 8, 16, 32}:
//   * tile t holds elements [t*64*sigma, (t+1)*64*sigma); lane l owns
//     elements t*64*sigma + l*sigma + i (i < sigma), stored at
//     t*64*sigma + i*64 + l, so each i is one coalesced load per wave; the
//     last tile is padded (value 0, column 0, masked by nnz);
//   * flags[t*64+l]: bit i set when element i of the lane starts a non-empty
//     row; yoff[t*64+l]: row starts in lanes before l; tile_base[t]: row starts
//     before the tile. The k-th row start of the tile is row
//     nerows[tile_base[t] + k], nerows listing the non-empty rows in order
//     (pava keeps y offsets relative to the tile's first row plus an extra
//     offset table for tiles holding empty rows; one list serves all tiles);
//   * a row crossing lanes is its owner lane's open segment plus the whole
//     lanes that follow plus the leading segment of the lane where it ends, in
//     lane order (pava: the segmented sum with scansum offsets); a row
//     crossing tiles is the tail of its first tile plus the whole tiles plus
//     the head of its last tile, in tile order (pava: the calibrator, added
//     with races resolved by thread order) -- deterministic, and bit-identical
//     to the serial sum for rows that stay inside one lane.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "cvr_kernels.h"
#include "pava_hip.h"
#include "pava_internal.h"

using pava::check_gpu_sizes;
using pava::set_error;

namespace csr5 {

constexpr int kWaves = 4;  // tiles (wavefronts) per 256-thread workgroup

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// first index r in [0, M] with rp[r] >= e
__device__ __forceinline__ int lower_row(const int *__restrict__ rp, int M, long long e) {
    int lo = 0, hi = M;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((long long)rp[mid] >= e) hi = mid; else lo = mid + 1;
    }
    return lo;
}

// row holding element e (0 <= e < nnz): the last r with rp[r] <= e
__device__ __forceinline__ int row_of(const int *__restrict__ rp, int M, long long e) {
    int lo = 0, hi = M - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((long long)rp[mid] <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
}

struct BuildArgs {
    const int *rp;
    const int *ci;
    const double *va;
    double *vals;
    int *cols;
    unsigned *flags;
    int *yoff;
    int *tile_count;
    const int *tile_base;
    int *nerows;
    int *fix;
    unsigned char *cont;
    int M, nnz, sigma, ntiles;
};

// Pass 1, one wave per tile: transposed copy, row-start flags, per-lane
// offsets, the tile's row-start count, rows crossing tiles.
__global__ __launch_bounds__(64 * kWaves) void flags_kernel(BuildArgs a) {
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (t >= a.ntiles) return;
    const long long T = 64LL * a.sigma;
    const long long e0 = t * T + (long long)lane * a.sigma;
    for (int i = 0; i < a.sigma; ++i) {
        const long long e = e0 + i;
        const long long dst = t * T + (long long)i * 64 + lane;
        a.vals[dst] = e < a.nnz ? a.va[e] : 0.0;
        a.cols[dst] = e < a.nnz ? a.ci[e] : 0;
    }
    unsigned f = 0;
    for (int r = lower_row(a.rp, a.M, e0); r < a.M && (long long)a.rp[r] < e0 + a.sigma && a.rp[r] < a.nnz; ++r)
        if (a.rp[r + 1] > a.rp[r]) f |= 1u << (int)(a.rp[r] - e0);
    const int n = __popc(f);
    // exclusive scan of the lanes' counts
    int incl = n;
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off);
        if (lane >= off) incl += v;
    }
    a.flags[(size_t)t * 64 + lane] = f;
    a.yoff[(size_t)t * 64 + lane] = incl - n;
    if (lane == 63) a.tile_count[t] = incl;
    if (lane == 0) {
        const long long t1e = (t + 1) * T < (long long)a.nnz ? (t + 1) * T : (long long)a.nnz;
        const int rf = row_of(a.rp, a.M, t * T);
        int fr = -1, ft = -1;
        if ((long long)a.rp[rf] < t * T && (long long)a.rp[rf + 1] <= t1e) fr = rf, ft = (int)(a.rp[rf] / T);
        a.fix[2 * t] = fr;
        a.fix[2 * t + 1] = ft;
        const int rl = row_of(a.rp, a.M, t1e - 1);
        a.cont[t] = (unsigned char)((long long)a.rp[rl + 1] > t1e);
    }
}

// Pass 2 (tile_base known): the list of non-empty rows.
__global__ __launch_bounds__(64 * kWaves) void rows_kernel(BuildArgs a) {
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (t >= a.ntiles) return;
    const long long e0 = t * 64LL * a.sigma + (long long)lane * a.sigma;
    int k = a.tile_base[t] + a.yoff[(size_t)t * 64 + lane];
    for (int r = lower_row(a.rp, a.M, e0); r < a.M && (long long)a.rp[r] < e0 + a.sigma && a.rp[r] < a.nnz; ++r)
        if (a.rp[r + 1] > a.rp[r]) a.nerows[k++] = r;
}

struct SpmvArgs {
    const double *vals;
    const int *cols;
    const unsigned *flags;
    const int *yoff;
    const int *tile_base;
    const int *nerows;
    const unsigned char *cont;
    const unsigned *empty_bits;
    const double *x;
    double *y;
    double *head;
    double *tail;
    int M, nnz, ntiles, has_empty;
};

template <int SIGMA>
__global__ __launch_bounds__(64 * kWaves) void spmv_kernel(SpmvArgs a) {
    __shared__ double first_lds[kWaves][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int t = blockIdx.x * kWaves + wv;
    if (t >= a.ntiles) return;
    constexpr long long T = 64LL * SIGMA;
    const size_t base = (size_t)t * T;
    // elements of this lane that exist (the last tile is padded)
    const long long rem = (long long)a.nnz - ((long long)t * T + (long long)lane * SIGMA);
    const int nvalid = rem >= SIGMA ? SIGMA : (rem > 0 ? (int)rem : 0);
    const unsigned f = a.flags[(size_t)t * 64 + lane];
    int k = a.tile_base[t] + a.yoff[(size_t)t * 64 + lane];

    double v[SIGMA], xv[SIGMA];
    int c[SIGMA];
#pragma unroll
    for (int i = 0; i < SIGMA; ++i) {
        v[i] = __builtin_nontemporal_load(&a.vals[base + (size_t)i * 64 + lane]);
        c[i] = __builtin_nontemporal_load(&a.cols[base + (size_t)i * 64 + lane]);
    }
#pragma unroll
    for (int i = 0; i < SIGMA; ++i) xv[i] = a.x[c[i]];
    // the rows of the lane's first two row starts, loaded with the gathers
    // (most lanes start at most two rows; later ones are loaded when reached)
    const int nst = __popc(f);
    const int r0 = nst > 0 ? a.nerows[k] : 0;
    const int r1 = nst > 1 ? a.nerows[k + 1] : 0;

    double acc = 0.0, first = 0.0;
    int nseg = 0, row = -1;
#pragma unroll
    for (int i = 0; i < SIGMA; ++i) {
        if (i < nvalid) {
            if ((f >> i) & 1u) {
                if (nseg == 0)
                    first = acc;
                else
                    a.y[row] = acc;  // a row wholly inside the lane
                row = nseg == 0 ? r0 : nseg == 1 ? r1 : a.nerows[k + nseg];
                ++nseg;
                acc = 0.0;
            }
            acc = __builtin_fma(v[i], xv[i], acc);
        }
    }
    if (nseg == 0) first = acc;
    first_lds[wv][lane] = first;
    const unsigned long long has = __ballot(nseg > 0);
    wave_sync_lds();
    // rows crossing lanes: the owner adds the following lanes in lane order
    if (nseg > 0) {
        double s = acc;
        int j = lane + 1;
        while (j < 64 && !((has >> j) & 1ull)) s += first_lds[wv][j++];
        if (j < 64) {
            s += first_lds[wv][j];
            a.y[row] = s;
        } else if (a.cont[t]) {
            a.tail[t] = s;
        } else {
            a.y[row] = s;
        }
    }
    // the row open at the tile start
    if (lane == 0 && !(f & 1u)) {
        double h = first;
        if (!(has & 1ull)) {
            int j = 1;
            while (j < 64 && !((has >> j) & 1ull)) h += first_lds[wv][j++];
            if (j < 64) h += first_lds[wv][j];
        }
        a.head[t] = h;
    }
    // empty rows: the bitmap words split evenly over the tiles
    if (a.has_empty) {
        const int nw = (a.M + 31) >> 5;
        const int w0 = (int)((long long)nw * t / a.ntiles), w1 = (int)((long long)nw * (t + 1) / a.ntiles);
        for (int w = w0 + lane; w < w1; w += 64) {
            unsigned bits = a.empty_bits[w];
            while (bits) {
                const int j = __ffs(bits) - 1;
                a.y[(w << 5) + j] = 0.0;
                bits &= bits - 1;
            }
        }
    }
}

// Rows crossing tiles: tail of the first tile + whole tiles + head of the last, in tile order.
__global__ void calibrate_kernel(const int *__restrict__ fix, const double *__restrict__ head,
                                 const double *__restrict__ tail, double *__restrict__ y, int ntiles) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int r = fix[2 * t];
    if (r < 0) return;
    const int t0 = fix[2 * t + 1];
    double s = tail[t0];
    for (int u = t0 + 1; u < t; ++u) s += head[u];
    s += head[t];
    y[r] = s;
}

__global__ void zero_kernel(double *y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = 0.0;
}

}  // namespace csr5

struct csr5_hip_handle {
    int device = 0, M = 0, N = 0, nnz = 0, sigma = 16, ntiles = 0, has_empty = 0;
    hipStream_t stream = nullptr;
    double *d_vals = nullptr, *d_head = nullptr, *d_tail = nullptr;
    int *d_cols = nullptr, *d_yoff = nullptr, *d_tile_base = nullptr, *d_nerows = nullptr, *d_fix = nullptr;
    unsigned *d_flags = nullptr, *d_empty = nullptr;
    unsigned char *d_cont = nullptr;
    void *d_hx = nullptr, *d_hy = nullptr;
    double convert_ms = 0.0;
    std::vector<void *> allocs;
};

#define CSR5_TRY(expr)                                                                                      \
    do {                                                                                                    \
        hipError_t _e = (expr);                                                                             \
        if (_e != hipSuccess)                                                                               \
            return set_error(PAVA_HIP_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                             __LINE__);                                                                     \
    } while (0)

static int csr5_alloc(csr5_hip_handle *h, void **p, size_t bytes) {
    if (hipMalloc(p, bytes ? bytes : 256) != hipSuccess) {
        *p = nullptr;
        return set_error(PAVA_HIP_ENOMEM, "device allocation of %zu bytes failed", bytes);
    }
    h->allocs.push_back(*p);
    return 0;
}

extern "C" int csr5_hip_destroy(csr5_hip_handle *h) {
    if (!h) return 0;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void *p : h->allocs) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

static int csr5_build(const pava_csr_matrix *A, csr5_hip_handle *h) {
    const int M = h->M, nnz = h->nnz;
    const long long T = 64LL * h->sigma;
    h->ntiles = nnz > 0 ? (int)((nnz + T - 1) / T) : 0;
    const size_t L = (size_t)h->ntiles * (size_t)T;
    int *d_rp = nullptr, *d_ci = nullptr, *d_status = nullptr, *d_count = nullptr;
    double *d_va = nullptr;
    int rc;
    if ((rc = csr5_alloc(h, (void **)&h->d_vals, L * 8)) || (rc = csr5_alloc(h, (void **)&h->d_cols, L * 4)) ||
        (rc = csr5_alloc(h, (void **)&h->d_flags, (size_t)h->ntiles * 256)) ||
        (rc = csr5_alloc(h, (void **)&h->d_yoff, (size_t)h->ntiles * 256)) ||
        (rc = csr5_alloc(h, (void **)&h->d_tile_base, (size_t)h->ntiles * 4 + 4)) ||
        (rc = csr5_alloc(h, (void **)&h->d_nerows, (size_t)M * 4 + 4)) ||
        (rc = csr5_alloc(h, (void **)&h->d_fix, (size_t)h->ntiles * 8 + 8)) ||
        (rc = csr5_alloc(h, (void **)&h->d_cont, (size_t)h->ntiles + 1)) ||
        (rc = csr5_alloc(h, (void **)&h->d_head, (size_t)h->ntiles * 8 + 8)) ||
        (rc = csr5_alloc(h, (void **)&h->d_tail, (size_t)h->ntiles * 8 + 8)) ||
        (rc = csr5_alloc(h, (void **)&h->d_empty, (size_t)((M + 31) / 32) * 4 + 4)))
        return rc;
    // the CSR is only needed while converting
    if (hipMalloc(&d_rp, (size_t)(M + 1) * 4) != hipSuccess || hipMalloc(&d_ci, (size_t)nnz * 4 + 4) != hipSuccess ||
        hipMalloc(&d_va, (size_t)nnz * 8 + 8) != hipSuccess || hipMalloc(&d_status, 64) != hipSuccess ||
        hipMalloc(&d_count, (size_t)h->ntiles * 4 + 4) != hipSuccess) {
        (void)hipFree(d_rp), (void)hipFree(d_ci), (void)hipFree(d_va), (void)hipFree(d_status), (void)hipFree(d_count);
        return set_error(PAVA_HIP_ENOMEM, "device allocation for the CSR5 conversion failed");
    }
    auto done = [&](int r) {
        if (hipStreamSynchronize(h->stream) != hipSuccess && r == 0)
            r = set_error(PAVA_HIP_EDEVICE, "CSR5 conversion failed: %s", hipGetErrorString(hipGetLastError()));
        (void)hipFree(d_rp), (void)hipFree(d_ci), (void)hipFree(d_va), (void)hipFree(d_status), (void)hipFree(d_count);
        return r;
    };
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return done(set_error(PAVA_HIP_EDEVICE, "hipEventCreate failed"));
    bool ok = hipMemcpyAsync(d_rp, A->rowOffsets, (size_t)(M + 1) * 4, hipMemcpyHostToDevice, h->stream) == hipSuccess &&
              (nnz == 0 || (hipMemcpyAsync(d_ci, A->cols, (size_t)nnz * 4, hipMemcpyHostToDevice, h->stream) == hipSuccess &&
                            hipMemcpyAsync(d_va, A->vals, (size_t)nnz * 8, hipMemcpyHostToDevice, h->stream) == hipSuccess)) &&
              hipMemsetAsync(d_status, 0, 64, h->stream) == hipSuccess &&
              hipEventRecord(e0, h->stream) == hipSuccess;
    csr5::BuildArgs b{};
    b.rp = d_rp, b.ci = d_ci, b.va = d_va, b.vals = h->d_vals, b.cols = h->d_cols, b.flags = h->d_flags;
    b.yoff = h->d_yoff, b.tile_count = d_count, b.tile_base = h->d_tile_base, b.nerows = h->d_nerows;
    b.fix = h->d_fix, b.cont = h->d_cont, b.M = M, b.nnz = nnz, b.sigma = h->sigma, b.ntiles = h->ntiles;
    const dim3 grid((h->ntiles + csr5::kWaves - 1) / csr5::kWaves), block(64 * csr5::kWaves);
    if (ok && (M + 31) / 32 > 0) ok = cvr::launch_empty_rows(d_rp, M, h->d_empty, d_status, h->stream) == hipSuccess;
    std::vector<int> cnt((size_t)h->ntiles + 1, 0);
    if (ok && h->ntiles > 0) {
        hipLaunchKernelGGL(csr5::flags_kernel, grid, block, 0, h->stream, b);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(cnt.data(), d_count, (size_t)h->ntiles * 4, hipMemcpyDeviceToHost, h->stream) == hipSuccess &&
             hipStreamSynchronize(h->stream) == hipSuccess;
        // exclusive scan of the tiles' row-start counts
        int s = 0;
        for (int t = 0; t < h->ntiles; ++t) {
            const int c = cnt[(size_t)t];
            cnt[(size_t)t] = s;
            s += c;
        }
        ok = ok && hipMemcpyAsync(h->d_tile_base, cnt.data(), (size_t)h->ntiles * 4, hipMemcpyHostToDevice, h->stream) == hipSuccess;
        if (ok) {
            hipLaunchKernelGGL(csr5::rows_kernel, grid, block, 0, h->stream, b);
            ok = hipGetLastError() == hipSuccess;
        }
    }
    ok = ok && hipEventRecord(e1, h->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float ms = 0.f;
    if (ok) ok = hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ok) return done(set_error(PAVA_HIP_EDEVICE, "CSR5 conversion failed"));
    int status[16] = {0};
    if (hipMemcpy(status, d_status, sizeof status, hipMemcpyDeviceToHost) != hipSuccess)
        return done(set_error(PAVA_HIP_EDEVICE, "status download failed"));
    h->has_empty = status[4] ? 1 : 0;
    h->convert_ms = ms;
    return done(0);
}

extern "C" int csr5_hip_create(const pava_csr_matrix *A, int sigma, int device, csr5_hip_handle **out) {
    if (!out) return set_error(PAVA_HIP_EINVAL, "null output handle");
    *out = nullptr;
    int rc = 0;
    if (A && (rc = check_gpu_sizes(A->numRows, A->numCols, A->nnz))) return rc;
    if ((rc = pava_csr_validate(A))) return rc;
    if (sigma == 0) sigma = 16;
    if (sigma != 4 && sigma != 8 && sigma != 16 && sigma != 32)
        return set_error(PAVA_HIP_EINVAL, "sigma must be 4, 8, 16 or 32, got %d", sigma);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(PAVA_HIP_EDEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return set_error(PAVA_HIP_EINVAL, "device %d out of range", device);
    CSR5_TRY(hipSetDevice(device));
    csr5_hip_handle *h = new csr5_hip_handle();
    h->device = device, h->M = A->numRows, h->N = A->numCols, h->nnz = A->nnz, h->sigma = sigma;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return set_error(PAVA_HIP_EDEVICE, "hipStreamCreate failed");
    }
    if ((rc = csr5_build(A, h))) {
        csr5_hip_destroy(h);
        return rc;
    }
    *out = h;
    return 0;
}

static int csr5_enqueue(csr5_hip_handle *h, const double *d_x, double *d_y, hipStream_t st) {
    if (h->ntiles == 0) {
        if (h->M > 0) hipLaunchKernelGGL(csr5::zero_kernel, dim3((h->M + 255) / 256), dim3(256), 0, st, d_y, h->M);
        CSR5_TRY(hipGetLastError());
        return 0;
    }
    csr5::SpmvArgs a{};
    a.vals = h->d_vals, a.cols = h->d_cols, a.flags = h->d_flags, a.yoff = h->d_yoff, a.tile_base = h->d_tile_base;
    a.nerows = h->d_nerows, a.cont = h->d_cont, a.empty_bits = h->d_empty, a.x = d_x, a.y = d_y;
    a.head = h->d_head, a.tail = h->d_tail, a.M = h->M, a.nnz = h->nnz, a.ntiles = h->ntiles;
    a.has_empty = h->has_empty;
    const dim3 grid((h->ntiles + csr5::kWaves - 1) / csr5::kWaves), block(64 * csr5::kWaves);
    switch (h->sigma) {
    case 4: hipLaunchKernelGGL(csr5::spmv_kernel<4>, grid, block, 0, st, a); break;
    case 8: hipLaunchKernelGGL(csr5::spmv_kernel<8>, grid, block, 0, st, a); break;
    case 16: hipLaunchKernelGGL(csr5::spmv_kernel<16>, grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL(csr5::spmv_kernel<32>, grid, block, 0, st, a); break;
    }
    hipLaunchKernelGGL(csr5::calibrate_kernel, dim3((h->ntiles + 255) / 256), dim3(256), 0, st, h->d_fix, h->d_head,
                       h->d_tail, d_y, h->ntiles);
    CSR5_TRY(hipGetLastError());
    return 0;
}

extern "C" int csr5_hip_spmv(csr5_hip_handle *h, const double *d_x, double *d_y, void *stream) {
    if (!h) return set_error(PAVA_HIP_EINVAL, "null handle");
    if (h->M > 0 && !d_y) return set_error(PAVA_HIP_EINVAL, "null y");
    if (h->nnz > 0 && !d_x) return set_error(PAVA_HIP_EINVAL, "null x");
    CSR5_TRY(hipSetDevice(h->device));
    return csr5_enqueue(h, d_x, d_y, (hipStream_t)stream);
}

extern "C" int csr5_hip_spmv_host(csr5_hip_handle *h, const double *x, double *y) {
    if (!h || (h->N > 0 && !x) || (h->M > 0 && !y)) return set_error(PAVA_HIP_EINVAL, "null argument");
    CSR5_TRY(hipSetDevice(h->device));
    int rc;
    if (!h->d_hx && ((rc = csr5_alloc(h, &h->d_hx, (size_t)h->N * 8 + 8)) ||
                     (rc = csr5_alloc(h, &h->d_hy, (size_t)h->M * 8 + 8))))
        return rc;
    CSR5_TRY(hipMemcpyAsync(h->d_hx, x, (size_t)h->N * 8, hipMemcpyHostToDevice, h->stream));
    if ((rc = csr5_enqueue(h, (const double *)h->d_hx, (double *)h->d_hy, h->stream))) return rc;
    CSR5_TRY(hipMemcpyAsync(y, h->d_hy, (size_t)h->M * 8, hipMemcpyDeviceToHost, h->stream));
    CSR5_TRY(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int csr5_hip_get_info(const csr5_hip_handle *h, int *ntiles, int *sigma, double *convert_ms) {
    if (!h) return set_error(PAVA_HIP_EINVAL, "null handle");
    if (ntiles) *ntiles = h->ntiles;
    if (sigma) *sigma = h->sigma;
    if (convert_ms) *convert_ms = h->convert_ms;
    return 0;
}

extern "C" int csr5_hip_export_layout(const csr5_hip_handle *h, double *vals, int *cols, unsigned *flags, int *yoff,
                                      int *tile_base, int *nerows, int *fix, unsigned char *cont) {
    if (!h) return set_error(PAVA_HIP_EINVAL, "null handle");
    CSR5_TRY(hipSetDevice(h->device));
    CSR5_TRY(hipStreamSynchronize(h->stream));
    const size_t nt = (size_t)h->ntiles, L = nt * 64 * (size_t)h->sigma;
    if (nt == 0) return 0;
    int ne = 0;
    CSR5_TRY(hipMemcpy(&ne, h->d_tile_base + nt - 1, 4, hipMemcpyDeviceToHost));
    std::vector<unsigned> last(64);
    CSR5_TRY(hipMemcpy(last.data(), h->d_flags + (nt - 1) * 64, 256, hipMemcpyDeviceToHost));
    for (unsigned f : last) ne += __builtin_popcount(f);
    CSR5_TRY(hipMemcpy(vals, h->d_vals, L * 8, hipMemcpyDeviceToHost));
    CSR5_TRY(hipMemcpy(cols, h->d_cols, L * 4, hipMemcpyDeviceToHost));
    CSR5_TRY(hipMemcpy(flags, h->d_flags, nt * 256, hipMemcpyDeviceToHost));
    CSR5_TRY(hipMemcpy(yoff, h->d_yoff, nt * 256, hipMemcpyDeviceToHost));
    CSR5_TRY(hipMemcpy(tile_base, h->d_tile_base, nt * 4, hipMemcpyDeviceToHost));
    if (ne > 0) CSR5_TRY(hipMemcpy(nerows, h->d_nerows, (size_t)ne * 4, hipMemcpyDeviceToHost));
    CSR5_TRY(hipMemcpy(fix, h->d_fix, nt * 8, hipMemcpyDeviceToHost));
    CSR5_TRY(hipMemcpy(cont, h->d_cont, nt, hipMemcpyDeviceToHost));
    return 0;
}

// The CSR5 block of pava's driver (csr5/csr5.h:72-179, main.cpp:578-587):
// conversion (timed), one verification SpMV into y, numIterations timed SpMVs,
// the report line " Performance <matrix> CSR5_HIP <tiles> <convert_s> <exec_s> ".
extern "C" int conduct_csr5_hip(int numRows, int numCols, int nnz, int *rowOffsets, int *cols, double *vals,
                                double *x, double *y, double alpha, char *matrixName, int numIterations) {
    (void)alpha;
    pava_csr_matrix A{nnz, numRows, numCols, vals, cols, rowOffsets};
    csr5_hip_handle *h = nullptr;
    const double t1 = pava_microtime();
    int rc = csr5_hip_create(&A, 16, 0, &h);
    const double t2 = pava_microtime();
    if (rc) return rc;
    double *d_x = nullptr, *d_y = nullptr;
    auto fin = [&](int r) {
        (void)hipFree(d_x);
        (void)hipFree(d_y);
        csr5_hip_destroy(h);
        return r;
    };
    if (hipMalloc(&d_x, (size_t)numCols * 8 + 8) != hipSuccess || hipMalloc(&d_y, (size_t)numRows * 8 + 8) != hipSuccess)
        return fin(set_error(PAVA_HIP_ENOMEM, "x/y allocation failed"));
    if (hipMemcpy(d_x, x, (size_t)numCols * 8, hipMemcpyHostToDevice) != hipSuccess)
        return fin(set_error(PAVA_HIP_EDEVICE, "x upload failed"));
    if ((rc = csr5_enqueue(h, d_x, d_y, nullptr))) return fin(rc);
    if (hipMemcpy(y, d_y, (size_t)numRows * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fin(set_error(PAVA_HIP_EDEVICE, "y download failed"));
    if (hipDeviceSynchronize() != hipSuccess) return fin(set_error(PAVA_HIP_EDEVICE, "verification SpMV failed"));
    const double t3 = pava_microtime();
    for (int it = 0; it < numIterations; ++it)
        if ((rc = csr5_enqueue(h, d_x, d_y, nullptr))) return fin(rc);
    if (hipDeviceSynchronize() != hipSuccess) return fin(set_error(PAVA_HIP_EDEVICE, "SpMV loop failed"));
    const double t4 = pava_microtime();
    pava_print_performance(matrixName ? matrixName : "?", "CSR5_HIP", h->ntiles, t2 - t1,
                           numIterations > 0 ? (t4 - t3) / numIterations : 0.0);
    return fin(0);
}
