// cvr_spmm.hip -- multi-RHS CVR-64 SpMM (Y = A X, k <= 32 right-hand sides)
// for gfx950, single matrix or a persistent batched launch.
//
// Same CVR-64 layout and schedule as the SpMV (cvr_spmv.hip); each nonzero
// scales one row of X (k contiguous FP32 values, fetched with 16-byte vector
// loads) into the lane's k accumulators. X and Y are FP32, row-major
// (X: numCols x k, Y: numRows x k). Accumulation is FP64: the product of two
// FP32 numbers is exact in FP64 and the FP64 FMA issues at the same VALU rate
// as the (non-packed) FP32 FMA on CDNA4, so every output is the FP64 sum of
// exact products rounded once to FP32 -- no recompute pass is needed to stay
// within 1e-5 of an FP64 evaluation on the same inputs.
//
// Batched path: many matrices' chunks form one work queue; a persistent grid
// (sized to the resident capacity of the GPU) pops chunk indices with one
// atomic per chunk until the queue is empty. The queue counter is reset by the
// fixup kernel that follows, so the pair can be captured in a hipGraph and
// replayed with no memset node.

#include <hip/hip_runtime.h>

#include "cvr_kernels.h"

namespace cvr {

__device__ __forceinline__ int mbcnt64m(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ void wave_sync_lds_m() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename VT>
__device__ __forceinline__ double val_f64(VT v);
template <>
__device__ __forceinline__ double val_f64<double>(double v) { return v; }
template <>
__device__ __forceinline__ double val_f64<float>(float v) { return (double)v; }
template <>
__device__ __forceinline__ double val_f64<unsigned short>(unsigned short v) {
    return (double)__uint_as_float((unsigned)v << 16);
}

// Loads row `col` of X (k <= KP floats) into xv[0..KP).
template <int KP>
__device__ __forceinline__ void load_xrow(const float *__restrict__ X, int k, int col, float (&xv)[KP]) {
    const float *p = X + (size_t)col * k;
    if (k == KP) {
#pragma unroll
        for (int j = 0; j < KP; j += 4) {
            const float4 q = *reinterpret_cast<const float4 *>(p + j);
            xv[j] = q.x, xv[j + 1] = q.y, xv[j + 2] = q.z, xv[j + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < KP; ++j) xv[j] = (j < k) ? p[j] : 0.f;
    }
}

// One chunk of one matrix. `slots` is this wave's LDS area of
// max_slots * KP doubles.
template <typename VT, int KP>
__device__ void spmm_chunk(const SpmmDesc &a, int c, int lane, double *slots) {
    const long long C = 64LL * a.S;
    const size_t base = (size_t)C * (size_t)c;
    const int steps = (c == a.nchunks - 1) ? (int)((a.layout_elems - (long long)base) >> 6) : a.S;
    const int zr = a.chunk_rows[c];
    const int *__restrict__ segs = a.seg + zr + kSegExtra * c;
    const int k = a.k;
    int w0 = segs[lane], w1 = segs[64 + lane];
    int wbase = 0, segp = 0;
    const VT *__restrict__ vp = reinterpret_cast<const VT *>(a.vals) + base + lane;
    const int *__restrict__ cp = a.cols + base + lane;

    double acc[KP];
#pragma unroll
    for (int j = 0; j < KP; ++j) acc[j] = 0.0;

    constexpr int U = (KP >= 32) ? 2 : 4;  // steps whose X rows are in flight together
    VT vb[U];
    int cb[U];
#pragma unroll
    for (int i = 0; i < U; ++i) {
        vb[i] = vp[i * 64];
        cb[i] = cp[i * 64];
    }
    for (int g = 0; g < steps; g += U) {
        float xv[U][KP];
#pragma unroll
        for (int i = 0; i < U; ++i) {
            const int cv = cb[i];
            const int col = (cv & kHotFlag) ? a.hot_cols[cv & kColMask] : (cv & kColMask);
            load_xrow<KP>(a.X, k, col, xv[i]);
        }
        VT vn[U];
        int cn[U];
        const bool more = (g + U) < steps;
        if (more) {
#pragma unroll
            for (int i = 0; i < U; ++i) {
                vn[i] = vp[(size_t)(g + U + i) * 64];
                cn[i] = cp[(size_t)(g + U + i) * 64];
            }
        }
#pragma unroll
        for (int i = 0; i < U; ++i) {
            const double v = val_f64<VT>(vb[i]);
#pragma unroll
            for (int j = 0; j < KP; ++j) acc[j] = __builtin_fma(v, (double)xv[i][j], acc[j]);
            const bool end = cb[i] < 0;
            const unsigned long long m = __ballot(end);
            if (m) {
                const int rel = segp - wbase + mbcnt64m(m);
                const int d0 = __shfl(w0, rel & 63);
                const int d1 = __shfl(w1, rel & 63);
                const int d = rel < 64 ? d0 : d1;
                if (end) {
                    if (d >= 0) {
                        float *yr = a.Y + (size_t)d * k;
                        if (k == KP) {
#pragma unroll
                            for (int j = 0; j < KP; j += 4)
                                *reinterpret_cast<float4 *>(yr + j) =
                                    make_float4((float)acc[j], (float)acc[j + 1], (float)acc[j + 2], (float)acc[j + 3]);
                        } else {
#pragma unroll
                            for (int j = 0; j < KP; ++j)
                                if (j < k) yr[j] = (float)acc[j];
                        }
                    } else {
                        double *sl = slots + (size_t)(~d) * KP;
#pragma unroll
                        for (int j = 0; j < KP; ++j) sl[j] = acc[j];
                    }
#pragma unroll
                    for (int j = 0; j < KP; ++j) acc[j] = 0.0;
                }
                segp += __popcll(m);
                if (segp - wbase >= 64) {
                    w0 = w1;
                    w1 = segs[wbase + 128 + lane];
                    wbase += 64;
                }
            }
        }
        if (more) {
#pragma unroll
            for (int i = 0; i < U; ++i) {
                vb[i] = vn[i];
                cb[i] = cn[i];
            }
        }
    }

    // split rows: (group, column) pairs over the lanes, slots summed in element order
    wave_sync_lds_m();
    const int ns = a.split_n[c];
    const int2 *tab = a.split_tab + (size_t)c * kSplitCap;
    for (int t = lane; t < ns * KP; t += 64) {
        const int gi = t / KP, j = t % KP;
        if (j >= k) continue;
        const int2 e = tab[gi];
        const int se = tab[gi + 1].y;
        double s = slots[(size_t)e.y * KP + j];
        for (int q = e.y + 1; q < se; ++q) s += slots[(size_t)q * KP + j];
        if (e.x >= 0)
            a.Y[(size_t)e.x * k + j] = (float)s;
        else if (e.x == kCodeHead)
            a.carry[(size_t)(2 * c) * k + j] = s;
        else if (e.x == kCodeTail)
            a.carry[(size_t)(2 * c + 1) * k + j] = s;
    }
    // empty rows of the chunk's row range
    if (a.has_empty) {
        const int r0 = zr, r1 = a.chunk_rows[c + 1];
        if (r1 > r0) {
            const int wlo = r0 >> 5, whi = (r1 - 1) >> 5;
            for (int w = wlo + lane; w <= whi; w += 64) {
                unsigned bits = a.empty_bits[w];
                const int rb = w << 5;
                if (rb < r0) bits &= ~0u << (r0 - rb);
                if (rb + 32 > r1) bits &= (r1 - rb >= 32) ? ~0u : ((1u << (r1 - rb)) - 1u);
                while (bits) {
                    const int jj = __ffs(bits) - 1;
                    float *yr = a.Y + (size_t)(rb + jj) * k;
                    for (int j = 0; j < k; ++j) yr[j] = 0.f;
                    bits &= bits - 1;
                }
            }
        }
    }
    // the LDS slots are reused by the wave's next chunk (batched path)
    wave_sync_lds_m();
}

template <typename VT, int KP>
__global__ __launch_bounds__(64) void spmm_kernel(SpmmDesc a) {
    extern __shared__ double smem[];
    const int lane = threadIdx.x & 63;
    const int nb = gridDim.x;
    // XCD-aware: each XCD gets a contiguous chunk range
    const int q = nb >> 3, r = nb & 7, x = blockIdx.x & 7, i = blockIdx.x >> 3;
    const int c = (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
    if (c >= a.nchunks) return;
    spmm_chunk<VT, KP>(a, c, lane, smem);
}

// Persistent batched SpMM: waves pop (matrix, chunk) work items until the
// queue is exhausted.
template <typename VT, int KP>
__global__ __launch_bounds__(64) void spmm_batch_kernel(const SpmmDesc *__restrict__ descs,
                                                        const int2 *__restrict__ queue, int total,
                                                        int *__restrict__ counter) {
    extern __shared__ double smem[];
    const int lane = threadIdx.x & 63;
    for (;;) {
        int idx = 0;
        if (lane == 0) idx = atomicAdd(counter, 1);
        idx = __shfl(idx, 0);
        if (idx >= total) break;
        const int2 w = queue[idx];
        spmm_chunk<VT, KP>(descs[w.x], w.y, lane, smem);
    }
}

// Cross-chunk rows of one matrix, all k columns; chunk order.
__device__ __forceinline__ void spmm_fix_chunk(const SpmmDesc &a, int c) {
    const int row = a.fix[2 * c];
    if (row < 0) return;
    const int c0 = a.fix[2 * c + 1];
    const int k = a.k;
    for (int j = 0; j < k; ++j) {
        double s = a.carry[(size_t)(2 * c0 + 1) * k + j];
        for (int t = c0 + 1; t < c; ++t) s += a.carry[(size_t)(2 * t + 1) * k + j];
        s += a.carry[(size_t)(2 * c) * k + j];
        a.Y[(size_t)row * k + j] = (float)s;
    }
}

__global__ void spmm_fixup_kernel(SpmmDesc a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < a.nchunks) spmm_fix_chunk(a, c);
}

__global__ void spmm_batch_fixup_kernel(const SpmmDesc *__restrict__ descs, const int2 *__restrict__ queue, int total,
                                        int *__restrict__ counter) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *counter = 0;  // the queue kernel has finished (stream order): rearm it
    if (t < total) {
        const int2 w = queue[t];
        spmm_fix_chunk(descs[w.x], w.y);
    }
}

// ---------------------------------------------------------------- launchers
template <typename VT>
static hipError_t launch_spmm_vt(const SpmmDesc &a, int kp, size_t lds, hipStream_t st) {
    dim3 grid(a.nchunks), block(64);
    switch (kp) {
    case 8: hipLaunchKernelGGL((spmm_kernel<VT, 8>), grid, block, lds, st, a); break;
    case 16: hipLaunchKernelGGL((spmm_kernel<VT, 16>), grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL((spmm_kernel<VT, 32>), grid, block, lds, st, a); break;
    }
    return hipGetLastError();
}

int spmm_kp(int k) { return k <= 8 ? 8 : k <= 16 ? 16 : 32; }

size_t spmm_lds_bytes(int max_slots, int k) { return (size_t)(max_slots > 0 ? max_slots : 1) * spmm_kp(k) * 8; }

hipError_t launch_spmm(const SpmmDesc &a, int prec, int max_slots, hipStream_t st) {
    if (a.nchunks == 0) return hipSuccess;
    const int kp = spmm_kp(a.k);
    const size_t lds = spmm_lds_bytes(max_slots, a.k);
    hipError_t e;
    if (prec == 0)
        e = launch_spmm_vt<double>(a, kp, lds, st);
    else if (prec == 1)
        e = launch_spmm_vt<float>(a, kp, lds, st);
    else
        e = launch_spmm_vt<unsigned short>(a, kp, lds, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(spmm_fixup_kernel, dim3((a.nchunks + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <typename VT, int KP>
static const void *batch_kernel_ptr() {
    return reinterpret_cast<const void *>(&spmm_batch_kernel<VT, KP>);
}

hipError_t spmm_batch_grid(int prec, int k, size_t lds, int *grid) {
    const int kp = spmm_kp(k);
    const void *f = nullptr;
    if (prec == 1)
        f = kp == 8 ? batch_kernel_ptr<float, 8>() : kp == 16 ? batch_kernel_ptr<float, 16>() : batch_kernel_ptr<float, 32>();
    else if (prec == 0)
        f = kp == 8 ? batch_kernel_ptr<double, 8>() : kp == 16 ? batch_kernel_ptr<double, 16>() : batch_kernel_ptr<double, 32>();
    else
        f = kp == 8 ? batch_kernel_ptr<unsigned short, 8>()
                    : kp == 16 ? batch_kernel_ptr<unsigned short, 16>() : batch_kernel_ptr<unsigned short, 32>();
    int per_cu = 0, dev = 0, ncu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 64, lds);
    if (e != hipSuccess) return e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    *grid = (per_cu > 0 ? per_cu : 1) * (ncu > 0 ? ncu : 1);
    return hipSuccess;
}

hipError_t launch_spmm_batch(const SpmmDesc *d_descs, const int2 *d_queue, int total, int *d_counter, int prec, int k,
                             size_t lds, int grid, hipStream_t st) {
    if (total == 0) return hipSuccess;
    const int kp = spmm_kp(k);
    dim3 g(grid), b(64);
#define PAVA_BATCH(VT)                                                                                          \
    switch (kp) {                                                                                              \
    case 8: hipLaunchKernelGGL((spmm_batch_kernel<VT, 8>), g, b, lds, st, d_descs, d_queue, total, d_counter); break;  \
    case 16: hipLaunchKernelGGL((spmm_batch_kernel<VT, 16>), g, b, lds, st, d_descs, d_queue, total, d_counter); break; \
    default: hipLaunchKernelGGL((spmm_batch_kernel<VT, 32>), g, b, lds, st, d_descs, d_queue, total, d_counter); break; \
    }
    if (prec == 1) {
        PAVA_BATCH(float)
    } else if (prec == 0) {
        PAVA_BATCH(double)
    } else {
        PAVA_BATCH(unsigned short)
    }
#undef PAVA_BATCH
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(spmm_batch_fixup_kernel, dim3((total + 255) / 256), dim3(256), 0, st, d_descs, d_queue, total,
                       d_counter);
    return hipGetLastError();
}

}  // namespace cvr
