// cvr_spmm.hip -- multi-RHS CVR-64 SpMM (Y = A X, k <= 32 right-hand sides)
// for gfx950, single matrix or a persistent batched launch.
//
// Same CVR-64 layout and schedule as the SpMV (cvr_spmv.hip); each nonzero
// scales one row of X (k contiguous FP32 values) into k accumulators, held 4
// per thread by the KP/4 consecutive threads that serve a CVR lane (KP = k
// rounded up to 8, 16 or 32): a row of X is one coalesced request, and a chunk
// runs on the KP/4 waves of a workgroup. X and Y are FP32, row-major
// (X: numCols x k, Y: numRows x k). Accumulation is FP64: the product of two
// FP32 numbers is exact in FP64 and the FP64 FMA issues at the same VALU rate
// as the (non-packed) FP32 FMA on CDNA4, so every output is the FP64 sum of
// exact products rounded once to FP32 -- no recompute pass is needed to stay
// within 1e-5 of an FP64 evaluation on the same inputs.
//
// Batched path: many matrices' chunks form one work queue, cut into one range
// per XCD; a persistent grid (sized to the resident capacity of the GPU) pops
// chunk indices with one atomic per chunk until the queue is empty. The queue
// counters are reset by the fixup kernel that follows, so the pair can be
// captured in a hipGraph and replayed with no memset node.

#include <hip/hip_runtime.h>

#include <algorithm>

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


// The layout, X and Y live in global memory; pointers read from a descriptor
// in memory (the batched path) go through to_global.
typedef float vfloat4 __attribute__((ext_vector_type(4)));  // 16-byte loads and stores of X and Y rows
// LDS pointers (the X window): address space 3, so reads are ds_read, not flat
typedef __attribute__((address_space(3))) const float lds_cf32;
typedef __attribute__((address_space(3))) const vfloat4 lds_cv4;

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

// One chunk of one matrix, processed by the KP/4 waves of a workgroup. Each
// CVR lane is served by KP/4 consecutive threads, thread `part` owning columns
// [4 part, 4 part + 4) of the right-hand sides: the X row of a nonzero is read
// as KP/4 consecutive 16-byte pieces by consecutive threads (one request per
// row), and the 4 accumulators of a thread are FP64. Every wave reads the
// whole step's columns (the end flags of all 64 CVR lanes give each ending
// segment its rank in the chunk's segment list) and a private copy of each
// 8-step group's segment entries, staged in LDS one group ahead as in the
// SpMV. `slots` (max_slots x KP doubles) is shared by the workgroup.
// Columns [j0, j0 + 4) of an X row when k is not the block width: element by
// element (kept out of line so the 16-byte load of full rows stays one load).
__device__ __noinline__ float4 xpart_partial(gptr<const float> p, int j0, int k) {
    float4 q;
    q.x = j0 < k ? p[0] : 0.f;
    q.y = j0 + 1 < k ? p[1] : 0.f;
    q.z = j0 + 2 < k ? p[2] : 0.f;
    q.w = j0 + 3 < k ? p[3] : 0.f;
    return q;
}

// Segment-end mask of step i of a group: the group's masks arrive in one
// vector load, mask i in lane i, read out with readlane.
__device__ __forceinline__ unsigned long long group_step_mask(unsigned long long em, int i) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)em, i);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(em >> 32), i);
    return ((unsigned long long)hi << 32) | lo;
}

// C16: 16-bit layouts (chunk-relative column offsets, end flags in a mask per
// step) instead of 32-bit columns whose sign bit flags a segment end.
// WIN: rows [w0, w0 + W) of X are staged in LDS (`win`, KP floats per row):
// a nonzero whose column falls there reads its X row from LDS. ltid / lnt:
// this chunk's thread index and thread count for the work spread over the
// chunk's threads (split rows, empty rows).
template <typename VT, int KP, bool C16, bool WIN = false>
__device__ void spmm_chunk(const SpmmDesc &a, const int c, const int wv, const int lane, double *slots,
                           int *segbuf, lds_cf32 *win = nullptr, const int w0 = 0, const int W = 0,
                           const int ltid = -1, const int lnt = 0) {
    constexpr int TPL = KP / 4;   // threads per CVR lane
    constexpr int VL = 64 / TPL;  // CVR lanes per wave
    // steps per pipeline group: 8 for k <= 8; 4 for wider blocks, whose X-row
    // buffers would otherwise cost occupancy (measured: batch of 512, k = 32
    // 11.4 -> 8.7 ms, k = 16 6.7 -> 6.0 ms; k = 8 is faster with 8)
    constexpr int G = KP >= 16 ? kStepGroup / 2 : kStepGroup;
    const int part = lane % TPL;
    const int cl = wv * VL + lane / TPL;  // this thread's CVR lane
    const int j0 = 4 * part;
    const long long C = 64LL * a.S;
    const size_t base = (size_t)C * (size_t)c;
    const int steps = (c == a.nchunks - 1) ? (int)((a.layout_elems - (long long)base) >> 6) : a.S;
    const int zr = a.chunk_rows[c];
    const gptr<const int> segs = to_global(a.seg + zr + kSegExtra * c);
    const gptr<const VT> vp = to_global(reinterpret_cast<const VT *>(a.vals) + base + cl);
    const gptr<const int> cpa = to_global(a.cols + base + lane);  // the step's 64 columns, one per thread
    const gptr<const unsigned short> cp16 =
        to_global(reinterpret_cast<const unsigned short *>(a.cols) + base + lane);
    const gptr<const uint64_t> emp = to_global(a.end_mask + (base >> 6));
    const int cbase = C16 ? a.col_base[c] : 0;
    const gptr<const float> X = to_global(a.X);
    const gptr<const int> hot_cols = to_global(a.hot_cols);
    const gptr<float> Y = to_global(a.Y);
    const int k = a.k;
    const bool full = (k == KP);

    auto load_cols = [&](int g0, int (&cc)[G], unsigned long long &em) {
        if constexpr (C16) {
#pragma unroll
            for (int i = 0; i < G; ++i) cc[i] = (int)cp16[(size_t)(g0 + i) * 64];
            em = emp[g0 + (lane & (G - 1))];
        } else {
#pragma unroll
            for (int i = 0; i < G; ++i) cc[i] = __builtin_nontemporal_load(&cpa[(size_t)(g0 + i) * 64]);
        }
    };
    auto step_ends = [&](int cv, unsigned long long em, int i) -> unsigned long long {
        if constexpr (C16)
            return group_step_mask(em, i);
        else
            return __ballot(cv < 0);
    };
    // this thread's 4 values of the X rows of a group's stored columns: the
    // hot-column lookups (layouts with hot-x renumbering) are issued first, so
    // the X loads wait for them once, not each for the loads before it
    const bool has_hot = a.hot_cols != nullptr;
    auto xgroup = [&](const int (&cc)[G], float4 (&xo)[G]) {
        int col[G];
#pragma unroll
        for (int i = 0; i < G; ++i) col[i] = __shfl(cc[i], cl);
        if constexpr (C16) {
#pragma unroll
            for (int i = 0; i < G; ++i) col[i] += cbase;
        } else if (has_hot) {
#pragma unroll
            for (int i = 0; i < G; ++i)
                col[i] = (col[i] & kHotFlag) ? hot_cols[col[i] & kColMask] : (col[i] & kColMask);
        } else {
#pragma unroll
            for (int i = 0; i < G; ++i) col[i] &= kColMask;
        }
#pragma unroll
        for (int i = 0; i < G; ++i) {
            if constexpr (WIN) {
                const unsigned rel = (unsigned)(col[i] - w0);
                if (rel < (unsigned)W) {
                    const vfloat4 q = *reinterpret_cast<lds_cv4 *>(win + rel * KP + j0);
                    xo[i] = make_float4(q.x, q.y, q.z, q.w);
                    continue;
                }
            }
            const gptr<const float> p = X + (size_t)col[i] * k + j0;
            if (full) {
                const vfloat4 q = *reinterpret_cast<gptr<const vfloat4>>(p);
                xo[i] = make_float4(q.x, q.y, q.z, q.w);
            } else {
                xo[i] = xpart_partial(p, j0, k);
            }
        }
    };
    auto group_ends = [&](const int (&cc)[G], unsigned long long em) -> int {
        int e = 0;
#pragma unroll
        for (int i = 0; i < G; ++i) e += __popcll(step_ends(cc[i], em, i));
        return e;
    };
    auto load_segs = [&](int sb, int ne, int (&r)[G]) {
#pragma unroll
        for (int j = 0; j < G; ++j)
            if (j * 64 < ne) r[j] = __builtin_nontemporal_load(&segs[sb + j * 64 + lane]);
    };

    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    VT vb[G], vn[G];
    int cb[G], cn[G], cnn[G];
    unsigned long long eb = 0, en = 0, enn = 0;
    float4 xb[G], xn[G];
    int sc[G], sn[G];

    // prologue: group 0 issued, columns of group 1 in flight
    load_cols(0, cb, eb);
#pragma unroll
    for (int i = 0; i < G; ++i) vb[i] = __builtin_nontemporal_load(&vp[(size_t)i * 64]);
    xgroup(cb, xb);
    int ecur = group_ends(cb, eb);
    load_segs(0, ecur, sc);
    if (G < steps) load_cols(G, cn, en);
    // the loop is entered with nothing in flight (its wait counts then follow
    // its own loads only)
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    int sbase = 0, segp = 0;

    for (int g = 0; g < steps; g += G) {
        const bool has1 = (g + G) < steps, has2 = (g + 2 * G) < steps;
        const int sbase_next = sbase + ecur;
        int enext = 0;
        if (has1) {
            xgroup(cn, xn);
            enext = group_ends(cn, en);
            load_segs(sbase_next, enext, sn);
#pragma unroll
            for (int i = 0; i < G; ++i) vn[i] = __builtin_nontemporal_load(&vp[(size_t)(g + G + i) * 64]);
            if (has2) load_cols(g + 2 * G, cnn, enn);
        }
        // this group's destinations -> this wave's LDS copy
#pragma unroll
        for (int j = 0; j < G; ++j)
            if (j * 64 < ecur) segbuf[j * 64 + lane] = sc[j];
        wave_sync_lds_m();
#pragma unroll
        for (int i = 0; i < G; ++i) {
            const double v = val_f64<VT>(vb[i]);
            acc[0] = __builtin_fma(v, (double)xb[i].x, acc[0]);
            acc[1] = __builtin_fma(v, (double)xb[i].y, acc[1]);
            acc[2] = __builtin_fma(v, (double)xb[i].z, acc[2]);
            acc[3] = __builtin_fma(v, (double)xb[i].w, acc[3]);
            const unsigned long long m = step_ends(cb[i], eb, i);
            if (m) {
                if ((m >> cl) & 1ull) {
                    const int rank = __popcll(m & ((1ull << cl) - 1ull));
                    const int d = segbuf[segp - sbase + rank];
                    if (d >= 0) {
                        const gptr<float> yr = Y + (size_t)d * k + j0;
                        if (full) {
                            const vfloat4 q = {(float)acc[0], (float)acc[1], (float)acc[2], (float)acc[3]};
                            *reinterpret_cast<gptr<vfloat4>>(yr) = q;
                        } else {
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                if (j0 + j < k) yr[j] = (float)acc[j];
                        }
                    } else {
                        double *sl = slots + (size_t)(~d) * KP + j0;
#pragma unroll
                        for (int j = 0; j < 4; ++j) sl[j] = acc[j];
                    }
                    acc[0] = acc[1] = acc[2] = acc[3] = 0.0;
                }
                segp += __popcll(m);
            }
        }
        if (has1) {
#pragma unroll
            for (int i = 0; i < G; ++i) {
                vb[i] = vn[i];
                cb[i] = cn[i];
                xb[i] = xn[i];
                cn[i] = cnn[i];
                sc[i] = sn[i];
            }
            eb = en;
            en = enn;
            sbase = sbase_next;
            ecur = enext;
        }
        wave_sync_lds_m();  // the next group overwrites segbuf
    }

    // split rows: (group, column) pairs over the workgroup, slots summed in element order
    __syncthreads();
    const int tid = ltid >= 0 ? ltid : (int)threadIdx.x, nthreads = ltid >= 0 ? lnt : (int)blockDim.x;
    const int ns = a.split_n[c];
    const int2 *tab = a.split_tab + (size_t)c * kSplitCap;
    for (int t = tid; t < ns * KP; t += nthreads) {
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
            for (int w = wlo + tid; w <= whi; w += nthreads) {
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
    // the slots are reused by the workgroup's next chunk (batched path)
    __syncthreads();
}

// Threads per workgroup: KP/4 waves, one chunk.
template <int KP>
constexpr int spmm_threads() { return 16 * KP; }
static inline int spmm_block(int kp) { return 16 * kp; }

template <typename VT, int KP>
__device__ __forceinline__ void spmm_any_chunk(const SpmmDesc &a, int c, double *smem, int *segbuf) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (a.col16)
        spmm_chunk<VT, KP, true>(a, c, wv, lane, smem, segbuf + wv * 64 * kStepGroup);
    else
        spmm_chunk<VT, KP, false>(a, c, wv, lane, smem, segbuf + wv * 64 * kStepGroup);
}

template <typename VT, int KP>
__global__ __launch_bounds__(spmm_threads<KP>()) void spmm_kernel(SpmmDesc a) {
    extern __shared__ double smem[];
    __shared__ int segbuf[KP / 4][64 * kStepGroup];
    const int nb = gridDim.x;
    // XCD-aware: each XCD gets a contiguous chunk range
    const int q = nb >> 3, r = nb & 7, x = blockIdx.x & 7, i = blockIdx.x >> 3;
    const int c = (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
    if (c >= a.nchunks) return;
    spmm_any_chunk<VT, KP>(a, c, smem, &segbuf[0][0]);
}

// Persistent batched SpMM: workgroups pop (matrix, chunk) work items until the
// queue is exhausted. The queue (matrices in order, chunks in order) is cut
// into kBatchQueues contiguous ranges, one per XCD (workgroups reach the XCDs
// round-robin): a workgroup drains its own XCD's range first, so the chunks of
// one matrix, and the rows of X they share, stay in one L2; then it helps
// with the other ranges.
template <typename VT, int KP>
__global__ __launch_bounds__(spmm_threads<KP>()) void spmm_batch_kernel(const SpmmDesc *__restrict__ descs,
                                                                        const int2 *__restrict__ queue, int total,
                                                                        int *__restrict__ counter) {
    extern __shared__ double smem[];
    __shared__ int segbuf[KP / 4][64 * kStepGroup];
    __shared__ int item;
    const int nq = (int)gridDim.x < kBatchQueues ? (int)gridDim.x : kBatchQueues;
    const int own = (int)blockIdx.x % nq;
    for (int r = 0; r < nq; ++r) {
        const int q = (own + r) % nq;
        const int q0 = (int)((long long)total * q / nq), q1 = (int)((long long)total * (q + 1) / nq);
        for (;;) {
            if (threadIdx.x == 0) item = q0 + atomicAdd(&counter[q], 1);
            __syncthreads();
            const int idx = item;
            __syncthreads();  // every thread has read `item` before it is overwritten
            if (idx >= q1) break;
            const int2 w = queue[idx];
            // a register copy: fields read through a pointer into global
            // memory would be reloaded after every store (and be flat loads)
            const SpmmDesc d = descs[w.x];
            spmm_any_chunk<VT, KP>(d, w.y, smem, &segbuf[0][0]);
        }
    }
}



// Persistent batched SpMM with an LDS window of X (k <= 8): a workgroup of
// 8 waves pops groups of kWinChunks consecutive chunks of one matrix, stages
// the rows [w0, w0 + W) of that matrix's X in LDS -- a window centred on the
// group's rows, scaled to the column range, which holds the X rows of a
// banded matrix's nonzeros -- and runs the group's chunks on its wave pairs
// side by side. Nonzeros whose column falls in the window read their X row
// from LDS instead of fetching a line through the L1; the rest read X as in
// spmm_batch_kernel. The dynamic LDS holds kWinChunks slot arrays
// (slot_doubles each) and then the window (W rows x 8 floats, zero beyond k).
template <typename VT>
__global__ __launch_bounds__(64 * 2 * kWinChunks) void spmm_batch_win_kernel(const SpmmDesc *__restrict__ descs,
                                                                           const int2 *__restrict__ queue,
                                                                           int total, int *__restrict__ counter,
                                                                           int slot_doubles, int win_rows) {
    extern __shared__ __align__(16) double smem[];
    __shared__ int segbuf[2 * kWinChunks][64 * kStepGroup];
    __shared__ int item;
    constexpr int KP = 8;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = wv >> 1;
    float *win = reinterpret_cast<float *>(smem + (size_t)kWinChunks * slot_doubles);
    const int nq = (int)gridDim.x < kBatchQueues ? (int)gridDim.x : kBatchQueues;
    const int own = (int)blockIdx.x % nq;
    for (int r = 0; r < nq; ++r) {
        const int q = (own + r) % nq;
        const int q0 = (int)((long long)total * q / nq), q1 = (int)((long long)total * (q + 1) / nq);
        for (;;) {
            if (threadIdx.x == 0) item = q0 + atomicAdd(&counter[q], 1);
            __syncthreads();
            const int idx = item;
            __syncthreads();  // every thread has read `item` before it is overwritten
            if (idx >= q1) break;
            const int2 w = queue[idx];
            const SpmmDesc d = descs[w.x];
            const int c0 = w.y;
            const int g = d.nchunks - c0 < kWinChunks ? d.nchunks - c0 : kWinChunks;
            // the window: W rows of X centred on the group's rows (scaled to the columns)
            const int W = win_rows < d.N ? win_rows : d.N;
            const long long rmid = ((long long)d.chunk_rows[c0] + d.chunk_rows[c0 + g]) / 2;
            long long w0 = (d.M > 0 ? rmid * d.N / d.M : 0) - W / 2;
            if (w0 > (long long)d.N - W) w0 = (long long)d.N - W;
            if (w0 < 0) w0 = 0;
            const gptr<const float> X = to_global(d.X);
            for (int t = threadIdx.x; t < 2 * W; t += blockDim.x) {
                const int row = t >> 1, j0 = 4 * (t & 1);
                const gptr<const float> p = X + (size_t)(w0 + row) * d.k + j0;
                float4 v;
                if (d.k == KP) {
                    const vfloat4 qv = *reinterpret_cast<gptr<const vfloat4>>(p);
                    v = make_float4(qv.x, qv.y, qv.z, qv.w);
                } else {
                    v = xpart_partial(p, j0, d.k);
                }
                *reinterpret_cast<vfloat4 *>(win + (size_t)row * KP + j0) = vfloat4{v.x, v.y, v.z, v.w};
            }
            __syncthreads();
            double *slots = smem + (size_t)slot * slot_doubles;
            lds_cf32 *lwin = (lds_cf32 *)win;
            const int ltid = threadIdx.x & 127;
            if (slot < g) {
                if (d.col16)
                    spmm_chunk<VT, KP, true, true>(d, c0 + slot, wv & 1, lane, slots, &segbuf[wv][0], lwin, (int)w0,
                                                   W, ltid, 128);
                else
                    spmm_chunk<VT, KP, false, true>(d, c0 + slot, wv & 1, lane, slots, &segbuf[wv][0], lwin, (int)w0,
                                                    W, ltid, 128);
            } else {
                __syncthreads();  // the two workgroup barriers of spmm_chunk
                __syncthreads();
            }
            __syncthreads();  // the window is refilled for the next group
        }
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
    if (t < kBatchQueues) {  // the queue kernel has finished (stream order): keep its pop count, rearm it
        counter[kBatchQueues + t] = counter[t];
        counter[t] = 0;
    }
    if (t < total) {
        const int2 w = queue[t];
        spmm_fix_chunk(descs[w.x], w.y);
    }
}

// ---------------------------------------------------------------- launchers
template <typename VT>
static hipError_t launch_spmm_vt(const SpmmDesc &a, int kp, size_t lds, hipStream_t st) {
    dim3 grid(a.nchunks), block(spmm_block(kp));
    switch (kp) {
    case 8: hipLaunchKernelGGL((spmm_kernel<VT, 8>), grid, block, lds, st, a); break;
    case 16: hipLaunchKernelGGL((spmm_kernel<VT, 16>), grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL((spmm_kernel<VT, 32>), grid, block, lds, st, a); break;
    }
    return hipGetLastError();
}

int spmm_kp(int k) { return k <= 8 ? 8 : k <= 16 ? 16 : 32; }

// Y = 0 for a matrix without nonzeros (a kernel, not a memset node, so that
// the call can be captured in a graph)
__global__ void zero_f32_kernel(float *__restrict__ y, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = 0.f;
}

hipError_t launch_zero_f32(float *y, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const long long blocks = std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, y, n);
    return hipGetLastError();
}

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
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, spmm_block(kp), lds);
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
    dim3 g(grid), b(spmm_block(kp));
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

template <typename VT>
static const void *win_kernel_ptr() {
    return reinterpret_cast<const void *>(&spmm_batch_win_kernel<VT>);
}
static const void *win_kernel_for(int prec) {
    return prec == 1 ? win_kernel_ptr<float>() : prec == 0 ? win_kernel_ptr<double>() : win_kernel_ptr<unsigned short>();
}

hipError_t spmm_batch_win_plan(int prec, int max_slots, int wgs_per_cu, int *win_rows, size_t *lds, int *grid) {
    *win_rows = 0;
    *lds = 0;
    *grid = 0;
    if (wgs_per_cu < 1) return hipSuccess;
    const void *f = win_kernel_for(prec);
    hipError_t e;
    int dev = 0, max_lds = 0, ncu = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev)) != hipSuccess)
        return e;
    if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    hipFuncAttributes fa{};
    if ((e = hipFuncGetAttributes(&fa, f)) != hipSuccess) return e;
    const long long slot_bytes = (long long)kWinChunks * (max_slots > 0 ? max_slots : 1) * 8 * 8;
    const long long budget = (long long)max_lds / wgs_per_cu - (long long)fa.sharedSizeBytes - slot_bytes;
    long long rows = budget / (8 * 4);
    if (rows > kWinMaxRows) rows = kWinMaxRows;
    rows = rows / 64 * 64;
    if (rows < kWinMinRows) return hipSuccess;
    const size_t bytes = (size_t)slot_bytes + (size_t)rows * 8 * 4;
    if ((e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)) != hipSuccess) return e;
    int per_cu = 0;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 64 * 2 * kWinChunks, bytes)) != hipSuccess)
        return e;
    if (per_cu < 1) return hipSuccess;
    *win_rows = (int)rows;
    *lds = bytes;
    *grid = per_cu * (ncu > 0 ? ncu : 1);
    return hipSuccess;
}

hipError_t launch_spmm_batch_win(const SpmmDesc *d_descs, const int2 *d_gqueue, int total_g, const int2 *d_queue,
                                 int total, int *d_counter, int prec, int max_slots, int win_rows, size_t lds, int grid,
                                 hipStream_t st) {
    if (total == 0) return hipSuccess;
    const int slot_doubles = (max_slots > 0 ? max_slots : 1) * 8;
    dim3 g(grid), b(64 * 2 * kWinChunks);
    if (prec == 1)
        hipLaunchKernelGGL(spmm_batch_win_kernel<float>, g, b, lds, st, d_descs, d_gqueue, total_g, d_counter,
                           slot_doubles, win_rows);
    else if (prec == 0)
        hipLaunchKernelGGL(spmm_batch_win_kernel<double>, g, b, lds, st, d_descs, d_gqueue, total_g, d_counter,
                           slot_doubles, win_rows);
    else
        hipLaunchKernelGGL(spmm_batch_win_kernel<unsigned short>, g, b, lds, st, d_descs, d_gqueue, total_g,
                           d_counter, slot_doubles, win_rows);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(spmm_batch_fixup_kernel, dim3((total + 255) / 256), dim3(256), 0, st, d_descs, d_queue, total,
                       d_counter);
    return hipGetLastError();
}

}  // namespace cvr
