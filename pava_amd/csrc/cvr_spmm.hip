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
typedef float vfloat2 __attribute__((ext_vector_type(2)));  // 8-byte pieces (MFMA engine, k <= 8)
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

// The end of a chunk, after its segments are stored: split rows, as (group,
// column) pairs over the chunk's threads with each row's slots summed in
// element order, then the empty rows of the chunk's row range.
template <int KP>
__device__ void spmm_chunk_tail(const SpmmDesc &a, const int c, const int zr, double *slots, const int ltid,
                                const int lnt) {
    __syncthreads();
    const int tid = ltid >= 0 ? ltid : (int)threadIdx.x, nthreads = ltid >= 0 ? lnt : (int)blockDim.x;
    const int k = a.k;
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

    // split rows and empty rows
    spmm_chunk_tail<KP>(a, c, zr, slots, ltid, lnt);
}

// ------------------------------------------------------------ MFMA engine
// The multi-RHS chunk on the matrix cores. v_mfma_f64_4x4x4f64 computes, in
// each of 4 blocks, D(4x4) += A(4x4) B(4x4) in FP64. Its lane layout
// (measured on gfx950): block b = (lane >> 2) & 3; A[i][t] in lane
// 16 t + 4 b + i, B[t][j] in lane 16 t + 4 b + j, D[i][j] in lane
// 16 i + 4 b + j.
//
// One instruction takes a window of 4 steps of 4 CVR lanes (a quad), block b
// serving CVR lane 4 q + b: B[t][j] is one column of the X row of that lane's
// element at step t, and A[i][t] is the element's value in row i if the
// element belongs to the segment held in row i (0 in the other rows). So each
// nonzero scales a tile of its X row into its segment's row of D, and the
// rows of D accumulate up to 4 segments; a quarter of the products are
// nonzero. Segments take rows in turn (segment n of a CVR lane in row
// n mod 4): the segment still open at a window's end keeps its row into the
// next window, and a row is cleared once its segment is stored, so no data
// moves between lanes.
//
// Right-hand-side columns: R = 2 (KP = 8) or 4 per lane and piece, KP / (4 R)
// pieces. Instruction n = piece * R + g gives block column j the column
// piece * 4 R + R j + g: a B lane loads R consecutive X values once per
// window and a D lane stores R consecutive Y values.
//
// Waves: KP / 4 per chunk, as with the FMA engine; wave wv serves quads
// [wv QW, wv QW + QW), QW = 64 / KP. Per window every wave reads the 4 steps'
// columns of all 64 CVR lanes (their end flags rank the ending segments in the
// chunk's segment list) and stages the decoded columns in LDS (`colbuf`, 256
// ints), from which each B lane reads the column of its (step, CVR lane); the
// next window's columns are in flight while the current window runs. X rows
// and values are not double-buffered: without the second set of registers the
// kernels fit 4 waves per SIMD for k > 8 (3 for k <= 8), and occupancy hides
// the latency better (measured on the batching workload, k = 32: 2.70 ->
// 1.72 ms per launch). `segbuf` (256 ints) holds the window's destinations.
// Accumulation is FP64 over exact products of FP32 (or FP64) inputs, as in
// the FMA engine; only the order of the four additions inside an instruction
// may differ.
template <typename VT, int KP, bool C16>
__device__ void spmm_chunk_mfma(const SpmmDesc &a, const int c, const int wv, const int lane, double *slots,
                                int *segbuf, int *colbuf) {
    constexpr int QW = 64 / KP;        // quads per wave
    constexpr int R = KP == 8 ? 2 : 4; // consecutive right-hand sides per lane and piece
    constexpr int NPC = KP / (4 * R);  // pieces
    constexpr int NI = NPC * R;        // instructions per quad and window (KP / 4)
    constexpr int WS = 4;              // steps per window (the instruction's K)
    const int lo = lane & 3, hi = lane >> 4;  // hi: step (A, B lanes) or row (D lanes)
    const long long C = 64LL * a.S;
    const size_t base = (size_t)C * (size_t)c;
    const int steps = (c == a.nchunks - 1) ? (int)((a.layout_elems - (long long)base) >> 6) : a.S;
    const int zr = a.chunk_rows[c];
    const gptr<const int> segs = to_global(a.seg + zr + kSegExtra * c);
    const gptr<const VT> vp = to_global(reinterpret_cast<const VT *>(a.vals) + base);
    const gptr<const int> cpa = to_global(a.cols + base + lane);
    const gptr<const unsigned short> cp16 =
        to_global(reinterpret_cast<const unsigned short *>(a.cols) + base + lane);
    const gptr<const uint64_t> emp = to_global(a.end_mask + (base >> 6));
    const int cbase = C16 ? a.col_base[c] : 0;
    const gptr<const float> X = to_global(a.X);
    const gptr<const int> hot_cols = to_global(a.hot_cols);
    const bool has_hot = a.hot_cols != nullptr;
    const gptr<float> Y = to_global(a.Y);
    const int k = a.k;
    const bool full = (k == KP);
    // CVR lane of this lane's block in quad t: 4 (wv QW + t) + b
    const int lq0 = 4 * wv * QW + ((lane >> 2) & 3);

    auto load_cols = [&](int w0, int (&cc)[WS], unsigned long long &em) {
        if constexpr (C16) {
#pragma unroll
            for (int s = 0; s < WS; ++s) cc[s] = (int)cp16[(size_t)(w0 + s) * 64];
            em = emp[w0 + (lane & (WS - 1))];
        } else {
#pragma unroll
            for (int s = 0; s < WS; ++s) cc[s] = __builtin_nontemporal_load(&cpa[(size_t)(w0 + s) * 64]);
        }
    };
    auto step_ends = [&](int cv, unsigned long long em, int s) -> unsigned long long {
        if constexpr (C16)
            return group_step_mask(em, s);
        else
            return __ballot(cv < 0);
    };
    // A window's end masks, destinations and (in flight) X tiles and values.
    // sb: index of the window's first segment end in the chunk's list.
    auto stage = [&](int w0, int sb, const int (&cc)[WS], unsigned long long em, unsigned long long (&m)[WS],
                     int &ne, float (&xs)[QW][NI], VT (&vs)[QW], int (&sg)[WS]) {
        ne = 0;
#pragma unroll
        for (int s = 0; s < WS; ++s) {
            m[s] = step_ends(cc[s], em, s);
            ne += __popcll(m[s]);
        }
        int dc[WS];
#pragma unroll
        for (int s = 0; s < WS; ++s) {
            if constexpr (C16)
                dc[s] = cc[s] + cbase;
            else
                dc[s] = cc[s] & kColMask;
        }
        if constexpr (!C16) {
            if (has_hot) {
#pragma unroll
                for (int s = 0; s < WS; ++s)
                    if (cc[s] & kHotFlag) dc[s] = hot_cols[dc[s]];
            }
        }
#pragma unroll
        for (int s = 0; s < WS; ++s) colbuf[s * 64 + lane] = dc[s];
        wave_sync_lds_m();
#pragma unroll
        for (int t = 0; t < QW; ++t) {
            const int L = lq0 + 4 * t;
            const int col = colbuf[hi * 64 + L];
            const gptr<const float> p = X + (size_t)col * k + R * lo;
#pragma unroll
            for (int pc = 0; pc < NPC; ++pc) {
                const int j0 = pc * 4 * R + R * lo;
                if (full) {
                    if constexpr (R == 4) {
                        const vfloat4 q = *reinterpret_cast<gptr<const vfloat4>>(p + pc * 4 * R);
                        xs[t][pc * R + 0] = q.x;
                        xs[t][pc * R + 1] = q.y;
                        xs[t][pc * R + 2] = q.z;
                        xs[t][pc * R + 3] = q.w;
                    } else {
                        const vfloat2 q = *reinterpret_cast<gptr<const vfloat2>>(p + pc * 4 * R);
                        xs[t][pc * R + 0] = q.x;
                        xs[t][pc * R + 1] = q.y;
                    }
                } else {
#pragma unroll
                    for (int g = 0; g < R; ++g) xs[t][pc * R + g] = j0 + g < k ? p[pc * 4 * R + g] : 0.f;
                }
            }
            vs[t] = __builtin_nontemporal_load(&vp[(size_t)(w0 + hi) * 64 + L]);
        }
#pragma unroll
        for (int j = 0; j < WS; ++j)
            if (j * 64 < ne) sg[j] = __builtin_nontemporal_load(&segs[sb + j * 64 + lane]);
    };

    double D[QW][NI];
    int sig[QW];  // segments ended so far by the CVR lane of this lane's block, per quad
#pragma unroll
    for (int t = 0; t < QW; ++t) {
        sig[t] = 0;
#pragma unroll
        for (int n = 0; n < NI; ++n) D[t][n] = 0.0;
    }
    unsigned long long mc[WS];
    int nec = 0;
    float xc[QW][NI];
    VT vc[QW];
    int sgc[WS];
    int cn[WS];
    unsigned long long en = 0;
    load_cols(0, cn, en);
    int segp = 0;  // segment ends before the current window

    for (int w = 0; w < steps; w += WS) {
        const bool has1 = (w + WS) < steps;
        stage(w, segp, cn, en, mc, nec, xc, vc, sgc);
        if (has1) load_cols(w + WS, cn, en);
        // this window's destinations -> LDS
#pragma unroll
        for (int j = 0; j < WS; ++j)
            if (j * 64 < nec) segbuf[j * 64 + lane] = sgc[j];
        wave_sync_lds_m();
        int pre[WS];
        pre[0] = 0;
#pragma unroll
        for (int s = 1; s < WS; ++s) pre[s] = pre[s - 1] + __popcll(mc[s - 1]);
        // end flags of this lane's CVR lanes (lq0 + 4 t), bit 4 t of ms[s]
        unsigned ms[WS];
#pragma unroll
        for (int s = 0; s < WS; ++s) ms[s] = (unsigned)(mc[s] >> lq0);
#pragma unroll
        for (int t = 0; t < QW; ++t) {
            const int L = lq0 + 4 * t;
            int e4 = 0;
#pragma unroll
            for (int s = 0; s < WS; ++s) e4 |= (int)((ms[s] >> (4 * t)) & 1u) << s;
            // A: the value of the element at step hi, in the row of its segment
            const int rowA = (sig[t] + __popc(e4 & ((1 << hi) - 1))) & 3;
            const double av = rowA == lo ? val_f64<VT>(vc[t]) : 0.0;
#pragma unroll
            for (int n = 0; n < NI; ++n)
                D[t][n] = __builtin_amdgcn_mfma_f64_4x4x4f64(av, (double)xc[t][n], D[t][n], 0, 0, 0);
            // D row hi holds a segment that ended in this window: store it, clear the row
            const int ne = __popc(e4);
            const int slot = (hi - sig[t]) & 3;
            if (slot < ne) {
                int e = e4;
                for (int q = 0; q < slot; ++q) e &= e - 1;
                const int ks = __ffs(e) - 1;
                const unsigned long long mk = ks == 0 ? mc[0] : ks == 1 ? mc[1] : ks == 2 ? mc[2] : mc[3];
                const int pk = ks == 0 ? pre[0] : ks == 1 ? pre[1] : ks == 2 ? pre[2] : pre[3];
                const int d = segbuf[pk + __popcll(mk & ((1ull << L) - 1ull))];
                if (d >= 0) {
                    const gptr<float> yr = Y + (size_t)d * k;
#pragma unroll
                    for (int pc = 0; pc < NPC; ++pc) {
                        const int j0 = pc * 4 * R + R * lo;
                        if (full) {
                            if constexpr (R == 4) {
                                const vfloat4 q = {(float)D[t][pc * R + 0], (float)D[t][pc * R + 1],
                                                   (float)D[t][pc * R + 2], (float)D[t][pc * R + 3]};
                                *reinterpret_cast<gptr<vfloat4>>(yr + j0) = q;
                            } else {
                                const vfloat2 q = {(float)D[t][pc * R + 0], (float)D[t][pc * R + 1]};
                                *reinterpret_cast<gptr<vfloat2>>(yr + j0) = q;
                            }
                        } else {
#pragma unroll
                            for (int g = 0; g < R; ++g)
                                if (j0 + g < k) yr[j0 + g] = (float)D[t][pc * R + g];
                        }
                    }
                } else {
                    double *sl = slots + (size_t)(~d) * KP;
#pragma unroll
                    for (int pc = 0; pc < NPC; ++pc)
#pragma unroll
                        for (int g = 0; g < R; ++g) sl[pc * 4 * R + R * lo + g] = D[t][pc * R + g];
                }
#pragma unroll
                for (int n = 0; n < NI; ++n) D[t][n] = 0.0;
            }
            sig[t] += ne;
        }
        segp += nec;
        wave_sync_lds_m();  // the next window overwrites segbuf and colbuf
    }
    spmm_chunk_tail<KP>(a, c, zr, slots, -1, 0);
}

// Threads per workgroup: KP/4 waves, one chunk.
template <int KP>
constexpr int spmm_threads() { return 16 * KP; }
static inline int spmm_block(int kp) { return 16 * kp; }

// MF: the MFMA engine (spmm_chunk_mfma), else per-lane FMA (spmm_chunk). Each
// wave's 64 * kStepGroup ints of segbuf: the FMA engine's group destinations,
// or the MFMA engine's window destinations and staged columns.
template <typename VT, int KP, bool MF>
__device__ __forceinline__ void spmm_any_chunk(const SpmmDesc &a, int c, double *smem, int *segbuf) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int *sb = segbuf + wv * 64 * kStepGroup;
    if constexpr (MF) {
        static_assert(64 * kStepGroup >= 2 * 256, "segbuf holds the window's destinations and columns");
        if (a.col16)
            spmm_chunk_mfma<VT, KP, true>(a, c, wv, lane, smem, sb, sb + 256);
        else
            spmm_chunk_mfma<VT, KP, false>(a, c, wv, lane, smem, sb, sb + 256);
    } else {
        if (a.col16)
            spmm_chunk<VT, KP, true>(a, c, wv, lane, smem, sb);
        else
            spmm_chunk<VT, KP, false>(a, c, wv, lane, smem, sb);
    }
}

template <typename VT, int KP, bool MF>
__global__ __launch_bounds__(spmm_threads<KP>()) __attribute__((amdgpu_waves_per_eu(MF && KP > 8 ? 4 : 1)))
void spmm_kernel(SpmmDesc a) {
    extern __shared__ double smem[];
    __shared__ int segbuf[KP / 4][64 * kStepGroup];
    const int nb = gridDim.x;
    // XCD-aware: each XCD gets a contiguous chunk range
    const int q = nb >> 3, r = nb & 7, x = blockIdx.x & 7, i = blockIdx.x >> 3;
    const int c = (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
    if (c >= a.nchunks) return;
    spmm_any_chunk<VT, KP, MF>(a, c, smem, &segbuf[0][0]);
}

// Persistent batched SpMM: workgroups pop (matrix, chunk) work items until the
// queue is exhausted. The queue (matrices in order, chunks in order) is cut
// into kBatchQueues contiguous ranges, one per XCD (workgroups reach the XCDs
// round-robin): a workgroup drains its own XCD's range first, so the chunks of
// one matrix, and the rows of X they share, stay in one L2; then it helps
// with the other ranges.
// (MFMA engine, k > 8: at least 4 waves per SIMD, i.e. 2 workgroups of 8 waves
// per CU at k = 32.)
template <typename VT, int KP, bool MF>
__global__ __launch_bounds__(spmm_threads<KP>()) __attribute__((amdgpu_waves_per_eu(MF && KP > 8 ? 4 : 1)))
void spmm_batch_kernel(const SpmmDesc *__restrict__ descs, const int2 *__restrict__ queue, int total,
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
            spmm_any_chunk<VT, KP, MF>(d, w.y, smem, &segbuf[0][0]);
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
template <typename VT, bool MF>
static hipError_t launch_spmm_vt(const SpmmDesc &a, int kp, size_t lds, hipStream_t st) {
    dim3 grid(a.nchunks), block(spmm_block(kp));
    switch (kp) {
    case 8: hipLaunchKernelGGL((spmm_kernel<VT, 8, MF>), grid, block, lds, st, a); break;
    case 16: hipLaunchKernelGGL((spmm_kernel<VT, 16, MF>), grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL((spmm_kernel<VT, 32, MF>), grid, block, lds, st, a); break;
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

hipError_t launch_spmm(const SpmmDesc &a, int prec, int max_slots, int engine, hipStream_t st) {
    if (a.nchunks == 0) return hipSuccess;
    const int kp = spmm_kp(a.k);
    const size_t lds = spmm_lds_bytes(max_slots, a.k);
    const bool mf = engine == kSpmmMfma;
    hipError_t e;
    if (prec == 0)
        e = mf ? launch_spmm_vt<double, true>(a, kp, lds, st) : launch_spmm_vt<double, false>(a, kp, lds, st);
    else if (prec == 1)
        e = mf ? launch_spmm_vt<float, true>(a, kp, lds, st) : launch_spmm_vt<float, false>(a, kp, lds, st);
    else
        e = mf ? launch_spmm_vt<unsigned short, true>(a, kp, lds, st)
               : launch_spmm_vt<unsigned short, false>(a, kp, lds, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(spmm_fixup_kernel, dim3((a.nchunks + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <typename VT, int KP>
static const void *batch_kernel_ptr(bool mf) {
    return mf ? reinterpret_cast<const void *>(&spmm_batch_kernel<VT, KP, true>)
              : reinterpret_cast<const void *>(&spmm_batch_kernel<VT, KP, false>);
}
template <typename VT>
static const void *batch_kernel_for(int kp, bool mf) {
    return kp == 8 ? batch_kernel_ptr<VT, 8>(mf) : kp == 16 ? batch_kernel_ptr<VT, 16>(mf) : batch_kernel_ptr<VT, 32>(mf);
}

hipError_t spmm_batch_grid(int prec, int k, size_t lds, int engine, int *grid) {
    const int kp = spmm_kp(k);
    const bool mf = engine == kSpmmMfma;
    const void *f = prec == 1 ? batch_kernel_for<float>(kp, mf)
                    : prec == 0 ? batch_kernel_for<double>(kp, mf) : batch_kernel_for<unsigned short>(kp, mf);
    int per_cu = 0, dev = 0, ncu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, spmm_block(kp), lds);
    if (e != hipSuccess) return e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    *grid = (per_cu > 0 ? per_cu : 1) * (ncu > 0 ? ncu : 1);
    return hipSuccess;
}

template <typename VT, bool MF>
static void launch_batch_vt(int kp, dim3 g, dim3 b, size_t lds, hipStream_t st, const SpmmDesc *d_descs,
                            const int2 *d_queue, int total, int *d_counter) {
    switch (kp) {
    case 8: hipLaunchKernelGGL((spmm_batch_kernel<VT, 8, MF>), g, b, lds, st, d_descs, d_queue, total, d_counter); break;
    case 16: hipLaunchKernelGGL((spmm_batch_kernel<VT, 16, MF>), g, b, lds, st, d_descs, d_queue, total, d_counter); break;
    default: hipLaunchKernelGGL((spmm_batch_kernel<VT, 32, MF>), g, b, lds, st, d_descs, d_queue, total, d_counter); break;
    }
}

hipError_t launch_spmm_batch(const SpmmDesc *d_descs, const int2 *d_queue, int total, int *d_counter, int prec, int k,
                             size_t lds, int grid, int engine, hipStream_t st) {
    if (total == 0) return hipSuccess;
    const int kp = spmm_kp(k);
    const bool mf = engine == kSpmmMfma;
    const dim3 g(grid), b(spmm_block(kp));
    if (prec == 1)
        mf ? launch_batch_vt<float, true>(kp, g, b, lds, st, d_descs, d_queue, total, d_counter)
           : launch_batch_vt<float, false>(kp, g, b, lds, st, d_descs, d_queue, total, d_counter);
    else if (prec == 0)
        mf ? launch_batch_vt<double, true>(kp, g, b, lds, st, d_descs, d_queue, total, d_counter)
           : launch_batch_vt<double, false>(kp, g, b, lds, st, d_descs, d_queue, total, d_counter);
    else
        mf ? launch_batch_vt<unsigned short, true>(kp, g, b, lds, st, d_descs, d_queue, total, d_counter)
           : launch_batch_vt<unsigned short, false>(kp, g, b, lds, st, d_descs, d_queue, total, d_counter);
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
