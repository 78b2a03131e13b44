// cvr_convert.hip -- GPU CSR -> CVR-64 conversion for gfx950.
//
// The layout re-derives pava's CVR (cvr/cvr.h:583-1060, pre_processing) for a
// 64-lane wavefront:
//   * the nonzeros are cut into chunks of C = 64*S consecutive CSR elements
//     (pava: per-thread nnz-balanced slices, cvr/cvr.h:602-648); one wave owns
//     one chunk;
//   * each of the 64 lanes streams one row at a time; when a lane finishes its
//     row, the next unassigned non-empty row of the chunk goes to it, lanes
//     served in ascending order (cvr/cvr.h:839-880);
//   * once the rows run out, an idle lane takes the LAST R elements of the
//     lowest lane holding more than R, where R is the number of steps left
//     (pava takes the first ave = R elements, cvr/cvr.h:909-978); every lane
//     therefore ends on the chunk's last step;
//   * values and columns are stored step-major, lane-minor (element (s, l) at
//     chunk_base + 64*s + l), so each step is one coalesced 512 B (FP64) load;
//   * the sign bit of a column marks the last element of a lane segment
//     (replaces pava's record pairs, cvr/cvr.h:861-873, 919-955);
//   * seg[] lists, in the order segments end (step, then lane), where each
//     segment's sum goes: a row id, or ~slot for segments of split rows whose
//     partial sums are reduced at the end of the chunk (pava's t_rets /
//     final_2 / split arrays, cvr/cvr.h:884-894, 1823-1868).
// A segment never jumps: a thief's elements are the tail of the victim's
// remaining run, so every segment is a contiguous CSR range and partial sums
// of a split row are added in element order.
//
// One wavefront simulates its chunk's schedule with ballots: the control is a
// scan over steps with wave-uniform state, the data movement is batched 8
// steps at a time so the gathers of a batch are in flight together.

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <utility>
#include <vector>

#include "cvr_kernels.h"

namespace cvr {

__device__ __forceinline__ int mbcnt64(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned short f32_to_bf16_rne(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

template <typename VT>
__device__ __forceinline__ VT cvt_val(double v);
template <>
__device__ __forceinline__ double cvt_val<double>(double v) { return v; }
template <>
__device__ __forceinline__ float cvt_val<float>(double v) { return (float)v; }
template <>
__device__ __forceinline__ unsigned short cvt_val<unsigned short>(double v) { return f32_to_bf16_rne((float)v); }

// Largest r in [0, M) with rp[r] <= b (b < nnz), i.e. the row holding element b.
__device__ int row_of_element(const int *rp, int M, int b) {
    int lo = 0, hi = M;  // invariant: rp[lo] <= b < rp[hi]
    while (hi - lo > 1) {
        int mid = lo + ((hi - lo) >> 1);
        if (rp[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Empty-row bitmap: bit j of word w is set when row 32w+j has no entries.
__global__ void empty_rows_kernel(const int *__restrict__ rp, int M, unsigned *__restrict__ bits,
                                  int *__restrict__ status) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    int nw = (M + 31) >> 5;
    if (w >= nw) return;
    unsigned m = 0;
    int r0 = w << 5;
    for (int j = 0; j < 32; ++j) {
        int r = r0 + j;
        if (r < M && rp[r + 1] == rp[r]) m |= 1u << j;
    }
    bits[w] = m;
    if (m) atomicOr(&status[4], 1);
}

template <typename VT, bool COL16>
__global__ __launch_bounds__(256) void convert_kernel(ConvertArgs a) {
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int c = blockIdx.x * kWavesPerBlock + wv;
    if (c >= a.nchunks) return;

    __shared__ int q_start[kWavesPerBlock][kLanes];
    __shared__ int q_len[kWavesPerBlock][kLanes];
    __shared__ int q_row[kWavesPerBlock][kLanes];
    __shared__ int q_kind[kWavesPerBlock][kLanes];
    __shared__ int sp_start[kWavesPerBlock][kSlotCap];
    __shared__ int sp_idx[kWavesPerBlock][kSlotCap];
    __shared__ int sp_row[kWavesPerBlock][kSlotCap];
    __shared__ int sp_kind[kWavesPerBlock][kSlotCap];
    __shared__ int srt_row[kWavesPerBlock][kSlotCap];
    __shared__ int srt_kind[kWavesPerBlock][kSlotCap];

    const int *__restrict__ rp = a.rp;
    const long long C = 64LL * a.S;
    const int b = (int)(C * c);
    const int e = (int)min<long long>((long long)b + C, (long long)a.nnz);
    const int n = e - b;
    int steps = a.S;
    if (c == a.nchunks - 1) {
        steps = (n + 63) / 64;
        steps = (steps + kStepGroup - 1) / kStepGroup * kStepGroup;
    }
    const int pad = 64 * steps - n;

    const int row_lo = row_of_element(rp, a.M, b);
    const int zr = (c == 0) ? 0 : row_lo;
    const int seg_base = zr + kSegExtra * c;
    if (lane == 0) {
        a.chunk_rows[c] = zr;
        if (c == a.nchunks - 1) a.chunk_rows[a.nchunks] = a.M;
        const bool head = rp[row_lo] < b;
        const bool tail = rp[row_lo + 1] > e;
        if (head && !tail) {
            a.fix[2 * c] = row_lo;
            a.fix[2 * c + 1] = (int)(rp[row_lo] / C);
            atomicAdd(&a.status[3], 1);
        } else {
            a.fix[2 * c] = -1;
            a.fix[2 * c + 1] = -1;
        }
    }

    // 16-bit layouts: column base = smallest column of the chunk
    int cbase = 0;
    if (COL16) {
        int mn = 0x7fffffff, mx = -1;
        for (int k = b + lane; k < e; k += 64) {
            int cc = a.ci[k];
            mn = min(mn, cc);
            mx = max(mx, cc);
        }
        for (int off = 32; off > 0; off >>= 1) {
            mn = min(mn, __shfl_xor(mn, off));
            mx = max(mx, __shfl_xor(mx, off));
        }
        cbase = mn;
        if (lane == 0) {
            a.col_base[c] = cbase;
            if (mx - mn > 0xffff) atomicOr(&a.status[2], 1);
        }
    }

    // ---------------------------------------------------------- state
    int cnt = 0, pos = 0, prow = 0, pkind = 0, sstart = 0;
    bool split = false;
    int qh = 0, qn = 0;              // piece queue head / size (wave-uniform)
    int next_row = row_lo;           // next row to scan for pieces
    bool rows_done = false, pad_done = (pad == 0);
    int segcount = 0, nsp = 0;

    VT *__restrict__ ov = reinterpret_cast<VT *>(a.ovals) + (size_t)b;

    for (int s0 = 0; s0 < steps; s0 += kStepGroup) {
        int P[kStepGroup];
        bool E[kStepGroup];
#pragma unroll
        for (int i = 0; i < kStepGroup; ++i) {
            const int s = s0 + i;
            // ---- assignment of idle lanes
            unsigned long long idle = __ballot(cnt == 0);
            while (idle) {
                if (qh == qn) {
                    // refill the piece queue
                    qh = qn = 0;
                    while (qn == 0 && !rows_done) {
                        int r = next_row + lane;
                        bool inrange = (r < a.M) && (rp[min(r, a.M)] < e);
                        int st = 0, ln = 0, kd = 0;
                        if (inrange) {
                            int r0 = rp[r], r1 = rp[r + 1];
                            st = max(r0, b);
                            ln = min(r1, e) - st;
                            kd = (r0 < b ? kHead : 0) | (r1 > e ? kTail : 0);
                        }
                        bool ne = inrange && ln > 0;
                        unsigned long long m = __ballot(ne);
                        if (ne) {
                            int k = mbcnt64(m);
                            q_start[wv][k] = st;
                            q_len[wv][k] = ln;
                            q_row[wv][k] = r;
                            q_kind[wv][k] = kd;
                        }
                        qn = __popcll(m);
                        unsigned long long in = __ballot(inrange);
                        if (in != ~0ull) rows_done = true;  // a row at/after the chunk end was seen
                        next_row += 64;
                    }
                    if (qn == 0 && !pad_done) {
                        if (lane == 0) {
                            q_start[wv][0] = a.nnz;
                            q_len[wv][0] = pad;
                            q_row[wv][0] = -1;
                            q_kind[wv][0] = kPad;
                        }
                        qn = 1;
                        pad_done = true;
                    }
                    wave_lds_sync();
                    if (qn == 0) break;  // no pieces left: tail
                }
                const int avail = qn - qh;
                const bool me_idle = (idle >> lane) & 1ull;
                const int rank = mbcnt64(idle);
                if (me_idle && rank < avail) {
                    int k = qh + rank;
                    pos = q_start[wv][k];
                    cnt = q_len[wv][k];
                    prow = q_row[wv][k];
                    pkind = q_kind[wv][k];
                    split = (pkind != 0);
                    sstart = pos;
                }
                unsigned long long taken = __ballot(me_idle && rank < avail);
                qh += __popcll(taken);
                idle &= ~taken;
            }
            // ---- tail: steal for the remaining idle lanes, in lane order
            while (idle) {
                const int l = __ffsll((long long)idle) - 1;
                const int R = steps - s;
                unsigned long long cand = __ballot(cnt > R);
                if (cand == 0) {
                    // unreachable (sum(cnt) = 64 R while lane l holds 0); keep every
                    // position inside the padding range and report it
                    if (lane == 0) atomicOr(&a.status[1], 4);
                    if (lane == l) {
                        pos = a.nnz;
                        cnt = R;
                        prow = -1;
                        pkind = kPad;
                        split = true;
                        sstart = pos;
                    }
                    idle &= idle - 1;
                    continue;
                }
                const int v = __ffsll((long long)cand) - 1;
                const int pv = __shfl(pos, v), cv = __shfl(cnt, v);
                const int rv = __shfl(prow, v), kv = __shfl(pkind, v);
                if (lane == l) {
                    pos = pv + cv - R;
                    cnt = R;
                    prow = rv;
                    pkind = kv;
                    split = true;
                    sstart = pos;
                }
                if (lane == v) {
                    cnt = cv - R;
                    split = true;
                }
                idle &= idle - 1;
            }
            // ---- emit one element per lane
            P[i] = pos;
            E[i] = (cnt == 1);
            ++pos;
            --cnt;
            // ---- segment ends, in lane order
            const bool ended = (cnt == 0);
            unsigned long long mend = __ballot(ended);
            if (mend) {
                unsigned long long msp = __ballot(ended && split);
                if (ended) {
                    int idx = segcount + mbcnt64(mend);
                    if (!split) {
                        a.seg[seg_base + idx] = prow;
                    } else {
                        int k = nsp + mbcnt64(msp);
                        if (k < kSlotCap) {
                            sp_start[wv][k] = sstart;
                            sp_idx[wv][k] = idx;
                            sp_row[wv][k] = prow;
                            sp_kind[wv][k] = pkind;
                        }
                    }
                }
                segcount += __popcll(mend);
                nsp += __popcll(msp);
            }
        }
        // ---- data movement for the batch
        int cv[kStepGroup];
        double vv[kStepGroup];
#pragma unroll
        for (int i = 0; i < kStepGroup; ++i) {
            const bool real = P[i] < a.nnz;
            // padding elements read a column inside the chunk's range (value 0,
            // sum discarded): column 0 for 32-bit layouts, offset 0 for 16-bit
            cv[i] = real ? a.ci[P[i]] : (COL16 ? cbase : 0);
            vv[i] = real ? reinterpret_cast<const double *>(a.va)[P[i]] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < kStepGroup; ++i) {
            const int s = s0 + i;
            ov[(size_t)s * 64 + lane] = cvt_val<VT>(vv[i]);
            if (COL16) {
                reinterpret_cast<unsigned short *>(a.ocols)[(size_t)b + (size_t)s * 64 + lane] =
                    (unsigned short)(cv[i] - cbase);
                unsigned long long em = __ballot(E[i]);
                if (lane == 0) a.end_mask[(size_t)b / 64 + s] = em;
            } else {
                reinterpret_cast<int *>(a.ocols)[(size_t)b + (size_t)s * 64 + lane] = cv[i] | (E[i] ? kEndFlag : 0);
            }
        }
    }

    // ---------------------------------------------------- finalisation
    if (nsp > kSlotCap) {
        if (lane == 0) atomicOr(&a.status[1], 1);
        return;
    }
    if (lane == 0) {
        a.seg_count[c] = segcount;
        atomicMax(&a.status[0], nsp);
    }
    wave_lds_sync();
    // rank of each split segment by start position (starts are distinct)
    for (int k = lane; k < nsp; k += 64) {
        const int sk = sp_start[wv][k];
        int rank = 0;
        for (int j = 0; j < nsp; ++j) rank += (sp_start[wv][j] < sk);
        a.seg[seg_base + sp_idx[wv][k]] = ~rank;
        srt_row[wv][rank] = sp_row[wv][k];
        srt_kind[wv][rank] = sp_kind[wv][k];
    }
    wave_lds_sync();
    int2 *tab = a.split_tab + (size_t)c * kSplitCap;
    int gbase = 0;
    for (int i0 = 0; i0 < nsp; i0 += 64) {
        const int i = i0 + lane;
        bool newg = false;
        int code = 0;
        if (i < nsp) {
            newg = (i == 0) || (srt_row[wv][i] != srt_row[wv][i - 1]);
            const int kd = srt_kind[wv][i];
            code = (kd & kPad) ? kCodeDiscard : (kd & kTail) ? kCodeTail : (kd & kHead) ? kCodeHead : srt_row[wv][i];
        }
        unsigned long long mg = __ballot(newg);
        if (newg) {
            const int g = gbase + mbcnt64(mg);
            if (g < kSplitCap - 1) tab[g] = make_int2(code, i);
        }
        gbase += __popcll(mg);
    }
    if (lane == 0) {
        if (gbase > kSplitCap - 1) {
            atomicOr(&a.status[1], 2);
            gbase = kSplitCap - 1;
        }
        tab[gbase] = make_int2(0, nsp);
        a.split_n[c] = gbase;
    }
}

// ---------------------------------------------------------------- hot columns
// Column use counts (GPU), the most used columns in descending use order, and
// the layout rewrite that renumbers those columns into the dense hot-x buffer
// (bit 30 of a column marks them). Hot number 0 is the most used column, so
// the first entries of the buffer are the ones worth keeping in LDS.
// col_hist_kernel (matrices wider than the banded count below handles): counts
// go through a direct-mapped LDS table first (a column claims a free slot of
// its hash once per workgroup), so a heavily used column costs one global
// atomic per workgroup instead of one per use; a column whose slot is held by
// another column counts globally.
constexpr int kHistSlots = 4096;

__global__ __launch_bounds__(256) void col_hist_kernel(const int *__restrict__ ci, long long nnz, int *__restrict__ cnt) {
    __shared__ int key[kHistSlots];
    __shared__ int val[kHistSlots];
    for (int i = threadIdx.x; i < kHistSlots; i += blockDim.x) key[i] = -1, val[i] = 0;
    __syncthreads();
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
         k += (long long)gridDim.x * blockDim.x) {
        const int col = ci[k];
        const int slot = (int)(((unsigned)col * 0x9E3779B1u) >> 20);  // 12-bit multiplicative hash
        int held = key[slot];
        if (held == -1) {
            held = atomicCAS(&key[slot], -1, col);
            if (held == -1) held = col;
        }
        if (held == col)
            atomicAdd(&val[slot], 1);
        else
            atomicAdd(&cnt[col], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHistSlots; i += blockDim.x)
        if (val[i]) atomicAdd(&cnt[key[i]], val[i]);
}

// Histogram of the column counts: bin v counts the columns used v times
// (bins saturate at kCountBins - 1). Small counts (most columns) go to an LDS
// copy per workgroup, larger ones straight to global memory.
constexpr int kCountBins = 65536;
constexpr int kCountLds = 2048;

__global__ __launch_bounds__(256) void count_hist_kernel(const int *__restrict__ cnt, int N, int *__restrict__ hist) {
    __shared__ int lh[kCountLds];
    for (int i = threadIdx.x; i < kCountLds; i += blockDim.x) lh[i] = 0;
    __syncthreads();
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
        const int v = min(cnt[c], kCountBins - 1);
        if (v < kCountLds)
            atomicAdd(&lh[v], 1);
        else
            atomicAdd(&hist[v], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kCountLds; i += blockDim.x)
        if (lh[i]) atomicAdd(&hist[i], lh[i]);
}

// Order-preserving compaction of the columns used at least `lo` times:
// pass 1 counts per block of kCompact columns, pass 2 (after an exclusive
// scan of the block counts) writes (column, count) in column order.
constexpr int kCompact = 1024;

__global__ __launch_bounds__(256) void compact_count_kernel(const int *__restrict__ cnt, int N, int lo,
                                                            int *__restrict__ bcount) {
    __shared__ int part[4];
    const int c0 = blockIdx.x * kCompact;
    int k = 0;
    for (int i = threadIdx.x; i < kCompact; i += blockDim.x) k += (c0 + i < N && cnt[c0 + i] >= lo) ? 1 : 0;
    for (int off = 32; off > 0; off >>= 1) k += __shfl_xor(k, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = k;
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

__global__ __launch_bounds__(256) void compact_write_kernel(const int *__restrict__ cnt, int N, int lo,
                                                            const int *__restrict__ boff, int2 *__restrict__ out) {
    __shared__ int wsum[4];
    const int c0 = blockIdx.x * kCompact;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int base = boff[blockIdx.x];
    // four rounds of 256 columns, each in thread order
    for (int r = 0; r < kCompact / 256; ++r) {
        const int c = c0 + r * 256 + threadIdx.x;
        const int v = c < N ? cnt[c] : 0;
        const bool f = c < N && v >= lo;
        const unsigned long long m = __ballot(f);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wv] = __popcll(m);
        __syncthreads();
        int wbase = base;
        for (int w = 0; w < wv; ++w) wbase += wsum[w];
        if (f) out[wbase + before] = make_int2(c, v);
        base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
}

// Column counts without global atomics, for N <= kMaxBands * kBand columns.
// (1) Each of kBandBlocks element ranges counts its elements per band of kBand
// columns (LDS); (2) an exclusive scan over (band, range) gives every range
// its write offset per band; (3) the ranges scatter their columns, as 16-bit
// offsets within the band, into band order; (4) one workgroup per slice of a
// band histograms its elements in LDS (all kBand counters of the band) and
// writes them out; (5) the slices of each band are added per column.
constexpr int kBandBits = 15;
constexpr int kBand = 1 << kBandBits;  // columns per band: 128 KiB of LDS counters
constexpr int kBandBlocks = 1024;      // element ranges of passes 1 and 3
constexpr int kMaxBands = 4096;
constexpr long long kSlice = 1 << 18;  // elements per workgroup of pass 4

__global__ __launch_bounds__(256) void band_count_kernel(const int *__restrict__ ci, long long nnz, int nb,
                                                         int *__restrict__ bcnt) {
    extern __shared__ int bh[];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) bh[i] = 0;
    __syncthreads();
    const long long per = (nnz + gridDim.x - 1) / gridDim.x;
    const long long b0 = per * blockIdx.x, b1 = b0 + per < nnz ? b0 + per : nnz;
    for (long long e = b0 + threadIdx.x; e < b1; e += blockDim.x) atomicAdd(&bh[ci[e] >> kBandBits], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) bcnt[(long long)i * gridDim.x + blockIdx.x] = bh[i];
}

// In-place exclusive scan of n ints by one workgroup (1024 threads, each a
// contiguous run; the 1024 run sums are scanned serially).
__global__ __launch_bounds__(1024) void exclusive_scan_kernel(int *__restrict__ a, long long n) {
    __shared__ long long part[1024];
    const long long per = (n + 1023) / 1024;
    const long long b0 = per * threadIdx.x, b1 = b0 + per < n ? b0 + per : n;
    long long s = 0;
    for (long long i = b0; i < b1; ++i) s += a[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long acc = 0;
        for (int i = 0; i < 1024; ++i) {
            const long long v = part[i];
            part[i] = acc;
            acc += v;
        }
    }
    __syncthreads();
    long long acc = part[threadIdx.x];
    for (long long i = b0; i < b1; ++i) {
        const int v = a[i];
        a[i] = (int)acc;
        acc += v;
    }
}

__global__ __launch_bounds__(256) void band_scatter_kernel(const int *__restrict__ ci, long long nnz, int nb,
                                                           const int *__restrict__ off, unsigned short *__restrict__ loc) {
    extern __shared__ int cur[];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) cur[i] = off[(long long)i * gridDim.x + blockIdx.x];
    __syncthreads();
    const long long per = (nnz + gridDim.x - 1) / gridDim.x;
    const long long b0 = per * blockIdx.x, b1 = b0 + per < nnz ? b0 + per : nnz;
    for (long long e = b0 + threadIdx.x; e < b1; e += blockDim.x) {
        const int c = ci[e];
        const int pos = atomicAdd(&cur[c >> kBandBits], 1);
        loc[pos] = (unsigned short)(c & (kBand - 1));
    }
}

__global__ __launch_bounds__(256) void band_hist_kernel(const unsigned short *__restrict__ loc,
                                                        const long long *__restrict__ sb, int *__restrict__ partial) {
    extern __shared__ int sh[];
    for (int i = threadIdx.x; i < kBand; i += blockDim.x) sh[i] = 0;
    __syncthreads();
    const long long e0 = sb[2 * blockIdx.x], e1 = sb[2 * blockIdx.x + 1];
    for (long long e = e0 + threadIdx.x; e < e1; e += blockDim.x) atomicAdd(&sh[loc[e]], 1);
    __syncthreads();
    int *__restrict__ out = partial + (long long)blockIdx.x * kBand;
    for (int i = threadIdx.x; i < kBand; i += blockDim.x) out[i] = sh[i];
}

__global__ void band_reduce_kernel(const int *__restrict__ partial, const int *__restrict__ first_slice, int N,
                                   int *__restrict__ cnt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const int b = c >> kBandBits, l = c & (kBand - 1);
    int s = 0;
    for (int k = first_slice[b]; k < first_slice[b + 1]; ++k) s += partial[(long long)k * kBand + l];
    cnt[c] = s;
}

// cnt[c] = uses of column c among ci[0, nnz).
static hipError_t column_counts(const int *ci, long long nnz, int N, int *cnt, hipStream_t st) {
    hipError_t e;
    const int nb = (int)(((long long)N + kBand - 1) >> kBandBits);
    if (nb > kMaxBands || nnz <= 0) {  // very wide matrices: the hashed atomic count
        if ((e = hipMemsetAsync(cnt, 0, (size_t)N * 4, st)) != hipSuccess) return e;
        hipLaunchKernelGGL(col_hist_kernel, dim3(2048), dim3(256), 0, st, ci, nnz, cnt);
        return hipGetLastError();
    }
    // the pass-4 workgroups hold a band's counters in LDS
    if ((e = hipFuncSetAttribute(reinterpret_cast<const void *>(&band_hist_kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, kBand * 4)) != hipSuccess)
        return e;
    const long long nbc = (long long)nb * kBandBlocks;
    int *bcnt = nullptr, *partial = nullptr, *first = nullptr;
    unsigned short *loc = nullptr;
    long long *sb = nullptr;
    auto done = [&](hipError_t r) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(bcnt);
        (void)hipFree(loc);
        (void)hipFree(partial);
        (void)hipFree(first);
        (void)hipFree(sb);
        return r;
    };
    if ((e = hipMalloc(&bcnt, (size_t)nbc * 4)) != hipSuccess) return done(e);
    if ((e = hipMalloc(&loc, (size_t)nnz * 2)) != hipSuccess) return done(e);
    hipLaunchKernelGGL(band_count_kernel, dim3(kBandBlocks), dim3(256), (size_t)nb * 4, st, ci, nnz, nb, bcnt);
    hipLaunchKernelGGL(exclusive_scan_kernel, dim3(1), dim3(1024), 0, st, bcnt, nbc);
    if ((e = hipGetLastError()) != hipSuccess) return done(e);
    // band starts: the offsets of range 0 of every band
    std::vector<int> off((size_t)nbc);
    if ((e = hipMemcpyAsync(off.data(), bcnt, (size_t)nbc * 4, hipMemcpyDeviceToHost, st)) != hipSuccess) return done(e);
    hipLaunchKernelGGL(band_scatter_kernel, dim3(kBandBlocks), dim3(256), (size_t)nb * 4, st, ci, nnz, nb, bcnt, loc);
    if ((e = hipGetLastError()) != hipSuccess) return done(e);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return done(e);
    std::vector<long long> slices;  // (begin, end) per slice, bands in order
    std::vector<int> first_h((size_t)nb + 1, 0);
    for (int b = 0; b < nb; ++b) {
        const long long s0 = off[(size_t)b * kBandBlocks];
        const long long s1 = b + 1 < nb ? off[(size_t)(b + 1) * kBandBlocks] : nnz;
        first_h[(size_t)b] = (int)(slices.size() / 2);
        for (long long x0 = s0; x0 < s1; x0 += kSlice) {
            slices.push_back(x0);
            slices.push_back(x0 + kSlice < s1 ? x0 + kSlice : s1);
        }
    }
    const int ns = (int)(slices.size() / 2);
    first_h[(size_t)nb] = ns;
    if ((e = hipMalloc(&partial, (size_t)std::max(ns, 1) * kBand * 4)) != hipSuccess) return done(e);
    if ((e = hipMalloc(&first, ((size_t)nb + 1) * 4)) != hipSuccess) return done(e);
    if ((e = hipMalloc(&sb, (size_t)std::max(ns, 1) * 16)) != hipSuccess) return done(e);
    if ((e = hipMemcpyAsync(first, first_h.data(), ((size_t)nb + 1) * 4, hipMemcpyHostToDevice, st)) != hipSuccess)
        return done(e);
    if (ns > 0) {
        if ((e = hipMemcpyAsync(sb, slices.data(), (size_t)ns * 16, hipMemcpyHostToDevice, st)) != hipSuccess)
            return done(e);
        hipLaunchKernelGGL(band_hist_kernel, dim3(ns), dim3(256), (size_t)kBand * 4, st, loc, sb, partial);
    }
    hipLaunchKernelGGL(band_reduce_kernel, dim3((N + 255) / 256), dim3(256), 0, st, partial, first, N, cnt);
    return done(hipGetLastError());
}

__global__ void hot_scatter_kernel(const int *__restrict__ hot_cols, int n, int *__restrict__ colmap) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) colmap[hot_cols[k]] = k;
}

__global__ void remap_hot_kernel(int *__restrict__ cols, long long L, const int *__restrict__ colmap) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < L; e += (long long)gridDim.x * blockDim.x) {
        const int v = cols[e];
        const int m = colmap[v & kColMask];
        if (m >= 0) cols[e] = (v & kEndFlag) | kHotFlag | m;
    }
}

hipError_t launch_remap_hot(int *cols, long long n, const int *colmap, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(remap_hot_kernel, dim3(2048), dim3(256), 0, st, cols, n, colmap);
    return hipGetLastError();
}

// Selection of the max_hot most used columns (each used at least twice),
// ordered by count descending, then column ascending. From the number of
// columns per count value: every column with count >= t is taken, and of the
// columns with count part = t - 1 the first part_take in column order.
struct Selection {
    int t = 0, part = 0, n = 0;
    long long part_take = 0;
};

static Selection select_hot(const std::vector<long long> &nwith, int maxc, int max_hot) {
    Selection s;
    long long taken = 0;
    int t = maxc + 1;
    while (t > 2 && taken + nwith[(size_t)t - 1] <= max_hot) taken += nwith[(size_t)--t];
    s.t = t;
    s.part = t - 1;
    s.part_take = s.part >= 2 ? std::min<long long>(max_hot - taken, nwith[(size_t)s.part]) : 0;
    s.n = (int)(taken + s.part_take);
    return s;
}

// Skew of the column use: uses of the most used eighth of the used columns
// (at most the selection) against an even spread over the used columns.
static bool skewed(const std::vector<long long> &nwith, int maxc, int N, long long nnz, int n, double min_skew) {
    const long long used = (long long)N - nwith[0];
    const long long K = std::max<long long>(1, std::min<long long>(n, used / 8));
    long long left = K;
    double top_uses = 0;
    for (int v = maxc; v >= 1 && left > 0; --v) {
        const long long k = std::min<long long>(left, nwith[(size_t)v]);
        top_uses += (double)v * (double)k;
        left -= k;
    }
    return top_uses * (double)used >= min_skew * (double)K * (double)nnz;
}

// Upload the ordered hot columns and renumber them in the layout.
static hipError_t finish_hot(const std::vector<int> &hc, int N, int *colmap, int *hot_cols, int *cols, long long L,
                             int *n_hot, hipStream_t st) {
    const int n = (int)hc.size();
    hipError_t e;
    if ((e = hipMemcpyAsync(hot_cols, hc.data(), (size_t)n * 4, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(colmap, 0xff, (size_t)N * 4, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(hot_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, st, hot_cols, n, colmap);
    hipLaunchKernelGGL(remap_hot_kernel, dim3(2048), dim3(256), 0, st, cols, L, colmap);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    *n_hot = n;
    return hipSuccess;
}

// The same selection from all the counts copied to the host (small matrices,
// and counts beyond the histogram's range).
static hipError_t hot_columns_host(const int *cnt, int N, int max_hot, double min_skew, long long nnz, int *colmap,
                                   int *hot_cols, int *cols, long long L, int *n_hot, hipStream_t st) {
    hipError_t e;
    std::vector<int> h((size_t)N);
    if ((e = hipMemcpyAsync(h.data(), cnt, (size_t)N * 4, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    int maxc = 0;
    for (int c = 0; c < N; ++c) maxc = std::max(maxc, h[(size_t)c]);
    std::vector<long long> nwith((size_t)maxc + 2, 0);
    for (int c = 0; c < N; ++c) ++nwith[(size_t)h[(size_t)c]];
    const Selection sel = select_hot(nwith, maxc, max_hot);
    if (sel.n == 0) return hipSuccess;
    if (min_skew > 0 && !skewed(nwith, maxc, N, nnz, sel.n, min_skew)) return hipSuccess;
    std::vector<long long> pos((size_t)maxc + 2, 0);
    long long acc = 0;
    for (int v = maxc; v >= std::max(sel.part, 2); --v) pos[(size_t)v] = acc, acc += nwith[(size_t)v];
    std::vector<int> hc((size_t)sel.n);
    long long part_used = 0;
    for (int c = 0; c < N; ++c) {
        const int v = h[(size_t)c];
        if (v >= sel.t || (v == sel.part && v >= 2 && part_used < sel.part_take)) {
            if (v == sel.part && v < sel.t) ++part_used;
            hc[(size_t)pos[(size_t)v]++] = c;
        }
    }
    return finish_hot(hc, N, colmap, hot_cols, cols, L, n_hot, st);
}

hipError_t launch_hot_columns(const int *ci, long long nnz, int N, int max_hot, double min_skew, int *cnt, int *colmap,
                              int *hot_cols, int *cols, long long L, int *n_hot, hipStream_t st) {
    *n_hot = 0;
    hipError_t e;
    if ((e = column_counts(ci, nnz, N, cnt, st)) != hipSuccess) return e;
    // histogram of the counts (colmap is free until the scatter below)
    int *hist = colmap;
    if (N < kCountBins) return hot_columns_host(cnt, N, max_hot, min_skew, nnz, colmap, hot_cols, cols, L, n_hot, st);
    if ((e = hipMemsetAsync(hist, 0, (size_t)kCountBins * 4, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(count_hist_kernel, dim3(1024), dim3(256), 0, st, cnt, N, hist);
    std::vector<long long> nwith(kCountBins, 0);
    {
        std::vector<int> h32(kCountBins);
        if ((e = hipMemcpyAsync(h32.data(), hist, (size_t)kCountBins * 4, hipMemcpyDeviceToHost, st)) != hipSuccess)
            return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        for (int v = 0; v < kCountBins; ++v) nwith[(size_t)v] = h32[(size_t)v];
    }
    int maxc = kCountBins - 1;
    while (maxc > 0 && nwith[(size_t)maxc] == 0) --maxc;
    // columns used kCountBins - 1 times or more share the last bin: they are
    // few, and are taken first, ordered by their own counts below
    const Selection sel = select_hot(nwith, maxc, max_hot);
    if (sel.n == 0) return hipSuccess;
    if (sel.t >= kCountBins || sel.part >= kCountBins - 1)  // the last bin not taken whole
        return hot_columns_host(cnt, N, max_hot, min_skew, nnz, colmap, hot_cols, cols, L, n_hot, st);
    if (min_skew > 0 && !skewed(nwith, maxc, N, nnz, sel.n, min_skew)) return hipSuccess;
    // the candidates (count >= the lowest count taken), in column order
    const int lo = sel.part_take > 0 ? sel.part : sel.t;
    const int nb = (N + kCompact - 1) / kCompact;
    int *bcount = hist;  // reuse: nb <= N / 1024 + 1 ints
    hipLaunchKernelGGL(compact_count_kernel, dim3(nb), dim3(256), 0, st, cnt, N, lo, bcount);
    std::vector<int> boff((size_t)nb + 1, 0);
    if ((e = hipMemcpyAsync(boff.data() + 1, bcount, (size_t)nb * 4, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    for (int b = 0; b < nb; ++b) boff[(size_t)b + 1] += boff[(size_t)b];
    const int m = boff[(size_t)nb];
    int2 *cand = nullptr;
    if ((e = hipMalloc(&cand, (size_t)std::max(m, 1) * 8)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(bcount, boff.data(), (size_t)nb * 4, hipMemcpyHostToDevice, st)) == hipSuccess) {
        hipLaunchKernelGGL(compact_write_kernel, dim3(nb), dim3(256), 0, st, cnt, N, lo, bcount, cand);
        e = hipGetLastError();
    }
    std::vector<int2> hcand((size_t)m);
    if (e == hipSuccess) e = hipMemcpyAsync(hcand.data(), cand, (size_t)m * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(cand);
    if (e != hipSuccess) return e;
    // stable counting sort of the candidates by count, descending; the
    // saturated-bin columns first, sorted by (count descending, column)
    std::vector<int2> big;
    for (const int2 &cv : hcand)
        if (cv.y >= kCountBins - 1) big.push_back(cv);
    std::sort(big.begin(), big.end(), [](const int2 &a, const int2 &b) { return a.y != b.y ? a.y > b.y : a.x < b.x; });
    std::vector<long long> pos((size_t)kCountBins, 0);
    long long acc = (long long)big.size();
    for (int v = std::min(maxc, kCountBins - 2); v >= lo; --v) pos[(size_t)v] = acc, acc += nwith[(size_t)v];
    std::vector<int> hc((size_t)sel.n);
    for (size_t i = 0; i < big.size(); ++i) hc[i] = big[i].x;
    long long part_used = 0;
    for (const int2 &cv : hcand) {
        const int v = cv.y;
        if (v >= kCountBins - 1) continue;
        if (v >= sel.t) {
            hc[(size_t)pos[(size_t)v]++] = cv.x;
        } else if (part_used < sel.part_take) {  // v == part: the first part_take in column order
            ++part_used;
            hc[(size_t)pos[(size_t)v]++] = cv.x;
        }
    }
    return finish_hot(hc, N, colmap, hot_cols, cols, L, n_hot, st);
}

// Automatic column width: *flag becomes nonzero if the nonzeros of some chunk
// (the CSR elements [C c, C (c+1)), one workgroup per chunk) span 65,536
// columns or more, so that 16-bit chunk-relative offsets cannot encode them.
__global__ __launch_bounds__(256) void chunk_span_kernel(const int *__restrict__ ci, long long nnz, long long C,
                                                         int *__restrict__ flag) {
    __shared__ int smn[4], smx[4], done;
    // a chunk already found too wide settles the answer: later workgroups stop
    if (threadIdx.x == 0) done = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (done) return;
    const long long b = C * blockIdx.x, e = b + C < nnz ? b + C : nnz;
    int mn = 0x7fffffff, mx = -1;
    for (long long k = b + threadIdx.x; k < e; k += blockDim.x) {
        const int v = ci[k];
        mn = min(mn, v);
        mx = max(mx, v);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        mn = min(mn, __shfl_xor(mn, d));
        mx = max(mx, __shfl_xor(mx, d));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) smn[w] = mn, smx[w] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 4; ++i) mn = min(mn, smn[i]), mx = max(mx, smx[i]);
        if (mx - mn >= 65536) atomicOr(flag, 1);
    }
}

// Host-side launchers --------------------------------------------------------
hipError_t launch_chunk_span(const int *ci, long long nnz, long long C, int nchunks, int *flag, hipStream_t st) {
    if (nchunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(chunk_span_kernel, dim3((unsigned)nchunks), dim3(256), 0, st, ci, nnz, C, flag);
    return hipGetLastError();
}

// A CSR handed over in device memory: rp[0] == 0, rp[M] == nnz, rp monotone,
// every column in [0, N). *bad |= 1 (offsets) or 2 (columns); *bad is
// cleared by the caller.
__global__ void validate_csr_kernel(const int *__restrict__ rp, const int *__restrict__ ci, int M, int N, long long nnz,
                                    int *bad) {
    int f = 0;
    const long long n = (long long)M > nnz ? (long long)M : nnz;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (i < M && rp[i + 1] < rp[i]) f |= 1;
        if (i < nnz) {
            const int c = ci[i];
            if (c < 0 || c >= N) f |= 2;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (rp[0] != 0 || (long long)rp[M] != nnz)) f |= 1;
    if (f) atomicOr(bad, f);
}

hipError_t launch_validate_csr(const int *rp, const int *ci, int M, int N, long long nnz, int *bad, hipStream_t st) {
    long long n = (long long)M > nnz ? (long long)M : nnz;
    long long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(validate_csr_kernel, dim3((unsigned)blocks), dim3(256), 0, st, rp, ci, M, N, nnz, bad);
    return hipGetLastError();
}

hipError_t launch_empty_rows(const int *rp, int M, unsigned *bits, int *status, hipStream_t st) {
    int nw = (M + 31) >> 5;
    if (nw == 0) return hipSuccess;
    hipLaunchKernelGGL(empty_rows_kernel, dim3((nw + 255) / 256), dim3(256), 0, st, rp, M, bits, status);
    return hipGetLastError();
}

hipError_t launch_convert(const ConvertArgs &a, hipStream_t st) {
    if (a.nchunks == 0) return hipSuccess;
    dim3 grid((a.nchunks + kWavesPerBlock - 1) / kWavesPerBlock), block(64 * kWavesPerBlock);
    if (a.col16) {
        switch (a.val_kind) {
        case 0: hipLaunchKernelGGL((convert_kernel<double, true>), grid, block, 0, st, a); break;
        case 1: hipLaunchKernelGGL((convert_kernel<float, true>), grid, block, 0, st, a); break;
        default: hipLaunchKernelGGL((convert_kernel<unsigned short, true>), grid, block, 0, st, a); break;
        }
    } else {
        switch (a.val_kind) {
        case 0: hipLaunchKernelGGL((convert_kernel<double, false>), grid, block, 0, st, a); break;
        case 1: hipLaunchKernelGGL((convert_kernel<float, false>), grid, block, 0, st, a); break;
        default: hipLaunchKernelGGL((convert_kernel<unsigned short, false>), grid, block, 0, st, a); break;
        }
    }
    return hipGetLastError();
}

}  // namespace cvr
