// cvr_spmv.hip -- CVR-64 SpMV kernels for gfx950.
//
// Replaces pava's spmv_compute_kernel (cvr/cvr.h:1062-1884): one wavefront
// walks one chunk of the CVR-64 layout step by step. Per step each lane loads
// one value and one column (coalesced: the step is 64 consecutive elements),
// gathers x, and does one FMA into its own accumulator (pava: the 8-lane
// AVX-512 FMA of phase 3, cvr/cvr.h:1512-1598). A lane whose column carries the
// end flag closes its segment: the wave ballots the flags, and each ending
// lane picks its destination from the chunk's segment list at
// (segments ended so far + its rank among this step's ending lanes) -- a row
// of y (pava's record-driven store, cvr/cvr.h:1303-1372) or an LDS slot for a
// segment of a split row. Each 8-step group's segment entries (exactly as
// many as its ends) are prefetched with the previous group's loads and staged
// in LDS. After the last step the wave sums the slots of each split row
// in element order (pava's t_rets / final_2 merge, cvr/cvr.h:1823-1868):
// rows wholly inside the chunk are written, the partial of a row shared with
// a neighbouring chunk goes to a per-chunk carry, which the fixup kernel adds
// in chunk order (pava uses omp atomic +=, cvr/cvr.h:1413-1421, 1717, 1855;
// here the SpMV is deterministic: the same inputs give bit-identical y).
// Finally the wave zeroes the empty rows of its share of the empty-row bitmap
// (pava leaves them unwritten and relies on the caller's zeroing,
// main.cpp:561).

#include <hip/hip_runtime.h>

#include "cvr_kernels.h"

namespace cvr {



__device__ __forceinline__ int mbcnt64s(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of chunks so neighbouring chunks (shared rows, nearby x)
// meet in the same L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    const int q = nblocks >> 3, r = nblocks & 7;
    const int x = bid & 7, i = bid >> 3;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

template <typename VT>
__device__ __forceinline__ float val_to_f32(VT v);
template <>
__device__ __forceinline__ float val_to_f32<float>(float v) { return v; }
template <>
__device__ __forceinline__ float val_to_f32<unsigned short>(unsigned short v) {
    return __uint_as_float((unsigned)v << 16);
}

// Accumulator: FP64 plain FMA chain (bit-identical to a sequential fma sum),
// or FP32 with compensated summation (TwoProduct by FMA + TwoSum, i.e. the
// Ogita-Rump-Oishi Dot2 scheme) for the reduced-precision storage.
template <typename VT, typename XT, bool COMP, bool KAPPA>
struct Acc;

template <>
struct Acc<double, double, false, false> {
    double s = 0.0;
    __device__ __forceinline__ void add(double v, double x) { s = __builtin_fma(v, x, s); }
    __device__ __forceinline__ double value() const { return s; }
    __device__ __forceinline__ double absval() const { return 0.0; }
    __device__ __forceinline__ void reset() { s = 0.0; }
};

// FP64 with serial parity: the same FMA chain, plus q = sum of the squared
// partial sums (the bound terms of the segment; see par_ok).
template <>
struct Acc<double, double, false, true> {
    double s = 0.0, q = 0.0;
    __device__ __forceinline__ void add(double v, double x) {
        s = __builtin_fma(v, x, s);
        q = __builtin_fma(s, s, q);
    }
    __device__ __forceinline__ double value() const { return s; }
    __device__ __forceinline__ double absval() const { return q; }
    __device__ __forceinline__ void reset() { s = 0.0, q = 0.0; }
};

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// x of a column of the FP64 CSR copy kept for the parity recompute. With hot
// staging (hot = 1) the copy's columns carry the layout's renumbering
// (kHotFlag | hot index, applied at handle creation), so that the recompute
// reads a hot column where the SpMV's own gathers do -- the LDS table (TAB:
// the LDS variant, hot numbers below n_lds) or the hot-x copy, whose lines
// the SpMV keeps in L2 -- instead of lines of x that only the recompute
// touches; the values are the same. A table column still issues its global
// load, to the hot-x copy's first element, which all such lanes of an
// instruction share: one line request instead of one per lane. Nothing is
// loaded under a branch (the compiler then waits for every load in flight at
// the join) and bases are chosen by a bit mask (a select of two pointers was
// lowered to an indexed scratch load). TAB: get2 returns both values and the
// choice, which the chain makes where it uses the value (a select right
// after the loads waits for them there). R-MAT scale 22 FP64 with parity:
// 0.4167 -> 0.4141 ms per step (five alternating timings).
template <bool TAB>
struct XSrcT {
    static constexpr bool kTab = TAB;
    uint64_t ax, ah;    // x, the hot-x copy
    const double *tab;  // TAB: the LDS table
    int n_lds, hot;
    __device__ __forceinline__ void get2(int cv, double &g, double &t, bool &in_tab) const {
        if (!hot) {
            g = load_global<double>(ax + (uint64_t)(unsigned)cv * sizeof(double));
            t = 0.0;
            in_tab = false;
            return;
        }
        static_assert(kHotFlag == (1 << 30), "bit 30 flags a hot column");
        const uint32_t idx = (uint32_t)(cv & kColMask);
        const uint32_t h = ((unsigned)cv >> 30) & 1u;
        const bool in = h && idx < (uint32_t)n_lds;
        t = tab[in ? idx : 0u];
        const uint64_t m = 0ull - (uint64_t)h;  // all ones for a hot column
        const uint64_t base = ax ^ ((ax ^ ah) & m);
        g = load_global<double>(base + (in ? 0ull : (uint64_t)idx * sizeof(double)));
        in_tab = in;
    }
    // !TAB: the value itself
    __device__ __forceinline__ double operator()(int cv) const {
        static_assert(!TAB, "the LDS variant uses get2");
        if (!hot) return load_global<double>(ax + (uint64_t)(unsigned)cv * sizeof(double));
        const uint64_t m = 0ull - (uint64_t)(((unsigned)cv >> 30) & 1u);
        return load_global<double>((ax ^ ((ax ^ ah) & m)) + (uint64_t)(uint32_t)(cv & kColMask) * sizeof(double));
    }
};

// The serial FMA sum of elements [b, e) of the FP64 CSR (the reference's
// order), by one wave. A round covers 64 * K consecutive elements, K per lane
// (lane l holds elements j0 + K l .. j0 + K l + K - 1). The chain runs lane
// by lane: all lanes apply their K FMAs to the running sum s (wave-uniform,
// an SGPR operand), and lane l's result becomes the next s, so the chain
// costs K dependent FMAs and one readlane per lane -- no LDS. While round r
// is chained, the x gathers and values of rounds up to r + D and the columns
// of round r + D + 1 are in flight; nothing waits on a load before the chain
// needs it. Elements past e are masked to (0, 0) when chained, which
// leaves s unchanged. All lanes must call it with the same b < e; the
// result is wave-uniform.
// POLL: `stop` is read once per round (the load overlaps the round's chain)
// and the chain ends early, its result unused, once kSpecOk is set there.
template <int K, int D, bool POLL = false, typename XS>
__device__ __forceinline__ double serial_fma_sum(const int *__restrict__ ci, const double *__restrict__ va,
                                                 const XS &xs, int b, int e, int lane, const int *stop = nullptr) {
    constexpr int R = 64 * K;
    double s = 0.0;
    int cn[K];                         // columns of round r + D + 1
    double vq[D + 1][K], xq[D + 1][K];  // values and x of rounds r .. r + D
    // TAB: the LDS table's values of rounds r .. r + D and whether each
    // element reads the table (XSrcT::get2)
    double tq[XS::kTab ? D + 1 : 1][K];
    bool iq[XS::kTab ? D + 1 : 1][K];
    auto at = [&](int j0, int k) -> unsigned {
        const int j = j0 + lane * K + k;
        return (unsigned)(j < e ? j : e - 1);
    };
    auto load_c = [&](int j0, int (&c)[K]) {
#pragma unroll
        for (int k = 0; k < K; ++k) c[k] = __builtin_nontemporal_load(&ci[at(j0, k)]);
    };
    auto load_v = [&](int j0, double (&v)[K]) {
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = __builtin_nontemporal_load(&va[at(j0, k)]);
    };
    auto gather = [&](const int (&c)[K], int d) {
        if constexpr (XS::kTab) {
#pragma unroll
            for (int k = 0; k < K; ++k) xs.get2(c[k], xq[d][k], tq[d][k], iq[d][k]);
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) xq[d][k] = xs(c[k]);
        }
    };
    // prologue: rounds 0 .. D loaded and gathered, the columns of round D + 1 loading
    static_assert(D >= 1 && D <= 3, "pipeline depth");
#pragma unroll
    for (int d = 0; d <= D; ++d) {
        load_c(b + d * R, cn);
        load_v(b + d * R, vq[d]);
        gather(cn, d);
    }
    load_c(b + (D + 1) * R, cn);
    for (int j0 = b; j0 < e; j0 += R) {
        int flag = 0;
        if constexpr (POLL) flag = __hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int nl = (e - j0 + K - 1) / K;  // lanes holding elements of this round
        const int nl64 = nl < 64 ? nl : 64;
        double vv[K], xx[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const bool in = j0 + lane * K + k < e;
            vv[k] = in ? vq[0][k] : 0.0;
            if constexpr (XS::kTab)
                xx[k] = in ? (iq[0][k] ? tq[0][k] : xq[0][k]) : 0.0;
            else
                xx[k] = in ? xq[0][k] : 0.0;
        }
        // a full round with the lane loop unrolled (no loop control between
        // the hops of the chain), a partial last round rolled
        auto hop = [&](int l) {
            double t = __builtin_fma(vv[0], xx[0], s);
#pragma unroll
            for (int k = 1; k < K; ++k) t = __builtin_fma(vv[k], xx[k], t);
            s = readlane_f64(t, l);
        };
        if (nl64 == 64) {
#pragma unroll
            for (int l = 0; l < 64; ++l) hop(l);
        } else {
            for (int l = 0; l < nl64; ++l) hop(l);
        }
        if constexpr (POLL) {
            if (__builtin_amdgcn_readfirstlane(flag) & kSpecOk) break;
        }
        // shift: round r + D + 1 gathered and its values loaded, the columns
        // of round r + D + 2 loading
        if (j0 + R < e) {
#pragma unroll
            for (int d = 0; d < D; ++d)
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    xq[d][k] = xq[d + 1][k];
                    vq[d][k] = vq[d + 1][k];
                    if constexpr (XS::kTab) {
                        tq[d][k] = tq[d + 1][k];
                        iq[d][k] = iq[d + 1][k];
                    }
                }
            gather(cn, D);
            load_v(j0 + (D + 1) * R, vq[D]);
            load_c(j0 + (D + 2) * R, cn);
        }
    }
    return s;
}

// serial_fma_sum for the plain kernel's long rows, kept out of line: its
// registers (8 elements per lane) are then allocated apart from the SpMV's
// main loop, which would otherwise spill at the plain kernel's 128-VGPR
// budget.
template <int K, int D, bool POLL = false>
__device__ __noinline__ double serial_fma_sum_ool(const int *__restrict__ ci, const double *__restrict__ va,
                                                  const XSrcT<false> xs, int b, int e, int lane, const int *stop = nullptr) {
    return serial_fma_sum<K, D, POLL>(ci, va, xs, b, e, lane, stop);
}

// The serial FMA sum of CSR elements [b, e) for the parity recompute. The LDS
// variant (up to 256 VGPRs) chains 8 elements per lane inline. The plain
// kernel chains rows of up to 2,048 elements with 4 per lane inline, longer
// ones with 8 per lane out of line: fewer lane hops per element (a chain alone
// in one wave: 13.5 cycles per element with 4, 10.1 with 8) where the chain
// sets the kernel's end, with the call's register saves only where they pay
// (selector-family matrix with a 20,000-element cancelling row spanning
// chunks: 0.270 -> 0.229 ms; the 27-point stencil unchanged).
template <bool LDSX, typename ARGS>
__device__ __forceinline__ double serial_chain(const ARGS &a, const XSrcT<LDSX> &xs, int b, int e, int lane) {
    if constexpr (LDSX)
        return serial_fma_sum<8, 1>(a.ci, a.va, xs, b, e, lane);
    else if (e - b > 2048)
        return serial_fma_sum_ool<8, 1>(a.ci, a.va, xs, b, e, lane);
    else
        return serial_fma_sum<4, 1>(a.ci, a.va, xs, b, e, lane);
}

// KAPPA also accumulates sum|v x| (the cancellation scale) for the FP64
// recompute decision.
template <typename VT, bool KAPPA>
struct Acc<VT, float, true, KAPPA> {
    float s = 0.f, c = 0.f, ab = 0.f;
    __device__ __forceinline__ void add(VT vr, float x) {
        const float v = val_to_f32<VT>(vr);
        const float p = v * x;
        const float pe = __builtin_fmaf(v, x, -p);
        const float t = s + p;
        const float z = t - s;
        const float te = (s - (t - z)) + (p - z);
        s = t;
        c += te + pe;
        if (KAPPA) ab += __builtin_fabsf(p);
    }
    __device__ __forceinline__ double value() const { return (double)s + (double)c; }
    __device__ __forceinline__ double absval() const { return (double)ab; }
    __device__ __forceinline__ void reset() { s = 0.f, c = 0.f, ab = 0.f; }
};


template <typename YT>
__device__ __forceinline__ void store_y(void *y, int r, double v) {
    reinterpret_cast<YT *>(y)[r] = (YT)v;
}

// FP64 parity carries (sum, bound terms). A row whose chunks are all run by
// one workgroup passes them through ordinary stores and loads, ordered by
// workgroup-scope fences; a row spanning workgroups passes them with
// agent-scope atomic stores and loads, which go past the XCD's L2, so its
// closer needs no L2 invalidate and its writers no L2 write-back.
__device__ __forceinline__ void put_carry(const SpmvArgs &a, int idx, double s, double A, double QN, bool coherent) {
    if (coherent) {
        double *bb = reinterpret_cast<double *>(a.carry_bnd + idx);
        __hip_atomic_store(&a.carry[idx], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&bb[0], A, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&bb[1], QN, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        a.carry[idx] = s;
        a.carry_bnd[idx] = make_double2(A, QN);
    }
}
__device__ __forceinline__ void get_carry(const SpmvArgs &a, int idx, double *s, double *A, double *QN,
                                          bool coherent) {
    if (coherent) {
        const double *bb = reinterpret_cast<const double *>(a.carry_bnd + idx);
        *s = __hip_atomic_load(&a.carry[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *A = __hip_atomic_load(&bb[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *QN = __hip_atomic_load(&bb[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        *s = a.carry[idx];
        const double2 b = a.carry_bnd[idx];
        *A = b.x, *QN = b.y;
    }
}

// LDSX: kWavesLds-wave workgroups (one per CU) first copy the head of the hot-x
// buffer -- the most used columns, hot numbers [0, n_lds) -- into LDS, then
// each wave walks chunks of the workgroup's contiguous chunk range; a gather
// of such a column is an LDS read instead of a divergent global load.
template <typename VT, typename XT, typename YT, bool COMP, bool COL16, bool KAPPA, bool LDSX>
// 32-bit layouts and FP64 16-bit layouts: >= 4 waves per SIMD (128 VGPRs, no
// spills); the LDS variant runs one workgroup of kWavesLds waves per CU (its
// table fills the LDS), so 2 waves per SIMD and up to 256 VGPRs; the
// compensated 16-bit layouts are left unconstrained
__global__ __launch_bounds__(LDSX ? 64 * kWavesLds : 64 * kWavesPerBlock,
                             LDSX ? kWavesLds / 4 : (COL16 && COMP ? 1 : 4)) void spmv_kernel(
    SpmvArgs a) {
    constexpr int WPB = LDSX ? kWavesLds : kWavesPerBlock;
    const int lane = threadIdx.x & 63;

    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int G = kStepGroup;
    // FP64 with serial parity: split rows carry bound terms, and the rows that
    // fail the bound are recomputed in serial order
    constexpr bool PAR = KAPPA && !COMP;

    __shared__ double slots[WPB][kSlotCap];
    // KAPPA: sum|a x| of each slot (mixed precision), or q|n packed (PAR)
    __shared__ double slots_abs[KAPPA ? WPB : 1][KAPPA ? kSlotCap : 1];
    // group segment entries in the main loop; then (PAR) the list of rows to
    // recompute (ints [0, 128))
    __shared__ __align__(16) int segbuf[WPB][64 * kStepGroup];
    __shared__ int fcnt[PAR ? WPB : 1];
    // PAR, plain variant: arrival counters of the rows whose chunks all run in
    // this workgroup (one per chunk of the workgroup, zero at launch)
    __shared__ int lcnt[PAR && !LDSX ? WPB : 1];
    extern __shared__ __align__(16) unsigned char lds_tab[];
    const XT *__restrict__ tab = reinterpret_cast<const XT *>(lds_tab);
    // PAR: where the serial recompute reads x
    XSrcT<LDSX> xs{};
    if constexpr (PAR)
        xs = XSrcT<LDSX>{reinterpret_cast<uint64_t>(a.x), reinterpret_cast<uint64_t>(a.xhot),
                  LDSX ? reinterpret_cast<const double *>(lds_tab) : nullptr, LDSX ? a.n_lds : 0, a.ci_hot};

    auto run_chunk = [&](const int c) {
        const long long C = 64LL * a.S;
        const size_t base = (size_t)C * (size_t)c;
        const int steps = (c == a.nchunks - 1) ? (int)((a.layout_elems - (long long)base) >> 6) : a.S;
        const int zr = a.chunk_rows[c];
        const int *__restrict__ segs = a.seg + zr + kSegExtra * c;
        const XT *__restrict__ x = reinterpret_cast<const XT *>(a.x);
        const VT *__restrict__ vp = reinterpret_cast<const VT *>(a.vals) + base + lane;
        const int *__restrict__ cp32 = reinterpret_cast<const int *>(a.cols) + base + lane;
        const unsigned short *__restrict__ cp16 = reinterpret_cast<const unsigned short *>(a.cols) + base + lane;
        const uint64_t *__restrict__ emp = a.end_mask + (base >> 6);
        const int cbase = COL16 ? a.col_base[c] : 0;
        // PAR: what the chunk's end needs (row closing, carry scope), read now
        int4 pi0 = make_int4(-1, 0, -1, 0), pi1 = make_int4(-1, 0, 0, 0);
        if constexpr (PAR) {
            pi0 = a.pinfo[2 * (size_t)c];
            pi1 = a.pinfo[2 * (size_t)c + 1];
        }
        // The chunk's end reads its split table and its words of the empty-row
        // bitmap; their first 64 entries are loaded now, with the chunk's first
        // loads, so that the end does not wait on two more dependent round
        // trips (R-MAT scale 22: ~2-3 us each under the SpMV's load, against
        // ~35 us of main loop). kSplitCap > 64, so lane + 1 stays inside the
        // chunk's table.
        const int ns = a.split_n[c];
        const int2 *stab = a.split_tab + (size_t)c * kSplitCap;
        const int2 tab_pre = stab[lane];
        const int se_pre = stab[lane + 1].y;
        const int nw_empty = (a.M + 31) >> 5;
        const int w0 = (int)((long long)nw_empty * c / a.nchunks),
                  w1 = (int)((long long)nw_empty * (c + 1) / a.nchunks);
        unsigned ebits_pre = 0;
        if (a.has_empty && w0 + lane < w1) ebits_pre = a.empty_bits[w0 + lane];

        // Pipeline, one 8-step group deep: while group g is processed, the x
        // gathers and the segment-list entries of group g+1 (exactly as many as
        // its row ends, <= 512, issued together with the gathers), the values of
        // group g+1 and the columns of group g+2 are in flight. A group's segment
        // entries are copied to LDS when its processing starts, so each ending
        // lane finds its destination with one LDS read and never waits on memory.
        // The layout streams (values, columns, segment entries) are read once:
        // non-temporal loads keep them from evicting x lines from the caches.
        const XT *__restrict__ xh = reinterpret_cast<const XT *>(a.xhot);
        // x element of a stored column: hot columns read the dense hot-x copy
        // (an explicit 64-bit address select: a select of two pointers would be
        // lowered to an indexed private array)
        const uint64_t ax = reinterpret_cast<uint64_t>(x), ah = reinterpret_cast<uint64_t>(xh);
        auto xval = [&](int cv) -> XT {
            if constexpr (COL16) {
                return load_global<XT>(ax + (uint64_t)(uint32_t)(cbase + cv) * sizeof(XT));
            } else {
                const uint32_t idx = (uint32_t)(cv & kColMask);
                if constexpr (LDSX) {
                    if ((cv & kHotFlag) && idx < (uint32_t)a.n_lds) return tab[idx];
                }
                const uint64_t b = (cv & kHotFlag) ? ah : ax;
                return load_global<XT>(b + (uint64_t)idx * sizeof(XT));
            }
        };
        // 16-bit layouts: the G end masks of a group arrive in one vector load,
        // mask i in lane i, and are read out with readlane when used (a
        // wave-uniform load would be turned into a readfirstlane right after
        // it, which waits for every load issued before it)
        auto step_mask = [&](uint64_t em, int i) -> uint64_t {
            const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)em, i);
            const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(em >> 32), i);
            return ((uint64_t)hi << 32) | lo;
        };
        auto load_cols = [&](int g0, int (&cc)[G], uint64_t &em) {
            if constexpr (COL16) {
#pragma unroll
                for (int i = 0; i < G; ++i) cc[i] = (int)cp16[(size_t)(g0 + i) * 64];
                em = emp[g0 + (lane & (G - 1))];
            } else {
#pragma unroll
                for (int i = 0; i < G; ++i) cc[i] = __builtin_nontemporal_load(&cp32[(size_t)(g0 + i) * 64]);
            }
        };
        auto group_ends = [&](const int (&cc)[G], uint64_t em) -> int {
            int e = 0;
            if constexpr (COL16) {
#pragma unroll
                for (int i = 0; i < G; ++i) e += __popcll(step_mask(em, i));
            } else {
#pragma unroll
                for (int i = 0; i < G; ++i) e += __popcll(__ballot(cc[i] < 0));
            }
            return e;
        };
        auto load_segs = [&](int sb, int ne, int (&r)[G]) {
#pragma unroll
            for (int j = 0; j < G; ++j)
                if (j * 64 < ne) r[j] = __builtin_nontemporal_load(&segs[sb + j * 64 + lane]);
        };

        Acc<VT, XT, COMP, KAPPA> acc;
        VT vb[G], vn[G];
        int cb[G], cn[G], cnn[G];
        uint64_t eb = 0, en = 0, enn = 0;
        XT xb[G], xn[G];
        int sc[G], sn[G];
        int st0 = 0;  // PAR: first step of the lane's current segment

        // prologue: group 0 fully issued, columns of group 1 in flight
        load_cols(0, cb, eb);
#pragma unroll
        for (int i = 0; i < G; ++i) vb[i] = __builtin_nontemporal_load(&vp[(size_t)i * 64]);
#pragma unroll
        for (int i = 0; i < G; ++i) xb[i] = xval(cb[i]);
        int ecur = group_ends(cb, eb);
        load_segs(0, ecur, sc);
        if (G < steps) load_cols(G, cn, en);
        // drain the prologue's loads: the loop is then entered with nothing in
        // flight, so the compiler's wait counts inside it follow the loop's own
        // loads only (otherwise the entry state makes the first group wait for
        // most of the next group's loads)
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
        int sbase = 0, segp = 0;

        // one group: its destinations to LDS, then its G steps
        auto run_group = [&](const int g) {
#pragma unroll
            for (int j = 0; j < G; ++j)
                if (j * 64 < ecur) segbuf[wv][j * 64 + lane] = sc[j];
            wave_sync_lds();
#pragma unroll
            for (int i = 0; i < G; ++i) {
                acc.add(vb[i], xb[i]);
                bool end;
                unsigned long long m;
                if constexpr (COL16) {
                    m = step_mask(eb, i);
                    end = (m >> lane) & 1ull;
                } else {
                    end = cb[i] < 0;
                    m = __ballot(end);
                }
                if (m) {
                    if (end) {
                        const int d = segbuf[wv][segp - sbase + mbcnt64s(m)];
                        const double v = acc.value();
                        if (d >= 0) {
                            store_y<YT>(a.y, d, v);
                            if (KAPPA && COMP && acc.absval() > (double)a.kappa * __builtin_fabs(v)) a.rflag[d] = 1;
                        } else {
                            slots[wv][~d] = v;
                            if (PAR)
                                slots_abs[wv][~d] = par_pack(acc.absval(), g + i + 1 - st0);
                            else if (KAPPA)
                                slots_abs[wv][~d] = acc.absval();
                        }
                        if (PAR) st0 = g + i + 1;
                        acc.reset();
                    }
                    segp += __popcll(m);
                }
            }
        };
        // Every group but the last issues the next group's loads before its
        // own steps and moves them into place after, unconditionally: with the
        // loads and the moves under two tests of one condition, the compiler's
        // wait counting saw a path on which a load stayed in flight into the
        // next group, and drained every load in flight (the next group's
        // gathers included) before reusing its register -- two memory round
        // trips per group instead of one. The last group runs after the loop.
        int g = 0;
        for (; g + G < steps; g += G) {
            const bool has2 = (g + 2 * G) < steps;
            const int sbase_next = sbase + ecur;
#pragma unroll
            for (int i = 0; i < G; ++i) xn[i] = xval(cn[i]);
            const int enext = group_ends(cn, en);
            load_segs(sbase_next, enext, sn);
#pragma unroll
            for (int i = 0; i < G; ++i) vn[i] = __builtin_nontemporal_load(&vp[(size_t)(g + G + i) * 64]);
            if (has2) load_cols(g + 2 * G, cnn, enn);
            run_group(g);
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
            wave_sync_lds();  // the next group overwrites segbuf
        }
        if (g < steps) run_group(g);
        wave_sync_lds();

        // split rows: sum slots in element order
        int *flist = &segbuf[wv][0];
        // PAR: 2 sqrt(n q) of every split slot, by all lanes at once, before the
        // merge: the correctly rounded square root is a long dependent
        // instruction sequence, and the merge is a serial chain per group (on
        // R-MAT scale 22 the longest group of a chunk has 27 segments on
        // average). Kept in segbuf after flist's 128 entries.
        double *troot = reinterpret_cast<double *>(&segbuf[wv][128]);
        static_assert(128 + 2 * kSlotCap <= 64 * kStepGroup, "segbuf holds flist and troot");
        if constexpr (PAR) {
            // slots in use: the split table's sentinel (all kSlotCap beyond 64
            // groups, which are rare)
            const int nsl = ns <= 0 ? 0 : ns <= 64 ? __builtin_amdgcn_readlane(se_pre, ns - 1) : kSlotCap;
            for (int i = lane; i < nsl; i += 64) {
                double qi;
                int ni;
                par_unpack(slots_abs[wv][i], &qi, &ni);
                troot[i] = 2.0 * __builtin_sqrt((double)ni * qi);
            }
        }
        if (PAR && lane == 0) fcnt[wv] = 0;
        wave_sync_lds();
        // one split group per lane: its slots merged in element order
        auto split_group = [&](const int2 t, const int se) {
            if constexpr (PAR) {
                // merged sum and the bound terms of par_ok
                double s = slots[wv][t.y], q0;
                int n0;
                par_unpack(slots_abs[wv][t.y], &q0, &n0);
                double A = troot[t.y], Q = 2.0 * q0;
                int N = n0;
                auto merge = [&](double sj, double pj, double tj) {
                    double qj;
                    int nj;
                    par_unpack(pj, &qj, &nj);
                    A += tj + (double)nj * __builtin_fabs(s);
                    Q += 2.0 * (double)nj * s * s + 3.0 * qj;
                    s += sj;
                    A += __builtin_fabs(s);
                    Q += s * s;
                    N += nj;
                };
                // slots read four at a time, so that a long row's merge waits
                // on LDS once per four segments (same order of operations)
                int j = t.y + 1;
                for (; j + 4 <= se; j += 4) {
                    double sv[4], pv[4], tv[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        sv[u] = slots[wv][j + u], pv[u] = slots_abs[wv][j + u], tv[u] = troot[j + u];
#pragma unroll
                    for (int u = 0; u < 4; ++u) merge(sv[u], pv[u], tv[u]);
                }
                for (; j < se; ++j) merge(slots[wv][j], slots_abs[wv][j], troot[j]);
                if (t.x >= 0) {
                    if (par_ok(s, A, Q, N, a.parity_tol))
                        store_y<YT>(a.y, t.x, s);
                    else
                        flist[atomicAdd(&fcnt[wv], 1)] = t.x;
                } else if (t.x == kCodeHead) {
                    // the row closed by this chunk
                    put_carry(a, 2 * c, s, A, par_pack(Q, N), !(pi0.w & kPiHeadIntra));
                } else if (t.x == kCodeTail) {
                    put_carry(a, 2 * c + 1, s, A, par_pack(Q, N), !(pi0.w & kPiTailIntra));
                }
                return;
            }
            double s = slots[wv][t.y];
            int j = t.y + 1;
            for (; j + 4 <= se; j += 4) {
                double sv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) sv[u] = slots[wv][j + u];
#pragma unroll
                for (int u = 0; u < 4; ++u) s += sv[u];
            }
            for (; j < se; ++j) s += slots[wv][j];
            double sa = 0.0;
            if (KAPPA)
                for (int j = t.y; j < se; ++j) sa += slots_abs[wv][j];
            if (t.x >= 0) {
                store_y<YT>(a.y, t.x, s);
                if (KAPPA && sa > (double)a.kappa * __builtin_fabs(s)) a.rflag[t.x] = 1;
            } else if (t.x == kCodeHead) {
                a.carry[2 * c] = s;
                if (KAPPA) a.carry_abs[2 * c] = sa;
            } else if (t.x == kCodeTail) {
                a.carry[2 * c + 1] = s;
                if (KAPPA) a.carry_abs[2 * c + 1] = sa;
            }
        };
        // the first 64 groups from the registers loaded at the chunk's start;
        // more (rare) are loaded here, in a loop of their own, so that the
        // common path has no load to wait for (a wait on a load also waits
        // for every store the main loop left in flight)
        if (lane < ns) split_group(tab_pre, se_pre);
        for (int k = lane + 64; k < ns; k += 64) split_group(stab[k], stab[k + 1].y);
        if constexpr (PAR) {
            // rows of this chunk whose CVR-order sum may differ from the
            // serial sum by more than parity_tol: one serial FMA chain each
            wave_sync_lds();
            const int nf = __builtin_amdgcn_readfirstlane(fcnt[wv]);
            for (int f = 0; f < nf; ++f) {
                const int row = __builtin_amdgcn_readfirstlane(flist[f]);
                const int b = a.rp[row], e = a.rp[row + 1];
                const double v = serial_chain<LDSX>(a, xs, b, e, lane);
                if (lane == 0) store_y<YT>(a.y, row, v);
            }
            if (lane == 0) a.rc_count[c] = nf;

            // rows spanning chunks: this chunk's head carry belongs to the row
            // the chunk closes, its tail carry to the row closed by chunk
            // pi0.z. Each contributing chunk counts itself in; the last one to
            // arrive sums the carries in chunk order with their bound terms
            // (carry: sum |S + w| <= N |S| + A, sum (S + w)^2 <= 2 N S^2 +
            // 2 Q) and writes the row, recomputing it serially when par_ok
            // fails. No wave waits for another; a global counter is reset by
            // its closer for the next SpMV.
            // the carries are in place before this chunk counts itself in:
            // the workgroup-scope release for rows inside the workgroup, and
            // the completion of the agent-scope carry stores for rows spanning
            // workgroups
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            {
                const bool span_head = pi0.x >= 0 && !(pi0.w & kPiHeadIntra);
                const bool span_tail = pi0.z >= 0 && !(pi0.w & kPiTailIntra);
                if (span_head || span_tail) __builtin_amdgcn_s_waitcnt(kWaitVm0);
            }
            // the chunk's two count-ins (head row: counter c, tail row: counter
            // pi0.z) are issued together, by lanes 0 and 1, so that they cost
            // one round trip; rows inside one workgroup of the plain variant
            // count in LDS (chunk k is then the workgroup's chunk k % WPB)
            auto close_chunk = [&](int t) { return t == 0 ? (pi0.x >= 0 ? c : -1) : pi0.z; };
            auto close_intra = [&](int t) { return (pi0.w & (t == 0 ? kPiHeadIntra : kPiTailIntra)) != 0; };
            int mine = 0;
            if (lane < 2) {
                const int k = close_chunk(lane);
                if (k >= 0)
                    mine = (!LDSX && close_intra(lane))
                               ? atomicAdd(&lcnt[k % WPB], 1)
                               : __hip_atomic_fetch_add(&a.close_cnt[k], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            const int olds[2] = {__builtin_amdgcn_readlane(mine, 0), __builtin_amdgcn_readlane(mine, 1)};
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int k = close_chunk(t);
                if (k < 0) continue;
                const int row = t == 0 ? pi0.x : pi1.x;
                const int c0 = t == 0 ? pi0.y : pi1.y;
                const bool intra = close_intra(t);
                const bool local = !LDSX && intra;
                if (olds[t] != k - c0) continue;
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                // the carries of chunks c0 .. k (tail carries, then chunk k's
                // head carry), up to 64 at a time one per lane, so that they
                // arrive in one round trip; then merged in chunk order
                double s = 0.0, A = 0.0, Q = 0.0;
                int N = 0;
                const int ncar = k - c0 + 1;
                for (int b0 = 0; b0 < ncar; b0 += 64) {
                    double cs = 0.0, cA = 0.0, cQN = 0.0;
                    if (b0 + lane < ncar) {
                        const int j = c0 + b0 + lane;
                        get_carry(a, j < k ? 2 * j + 1 : 2 * k, &cs, &cA, &cQN, !intra);
                    }
                    const int nb = ncar - b0 < 64 ? ncar - b0 : 64;
                    for (int i = 0; i < nb; ++i) {
                        const double sj = readlane_f64(cs, i), Aj = readlane_f64(cA, i);
                        double Qj;
                        int Nj;
                        par_unpack(readlane_f64(cQN, i), &Qj, &Nj);
                        if (b0 + i == 0) {
                            s = sj, A = Aj, Q = Qj, N = Nj;
                            continue;
                        }
                        A += Aj + (double)Nj * __builtin_fabs(s);
                        Q += 2.0 * Qj + 2.0 * (double)Nj * s * s;
                        s += sj;
                        A += __builtin_fabs(s);
                        Q += s * s;
                        N += Nj;
                    }
                }
                const bool ok = par_ok(s, A, Q, N, a.parity_tol);
                const int sp = LDSX ? 0 : (t == 0 ? pi1.w : pi1.z);  // speculative chain + 1
                if (sp > 0) {
                    // the row's serial chain started at kernel start: tell it the
                    // outcome; whichever side comes second writes y if the
                    // serial sum is needed and clears the state word
                    if (lane == 0) {
                        int *stp = a.spec_state + (sp - 1);
                        if (ok) store_y<YT>(a.y, row, s);
                        const int old = __hip_atomic_fetch_or(stp, ok ? kSpecOk : kSpecNeed, __ATOMIC_ACQ_REL,
                                                              __HIP_MEMORY_SCOPE_AGENT);
                        if (old & kSpecDone) {
                            if (!ok)
                                store_y<YT>(a.y, row,
                                            __hip_atomic_load(&a.spec_val[sp - 1], __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT));
                            __hip_atomic_store(stp, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        }
                    }
                } else {
                    if (!ok) s = serial_chain<LDSX>(a, xs, a.rp[row], a.rp[row + 1], lane);
                    if (lane == 0) store_y<YT>(a.y, row, s);
                }
                if (lane == 0) {
                    a.rc_count[a.nchunks + k] = ok ? 0 : 1;
                    if (!local) __hip_atomic_store(&a.close_cnt[k], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }

        // empty rows: any wave may zero any empty row, so the bitmap words are
        // split evenly over the waves (not by the chunks' row ranges, which are
        // very uneven on power-law matrices)
        if (a.has_empty) {
            // one row per lane, so that each store instruction covers 64
            // consecutive rows (lanes 0-31 the rows of word w, lanes 32-63 those
            // of word w + 1) instead of one scattered row per lane: the
            // memory pipeline is the SpMV's bottleneck, and every store
            // instruction queues in it
            auto zero_pair = [&](int w, unsigned lo, unsigned hi) {
                if (lo | hi) {
                    const int r = (w << 5) + lane;
                    if ((((lane < 32 ? lo : hi) >> (lane & 31)) & 1u) && r < a.M) store_y<YT>(a.y, r, 0.0);
                }
            };
            // words w0 .. w0 + 63 from the register loaded at the chunk's start
            // (zero past w1); more (over 2,048 rows per chunk on average)
            // loaded in a loop of their own, so that the common path has no
            // load to wait for
            const int nwc = w1 - w0;
            for (int i = 0; i < nwc && i < 64; i += 2)
                zero_pair(w0 + i, (unsigned)__builtin_amdgcn_readlane((int)ebits_pre, i),
                          (unsigned)__builtin_amdgcn_readlane((int)ebits_pre, i + 1));
            for (int w = w0 + 64; w < w1; w += 2) zero_pair(w, a.empty_bits[w], w + 1 < w1 ? a.empty_bits[w + 1] : 0u);
        }
        wave_sync_lds();  // the next chunk reuses this wave's slots
    };

    if constexpr (LDSX) {
        // Workgroup w's chunks: first its list wg_list[wg_start[w] ..
        // wg_start[w + 1]) -- a contiguous range, or with FP64 parity one that
        // begins with the chunks of the long rows assigned to it (long rows
        // are spread over the workgroups, so that their serial recomputes,
        // when needed, start early and in parallel) -- popped by its waves
        // from an LDS counter (a wave busy with a serial recompute leaves the
        // rest to the others); then the shared queue wg_list[wg_start[G] ..
        // nchunks) (the second half of every workgroup's range, interleaved),
        // popped by every workgroup from one global counter. Recomputes
        // cluster in some ranges (R-MAT scale 22 with parity: workgroups
        // finished between 365 and 443 us with private ranges only); the queue
        // lets the others take over the ends of those ranges. Each wave asks
        // for its next queue item before running the current one, so the
        // atomic's round trip overlaps the chunk. The last workgroup to finish
        // zeroes the two counters (wg_next[0]: queue items handed out,
        // wg_next[1]: workgroups finished) for the next SpMV.
        __shared__ int next_chunk;
        const int nwg = (int)gridDim.x;
        const int lb = xcd_remap(blockIdx.x, nwg);
        const int c_lo = a.wg_start[lb], c_end = a.wg_start[lb + 1];
        const int q0 = a.wg_start[nwg], nq = a.nchunks - q0;
        XT *t = reinterpret_cast<XT *>(lds_tab);
        const XT *__restrict__ src = reinterpret_cast<const XT *>(a.xhot);
        for (int i = threadIdx.x; i < a.n_lds; i += blockDim.x) t[i] = src[i];
        if (threadIdx.x == 0) next_chunk = c_lo;
        __syncthreads();
        for (;;) {
            int c = 0;
            if (lane == 0) c = atomicAdd(&next_chunk, 1);
            c = __builtin_amdgcn_readfirstlane(c);
            if (c >= c_end) break;
            run_chunk(a.wg_list[c]);
        }
        int q = 0;
        if (lane == 0) q = atomicAdd(&a.wg_next[0], 1);
        q = __builtin_amdgcn_readfirstlane(q);
        while (q < nq) {
            int qn = 0;
            if (lane == 0) qn = atomicAdd(&a.wg_next[0], 1);
            run_chunk(a.wg_list[q0 + q]);
            q = __builtin_amdgcn_readfirstlane(qn);
        }
        __syncthreads();
        if (threadIdx.x == 0 && atomicAdd(&a.wg_next[1], 1) == nwg - 1) {
            a.wg_next[0] = 0;
            a.wg_next[1] = 0;
        }
    } else {
        if constexpr (PAR) {
            // the first spec_blocks workgroups (dispatched first) run the
            // speculative serial chains, one row per workgroup: chains that
            // shared a CU ran up to 3.5x slower than alone (on a matrix with
            // 133 such rows of 20,000 elements: 0.193 ms with 4 per workgroup,
            // 0.172 with 1); the other waves of such a workgroup exit
            if ((int)blockIdx.x < a.spec_blocks) {
                const int i = (int)blockIdx.x;
                if (wv == 0 && i < a.n_spec) {
                    const int row = a.spec_rows[i];
                    int *stp = a.spec_state + i;
                    const double v = serial_fma_sum_ool<8, 1, true>(a.ci, a.va, xs, a.rp[row], a.rp[row + 1], lane, stp);
                    if (lane == 0) {
                        __hip_atomic_store(&a.spec_val[i], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        const int old = __hip_atomic_fetch_or(stp, kSpecDone, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
                        if (old & (kSpecNeed | kSpecOk)) {
                            if (old & kSpecNeed) store_y<YT>(a.y, row, v);
                            __hip_atomic_store(stp, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        }
                    }
                }
                return;  // the whole workgroup: no barrier follows for it
            }
        }
        const int sb = PAR ? a.spec_blocks : 0;  // a multiple of 8: the XCD of a block is unchanged
        const int g = xcd_remap((int)blockIdx.x - sb, (int)gridDim.x - sb) * WPB;
        const int c = g + wv;
        if constexpr (PAR) {
            if (threadIdx.x < WPB) lcnt[threadIdx.x] = 0;
            __syncthreads();
        }
        if (c < a.nchunks) run_chunk(c);
    }
}

// Rows spanning several chunks: y[row] = tail partial of its first chunk +
// partials of the chunks it covers entirely + head partial of its last chunk,
// added in chunk order.
template <typename YT>
__global__ void fixup_kernel(const int *__restrict__ fix, const double *__restrict__ carry, void *y, int nchunks) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    const int row = fix[2 * c];
    if (row < 0) return;
    const int c0 = fix[2 * c + 1];
    double s = carry[2 * c0 + 1];
    for (int k = c0 + 1; k < c; ++k) s += carry[2 * k + 1];
    s += carry[2 * c];
    reinterpret_cast<YT *>(y)[row] = (YT)s;
}

// len[c] = length of the row chunk c closes (rows spanning chunks), or -1
__global__ void closer_len_kernel(const int *__restrict__ fix, const int *__restrict__ rp, int *__restrict__ len,
                                  int nchunks) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    const int r = fix[2 * c];
    len[c] = r >= 0 ? rp[r + 1] - rp[r] : -1;
}

template <typename YT>
__global__ void zero_kernel(void *y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) reinterpret_cast<YT *>(y)[i] = (YT)0;
}

// Mixed precision, after the SpMV. Waves [0, ceil(nchunks / 64)) take 64
// chunks each, one per lane: each lane closes its chunk's cross-chunk row (as
// fixup_kernel), and the rows whose cancellation ratio exceeds kappa are
// recomputed in FP64. The following waves take kFinishRows rows each, recompute
// the rows the SpMV flagged from the resident FP64 CSR and FP64 x, and clear
// the flags for the next SpMV. Every reduction has a fixed order
// (deterministic).
__device__ __forceinline__ double wave_sum_f64(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Elements [b, e) of the FP64 CSR: lane-strided FMA chains (loads issued 8
// elements per lane at a time), then a fixed xor-butterfly reduction.
__device__ __forceinline__ double range_fp64(const int *__restrict__ ci, const double *__restrict__ va,
                                             const double *__restrict__ x64, int b, int e, int lane) {
    double s = 0.0;
    for (int j0 = b + lane; j0 < e; j0 += 8 * 64) {
        int cc[8];
        double vv[8], xx[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = j0 + u * 64;
            cc[u] = j < e ? ci[j] : 0;
            vv[u] = j < e ? va[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) xx[u] = x64[cc[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (j0 + u * 64 < e) s = __builtin_fma(vv[u], xx[u], s);
    }
    return wave_sum_f64(s);
}

constexpr int kStage = 512;  // elements of short rows staged in LDS per pass (8 per lane)

struct FinishLds {
    int flist[kFinishRows];  // flagged rows of the wave's block
    int soff[64];            // staged offset of each lane's row (exclusive prefix)
    int sb[64];              // first element of each lane's row
    double sv[kStage], sx[kStage];
};

// y[rows[i]] for i < n <= 64, in FP64. Rows of at most kLongRow elements: lane
// i owns rows[i]; the values and x of all of them are loaded side by side
// (kStage elements per pass) into LDS, then each owning lane runs its row's
// FMA chain in element order -- the sequential FMA sum of the row, rounded
// once to FP32 -- and clears the row's flag. Longer rows are left flagged (or
// flagged) for the long-row kernels. Returns the number of rows recomputed.
__device__ int recompute_rows(const FinishArgs &a, const int *rows, int n, int lane, FinishLds &L) {
    int row = -1, b = 0, len = 0;
    if (lane < n) {
        row = rows[lane];
        b = a.rp[row];
        len = a.rp[row + 1] - b;
    }
    const bool lng = len > kLongRow;
    const int sl = lng ? 0 : len;
    int inc = sl;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(inc, d);
        if (lane >= d) inc += t;
    }
    const int off = inc - sl;
    const int T = __shfl(inc, 63);
    L.soff[lane] = off;
    L.sb[lane] = b;
    wave_sync_lds();
    double s = 0.0;
    for (int p0 = 0; p0 < T; p0 += kStage) {
        int cc[8];
        double vv[8], xx[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = p0 + lane + 64 * u;
            cc[u] = 0;
            vv[u] = 0.0;
            if (e < T) {
                // owner: the last lane whose staged range starts at or before e
                int k = 0;
#pragma unroll
                for (int st = 32; st > 0; st >>= 1)
                    if (L.soff[k + st] <= e) k += st;
                const int j = L.sb[k] + (e - L.soff[k]);
                cc[u] = a.ci[j];
                vv[u] = a.va[j];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) xx[u] = a.x64[cc[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = p0 + lane + 64 * u;
            if (e < T) {
                L.sv[e - p0] = vv[u];
                L.sx[e - p0] = xx[u];
            }
        }
        wave_sync_lds();
        const int lo = off > p0 ? off : p0;
        const int hi = off + sl < p0 + kStage ? off + sl : p0 + kStage;
        int e = lo;
        for (; e + 4 <= hi; e += 4) {  // LDS reads of 4 elements issued together
            double v[4], w[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = L.sv[e + u - p0], w[u] = L.sx[e + u - p0];
#pragma unroll
            for (int u = 0; u < 4; ++u) s = __builtin_fma(v[u], w[u], s);
        }
        for (; e < hi; ++e) s = __builtin_fma(L.sv[e - p0], L.sx[e - p0], s);
        wave_sync_lds();  // the next pass overwrites the stage
    }
    if (lane < n) {
        if (lng) {
            a.rflag[row] = 1;
        } else {
            a.y[row] = (float)s;
            a.rflag[row] = 0;
        }
    }
    return __popcll(__ballot(lane < n && !lng));
}

__global__ __launch_bounds__(64 * kWavesPerBlock) void finish_kernel(FinishArgs a) {
    __shared__ FinishLds lds[kWavesPerBlock];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gw = blockIdx.x * kWavesPerBlock + wv;
    FinishLds &L = lds[wv];
    const int nfw = finish_fix_waves(a.nchunks);
    if (gw < nfw) {
        const int c = 64 * gw + lane;
        int frow = -1;
        double s = 0.0, sa = 0.0;
        if (c < a.nchunks) frow = a.fix[2 * c];
        if (frow >= 0) {
            const int c0 = a.fix[2 * c + 1];
            s = a.carry[2 * c0 + 1], sa = a.carry_abs[2 * c0 + 1];
            for (int k = c0 + 1; k < c; ++k) s += a.carry[2 * k + 1], sa += a.carry_abs[2 * k + 1];
            s += a.carry[2 * c];
            sa += a.carry_abs[2 * c];
        }
        const bool need = frow >= 0 && a.x64 && sa > (double)a.kappa * __builtin_fabs(s);
        if (frow >= 0 && !need) a.y[frow] = (float)s;
        const unsigned long long m = __ballot(need);
        if (need) L.flist[mbcnt64s(m)] = frow;
        const int nf = __popcll(m);
        wave_sync_lds();
        const int done = nf ? recompute_rows(a, L.flist, nf, lane, L) : 0;
        if (lane == 0) a.rc_count[gw] = done;
        return;
    }
    const int rw = gw - nfw;
    if (rw >= finish_row_blocks(a.M)) return;
    // flags of rows [r0, r0 + kFinishRows): 16 per lane in two loads (the flag
    // buffer is padded to a whole number of blocks)
    const int rb = rw * kFinishRows + 16 * lane;
    uint64_t f0 = 0, f1 = 0;
    if (rb < a.M) {
        f0 = *reinterpret_cast<const uint64_t *>(a.rflag + rb);
        f1 = *reinterpret_cast<const uint64_t *>(a.rflag + rb + 8);
    }
    unsigned bits = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        bits |= ((f0 >> (8 * i)) & 0xffull) ? (1u << i) : 0u;
        bits |= ((f1 >> (8 * i)) & 0xffull) ? (1u << (8 + i)) : 0u;
    }
    if (!a.x64) {  // nothing to recompute with: drop the flags
        if (f0) *reinterpret_cast<uint64_t *>(a.rflag + rb) = 0ull;
        if (f1) *reinterpret_cast<uint64_t *>(a.rflag + rb + 8) = 0ull;
        bits = 0;
    }
    const int cnt = __popc(bits);
    int inc = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(inc, d);
        if (lane >= d) inc += t;
    }
    const int nf = __shfl(inc, 63);
    int pos = inc - cnt;
    while (bits) {
        L.flist[pos++] = rb + __ffs(bits) - 1;
        bits &= bits - 1;
    }
    wave_sync_lds();
    int done = 0;
    for (int g = 0; g < nf; g += 64) done += recompute_rows(a, L.flist + g, nf - g < 64 ? nf - g : 64, lane, L);
    if (lane == 0) a.rc_count[gw] = done;
}

// Long flagged rows: one wave per piece of kPiece elements computes the
// piece's FP64 sum; then one thread per long row adds its pieces in order,
// stores the row and clears its flag.
__global__ __launch_bounds__(64 * kWavesPerBlock) void long_piece_kernel(LongArgs a) {
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (p >= a.n_pieces) return;
    const int li = a.piece_row[p];
    const int r = a.long_rows[li];
    if (!a.rflag[r]) return;
    const int b = a.rp[r] + (p - a.poff[li]) * kPiece;
    const int re = a.rp[r + 1];
    const double s = range_fp64(a.ci, a.va, a.x64, b, re - b > kPiece ? b + kPiece : re, lane);
    if (lane == 0) a.partial[p] = s;
}

__global__ void long_sum_kernel(LongArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_long) return;
    const int r = a.long_rows[t];
    int done = 0;
    if (a.rflag[r]) {
        double s = a.partial[a.poff[t]];
        for (int k = a.poff[t] + 1; k < a.poff[t + 1]; ++k) s += a.partial[k];
        a.y[r] = (float)s;
        a.rflag[r] = 0;
        done = 1;
    }
    a.rc_long[t] = done;
}

// xhot[k] = x[hot_cols[k]] (before each SpMV on layouts with hot columns)
template <typename XT>
__global__ void hot_gather_kernel(const XT *__restrict__ x, const int *__restrict__ hot_cols, XT *__restrict__ xhot,
                                  int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) xhot[k] = x[hot_cols[k]];
}

hipError_t launch_hot_gather(const void *x, const int *hot_cols, void *xhot, int n_hot, int xsize, hipStream_t st) {
    if (n_hot <= 0) return hipSuccess;
    dim3 grid((n_hot + 255) / 256), block(256);
    if (xsize == 8)
        hipLaunchKernelGGL(hot_gather_kernel<double>, grid, block, 0, st, (const double *)x, hot_cols, (double *)xhot, n_hot);
    else
        hipLaunchKernelGGL(hot_gather_kernel<float>, grid, block, 0, st, (const float *)x, hot_cols, (float *)xhot, n_hot);
    return hipGetLastError();
}

// --------------------------------------------------------------- launchers
template <typename VT, typename XT, typename YT, bool COMP, bool KAPPA>
static void launch_one(const SpmvArgs &a, dim3 grid, hipStream_t st) {
    if (a.n_lds > 0)
        hipLaunchKernelGGL((spmv_kernel<VT, XT, YT, COMP, false, KAPPA, true>), dim3(a.lds_blocks),
                           dim3(64 * kWavesLds), (size_t)a.n_lds * sizeof(XT), st, a);
    else
        hipLaunchKernelGGL((spmv_kernel<VT, XT, YT, COMP, false, KAPPA, false>), grid, dim3(64 * kWavesPerBlock), 0,
                           st, a);
}

template <typename VT, bool COL16>
static void launch_mixed(const SpmvArgs &a, dim3 grid, hipStream_t st) {
    if constexpr (COL16) {
        if (a.recompute)
            hipLaunchKernelGGL((spmv_kernel<VT, float, float, true, true, true, false>), grid,
                               dim3(64 * kWavesPerBlock), 0, st, a);
        else
            hipLaunchKernelGGL((spmv_kernel<VT, float, float, true, true, false, false>), grid,
                               dim3(64 * kWavesPerBlock), 0, st, a);
    } else {
        if (a.recompute)
            launch_one<VT, float, float, true, true>(a, grid, st);
        else
            launch_one<VT, float, float, true, false>(a, grid, st);
    }
}

hipError_t launch_spmv(const SpmvArgs &a, int prec, int col16, hipStream_t st) {
    if (a.nchunks == 0) {
        if (a.M > 0) {
            if (prec == 0)
                hipLaunchKernelGGL(zero_kernel<double>, dim3((a.M + 255) / 256), dim3(256), 0, st, a.y, a.M);
            else
                hipLaunchKernelGGL(zero_kernel<float>, dim3((a.M + 255) / 256), dim3(256), 0, st, a.y, a.M);
        }
        return hipGetLastError();
    }
    if (a.n_lds > 0 && (col16 || a.lds_blocks <= 0 || !a.wg_next || !a.wg_list || !a.wg_start))
        return hipErrorInvalidValue;
    // a renumbered CSR copy needs the hot-x copy its hot columns point into
    if (a.parity_tol > 0.0 && a.ci_hot && (!a.xhot || col16)) return hipErrorInvalidValue;
    if (a.n_spec > 0 && (a.n_lds > 0 || a.parity_tol <= 0.0 || prec != 0 || !a.spec_rows || !a.spec_state ||
                         !a.spec_val || a.spec_blocks < a.n_spec || a.spec_blocks % 8))
        return hipErrorInvalidValue;
    const int sblocks = a.n_spec > 0 ? a.spec_blocks : 0;
    dim3 grid((a.nchunks + kWavesPerBlock - 1) / kWavesPerBlock + sblocks);
    if (prec == 0) {
        const bool par = a.parity_tol > 0.0;
        if (col16 && par)
            hipLaunchKernelGGL((spmv_kernel<double, double, double, false, true, true, false>), grid,
                               dim3(64 * kWavesPerBlock), 0, st, a);
        else if (col16)
            hipLaunchKernelGGL((spmv_kernel<double, double, double, false, true, false, false>), grid,
                               dim3(64 * kWavesPerBlock), 0, st, a);
        else if (par)
            launch_one<double, double, double, false, true>(a, grid, st);
        else
            launch_one<double, double, double, false, false>(a, grid, st);
    } else if (prec == 1) {
        if (col16)
            launch_mixed<float, true>(a, grid, st);
        else
            launch_mixed<float, false>(a, grid, st);
    } else {
        if (col16)
            launch_mixed<unsigned short, true>(a, grid, st);
        else
            launch_mixed<unsigned short, false>(a, grid, st);
    }
    return hipGetLastError();
}

template <typename VT, typename XT, bool COMP, bool KAPPA>
static const void *lds_kernel() {
    return reinterpret_cast<const void *>(&spmv_kernel<VT, XT, XT, COMP, false, KAPPA, true>);
}

hipError_t plan_spmv_lds(int prec, int bound, int n_hot, long long max_bytes, int nchunks, int *n_lds,
                         int *lds_blocks) {
    *n_lds = 0;
    *lds_blocks = 0;
    if (n_hot <= 0 || max_bytes == 0 || nchunks <= 0) return hipSuccess;
    const bool recompute = bound != 0;
    const void *fn = prec == 0   ? (bound ? lds_kernel<double, double, false, true>()
                                          : lds_kernel<double, double, false, false>())
                     : prec == 1 ? (recompute ? lds_kernel<float, float, true, true>()
                                              : lds_kernel<float, float, true, false>())
                                 : (recompute ? lds_kernel<unsigned short, float, true, true>()
                                              : lds_kernel<unsigned short, float, true, false>());
    const int xsize = prec == 0 ? 8 : 4;
    hipError_t e;
    int dev = 0, max_lds = 0, ncu = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    hipFuncAttributes fa{};
    if ((e = hipFuncGetAttributes(&fa, fn)) != hipSuccess) return e;
    long long bytes = (long long)max_lds - (long long)fa.sharedSizeBytes;
    if (max_bytes > 0 && max_bytes < bytes) bytes = max_bytes;
    if ((long long)n_hot * xsize < bytes) bytes = (long long)n_hot * xsize;
    const int n = (int)(bytes / xsize);
    // when sized automatically, a table this small does not repay its load
    if (n < (max_bytes < 0 ? 1024 : 1)) return hipSuccess;
    if ((e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, n * xsize)) != hipSuccess) return e;
    int per_cu = 0;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * kWavesLds, (size_t)n * xsize)) !=
        hipSuccess)
        return e;
    if (per_cu < 1) return hipSuccess;
    // at least one workgroup per CU while there are chunks for it
    long long want = (nchunks + kWavesLds - 1) / kWavesLds;
    if (want < ncu) want = nchunks < ncu ? nchunks : ncu;
    const long long cap = (long long)per_cu * ncu;
    *lds_blocks = (int)(want < cap ? want : cap);
    *n_lds = n;
    return hipSuccess;
}

hipError_t launch_finish(const FinishArgs &a, hipStream_t st) {
    if (a.nchunks == 0) return hipSuccess;
    const long long waves = (long long)finish_fix_waves(a.nchunks) + finish_row_blocks(a.M);
    dim3 grid((unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock)), block(64 * kWavesPerBlock);
    hipLaunchKernelGGL(finish_kernel, grid, block, 0, st, a);
    return hipGetLastError();
}

hipError_t launch_long_rows(const LongArgs &a, hipStream_t st) {
    if (a.n_long == 0) return hipSuccess;
    hipLaunchKernelGGL(long_piece_kernel, dim3((a.n_pieces + kWavesPerBlock - 1) / kWavesPerBlock),
                       dim3(64 * kWavesPerBlock), 0, st, a);
    hipLaunchKernelGGL(long_sum_kernel, dim3((a.n_long + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_closer_len(const int *fix, const int *rp, int *len, int nchunks, hipStream_t st) {
    if (nchunks == 0) return hipSuccess;
    hipLaunchKernelGGL(closer_len_kernel, dim3((nchunks + 255) / 256), dim3(256), 0, st, fix, rp, len, nchunks);
    return hipGetLastError();
}

hipError_t launch_fixup(const int *fix, const double *carry, void *y, int nchunks, int prec, hipStream_t st) {
    if (nchunks == 0) return hipSuccess;
    dim3 grid((nchunks + 255) / 256), block(256);
    if (prec == 0)
        hipLaunchKernelGGL(fixup_kernel<double>, grid, block, 0, st, fix, carry, y, nchunks);
    else
        hipLaunchKernelGGL(fixup_kernel<float>, grid, block, 0, st, fix, carry, y, nchunks);
    return hipGetLastError();
}

}  // namespace cvr
