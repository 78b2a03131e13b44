/*
 * oracle.c -- CPU oracle for the CVR SpMV path.
 *
 * TEST INFRASTRUCTURE ONLY. Loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; the product (pava_amd/) never
 * links, loads or calls it.
 *
 * What each function restates and what pins it (tests/test_oracle.py):
 *   orc_csr_spmv_fma     referenceSpMV, utils/basicMatrix.cpp:538-556.
 *                        Pinned by the 5x5 known answer of the reference's
 *                        ESB example (esb/examples/spmv_new.c:73-93,
 *                        esb/examples/example.exe.out:16-21), Laplacian closed
 *                        forms, dense brute force, scipy.sparse fixtures.
 *   orc_csr_spmv_muladd  same order, unfused; pinned like the above (exact on
 *                        integer data) and by exact agreement with a
 *                        numpy sequential loop.
 *   orc_csr_spmv_dd      pinned by exact rational arithmetic (Python
 *                        fractions) on random rows: the result must be the
 *                        correctly rounded sum within 1 ulp.
 *   orc_row_abs          pinned by numpy |A| @ |x|.
 *   orc_check_results    checkResults, utils/basicMatrix.cpp:559-587; pinned
 *                        by hand-made vectors on both sides of the 1% bound.
 *   orc_cvr64_build/spmv CPU emulation of the GPU CVR-64 layout (a 64-lane
 *                        re-derivation of cvr/cvr.h:583-1060 and 1062-1884).
 *                        Pinned by layout invariants (every nonzero exactly
 *                        once, rows rebuilt equal the CSR, all lanes end on the
 *                        last step, record counts) and by its SpMV equalling
 *                        orc_csr_spmv_fma bit for bit on rows that are not
 *                        split and within the rounding bound on split rows.
 *                        The FP32/BF16 lane accumulation (Dot2: TwoProduct by
 *                        fmaf, TwoSum) is pinned against exact rational row
 *                        sums on strongly cancelling rows, within Dot2's
 *                        published bound, which an uncompensated sum and Dot2
 *                        without TwoProduct both break (tests/test_compensated.py).
 *   orc_csr5_build/spmv  CPU emulation of the GPU CSR5-64 tiles (csr5/); pinned
 *                        by tile invariants and by its SpMV against
 *                        orc_csr_spmv_fma within the rounding bound
 *                        (tests/test_csr5_layout.py).
 *
 * Build with -ffp-contract=off: every fused operation below is an explicit fma.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_csr_spmv_fma(int m, const int *rp, const int *ci, const double *va, const double *x, double *y) {
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) s = fma(va[j], x[ci[j]], s);
        y[i] = s;
    }
}

void orc_csr_spmv_muladd(int m, const int *rp, const int *ci, const double *va, const double *x, double *y) {
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            double p = va[j] * x[ci[j]];
            s = s + p;
        }
        y[i] = s;
    }
}

void orc_csr_spmv_dd(int m, const int *rp, const int *ci, const double *va, const double *x, double *y) {
    for (int i = 0; i < m; ++i) {
        double s = 0.0, c = 0.0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            double a = va[j], b = x[ci[j]];
            double p = a * b;
            double pe = fma(a, b, -p); /* TwoProduct */
            double t = s + p;          /* TwoSum */
            double z = t - s;
            double te = (s - (t - z)) + (p - z);
            s = t;
            c += te + pe;
        }
        y[i] = s + c;
    }
}

void orc_row_abs(int m, const int *rp, const int *ci, const double *va, const double *x, double *s) {
    for (int i = 0; i < m; ++i) {
        double a = 0.0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) a += fabs(va[j] * x[ci[j]]);
        s[i] = a;
    }
}

void orc_csr_spmv_omp(int m, const int *rp, const int *ci, const double *va, const double *x, double *y,
                      int threads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
#endif
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) s = fma(va[j], x[ci[j]], s);
        y[i] = s;
    }
    (void)threads;
}

int orc_check_results(int n, const double *y, const double *y_ref) {
    int errors = 0;
    for (int i = 0; i < n; ++i)
        if (fabs(y_ref[i] - y[i]) > 0.01 * fabs(y_ref[i])) ++errors;
    return errors;
}

/* ------------------------------------------------------------------ CVR-64 */
enum { LANES = 64, GROUP = 8, SPLIT_CAP = 72, SEG_EXTRA = 66, SLOT_CAP = 160 };
enum { K_HEAD = 1, K_TAIL = 2, K_PAD = 4 };
enum { C_HEAD = -1, C_TAIL = -2, C_DISCARD = -3 };

void orc_cvr64_sizes(int m, int nnz, int S, int *nchunks, long long *layout_elems, long long *seg_cap) {
    long long C = 64LL * S;
    int nc = nnz > 0 ? (int)((nnz + C - 1) / C) : 0;
    long long L = 0;
    if (nc > 0) {
        long long nlast = nnz - (long long)(nc - 1) * C;
        long long st = (nlast + 63) / 64;
        st = (st + GROUP - 1) / GROUP * GROUP;
        L = (long long)(nc - 1) * C + 64 * st;
    }
    *nchunks = nc;
    *layout_elems = L;
    *seg_cap = (long long)m + (long long)SEG_EXTRA * nc + 512;
}

static int row_of_element(const int *rp, int m, int b) {
    int lo = 0, hi = m;
    while (hi - lo > 1) {
        int mid = lo + (hi - lo) / 2;
        if (rp[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

static uint16_t f32_to_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

static float bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

typedef struct {
    int start, len, row, kind;
} piece_t;

typedef struct {
    int start, idx, row, kind;
} splitseg_t;

static int cmp_splitseg(const void *a, const void *b) {
    const splitseg_t *x = (const splitseg_t *)a, *y = (const splitseg_t *)b;
    return (x->start > y->start) - (x->start < y->start);
}

int orc_cvr64_build(int m, int n, int nnz, const int *rp, const int *ci, const double *va, int S, int prec,
                    void *ovals, int *ocols, int *chunk_rows, int *seg, int *seg_count, int *split_n,
                    int *split_tab, int *fix, int *info) {
    (void)n;
    int nchunks;
    long long L, segcap;
    orc_cvr64_sizes(m, nnz, S, &nchunks, &L, &segcap);
    const long long C = 64LL * S;
    int max_slots = 0, nfix = 0, has_empty = 0;
    for (int r = 0; r < m; ++r)
        if (rp[r + 1] == rp[r]) has_empty = 1;
    piece_t *pieces = (piece_t *)malloc(sizeof(piece_t) * (size_t)(C + 2));
    splitseg_t sp[SLOT_CAP + LANES];
    for (int c = 0; c < nchunks; ++c) {
        const int b = (int)(C * c);
        const int e = (int)((b + C < nnz) ? b + C : nnz);
        const int nn = e - b;
        int steps = S;
        if (c == nchunks - 1) {
            steps = (nn + 63) / 64;
            steps = (steps + GROUP - 1) / GROUP * GROUP;
        }
        const int pad = 64 * steps - nn;
        const int row_lo = row_of_element(rp, m, b);
        const int zr = c == 0 ? 0 : row_lo;
        const int seg_base = zr + SEG_EXTRA * c;
        chunk_rows[c] = zr;
        if (c == nchunks - 1) chunk_rows[nchunks] = m;
        {
            int head = rp[row_lo] < b, tail = rp[row_lo + 1] > e;
            if (head && !tail) {
                fix[2 * c] = row_lo;
                fix[2 * c + 1] = (int)(rp[row_lo] / C);
                ++nfix;
            } else {
                fix[2 * c] = -1;
                fix[2 * c + 1] = -1;
            }
        }
        /* pieces of the chunk in CSR order, then the padding piece */
        int np = 0;
        for (int r = row_lo; r < m && rp[r] < e; ++r) {
            int st = rp[r] > b ? rp[r] : b;
            int en = rp[r + 1] < e ? rp[r + 1] : e;
            if (en - st <= 0) continue;
            pieces[np].start = st;
            pieces[np].len = en - st;
            pieces[np].row = r;
            pieces[np].kind = (rp[r] < b ? K_HEAD : 0) | (rp[r + 1] > e ? K_TAIL : 0);
            ++np;
        }
        if (pad > 0) {
            pieces[np].start = nnz;
            pieces[np].len = pad;
            pieces[np].row = -1;
            pieces[np].kind = K_PAD;
            ++np;
        }
        int cnt[LANES] = {0}, pos[LANES] = {0}, prow[LANES] = {0}, pkind[LANES] = {0}, sstart[LANES] = {0};
        int split[LANES] = {0};
        int next = 0, segcount = 0, nsp = 0;
        for (int s = 0; s < steps; ++s) {
            for (int l = 0; l < LANES; ++l) {
                if (cnt[l] != 0) continue;
                if (next < np) {
                    pos[l] = pieces[next].start;
                    cnt[l] = pieces[next].len;
                    prow[l] = pieces[next].row;
                    pkind[l] = pieces[next].kind;
                    split[l] = pkind[l] != 0;
                    sstart[l] = pos[l];
                    ++next;
                } else {
                    const int R = steps - s;
                    int v = -1;
                    for (int k = 0; k < LANES; ++k)
                        if (cnt[k] > R) {
                            v = k;
                            break;
                        }
                    if (v < 0) {
                        free(pieces);
                        return -1; /* cannot happen: sum(cnt) = 64 R */
                    }
                    pos[l] = pos[v] + cnt[v] - R;
                    cnt[l] = R;
                    prow[l] = prow[v];
                    pkind[l] = pkind[v];
                    split[l] = 1;
                    sstart[l] = pos[l];
                    cnt[v] -= R;
                    split[v] = 1;
                }
            }
            for (int l = 0; l < LANES; ++l) {
                const int p = pos[l];
                const int endf = cnt[l] == 1;
                const long long o = (long long)b + (long long)s * 64 + l;
                const double v = p < nnz ? va[p] : 0.0;
                const int cc = p < nnz ? ci[p] : 0;
                if (prec == 0)
                    ((double *)ovals)[o] = v;
                else if (prec == 1)
                    ((float *)ovals)[o] = (float)v;
                else
                    ((uint16_t *)ovals)[o] = f32_to_bf16((float)v);
                ocols[o] = (int)((uint32_t)cc | (endf ? 0x80000000u : 0u));
                ++pos[l];
                --cnt[l];
            }
            for (int l = 0; l < LANES; ++l) {
                if (cnt[l] != 0) continue;
                const int idx = segcount++;
                if (!split[l]) {
                    seg[seg_base + idx] = prow[l];
                } else {
                    if (nsp >= SLOT_CAP + LANES) {
                        free(pieces);
                        return -2;
                    }
                    sp[nsp].start = sstart[l];
                    sp[nsp].idx = idx;
                    sp[nsp].row = prow[l];
                    sp[nsp].kind = pkind[l];
                    ++nsp;
                }
            }
        }
        if (nsp > SLOT_CAP) {
            free(pieces);
            return -2;
        }
        seg_count[c] = segcount;
        if (nsp > max_slots) max_slots = nsp;
        qsort(sp, (size_t)nsp, sizeof sp[0], cmp_splitseg);
        int *tab = split_tab + (size_t)c * SPLIT_CAP * 2;
        int ng = 0;
        for (int i = 0; i < nsp; ++i) {
            seg[seg_base + sp[i].idx] = ~i;
            if (i == 0 || sp[i].row != sp[i - 1].row) {
                int kd = sp[i].kind;
                int code = (kd & K_PAD) ? C_DISCARD : (kd & K_TAIL) ? C_TAIL : (kd & K_HEAD) ? C_HEAD : sp[i].row;
                if (ng >= SPLIT_CAP - 1) {
                    free(pieces);
                    return -3;
                }
                tab[2 * ng] = code;
                tab[2 * ng + 1] = i;
                ++ng;
            }
        }
        tab[2 * ng] = 0;
        tab[2 * ng + 1] = nsp;
        split_n[c] = ng;
    }
    free(pieces);
    if (info) {
        info[0] = max_slots;
        info[1] = nfix;
        info[2] = has_empty;
    }
    return 0;
}

/* FP64 serial parity (the GPU's rule, restated): bound terms of a split row,
 * packing of (q, n) into one double, and the decision. */
static const double PAR_LAMBDA = 10.0;
enum { PAR_RIGOROUS_LEN = 4096 };
static const int64_t PAR_NMASK = 0xFFFFF;

static double par_pack(double q, int n) {
    int64_t b;
    memcpy(&b, &q, 8);
    b = (b & ~PAR_NMASK) | (int64_t)n;
    double r;
    memcpy(&r, &b, 8);
    return r;
}
static void par_unpack(double p, double *q, int *n) {
    int64_t b;
    memcpy(&b, &p, 8);
    *n = (int)(b & PAR_NMASK);
    b &= ~PAR_NMASK;
    memcpy(q, &b, 8);
}
/* rigorous running-error bound (rows <= PAR_RIGOROUS_LEN) or, for longer
 * rows, also the probabilistic one; 0 -> recompute in serial order */
static int par_ok(double s, double A, double Q, int N, double tol) {
    const double u = 0x1p-53;
    const double lim = tol * fabs(s);
    const double Ar = A + (double)N * 0x1p-1021;
    if (u * Ar * (1.0 + 2e-6) <= lim) return 1;
    if (N > PAR_RIGOROUS_LEN && Q > 0x1p-900) {
        const double p = PAR_LAMBDA * u;
        if (p * p * Q * (1.0 + 4e-6) <= lim * lim) return 1;
    }
    return 0;
}
static double serial_row(const int *rp, const int *ci, const double *va, const double *x, int r) {
    double s = 0.0;
    for (int j = rp[r]; j < rp[r + 1]; ++j) s = fma(va[j], x[ci[j]], s);
    return s;
}

void orc_cvr64_spmv(int m, int nnz, int S, int prec, const void *vals, const int *cols, const int *chunk_rows,
                    const int *seg, const int *split_n, const int *split_tab, const int *fix, const int *rp,
                    const void *x, void *y) {
    orc_cvr64_spmv_par(m, nnz, S, prec, vals, cols, chunk_rows, seg, split_n, split_tab, fix, rp, NULL, NULL, x, y,
                       0.0, NULL, NULL);
}

void orc_cvr64_spmv_par(int m, int nnz, int S, int prec, const void *vals, const int *cols, const int *chunk_rows,
                        const int *seg, const int *split_n, const int *split_tab, const int *fix, const int *rp,
                        const int *ci, const double *va, const void *x, void *y, double parity_tol, double *bound,
                        int *n_recomputed) {
    int nchunks;
    long long L, segcap;
    orc_cvr64_sizes(m, nnz, S, &nchunks, &L, &segcap);
    const long long C = 64LL * S;
    const int par = prec == 0 && (parity_tol > 0.0 || bound != NULL);
    int nrc = 0;
    if (bound)
        for (int r = 0; r < m; ++r) bound[r] = 0.0;
    double *carry = (double *)calloc((size_t)(2 * nchunks + 2), sizeof(double));
    double *cbA = (double *)calloc((size_t)(2 * nchunks + 2), sizeof(double));
    double *cbQ = (double *)calloc((size_t)(2 * nchunks + 2), sizeof(double));
#define STORE_Y(r, v)                               \
    do {                                            \
        if (prec == 0)                              \
            ((double *)y)[r] = (v);                 \
        else                                        \
            ((float *)y)[r] = (float)(v);           \
    } while (0)
    if (nchunks == 0)
        for (int r = 0; r < m; ++r) STORE_Y(r, 0.0);
    for (int c = 0; c < nchunks; ++c) {
        const long long base = C * c;
        const int steps = (c == nchunks - 1) ? (int)((L - base) / 64) : S;
        const int *segs = seg + chunk_rows[c] + SEG_EXTRA * c;
        double slots[SLOT_CAP], slotq[SLOT_CAP];
        double accd[LANES], accq[LANES];
        float accs[LANES], accc[LANES];
        int st0[LANES];
        for (int l = 0; l < LANES; ++l) accd[l] = 0.0, accq[l] = 0.0, accs[l] = 0.f, accc[l] = 0.f, st0[l] = 0;
        int segp = 0;
        for (int s = 0; s < steps; ++s) {
            for (int l = 0; l < LANES; ++l) {
                const long long o = base + (long long)s * 64 + l;
                const int col = cols[o] & 0x7fffffff;
                if (prec == 0) {
                    accd[l] = fma(((const double *)vals)[o], ((const double *)x)[col], accd[l]);
                    accq[l] = fma(accd[l], accd[l], accq[l]);
                } else {
                    float v = prec == 1 ? ((const float *)vals)[o] : bf16_to_f32(((const uint16_t *)vals)[o]);
                    float xv = ((const float *)x)[col];
                    float p = v * xv;
                    float pe = fmaf(v, xv, -p);
                    float t = accs[l] + p;
                    float z = t - accs[l];
                    float te = (accs[l] - (t - z)) + (p - z);
                    accs[l] = t;
                    accc[l] += te + pe;
                }
            }
            int k = 0;
            for (int l = 0; l < LANES; ++l) {
                const long long o = base + (long long)s * 64 + l;
                if (cols[o] >= 0) continue;
                const int d = segs[segp + k];
                ++k;
                double v = prec == 0 ? accd[l] : (double)accs[l] + (double)accc[l];
                if (d >= 0) {
                    STORE_Y(d, v);
                } else {
                    slots[~d] = v;
                    slotq[~d] = par_pack(accq[l], s + 1 - st0[l]);
                }
                st0[l] = s + 1;
                accd[l] = 0.0, accq[l] = 0.0, accs[l] = 0.f, accc[l] = 0.f;
            }
            segp += k;
        }
        const int *tab = split_tab + (size_t)c * SPLIT_CAP * 2;
        for (int g = 0; g < split_n[c]; ++g) {
            const int code = tab[2 * g], sb = tab[2 * g + 1], se = tab[2 * g + 3];
            double s = slots[sb];
            if (par) {
                double q0;
                int n0;
                par_unpack(slotq[sb], &q0, &n0);
                double A = 2.0 * sqrt((double)n0 * q0), Q = 2.0 * q0;
                int N = n0;
                for (int j = sb + 1; j < se; ++j) {
                    double qj;
                    int nj;
                    par_unpack(slotq[j], &qj, &nj);
                    A += 2.0 * sqrt((double)nj * qj) + (double)nj * fabs(s);
                    Q += 2.0 * (double)nj * s * s + 3.0 * qj;
                    s += slots[j];
                    A += fabs(s);
                    Q += s * s;
                    N += nj;
                }
                if (code >= 0) {
                    if (bound) bound[code] = 0x1p-53 * (A + (double)N * 0x1p-1021);
                    if (parity_tol > 0.0 && !par_ok(s, A, Q, N, parity_tol)) {
                        s = serial_row(rp, ci, va, (const double *)x, code);
                        ++nrc;
                    }
                    STORE_Y(code, s);
                } else if (code == C_HEAD) {
                    carry[2 * c] = s, cbA[2 * c] = A, cbQ[2 * c] = par_pack(Q, N);
                } else if (code == C_TAIL) {
                    carry[2 * c + 1] = s, cbA[2 * c + 1] = A, cbQ[2 * c + 1] = par_pack(Q, N);
                }
                continue;
            }
            for (int j = sb + 1; j < se; ++j) s += slots[j];
            if (code >= 0)
                STORE_Y(code, s);
            else if (code == C_HEAD)
                carry[2 * c] = s;
            else if (code == C_TAIL)
                carry[2 * c + 1] = s;
        }
        for (int r = chunk_rows[c]; r < chunk_rows[c + 1]; ++r)
            if (rp[r + 1] == rp[r]) STORE_Y(r, 0.0);
    }
    for (int c = 0; c < nchunks; ++c) {
        const int row = fix[2 * c];
        if (row < 0) continue;
        const int c0 = fix[2 * c + 1];
        double s = carry[2 * c0 + 1];
        if (par) {
            double A = cbA[2 * c0 + 1], Q;
            int N;
            par_unpack(cbQ[2 * c0 + 1], &Q, &N);
            for (int k = c0 + 1; k <= c; ++k) {
                const int idx = k < c ? 2 * k + 1 : 2 * c;
                double Qk;
                int Nk;
                par_unpack(cbQ[idx], &Qk, &Nk);
                A += cbA[idx] + (double)Nk * fabs(s);
                Q += 2.0 * Qk + 2.0 * (double)Nk * s * s;
                s += carry[idx];
                A += fabs(s);
                Q += s * s;
                N += Nk;
            }
            if (bound) bound[row] = 0x1p-53 * (A + (double)N * 0x1p-1021);
            if (parity_tol > 0.0 && !par_ok(s, A, Q, N, parity_tol)) {
                s = serial_row(rp, ci, va, (const double *)x, row);
                ++nrc;
            }
            STORE_Y(row, s);
            continue;
        }
        for (int k = c0 + 1; k < c; ++k) s += carry[2 * k + 1];
        s += carry[2 * c];
        STORE_Y(row, s);
    }
#undef STORE_Y
    free(carry);
    free(cbA);
    free(cbQ);
    if (n_recomputed) *n_recomputed = nrc;
}

/* ---------------------------------------------------------------- CSR5-64 */
int orc_csr5_ntiles(int nnz, int sigma) {
    const long long T = 64LL * sigma;
    return nnz > 0 ? (int)((nnz + T - 1) / T) : 0;
}

/* first row r with rp[r] > e (rows 0..m-1), i.e. the row holding element e is r-1 */
static int row_of(const int *rp, int m, long long e) {
    int lo = 0, hi = m;  /* answer in [1, m] for 0 <= e < nnz */
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((long long)rp[mid + 1] > e) hi = mid; else lo = mid + 1;
    }
    return lo;  /* row containing element e */
}

int orc_csr5_build(int m, int nnz, const int *rp, const int *ci, const double *va, int sigma, double *vals_t,
                   int *cols_t, uint32_t *flags, int *yoff, int *tile_base, int *nerows, int *fix,
                   unsigned char *cont) {
    if (sigma < 1 || sigma > 32) return -1;
    const int nt = orc_csr5_ntiles(nnz, sigma);
    const long long T = 64LL * sigma;
    /* element e starts a non-empty row: mark by walking the rows */
    unsigned char *start = (unsigned char *)calloc((size_t)(nt * T + 1), 1);
    int *rowstart = (int *)malloc(sizeof(int) * (size_t)(nt * T + 1));
    if (!start || !rowstart) { free(start); free(rowstart); return -2; }
    int ne = 0;
    for (int r = 0; r < m; ++r)
        if (rp[r + 1] > rp[r]) { start[rp[r]] = 1; rowstart[rp[r]] = r; nerows[ne++] = r; }
    int base = 0;
    for (int t = 0; t < nt; ++t) {
        tile_base[t] = base;
        int acc = 0;
        for (int l = 0; l < 64; ++l) {
            uint32_t f = 0;
            yoff[t * 64 + l] = acc;
            for (int i = 0; i < sigma; ++i) {
                const long long e = t * T + (long long)l * sigma + i;
                const long long dst = t * T + (long long)i * 64 + l;
                if (e < nnz) {
                    vals_t[dst] = va[e];
                    cols_t[dst] = ci[e];
                    if (start[e]) { f |= 1u << i; ++acc; }
                } else {
                    vals_t[dst] = 0.0;
                    cols_t[dst] = 0;
                }
            }
            flags[t * 64 + l] = f;
        }
        base += acc;
        /* the row open at the tile start, when it began earlier and ends here */
        const long long t0e = t * T, t1e = (t + 1) * T < nnz ? (t + 1) * T : nnz;
        const int rf = row_of(rp, m, t0e);
        fix[2 * t] = -1;
        fix[2 * t + 1] = -1;
        if (rp[rf] < t0e && rp[rf + 1] <= t1e) {
            fix[2 * t] = rf;
            fix[2 * t + 1] = (int)(rp[rf] / T);
        }
        const int rl = row_of(rp, m, t1e - 1);
        cont[t] = (unsigned char)(rp[rl + 1] > t1e);
    }
    free(start);
    free(rowstart);
    return 0;
}

void orc_csr5_spmv(int m, int nnz, int sigma, const double *vals_t, const int *cols_t, const uint32_t *flags,
                   const int *yoff, const int *tile_base, const int *nerows, const int *fix,
                   const unsigned char *cont, const double *x, double *y) {
    const int nt = orc_csr5_ntiles(nnz, sigma);
    const long long T = 64LL * sigma;
    double *head = (double *)calloc((size_t)(nt > 0 ? nt : 1), sizeof(double));
    double *tail = (double *)calloc((size_t)(nt > 0 ? nt : 1), sizeof(double));
    for (int r = 0; r < m; ++r) y[r] = 0.0;
    for (int t = 0; t < nt; ++t) {
        double first[64], open[64];
        int has[64], owner_row[64];
        for (int l = 0; l < 64; ++l) {
            const uint32_t f = flags[t * 64 + l];
            int k = tile_base[t] + yoff[t * 64 + l];
            double acc = 0.0;
            int nseg = 0; /* flags seen */
            int row = -1;
            for (int i = 0; i < sigma; ++i) {
                const long long e = t * T + (long long)l * sigma + i;
                if (e >= nnz) break;
                if ((f >> i) & 1u) {
                    if (nseg == 0) first[l] = acc;
                    else y[row] = acc; /* a row wholly inside the lane */
                    row = nerows[k++];
                    ++nseg;
                    acc = 0.0;
                }
                const long long src = t * T + (long long)i * 64 + l;
                acc = fma(vals_t[src], x[cols_t[src]], acc);
            }
            if (nseg == 0) first[l] = acc;
            has[l] = nseg > 0;
            open[l] = nseg > 0 ? acc : 0.0;
            owner_row[l] = row;
        }
        /* rows crossing lanes, in lane order */
        for (int l = 0; l < 64; ++l) {
            if (!has[l]) continue;
            double s = open[l];
            int j = l + 1;
            while (j < 64 && !has[j]) s += first[j++];
            if (j < 64) {
                s += first[j];
                y[owner_row[l]] = s;
            } else if (cont[t]) {
                tail[t] = s;
            } else {
                y[owner_row[l]] = s;
            }
        }
        /* the row open at the tile start */
        if (!(flags[t * 64] & 1u)) {
            double h = first[0];
            if (!has[0]) {
                int j = 1;
                while (j < 64 && !has[j]) h += first[j++];
                if (j < 64) h += first[j];
            }
            head[t] = h;
        }
    }
    /* rows crossing tiles, in tile order */
    for (int t = 0; t < nt; ++t) {
        const int r = fix[2 * t];
        if (r < 0) continue;
        const int t0 = fix[2 * t + 1];
        double s = tail[t0];
        for (int u = t0 + 1; u < t; ++u) s += head[u];
        s += head[t];
        y[r] = s;
    }
    free(head);
    free(tail);
}
