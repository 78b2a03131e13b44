// gen.cpp -- seeded synthetic matrices for the measured workloads:
//   * 2D 5-point Laplacian (diag 4, neighbours -1)
//   * 3D 27-point stencil (diag 26, off-diagonal -1 + U(-jitter, jitter))
//   * R-MAT (Chakrabarti, Zhan, Faloutsos, SDM 2004): recursive quadrant
//     choice with probabilities a, b, c, d per level, duplicates summed
//   * families with constant / Poisson / lognormal / power-law row lengths and
//     banded / block-diagonal / uniform / mixed column structure.
// Every draw is a pure function of (seed, index), so a matrix is identical for
// any thread count and on any host.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#include "pava_hip.h"
#include "pava_internal.h"

using namespace pava;

extern "C" int pava_gen_laplacian2d(int nx, int ny, pava_csr_matrix *out) {
    if (!out || nx <= 0 || ny <= 0) return set_error(PAVA_HIP_EINVAL, "bad laplacian size");
    memset(out, 0, sizeof *out);
    long long M = (long long)nx * ny;
    long long nnz = 5 * M - 2LL * nx - 2LL * ny;
    if (M > INT32_MAX || nnz > INT32_MAX) return set_error(PAVA_HIP_ERANGE, "laplacian too large");
    int *rp = (int *)aligned_alloc64(sizeof(int) * (M + 1));
    int *ci = (int *)aligned_alloc64(sizeof(int) * nnz);
    double *va = (double *)aligned_alloc64(sizeof(double) * nnz);
    if (!rp || !ci || !va) {
        free(rp);
        free(ci);
        free(va);
        return set_error(PAVA_HIP_ENOMEM, "alloc");
    }
    long long k = 0;
    rp[0] = 0;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            int i = iy * nx + ix;
            if (iy > 0) ci[k] = i - nx, va[k++] = -1.0;
            if (ix > 0) ci[k] = i - 1, va[k++] = -1.0;
            ci[k] = i, va[k++] = 4.0;
            if (ix < nx - 1) ci[k] = i + 1, va[k++] = -1.0;
            if (iy < ny - 1) ci[k] = i + nx, va[k++] = -1.0;
            rp[i + 1] = (int)k;
        }
    out->nnz = (int)k;
    out->numRows = out->numCols = (int)M;
    out->rowOffsets = rp;
    out->cols = ci;
    out->vals = va;
    return 0;
}

extern "C" int pava_gen_stencil27(int n, double jitter, unsigned long long seed, pava_csr_matrix *out) {
    if (!out || n <= 0) return set_error(PAVA_HIP_EINVAL, "bad stencil size");
    memset(out, 0, sizeof *out);
    long long M = (long long)n * n * n;
    long long nnz = (3LL * n - 2) * (3LL * n - 2) * (3LL * n - 2);
    if (M > INT32_MAX || nnz > INT32_MAX) return set_error(PAVA_HIP_ERANGE, "stencil too large");
    int *rp = (int *)aligned_alloc64(sizeof(int) * (M + 1));
    int *ci = (int *)aligned_alloc64(sizeof(int) * nnz);
    double *va = (double *)aligned_alloc64(sizeof(double) * nnz);
    if (!rp || !ci || !va) {
        free(rp);
        free(ci);
        free(va);
        return set_error(PAVA_HIP_ENOMEM, "alloc");
    }
    // row lengths are known in closed form: prod over dims of (#valid offsets)
    auto span = [n](int i) { return (i > 0) + 1 + (i < n - 1); };
    rp[0] = 0;
    for (long long r = 0; r < M; ++r) {
        int ix = (int)(r % n), iy = (int)((r / n) % n), iz = (int)(r / ((long long)n * n));
        rp[r + 1] = rp[r] + span(ix) * span(iy) * span(iz);
    }
#pragma omp parallel for schedule(static)
    for (long long r = 0; r < M; ++r) {
        int ix = (int)(r % n), iy = (int)((r / n) % n), iz = (int)(r / ((long long)n * n));
        long long k = rp[r];
        int kk = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx, ++kk) {
                    int jx = ix + dx, jy = iy + dy, jz = iz + dz;
                    if (jx < 0 || jy < 0 || jz < 0 || jx >= n || jy >= n || jz >= n) continue;
                    long long c = ((long long)jz * n + jy) * n + jx;
                    ci[k] = (int)c;
                    if (c == r)
                        va[k] = 26.0;
                    else
                        va[k] = -1.0 + jitter * (2.0 * u01(seed, (unsigned long long)(r * 27 + kk)) - 1.0);
                    ++k;
                }
    }
    out->nnz = (int)nnz;
    out->numRows = out->numCols = (int)M;
    out->rowOffsets = rp;
    out->cols = ci;
    out->vals = va;
    return 0;
}

extern "C" int pava_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                             unsigned long long seed, pava_csr_matrix *out) {
    if (!out || scale < 1 || scale > 30 || edge_factor < 1)
        return set_error(PAVA_HIP_EINVAL, "bad rmat parameters");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return set_error(PAVA_HIP_EINVAL, "bad rmat probabilities");
    memset(out, 0, sizeof *out);
    const long long M = 1LL << scale;
    const long long E = (long long)edge_factor * M;
    if (M > INT32_MAX || E > INT32_MAX) return set_error(PAVA_HIP_ERANGE, "rmat too large");
    // quadrant thresholds on 32-bit uniforms
    const uint64_t ta = (uint64_t)std::llround(a * 4294967296.0);
    const uint64_t tab = (uint64_t)std::llround((a + b) * 4294967296.0);
    const uint64_t tabc = (uint64_t)std::llround((a + b + c) * 4294967296.0);
    std::vector<int> er((size_t)E), ec((size_t)E);
    std::vector<double> ev((size_t)E);
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < E; ++e) {
        Stream s(seed, (uint64_t)e);
        uint32_t r = 0, cc = 0;
        uint64_t bits = 0;
        for (int lvl = 0; lvl < scale; ++lvl) {
            if ((lvl & 1) == 0) bits = s.next();
            uint64_t u = (lvl & 1) ? (bits >> 32) : (bits & 0xffffffffull);
            uint32_t rb, cb;
            if (u < ta)
                rb = 0, cb = 0;
            else if (u < tab)
                rb = 0, cb = 1;
            else if (u < tabc)
                rb = 1, cb = 0;
            else
                rb = 1, cb = 1;
            r = (r << 1) | rb;
            cc = (cc << 1) | cb;
        }
        er[(size_t)e] = (int)r;
        ec[(size_t)e] = (int)cc;
        ev[(size_t)e] = 1.0 - s.uniform();  // U(0, 1]
    }
    // bucket by row (order inside a bucket is irrelevant: rows are sorted by
    // (col, value) before duplicates are summed)
    std::vector<long long> start((size_t)M + 1, 0);
    {
        std::vector<std::atomic<int>> cnt((size_t)M);
        for (auto &x : cnt) x.store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
        for (long long e = 0; e < E; ++e) cnt[(size_t)er[(size_t)e]].fetch_add(1, std::memory_order_relaxed);
        for (long long r = 0; r < M; ++r) start[(size_t)r + 1] = start[(size_t)r] + cnt[(size_t)r].load();
    }
    std::vector<int> bcol((size_t)E);
    std::vector<double> bval((size_t)E);
    {
        std::vector<std::atomic<long long>> fill((size_t)M);
        for (long long r = 0; r < M; ++r) fill[(size_t)r].store(start[(size_t)r], std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
        for (long long e = 0; e < E; ++e) {
            long long d = fill[(size_t)er[(size_t)e]].fetch_add(1, std::memory_order_relaxed);
            bcol[(size_t)d] = ec[(size_t)e];
            bval[(size_t)d] = ev[(size_t)e];
        }
    }
    er.clear();
    er.shrink_to_fit();
    ec.clear();
    ec.shrink_to_fit();
    ev.clear();
    ev.shrink_to_fit();
    // canonical order inside each row: (col, value) so the duplicate sums are
    // independent of the scatter order
#pragma omp parallel for schedule(dynamic, 512)
    for (long long r = 0; r < M; ++r) {
        long long s = start[(size_t)r], e = start[(size_t)r + 1];
        if (e - s < 2) continue;
        std::vector<std::pair<int, double>> t((size_t)(e - s));
        for (long long k = s; k < e; ++k) t[(size_t)(k - s)] = {bcol[(size_t)k], bval[(size_t)k]};
        std::sort(t.begin(), t.end());
        for (long long k = s; k < e; ++k) {
            bcol[(size_t)k] = t[(size_t)(k - s)].first;
            bval[(size_t)k] = t[(size_t)(k - s)].second;
        }
    }
    return csr_from_buckets((int)M, (int)M, start, bcol, bval, out);
}

namespace {

// Poisson(lambda) by inversion of the CDF (lambda <= 500), normal
// approximation beyond.
static int draw_poisson(Stream &s, double lam) {
    if (lam <= 0) return 0;
    if (lam > 500) {
        double u1 = s.uniform(), u2 = s.uniform();
        double z = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
        return std::max(0, (int)std::lround(lam + std::sqrt(lam) * z));
    }
    double u = s.uniform();
    double p = std::exp(-lam), cdf = p;
    int k = 0;
    while (u > cdf && k < 100000) {
        ++k;
        p *= lam / k;
        cdf += p;
        if (p < 1e-300 && cdf < u) break;
    }
    return k;
}

static double draw_normal(Stream &s) {
    double u1 = s.uniform(), u2 = s.uniform();
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
}

static int draw_len(Stream &s, const pava_gen_params *p) {
    double v = 0;
    switch (p->row_dist) {
    case PAVA_RL_CONSTANT: v = p->len_param; break;
    case PAVA_RL_POISSON: v = draw_poisson(s, p->len_param); break;
    case PAVA_RL_LOGNORMAL: v = std::exp(std::log(std::max(p->len_param, 1e-9)) + p->shape * draw_normal(s)); break;
    case PAVA_RL_POWERLAW: {
        double u = s.uniform();
        double alpha = std::max(p->shape, 1.01);
        v = std::max(p->len_param, 1.0) * std::pow(1.0 - u, -1.0 / (alpha - 1.0));
        break;
    }
    default: v = p->len_param;
    }
    long long L = std::llround(v);
    if (L < 0) L = 0;
    long long cap = std::min<long long>(p->max_len > 0 ? p->max_len : p->rows, p->rows);
    if (L > cap) L = cap;
    return (int)L;
}

static int draw_col(Stream &s, const pava_gen_params *p, int row) {
    int n = p->rows;
    auto banded = [&]() {
        int hb = std::max(p->half_band, 0);
        long long lo = std::max<long long>(0, (long long)row - hb), hi = std::min<long long>(n - 1, (long long)row + hb);
        return (int)(lo + (long long)(s.uniform() * (double)(hi - lo + 1)));
    };
    switch (p->col_struct) {
    case PAVA_COL_BANDED: return banded();
    case PAVA_COL_BLOCKDIAG: {
        int B = std::max(p->block, 1);
        long long lo = (long long)(row / B) * B, hi = std::min<long long>(n, lo + B);
        return (int)(lo + (long long)(s.uniform() * (double)(hi - lo)));
    }
    case PAVA_COL_UNIFORM: return (int)(s.uniform() * (double)n);
    case PAVA_COL_MIXED:
    default:
        if (s.uniform() < p->band_frac) return banded();
        return (int)(s.uniform() * (double)n);
    }
}

// Columns a row may use: [lo, hi).
static void col_range(const pava_gen_params *p, int row, long long &lo, long long &hi) {
    int n = p->rows;
    switch (p->col_struct) {
    case PAVA_COL_BANDED: {
        int hb = std::max(p->half_band, 0);
        lo = std::max<long long>(0, (long long)row - hb);
        hi = std::min<long long>(n - 1, (long long)row + hb) + 1;
        return;
    }
    case PAVA_COL_BLOCKDIAG: {
        int B = std::max(p->block, 1);
        lo = (long long)(row / B) * B;
        hi = std::min<long long>(n, lo + B);
        return;
    }
    default: lo = 0; hi = n; return;
    }
}

static int col_space(const pava_gen_params *p, int row) {
    long long lo, hi;
    col_range(p, row, lo, hi);
    return (int)(hi - lo);
}

}  // namespace

extern "C" int pava_gen_synthetic(const pava_gen_params *p, pava_csr_matrix *out) {
    if (!p || !out || p->rows <= 0) return set_error(PAVA_HIP_EINVAL, "bad synthetic parameters");
    memset(out, 0, sizeof *out);
    const int n = p->rows;
    std::vector<int> len((size_t)n);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; ++r) {
        Stream s(p->seed, (uint64_t)r * 2);
        len[(size_t)r] = std::min(draw_len(s, p), col_space(p, r));
    }
    long long total = 0;
    std::vector<long long> start((size_t)n + 1, 0);
    for (int r = 0; r < n; ++r) start[(size_t)r + 1] = start[(size_t)r] + len[(size_t)r];
    total = start[(size_t)n];
    if (total > INT32_MAX) return set_error(PAVA_HIP_ERANGE, "synthetic matrix nnz beyond int32");
    std::vector<int> bcol((size_t)total);
    std::vector<double> bval((size_t)total);
#pragma omp parallel for schedule(dynamic, 256)
    for (int r = 0; r < n; ++r) {
        int L = len[(size_t)r];
        if (!L) continue;
        Stream s(p->seed, (uint64_t)r * 2 + 1);
        std::vector<int> cols;
        cols.reserve((size_t)L);
        // distinct columns: draw, sort, unique; top up until L are found
        int guard = 0;
        while ((int)cols.size() < L && guard < 64) {
            int need = L - (int)cols.size();
            for (int k = 0; k < need; ++k) cols.push_back(draw_col(s, p, r));
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
            ++guard;
        }
        if ((int)cols.size() < L) {  // nearly full rows: fill in order from the row's column range
            long long lo, hi;
            col_range(p, r, lo, hi);
            std::vector<int> extra;
            for (long long c = lo; c < hi && (int)(cols.size() + extra.size()) < L; ++c)
                if (!std::binary_search(cols.begin(), cols.end(), (int)c)) extra.push_back((int)c);
            cols.insert(cols.end(), extra.begin(), extra.end());
            std::sort(cols.begin(), cols.end());
        }
        long long b = start[(size_t)r];
        for (int k = 0; k < L; ++k) {
            bcol[(size_t)(b + k)] = cols[(size_t)k];
            bval[(size_t)(b + k)] = p->value_dist == 1 ? 1.0 - s.uniform() : draw_normal(s);
        }
    }
    return csr_from_buckets(n, n, start, bcol, bval, out);
}
