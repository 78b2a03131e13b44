// host_io.cpp -- host side of the cvr_hip plugin surface: Matrix Market input,
// COO -> CSR, the driver's vectors / verification / report line, and the
// seeded synthetic matrix generators.
//
// Reference behaviour followed (paths relative to the pava source tree):
//   readCOOMatrix     utils/basicMatrix.cpp:63-253   (banner checks and codes 84-132)
//   convertCOO2CSR    utils/basicMatrix.cpp:350-415  (MKL mkl_dcsrcoo in the reference)
//   initVectors       utils/basicMatrix.cpp:499-520
//   referenceSpMV     utils/basicMatrix.cpp:538-556
//   checkResults      utils/basicMatrix.cpp:559-587
//   printPerformance  utils/basicMatrix.cpp:491-495
//   microtime         utils/microtime.h:9-17
// Deliberate differences from the reference reader: indices are converted to
// 0-based, no padding entries are appended, and the matrix is never
// "square-ified" (utils/basicMatrix.cpp:150, 211-216, 228-241).

#include "pava_hip.h"
#include "pava_internal.h"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <sys/stat.h>
#include <sys/time.h>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace pava {

thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

// Allocation-failure injection (pava_hip_test_fail_alloc): the n-th host
// allocation through aligned_alloc64, or the n-th device allocation of the
// CVR handle code, counted from the call, fails as if memory had run out.
static std::atomic<int> g_fail_host{0}, g_fail_device{0};
static bool take_failure(std::atomic<int> &ctr) {
    int v = ctr.load();
    while (v > 0 && !ctr.compare_exchange_weak(v, v - 1)) {
    }
    return v == 1;
}
bool inject_device_failure() { return take_failure(g_fail_device); }

void *aligned_alloc64(size_t bytes) {
    if (take_failure(g_fail_host)) return nullptr;
    if (bytes == 0) bytes = 64;
    void *p = nullptr;
    if (posix_memalign(&p, 64, (bytes + 63) / 64 * 64) != 0) return nullptr;
    return p;
}

}  // namespace pava

using namespace pava;

extern "C" const char *pava_hip_last_error(void) { return g_last_error.c_str(); }

void pava::parallel_memcpy(void *dst, const void *src, size_t bytes) {
    constexpr size_t kBlock = size_t(1) << 20;
    const long long nb = (long long)((bytes + kBlock - 1) / kBlock);
#pragma omp parallel for schedule(static) num_threads(8) if (nb > 4)
    for (long long b = 0; b < nb; ++b) {
        const size_t off = (size_t)b * kBlock;
        std::memcpy(static_cast<char *>(dst) + off, static_cast<const char *>(src) + off,
                    std::min(kBlock, bytes - off));
    }
}

extern "C" void pava_hip_test_fail_alloc(int host_nth, int device_nth) {
    g_fail_host.store(host_nth > 0 ? host_nth : 0);
    g_fail_device.store(device_nth > 0 ? device_nth : 0);
}
extern "C" const char *pava_hip_version(void) { return "pava_hip 0.1 (gfx950, CVR-64)"; }

// ---------------------------------------------------------------- MM reader
namespace {

struct Banner {
    bool matrix = false, coordinate = false;
    bool real = false, integer = false, pattern = false, complex_ = false;
    bool symmetric = false, skew = false, hermitian = false;
};

static std::string lower(std::string s) {
    for (auto &c : s) c = (char)tolower((unsigned char)c);
    return s;
}

// Parses "%%MatrixMarket object format field symmetry" (utils/mmio.c:96).
static int parse_banner(const char *line, Banner &b) {
    char t[5][64] = {{0}};
    if (sscanf(line, "%63s %63s %63s %63s %63s", t[0], t[1], t[2], t[3], t[4]) != 5)
        return PAVA_MM_EBANNER;
    if (lower(t[0]) != "%%matrixmarket") return PAVA_MM_EBANNER;
    std::string obj = lower(t[1]), fmt = lower(t[2]), fld = lower(t[3]), sym = lower(t[4]);
    if (obj != "matrix") return PAVA_MM_ENOTMATRIX;
    b.matrix = true;
    if (fmt == "coordinate")
        b.coordinate = true;
    else if (fmt == "array")
        return PAVA_MM_ENOTSPARSE;
    else
        return PAVA_MM_EBANNER;
    if (fld == "real" || fld == "double")
        b.real = true;
    else if (fld == "integer")
        b.integer = true;
    else if (fld == "pattern")
        b.pattern = true;
    else if (fld == "complex")
        b.complex_ = true;
    else
        return PAVA_MM_EBANNER;
    if (sym == "general") {
    } else if (sym == "symmetric")
        b.symmetric = true;
    else if (sym == "skew-symmetric")
        b.skew = true;
    else if (sym == "hermitian")
        b.hermitian = true;
    else
        return PAVA_MM_EBANNER;
    return 0;
}

// Integer / float scanners over a NUL-terminated buffer.
static inline const char *skip_ws(const char *p) {
    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
    return p;
}

}  // namespace

extern "C" int pava_read_coo_matrix(const char *path, pava_coo_matrix *out) {
    if (!path || !out) return set_error(PAVA_HIP_EINVAL, "null argument");
    memset(out, 0, sizeof *out);
    FILE *f = fopen(path, "rb");
    if (!f) return set_error(PAVA_MM_EOPEN, "cannot open %s: %s", path, strerror(errno));
    std::vector<char> buf;
    {
        if (fseek(f, 0, SEEK_END) != 0) {
            fclose(f);
            return set_error(PAVA_MM_EOPEN, "cannot seek %s", path);
        }
        long sz = ftell(f);
        rewind(f);
        if (sz < 0) {
            fclose(f);
            return set_error(PAVA_MM_EOPEN, "cannot size %s", path);
        }
        buf.resize((size_t)sz + 1);
        size_t got = fread(buf.data(), 1, (size_t)sz, f);
        buf[got] = 0;
    }
    fclose(f);
    const char *p = buf.data();
    // banner line
    const char *eol = strchr(p, '\n');
    std::string first(p, eol ? (size_t)(eol - p) : strlen(p));
    Banner b;
    int rc = parse_banner(first.c_str(), b);
    if (rc != 0) return set_error(rc, "unsupported Matrix Market banner: %s", first.c_str());
    p = eol ? eol + 1 : p + first.size();
    // comments / blank lines, then the size line (utils/mmio.c:189)
    long long M = -1, N = -1, L = -1;
    while (*p) {
        const char *e = strchr(p, '\n');
        size_t len = e ? (size_t)(e - p) : strlen(p);
        std::string line(p, len);
        p = e ? e + 1 : p + len;
        const char *q = skip_ws(line.c_str());
        if (*q == 0 || *q == '%') continue;
        if (sscanf(q, "%lld %lld %lld", &M, &N, &L) != 3) M = -1;
        break;
    }
    if (M < 0 || N < 0 || L < 0)
        return set_error(PAVA_MM_ESIZE, "could not process matrix sizes");
    if (M > INT32_MAX || N > INT32_MAX) return set_error(PAVA_MM_ESIZE, "matrix dimension beyond int32");
    bool mirror = b.symmetric || b.skew || b.hermitian;
    long long cap = mirror ? 2 * L : L;
    if (cap > INT32_MAX) return set_error(PAVA_MM_ESIZE, "nnz beyond int32");
    // The body is parsed in parallel: the text is cut at line starts into one
    // piece per thread, each piece parsed into its own arrays, and the pieces
    // are concatenated in file order (so duplicates later sum in file order).
    const char *body = p;
    const size_t blen = strlen(body);
    int T = 1;
#ifdef _OPENMP
    if (blen >= ((size_t)1 << 22)) T = std::max(1, std::min(omp_get_max_threads(), 64));
#endif
    std::vector<size_t> cut((size_t)T + 1, blen);
    cut[0] = 0;
    for (int t = 1; t < T; ++t) {
        size_t q = std::max(cut[(size_t)t - 1], blen * (size_t)t / (size_t)T);
        while (q < blen && body[q - 1] != '\n') ++q;
        cut[(size_t)t] = q;
    }
    struct Part {
        std::vector<int> r, c;
        std::vector<double> v;
        long long bad = -1;  // local index of the first malformed entry
        long long oor = -1;  // local index of the first out-of-range entry
        long long br = 0, bc = 0;
    };
    std::vector<Part> parts((size_t)T);
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; ++t) {
        Part &P = parts[(size_t)t];
        const char *q = body + cut[(size_t)t];
        const char *qe = body + cut[(size_t)t + 1];
        while (q < qe) {
            q = skip_ws(q);
            if (q >= qe) break;
            if (*q == '%') {
                while (q < qe && *q != '\n') ++q;
                continue;
            }
            char *end;
            const long long r = strtoll(q, &end, 10);
            if (end == q) {
                P.bad = (long long)P.r.size();
                break;
            }
            q = end;
            const long long c = strtoll(q, &end, 10);
            if (end == q) {
                P.bad = (long long)P.r.size();
                break;
            }
            q = end;
            double v = 1.0;
            if (!b.pattern) {
                v = strtod(q, &end);
                if (end == q) {
                    P.bad = (long long)P.r.size();
                    break;
                }
                q = end;
                if (b.complex_) {
                    (void)strtod(q, &end);  // imaginary part ignored, as the reference does
                    if (end == q) {
                        P.bad = (long long)P.r.size();
                        break;
                    }
                    q = end;
                }
            }
            if ((r < 1 || r > M || c < 1 || c > N) && P.oor < 0) {
                P.oor = (long long)P.r.size();
                P.br = r;
                P.bc = c;
            }
            P.r.push_back((int)(r - 1));
            P.c.push_back((int)(c - 1));
            P.v.push_back(v);
        }
    }
    // first error in file order, entries beyond L ignored (the reference reads L)
    long long seen = 0;
    int nparts_used = T;
    for (int t = 0; t < T; ++t) {
        const Part &P = parts[(size_t)t];
        const long long n = (long long)P.r.size();
        const long long firstbad = P.bad;
        if (P.oor >= 0 && seen + P.oor < L && (firstbad < 0 || P.oor < firstbad)) {
            return set_error(PAVA_MM_EBODY, "entry %lld (%lld, %lld) outside %lldx%lld", seen + P.oor + 1, P.br, P.bc,
                             M, N);
        }
        if (firstbad >= 0 && seen + firstbad < L) {
            return set_error(PAVA_MM_EBODY, "malformed entry %lld", seen + firstbad + 1);
        }
        seen += n;
        if (seen >= L) {
            nparts_used = t + 1;
            break;
        }
    }
    if (seen < L) {
        return set_error(PAVA_MM_EBODY, "truncated body: %lld of %lld entries", seen, L);
    }
    // concatenate (with the mirrored halves) in file order
    std::vector<long long> in_off((size_t)nparts_used + 1, 0), out_off((size_t)nparts_used + 1, 0);
    for (int t = 0; t < nparts_used; ++t) {
        const Part &P = parts[(size_t)t];
        const long long take = std::min<long long>((long long)P.r.size(), L - in_off[(size_t)t]);
        in_off[(size_t)t + 1] = in_off[(size_t)t] + take;
        long long m = take;
        if (mirror)
            for (long long k = 0; k < take; ++k) m += (P.r[(size_t)k] != P.c[(size_t)k]);
        out_off[(size_t)t + 1] = out_off[(size_t)t] + m;
    }
    // the COO arrays are sized by the entries actually read, not by the size
    // line, so a header that overstates nnz cannot force a huge allocation
    const long long cnt = out_off[(size_t)nparts_used];
    int *rows = (int *)aligned_alloc64(sizeof(int) * (size_t)std::max<long long>(cnt, 1));
    int *cols = (int *)aligned_alloc64(sizeof(int) * (size_t)std::max<long long>(cnt, 1));
    double *vals = (double *)aligned_alloc64(sizeof(double) * (size_t)std::max<long long>(cnt, 1));
    if (!rows || !cols || !vals) {
        free(rows);
        free(cols);
        free(vals);
        return set_error(PAVA_MM_ENOMEM, "could not allocate COO arrays (nnz %lld)", cnt);
    }
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < nparts_used; ++t) {
        const Part &P = parts[(size_t)t];
        const long long take = in_off[(size_t)t + 1] - in_off[(size_t)t];
        long long d = out_off[(size_t)t];
        for (long long k = 0; k < take; ++k) {
            const int r = P.r[(size_t)k], c = P.c[(size_t)k];
            const double v = P.v[(size_t)k];
            rows[d] = r;
            cols[d] = c;
            vals[d] = v;
            ++d;
            if (mirror && r != c) {
                rows[d] = c;
                cols[d] = r;
                vals[d] = b.skew ? -v : v;
                ++d;
            }
        }
    }
    out->nnz = (int)cnt;
    out->numRows = (int)M;
    out->numCols = (int)N;
    out->rows = rows;
    out->cols = cols;
    out->vals = vals;
    return 0;
}

extern "C" void pava_delete_coo_matrix(pava_coo_matrix *m) {
    if (!m) return;
    free(m->rows);
    free(m->cols);
    free(m->vals);
    memset(m, 0, sizeof *m);
}

extern "C" void pava_delete_csr_matrix(pava_csr_matrix *m) {
    if (!m) return;
    free(m->rowOffsets);
    free(m->cols);
    free(m->vals);
    memset(m, 0, sizeof *m);
}

namespace pava {

// Builds a CSR from row-bucketed (col, val) pairs: each row's entries are
// stably sorted by column and duplicates summed in their incoming order.
int csr_from_buckets(int M, int N, std::vector<long long> &start, std::vector<int> &bcol,
                     std::vector<double> &bval, pava_csr_matrix *out) {
    std::vector<int> newlen(M);
#pragma omp parallel for schedule(dynamic, 256)
    for (int r = 0; r < M; ++r) {
        long long s = start[r], e = start[r + 1];
        long long n = e - s;
        if (n > 1) {
            // stable sort (col) keeping incoming order among equal columns
            std::vector<std::pair<int, long long>> idx((size_t)n);
            for (long long k = 0; k < n; ++k) idx[(size_t)k] = {bcol[s + k], k};
            std::sort(idx.begin(), idx.end());
            std::vector<int> c2;
            std::vector<double> v2;
            c2.reserve((size_t)n);
            v2.reserve((size_t)n);
            for (auto &pr : idx) {
                double v = bval[s + pr.second];
                if (!c2.empty() && c2.back() == pr.first)
                    v2.back() += v;
                else {
                    c2.push_back(pr.first);
                    v2.push_back(v);
                }
            }
            for (size_t k = 0; k < c2.size(); ++k) {
                bcol[s + (long long)k] = c2[k];
                bval[s + (long long)k] = v2[k];
            }
            newlen[r] = (int)c2.size();
        } else {
            newlen[r] = (int)n;
        }
    }
    long long total = 0;
    for (int r = 0; r < M; ++r) total += newlen[r];
    if (total > INT32_MAX) return set_error(PAVA_HIP_ERANGE, "nnz beyond int32");
    int *rp = (int *)aligned_alloc64(sizeof(int) * ((size_t)M + 1));
    int *ci = (int *)aligned_alloc64(sizeof(int) * (size_t)std::max<long long>(total, 1));
    double *va = (double *)aligned_alloc64(sizeof(double) * (size_t)std::max<long long>(total, 1));
    if (!rp || !ci || !va) {
        free(rp);
        free(ci);
        free(va);
        return set_error(PAVA_HIP_ENOMEM, "could not allocate CSR (nnz %lld)", total);
    }
    rp[0] = 0;
    for (int r = 0; r < M; ++r) rp[r + 1] = rp[r] + newlen[r];
#pragma omp parallel for schedule(dynamic, 1024)
    for (int r = 0; r < M; ++r) {
        long long s = start[r];
        int d = rp[r];
        for (int k = 0; k < newlen[r]; ++k) {
            ci[d + k] = bcol[s + k];
            va[d + k] = bval[s + k];
        }
    }
    out->nnz = (int)total;
    out->numRows = M;
    out->numCols = N;
    out->rowOffsets = rp;
    out->cols = ci;
    out->vals = va;
    return 0;
}

}  // namespace pava

extern "C" int pava_convert_coo_to_csr(const pava_coo_matrix *coo, pava_csr_matrix *out) {
    if (!coo || !out) return set_error(PAVA_HIP_EINVAL, "null argument");
    memset(out, 0, sizeof *out);
    int M = coo->numRows, N = coo->numCols;
    long long nnz = coo->nnz;
    if (M < 0 || N < 0 || nnz < 0) return set_error(PAVA_HIP_EINVAL, "negative size");
    std::vector<long long> start((size_t)M + 1, 0);
    for (long long k = 0; k < nnz; ++k) {
        int r = coo->rows[k], c = coo->cols[k];
        if (r < 0 || r >= M || c < 0 || c >= N)
            return set_error(PAVA_HIP_ERANGE, "COO entry %lld (%d,%d) out of range", k, r, c);
        start[(size_t)r + 1]++;
    }
    for (int r = 0; r < M; ++r) start[(size_t)r + 1] += start[r];
    std::vector<int> bcol((size_t)nnz);
    std::vector<double> bval((size_t)nnz);
    std::vector<long long> fill(start.begin(), start.end() - 1);
    for (long long k = 0; k < nnz; ++k) {  // stable: file order within a row
        long long d = fill[coo->rows[k]]++;
        bcol[d] = coo->cols[k];
        bval[d] = coo->vals[k];
    }
    return csr_from_buckets(M, N, start, bcol, bval, out);
}

// ---------------------------------------------------------- binary cache
// "<file>.csrbin": 16-byte header ("PAVACSR1"), int32 numRows, numCols, int64
// nnz, then rowOffsets (numRows+1 int32), cols (nnz int32), vals (nnz f64),
// little-endian. Read back with one fread per array.
static const char kCsrMagic[8] = {'P', 'A', 'V', 'A', 'C', 'S', 'R', '1'};

extern "C" int pava_write_csr_binary(const char *path, const pava_csr_matrix *A) {
    if (!path || !A) return set_error(PAVA_HIP_EINVAL, "null argument");
    int rc = pava_csr_validate(A);
    if (rc) return rc;
    std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return set_error(PAVA_HIP_EIO, "cannot open %s for writing", tmp.c_str());
    const int32_t dims[2] = {A->numRows, A->numCols};
    const int64_t nnz = A->nnz;
    bool ok = fwrite(kCsrMagic, 1, 8, f) == 8 && fwrite(dims, 4, 2, f) == 2 && fwrite(&nnz, 8, 1, f) == 1 &&
              fwrite(A->rowOffsets, 4, (size_t)A->numRows + 1, f) == (size_t)A->numRows + 1 &&
              fwrite(A->cols, 4, (size_t)nnz, f) == (size_t)nnz && fwrite(A->vals, 8, (size_t)nnz, f) == (size_t)nnz;
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path) != 0) {
        remove(tmp.c_str());
        return set_error(PAVA_HIP_EIO, "write failed: %s", path);
    }
    return 0;
}

extern "C" int pava_read_csr_binary(const char *path, pava_csr_matrix *out) {
    if (!path || !out) return set_error(PAVA_HIP_EINVAL, "null argument");
    memset(out, 0, sizeof *out);
    FILE *f = fopen(path, "rb");
    if (!f) return set_error(PAVA_MM_EOPEN, "cannot open %s", path);
    char magic[8];
    int32_t dims[2];
    int64_t nnz = 0;
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, kCsrMagic, 8) != 0 || fread(dims, 4, 2, f) != 2 ||
        fread(&nnz, 8, 1, f) != 1 || dims[0] < 0 || dims[1] < 0 || nnz < 0 || nnz > INT32_MAX) {
        fclose(f);
        return set_error(PAVA_MM_EBANNER, "%s is not a CSR cache file", path);
    }
    int *rp = (int *)aligned_alloc64(4 * ((size_t)dims[0] + 1));
    int *ci = (int *)aligned_alloc64(4 * (size_t)std::max<int64_t>(nnz, 1));
    double *va = (double *)aligned_alloc64(8 * (size_t)std::max<int64_t>(nnz, 1));
    bool ok = rp && ci && va && fread(rp, 4, (size_t)dims[0] + 1, f) == (size_t)dims[0] + 1 &&
              fread(ci, 4, (size_t)nnz, f) == (size_t)nnz && fread(va, 8, (size_t)nnz, f) == (size_t)nnz;
    fclose(f);
    if (!ok) {
        free(rp);
        free(ci);
        free(va);
        return set_error(PAVA_MM_EBODY, "truncated CSR cache %s", path);
    }
    out->numRows = dims[0];
    out->numCols = dims[1];
    out->nnz = (int)nnz;
    out->rowOffsets = rp;
    out->cols = ci;
    out->vals = va;
    int rc = pava_csr_validate(out);
    if (rc) pava_delete_csr_matrix(out);
    return rc;
}

extern "C" int pava_read_csr_matrix(const char *path, pava_csr_matrix *out) {
    if (!path || !out) return set_error(PAVA_HIP_EINVAL, "null argument");
    // a cache next to the file, newer than it, is used instead of parsing
    const std::string cache = std::string(path) + ".csrbin";
    struct stat st_m, st_c;
    if (stat(path, &st_m) == 0 && stat(cache.c_str(), &st_c) == 0 && st_c.st_mtime >= st_m.st_mtime)
        if (pava_read_csr_binary(cache.c_str(), out) == 0) return 0;
    pava_coo_matrix coo;
    int rc = pava_read_coo_matrix(path, &coo);
    if (rc != 0) return rc;
    rc = pava_convert_coo_to_csr(&coo, out);
    pava_delete_coo_matrix(&coo);
    return rc;
}

extern "C" int pava_write_matrix_market(const char *path, const pava_csr_matrix *A) {
    if (!path || !A) return set_error(PAVA_HIP_EINVAL, "null argument");
    FILE *f = fopen(path, "w");
    if (!f) return set_error(PAVA_HIP_EIO, "cannot open %s for writing", path);
    fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    fprintf(f, "%d %d %d\n", A->numRows, A->numCols, A->nnz);
    for (int r = 0; r < A->numRows; ++r)
        for (int k = A->rowOffsets[r]; k < A->rowOffsets[r + 1]; ++k)
            fprintf(f, "%d %d %.17g\n", r + 1, A->cols[k] + 1, A->vals[k]);
    int bad = ferror(f);
    if (fclose(f) != 0 || bad) return set_error(PAVA_HIP_EIO, "write failed: %s", path);
    return 0;
}

namespace pava {
int check_gpu_sizes(long long M, long long N, long long nnz) {
    if (M > PAVA_HIP_MAX_SIZE || N > PAVA_HIP_MAX_SIZE || nnz > PAVA_HIP_MAX_SIZE)
        return set_error(PAVA_HIP_ERANGE, "matrix %lld x %lld with %lld nonzeros beyond the GPU formats' limit %d", M,
                         N, nnz, PAVA_HIP_MAX_SIZE);
    return 0;
}
}  // namespace pava

extern "C" int pava_csr_validate(const pava_csr_matrix *A) {
    if (!A) return set_error(PAVA_HIP_EINVAL, "null matrix");
    if (A->numRows < 0 || A->numCols < 0 || A->nnz < 0)
        return set_error(PAVA_HIP_EINVAL, "negative size");
    if (A->numRows > 0 && !A->rowOffsets) return set_error(PAVA_HIP_EINVAL, "null rowOffsets");
    if (A->nnz > 0 && (!A->cols || !A->vals)) return set_error(PAVA_HIP_EINVAL, "null cols/vals");
    if (A->rowOffsets && A->rowOffsets[0] != 0) return set_error(PAVA_HIP_EINVAL, "rowOffsets[0] != 0");
    for (int r = 0; r < A->numRows; ++r)
        if (A->rowOffsets[r + 1] < A->rowOffsets[r])
            return set_error(PAVA_HIP_EINVAL, "rowOffsets not monotone at row %d", r);
    if (A->numRows > 0 && A->rowOffsets[A->numRows] != A->nnz)
        return set_error(PAVA_HIP_EINVAL, "rowOffsets[numRows]=%d != nnz=%d", A->rowOffsets[A->numRows],
                         A->nnz);
    for (int k = 0; k < A->nnz; ++k)
        if (A->cols[k] < 0 || A->cols[k] >= A->numCols)
            return set_error(PAVA_HIP_ERANGE, "column %d at %d out of range", A->cols[k], k);
    return 0;
}

// ------------------------------------------------------- driver utilities
extern "C" void pava_init_vectors(int numRows, int numCols, double *x, double *y, double *y_ref) {
    if (y)
        for (int i = 0; i < numRows; ++i) y[i] = 0;
    if (y_ref)
        for (int i = 0; i < numRows; ++i) y_ref[i] = 0;
    if (x)
        for (int i = 0; i < numCols; ++i) x[i] = i % 3;
}

extern "C" void pava_fill_uniform(double *x, long long n, unsigned long long seed, double lo, double hi) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) x[i] = lo + (hi - lo) * pava::u01(seed, (unsigned long long)i);
}

extern "C" void pava_reference_spmv(const pava_csr_matrix *A, const double *x, double *y_ref) {
    const int M = A->numRows;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i) {
        double yi = 0.0;
        for (int j = A->rowOffsets[i]; j < A->rowOffsets[i + 1]; ++j) yi = std::fma(A->vals[j], x[A->cols[j]], yi);
        y_ref[i] = yi;
    }
}

extern "C" int pava_check_results(int size, const double *y, const double *y_ref) {
    int errorCount = 0;
    for (int i = 0; i < size; ++i)
        if (!(std::fabs(y_ref[i] - y[i]) <= 0.01 * std::fabs(y_ref[i]))) ++errorCount;
    if (errorCount == 0)
        printf(" Checking Pass !  \n");
    else
        printf("WARNING! # %d out of %d # was wrong\n", errorCount, size);
    fflush(stdout);
    return errorCount;
}

extern "C" int pava_print_performance(const char *matrixName, const char *format, int numThreads,
                                      double convertTime, double executionTime) {
    int n = printf(" Performance %s %s %d %f %f \n", matrixName, format, numThreads, convertTime, executionTime);
    fflush(stdout);
    return n;
}

extern "C" double pava_microtime(void) {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return (double)tv.tv_sec + (double)tv.tv_usec / 1e6;
}
