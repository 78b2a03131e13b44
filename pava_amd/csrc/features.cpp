// features.cpp -- 16 cheap CSR features for the format / precision / tile
// selector. One pass over the row pointer and the column indices (OpenMP),
// plus a sampled locality probe; all values are scale-free or log-scaled so a
// small network sees comparable magnitudes across 1e4..5e6-row matrices.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "pava_hip.h"
#include "pava_internal.h"

using namespace pava;

extern "C" int pava_csr_features(const pava_csr_matrix *A, float *out) {
    if (!A || !out) return set_error(PAVA_HIP_EINVAL, "null argument");
    const int m = A->numRows, n = std::max(A->numCols, 1);
    const long long nnz = A->nnz;
    const int *rp = A->rowOffsets, *ci = A->cols;
    double sum = 0, sum2 = 0, mx = 0, empty = 0, shortr = 0, longmass = 0, dist = 0, band = 0, span = 0;
#pragma omp parallel for schedule(static) reduction(+ : sum, sum2, empty, shortr, longmass, dist, band, span) reduction(max : mx)
    for (int r = 0; r < m; ++r) {
        const int L = rp[r + 1] - rp[r];
        sum += L;
        sum2 += (double)L * L;
        mx = std::max(mx, (double)L);
        empty += (L == 0);
        shortr += (L < 4);
        if (L > 1024) longmass += L;
        int lo = INT32_MAX, hi = -1;
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
            const int c = ci[k];
            const double d = std::fabs((double)c - (double)r);
            dist += d;
            band += (d <= 1024.0);
            lo = std::min(lo, c);
            hi = std::max(hi, c);
        }
        if (L > 0) span += (double)(hi - lo) / n;
    }
    const double mean = m ? sum / m : 0.0;
    const double var = m ? std::max(0.0, sum2 / m - mean * mean) : 0.0;
    // Gini coefficient of the row lengths (sorted sample of up to 65536 rows)
    double gini = 0.0;
    {
        const int ns = std::min(m, 65536);
        std::vector<int> s((size_t)ns);
        for (int i = 0; i < ns; ++i) {
            const long long r = (long long)i * m / std::max(ns, 1);
            s[(size_t)i] = rp[r + 1] - rp[r];
        }
        std::sort(s.begin(), s.end());
        double acc = 0, tot = 0;
        for (int i = 0; i < ns; ++i) {
            acc += (double)(i + 1) * s[(size_t)i];
            tot += s[(size_t)i];
        }
        if (tot > 0) gini = (2.0 * acc) / (ns * tot) - (double)(ns + 1) / ns;
    }
    // gather locality: distinct 64-byte x lines per nonzero over windows of 64
    // consecutive rows (what one CVR wavefront step touches), sampled
    double lines_per_nnz = 0, share = 0;
    {
        const int windows = std::max(1, std::min(256, m / 64));
        double lsum = 0, nsum = 0, sh = 0, shn = 0;
        std::vector<int> buf;
        for (int w = 0; w < windows; ++w) {
            const long long r0 = (long long)w * std::max(m - 64, 0) / std::max(windows, 1);
            const long long r1 = std::min<long long>(m, r0 + 64);
            buf.clear();
            for (long long r = r0; r < r1; ++r)
                for (int k = rp[r]; k < rp[r + 1]; ++k) buf.push_back(ci[k] >> 3);
            if (buf.empty()) continue;
            nsum += buf.size();
            std::sort(buf.begin(), buf.end());
            lsum += (double)(std::unique(buf.begin(), buf.end()) - buf.begin());
            // neighbouring rows sharing a column
            for (long long r = r0; r + 1 < r1; ++r) {
                int a = rp[r], b = rp[r + 1], c = rp[r + 1], d = rp[r + 2];
                if (a == b || c == d) continue;
                bool hit = false;
                for (int k = a; k < b && !hit; ++k)
                    for (int q = c; q < d; ++q)
                        if (ci[k] == ci[q]) {
                            hit = true;
                            break;
                        }
                sh += hit;
                shn += 1;
                if (b - a > 64) break;  // bound the quadratic probe on long rows
            }
        }
        lines_per_nnz = nsum > 0 ? lsum / nsum : 1.0;
        share = shn > 0 ? sh / shn : 0.0;
    }
    out[0] = (float)std::log2(std::max(m, 1));
    out[1] = (float)std::log2((double)std::max<long long>(nnz, 1));
    out[2] = (float)std::log2(1.0 + mean);
    out[3] = (float)(mean > 0 ? std::sqrt(var) / mean : 0.0);
    out[4] = (float)std::log2(1.0 + mx);
    out[5] = (float)std::log2(1.0 + (mean > 0 ? mx / mean : 0.0));
    out[6] = (float)(m ? empty / m : 0.0);
    out[7] = (float)(m ? shortr / m : 0.0);
    out[8] = (float)(nnz ? longmass / nnz : 0.0);
    out[9] = (float)(nnz ? dist / nnz / n : 0.0);
    out[10] = (float)(nnz ? band / nnz : 0.0);
    out[11] = (float)(m ? span / m : 0.0);
    out[12] = (float)lines_per_nnz;
    out[13] = (float)share;
    out[14] = (float)gini;
    out[15] = (float)std::log2(n);
    return 0;
}
