// Sanitizer harness for the host side of the library (Matrix Market reader,
// COO->CSR, binary cache, writer, generators, features), built with
// -fsanitize=address,undefined by tests/test_sanitizers.py without any HIP
// code. argv[1] is a scratch directory. Exit status 0 = all good.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pava_hip.h"

static int fails = 0;
#define CHECK(c)                                                    \
    do {                                                            \
        if (!(c)) {                                                 \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++fails;                                                \
        }                                                           \
    } while (0)

static std::string write(const std::string &dir, const char *name, const std::string &text) {
    const std::string p = dir + "/" + name;
    FILE *f = std::fopen(p.c_str(), "wb");
    std::fwrite(text.data(), 1, text.size(), f);
    std::fclose(f);
    return p;
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    const char *banner = "%%MatrixMarket matrix coordinate real general\n";
    // a valid file with duplicates and an empty row
    std::string ok = std::string(banner) + "% comment\n4 5 6\n1 1 1.5\n1 2 -2\n3 5 4\n3 5 1\n4 1 7\n4 4 -1e-3\n";
    pava_csr_matrix A{};
    CHECK(pava_read_csr_matrix(write(dir, "ok.mtx", ok).c_str(), &A) == 0);
    CHECK(A.numRows == 4 && A.numCols == 5 && A.nnz == 5 && pava_csr_validate(&A) == 0);
    // binary cache round trip and the writer
    const std::string bin = dir + "/ok.csrbin";
    CHECK(pava_write_csr_binary(bin.c_str(), &A) == 0);
    pava_csr_matrix B{};
    CHECK(pava_read_csr_binary(bin.c_str(), &B) == 0 && B.nnz == A.nnz);
    CHECK(std::memcmp(A.vals, B.vals, sizeof(double) * (size_t)A.nnz) == 0);
    CHECK(pava_write_matrix_market((dir + "/w.mtx").c_str(), &A) == 0);
    pava_delete_csr_matrix(&B);
    float feat[PAVA_N_FEATURES];
    CHECK(pava_csr_features(&A, feat) == 0);
    pava_delete_csr_matrix(&A);
    // CRLF line endings, no final newline, an empty matrix
    CHECK(pava_read_csr_matrix(write(dir, "crlf.mtx", std::string("%%MatrixMarket matrix coordinate real general\r\n") +
                                                         "2 2 2\r\n1 1 1\r\n2 2 2").c_str(), &A) == 0 &&
          A.nnz == 2 && A.vals[1] == 2.0);
    pava_delete_csr_matrix(&A);
    CHECK(pava_read_csr_matrix(write(dir, "empty.mtx", std::string(banner) + "0 0 0\n").c_str(), &A) == 0 &&
          A.numRows == 0 && A.nnz == 0);
    pava_delete_csr_matrix(&A);
    // symmetric / pattern / integer fields
    const char *sym = "%%MatrixMarket matrix coordinate pattern symmetric\n3 3 3\n1 1\n2 1\n3 2\n";
    CHECK(pava_read_csr_matrix(write(dir, "sym.mtx", sym).c_str(), &A) == 0 && A.nnz == 5);
    pava_delete_csr_matrix(&A);
    // malformed inputs: every one must fail cleanly
    const std::vector<std::string> bad = {
        "",                                                        // empty file
        "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",  // dense array
        "%%MatrixMarket vector coordinate real general\n2 2 1\n1 1 1\n", // not a matrix
        std::string(banner) + "3 3\n",                            // short size line
        std::string(banner) + "3 3 4\n1 1 1\n2 2 2\n",            // truncated body
        std::string(banner) + "3 3 1\n4 1 1\n",                   // row out of range
        std::string(banner) + "3 3 1\n1 0 1\n",                   // column 0 (1-based)
        std::string(banner) + "3 3 1\n1 1 abc\n",                 // bad value
        std::string(banner) + "-3 3 1\n1 1 1\n",                  // negative size
        "%%MatrixMarket matrix coordinate real generic\n1 1 1\n1 1 1\n",  // unknown symmetry
        std::string(banner) + "3 3 2000000000\n1 1 1\n",       // size line overstates nnz
        "%%MatrixMarket matrix coordinate real symmetric\n3 3 1500000000\n1 1 1\n",  // mirrored nnz beyond int32
        std::string(banner) + "99999999999 3 1\n1 1 1\n",      // dimension beyond int32
        std::string(banner) + "3 3 1\n99999999999 1 1\n",      // huge row index
        std::string(banner) + "3 3 2\n1 1 1\n2 2",              // last entry cut mid-line
        "%%MatrixMarket",                                          // banner only, no newline
    };
    for (size_t i = 0; i < bad.size(); ++i) {
        pava_csr_matrix C{};
        const std::string p = write(dir, ("bad" + std::to_string(i) + ".mtx").c_str(), bad[i]);
        CHECK(pava_read_csr_matrix(p.c_str(), &C) < 0);
        pava_delete_csr_matrix(&C);
    }
    // a truncated binary cache
    FILE *f = std::fopen((dir + "/trunc.csrbin").c_str(), "wb");
    std::fwrite("PAVACSR1\x05\x00", 1, 10, f);
    std::fclose(f);
    pava_csr_matrix D{};
    CHECK(pava_read_csr_binary((dir + "/trunc.csrbin").c_str(), &D) < 0);
    // a cache whose row offsets decrease: read fully, then rejected by validation
    {
        const int32_t dims[2] = {2, 2};
        const int64_t nnz = 1;
        const int32_t rp[3] = {0, 1, 0}, ci[1] = {0};
        const double va[1] = {1.0};
        FILE *g = std::fopen((dir + "/badrp.csrbin").c_str(), "wb");
        std::fwrite("PAVACSR1", 1, 8, g), std::fwrite(dims, 4, 2, g), std::fwrite(&nnz, 8, 1, g);
        std::fwrite(rp, 4, 3, g), std::fwrite(ci, 4, 1, g), std::fwrite(va, 8, 1, g);
        std::fclose(g);
        CHECK(pava_read_csr_binary((dir + "/badrp.csrbin").c_str(), &D) < 0);
    }
    // generators and features on generated matrices
    pava_csr_matrix R{};
    CHECK(pava_gen_rmat(10, 8, 0.57, 0.19, 0.19, 3, &R) == 0 && pava_csr_validate(&R) == 0);
    CHECK(pava_csr_features(&R, feat) == 0);
    pava_delete_csr_matrix(&R);
    // a file above the parallel-parse threshold (several MB): write, read back, compare
    CHECK(pava_gen_rmat(14, 16, 0.57, 0.19, 0.19, 5, &R) == 0);
    CHECK(pava_write_matrix_market((dir + "/big.mtx").c_str(), &R) == 0);
    pava_csr_matrix E{};
    CHECK(pava_read_csr_matrix((dir + "/big.mtx").c_str(), &E) == 0 && E.nnz == R.nnz && E.numRows == R.numRows);
    if (E.nnz == R.nnz && E.numRows == R.numRows) {
        CHECK(std::memcmp(E.rowOffsets, R.rowOffsets, sizeof(int) * ((size_t)R.numRows + 1)) == 0);
        CHECK(std::memcmp(E.cols, R.cols, sizeof(int) * (size_t)R.nnz) == 0);
        CHECK(std::memcmp(E.vals, R.vals, sizeof(double) * (size_t)R.nnz) == 0);
    }
    pava_delete_csr_matrix(&E);
    pava_delete_csr_matrix(&R);
    for (int d = 0; d < 4; ++d) {
        for (int c = 0; c < 4; ++c) {
            pava_gen_params gp{};
            gp.rows = 2000, gp.row_dist = d, gp.len_param = 9.0, gp.shape = d == 3 ? 2.0 : 1.0, gp.col_struct = c;
            gp.half_band = 50, gp.block = 64, gp.band_frac = 0.7, gp.value_dist = 0, gp.max_len = 500, gp.seed = 7;
            pava_csr_matrix G{};
            CHECK(pava_gen_synthetic(&gp, &G) == 0 && pava_csr_validate(&G) == 0);
            CHECK(pava_csr_features(&G, feat) == 0);
            pava_delete_csr_matrix(&G);
        }
    }
    // allocation failures (readCOOMatrix's -5, utils/basicMatrix.cpp:84-132):
    // the n-th host allocation fails; every call reports it and, under the
    // leak checker, frees whatever it had taken
    {
        const std::string okp = write(dir, "ok2.mtx", ok);
        int reader_fails = 0;
        for (int n = 1; n <= 8; ++n) {
            pava_coo_matrix Co{};
            pava_hip_test_fail_alloc(n, 0);
            const int rc = pava_read_coo_matrix(okp.c_str(), &Co);
            pava_hip_test_fail_alloc(0, 0);
            if (rc == 0) {
                pava_delete_coo_matrix(&Co);
                break;
            }
            CHECK(rc == PAVA_MM_ENOMEM);
            CHECK(Co.vals == nullptr && Co.cols == nullptr && Co.rows == nullptr);
            ++reader_fails;
        }
        CHECK(reader_fails >= 1);
        pava_coo_matrix Co{};
        CHECK(pava_read_coo_matrix(okp.c_str(), &Co) == 0);
        int convert_fails = 0;
        for (int n = 1; n <= 8; ++n) {
            pava_csr_matrix Cs{};
            pava_hip_test_fail_alloc(n, 0);
            const int rc = pava_convert_coo_to_csr(&Co, &Cs);
            pava_hip_test_fail_alloc(0, 0);
            if (rc == 0) {
                pava_delete_csr_matrix(&Cs);
                break;
            }
            CHECK(rc == PAVA_HIP_ENOMEM);
            CHECK(Cs.vals == nullptr && Cs.cols == nullptr && Cs.rowOffsets == nullptr);
            ++convert_fails;
        }
        CHECK(convert_fails >= 1);
        pava_delete_coo_matrix(&Co);
        std::printf("allocation failures: reader %d, convert %d\n", reader_fails, convert_fails);
    }
    std::printf("host harness: %d failures\n", fails);
    return fails != 0;
}
