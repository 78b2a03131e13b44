// pava_internal.h -- shared helpers of the host library (not part of the C ABI).
#pragma once

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "pava_hip.h"

namespace pava {

extern thread_local std::string g_last_error;
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void *aligned_alloc64(size_t bytes);
// true if this device allocation is to fail (pava_hip_test_fail_alloc)
bool inject_device_failure();
// PAVA_HIP_ERANGE if a dimension or nnz exceeds PAVA_HIP_MAX_SIZE
int check_gpu_sizes(long long M, long long N, long long nnz);
// memcpy split over the host's OpenMP threads (large copies into pinned staging memory)
void parallel_memcpy(void *dst, const void *src, size_t bytes);

// Counter-based generator: every draw is a pure function of (key, index), so
// generation is deterministic whatever the thread count. The mixer is the
// SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014).
static inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint64_t rand64(uint64_t key, uint64_t i) {
    return mix64(key * 0xD1B54A32D192ED03ull ^ mix64(i));
}
// uniform in [0, 1) with 53 random bits
static inline double u01(uint64_t key, uint64_t i) { return (double)(rand64(key, i) >> 11) * 0x1.0p-53; }

// Sequential stream derived from a key: used per row / per edge.
struct Stream {
    uint64_t key, ctr;
    Stream(uint64_t k, uint64_t sub) : key(mix64(k ^ mix64(sub + 0x632BE59BD9B4E019ull))), ctr(0) {}
    uint64_t next() { return mix64(key + 0x9E3779B97F4A7C15ull * (++ctr)); }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
};

int csr_from_buckets(int M, int N, std::vector<long long> &start, std::vector<int> &bcol,
                     std::vector<double> &bval, pava_csr_matrix *out);

}  // namespace pava
