// Internal helpers shared by the host and device translation units of
// libkz_hip.so.  Not part of the C ABI.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "kz.h"

namespace kz {

// Per-thread error message returned by kz_last_error().
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
const char* get_error();

// Host and device (the GPU generators draw the same values).
#ifdef __HIPCC__
#define KZ_HD __host__ __device__
#else
#define KZ_HD
#endif

// splitmix64 output function.  A stream seeded with `key` yields
// mix(key + (i + 1) * golden) as its i-th value, which makes every draw
// addressable by its index (counter-based, thread-count independent).
KZ_HD static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;
KZ_HD static inline uint64_t stream_key(uint64_t seed, uint64_t tag) {
    return mix64(seed * kGolden + tag + 0x632BE59BD9B4E019ULL);
}
KZ_HD static inline uint64_t draw64(uint64_t key, uint64_t index) {
    return mix64(key + (index + 1) * kGolden);
}

// Owning allocation helpers for kz_csr arrays (malloc/free so that
// kz_csr_release can free arrays from any source).
bool csr_alloc(kz_csr* g, uint32_t n, uint64_t entries, bool labels);

// Chung-Lu weights w_i = max_weight * (i0 / (i + i0))^(1/(exponent-1)) with
// sum(w) = 2 * num_draws, as a cumulative table cdf[0..n] (cdf[0] = 0); a
// draw r32 picks the first i with cdf[i+1] > (r32 + 0.5) * cdf[n] / 2^32.
int chung_lu_cdf(uint32_t n, uint64_t num_draws, double exponent, uint32_t max_weight, std::vector<double>* cdf);

}  // namespace kz

#define KZ_FAIL(code, ...)           \
    do {                             \
        ::kz::set_error(__VA_ARGS__); \
        return (code);               \
    } while (0)
