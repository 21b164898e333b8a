// Synthetic-input helper: the reference's xorshift64* generator
// (src/matgen/mod.rs:63-96), used to draw the reproducible values of the
// benchmark matrices and right-hand sides.
#include <cstdint>

extern "C" void rsl_rng_uniform(uint64_t seed, double lo, double hi, int64_t count, double* out) {
  uint64_t x = (seed ^ 0x9E3779B97F4A7C15ULL) | 1ULL;
  for (int64_t k = 0; k < count; ++k) {
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    const uint64_t r = x * 0x2545F4914F6CDD1DULL;
    const double u = (double)(r >> 11) / (double)(1ULL << 53);
    out[k] = lo + (hi - lo) * u;
  }
}
