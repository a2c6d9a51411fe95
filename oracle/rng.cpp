// Oracle RNG helper -- TEST INFRASTRUCTURE ONLY.
//
// Fills a column-major buffer with the exact Gaussian stream that the
// reference's sketch generator produces: std::mt19937 seeded with `seed`,
// std::normal_distribution<double>(0, 1), one draw per element in column-major
// element order.  Follows src/svd_cpu_rsvd.cpp:10-22 and the identical
// gaussian_matrix_local at src/fastPLS.cpp:99-115.  Because the distribution
// is a header template of libstdc++, building this file with the same
// libstdc++ reproduces the reference's draws bit for bit.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// the shared library built from this file, and only as the checker.
#include <cstdint>
#include <random>

extern "C" void fpls_oracle_gaussian(double* out, int64_t n_elem, unsigned int seed) {
  std::mt19937 rng(seed);
  std::normal_distribution<double> norm(0.0, 1.0);
  for (int64_t i = 0; i < n_elem; ++i) out[i] = norm(rng);
}

// Raw 32-bit mt19937 outputs, used by the tests to pin the generator against
// the published MT19937 known-answer value (10000th output for seed 5489).
extern "C" void fpls_oracle_mt19937_raw(uint32_t* out, int64_t n, unsigned int seed) {
  std::mt19937 rng(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = static_cast<uint32_t>(rng());
}
