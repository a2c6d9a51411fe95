// Gaussian sketch streams on the device.
//
// The reference draws every sketch on the host with std::mt19937(seed) and
// std::normal_distribution<double>(0, 1) (src/fastPLS.cpp:99-115,
// src/svd_cpu_rsvd.cpp:10-22).  A SIMPLS fit draws p values (only component 0
// consumes random numbers, src/fastPLS.cpp:1124-1131), so a batch of F
// cross-validation fits over p = 20,000 features needs 10^6 draws: about 30 ms
// of one host core, more than the whole batched fit on the GPU.  This kernel
// draws the same streams on the device, one workgroup per stream:
//
//   * MT19937 (Matsumoto & Nishimura 1998): seeding recurrence
//     x[i] = 1812433253 (x[i-1] ^ (x[i-1] >> 30)) + i, then per block of 624
//     words the twist.  The twist's in-place sweep reads x[k+1] (not yet
//     updated), and x[k+397] (old, k < 227) or x[k-227] (updated, k >= 227);
//     it is evaluated in three parallel phases k in [0,227), [227,454),
//     [454,624) with the same values.  Tempering per word as published.
//   * Uniforms as libstdc++'s generate_canonical<double, 53> forms them from a
//     32-bit engine: two words w0, w1 -> (w0 + w1 2^32) / 2^64 in double
//     arithmetic (the product is exact, one rounding in the sum), clamped
//     below 1.
//   * Marsaglia's polar method as libstdc++'s normal_distribution applies it:
//     each attempt takes four words (x from the first two, y from the next
//     two), x = 2u - 1, y = 2v - 1, r2 = x x + y y (two roundings, no fused
//     multiply-add), rejected when r2 > 1 or r2 == 0; an accepted attempt
//     yields y m and then x m with m = sqrt(-2 log(r2) / r2).
//     624 words are exactly 156 attempts, so attempts never straddle blocks;
//     accepted attempts are compacted with a wave ballot prefix count.
//
// Everything above is exact integer work or correctly rounded IEEE double
// operations, so it matches the host bit for bit.  The one exception is log:
// glibc's log is accurate to about 0.52 ulp and is not correctly rounded on
// every argument, while log_cr below returns the correctly rounded value
// (double-double evaluation, relative error below 2^-86).  The two differ by
// one ulp on roughly 0.1% of the accepted attempts, which then changes both
// draws of that attempt by at most one or two ulps; acceptance and the stream
// position never differ.
#include <algorithm>
#include <cmath>

#include "model.h"

namespace fpls {

namespace {

constexpr int kMtN = 624;
constexpr uint32_t kMtA = 0x9908b0dfu, kUpper = 0x80000000u, kLower = 0x7fffffffu;

__device__ inline uint32_t mt_twist_word(uint32_t cur, uint32_t next, uint32_t far) {
  const uint32_t y = (cur & kUpper) | (next & kLower);
  return far ^ (y >> 1) ^ ((y & 1u) ? kMtA : 0u);
}

__device__ inline uint32_t mt_temper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

// ---- double-double arithmetic (Dekker / Knuth error-free transformations)
struct dd {
  double hi, lo;
};

__device__ inline dd two_sum(double a, double b) {
#pragma clang fp contract(off)
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}

__device__ inline dd quick_two_sum(double a, double b) {
#pragma clang fp contract(off)
  const double s = a + b;
  return {s, b - (s - a)};
}

__device__ inline dd two_prod(double a, double b) {
#pragma clang fp contract(off)
  const double p = a * b;
  return {p, __fma_rn(a, b, -p)};
}

__device__ inline dd dd_add(dd a, dd b) {
#pragma clang fp contract(off)
  dd s = two_sum(a.hi, b.hi);
  const dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}

__device__ inline dd dd_mul(dd a, dd b) {
#pragma clang fp contract(off)
  dd p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return quick_two_sum(p.hi, p.lo);
}

// a / b for a double a and a double-double b, three quotient terms.
__device__ inline dd dd_div(double a, dd b) {
#pragma clang fp contract(off)
  const double q1 = a / b.hi;
  dd r = dd_add({a, 0.0}, dd_mul({-q1, 0.0}, b));
  const double q2 = r.hi / b.hi;
  r = dd_add(r, dd_mul({-q2, 0.0}, b));
  const double q3 = r.hi / b.hi;
  const dd q = quick_two_sum(q1, q2);
  return dd_add(q, {q3, 0.0});
}

// 1 / d as a double-double: hi = RN(1/d), lo from the exact residual.
__device__ inline dd dd_recip(double d) {
  const double h = 1.0 / d;
  const double r = __fma_rn(-h, d, 1.0);  // exact: 1 - h d
  return {h, r / d};
}

// Correctly rounded natural logarithm of a positive normal double.
//   x = mm 2^e with mm in [sqrt(1/2), sqrt(2));  log x = e ln2 + 2 atanh(s),
//   s = (mm - 1) / (mm + 1), |s| <= 0.1716, z = s^2 <= 0.0295, and
//   2 atanh(s) = 2s + 2s z Q(z),  Q(z) = sum_k z^k / (2k + 3).
// Q's first five terms are summed in double-double, the rest (down to
// z^16 / 35 < 2^-85 relative) in double.  Against a quad-precision log the
// double-double result stays within 2^-86 relative, so rounding it gives the
// correctly rounded log except on arguments whose log lies closer than that
// to a rounding boundary.
__device__ double log_cr(double x) {
#pragma clang fp contract(off)
  int e = 0;
  double mm = frexp(x, &e);  // mm in [0.5, 1)
  if (mm < 0.70710678118654752440) {
    mm *= 2.0;
    e -= 1;
  }
  const double f = mm - 1.0;  // exact (Sterbenz)
  const dd den = two_sum(mm, 1.0);
  const dd s = dd_div(f, den);
  const dd z = dd_mul(s, s);
  // double tail: sum_{k=5..16} z^k / (2k + 3), Horner in z, times z^5 later
  double tail = 0.0;
  for (int k = 16; k >= 5; --k) tail = tail * z.hi + 1.0 / (double)(2 * k + 3);
  // dd Horner over k = 4..0 with the tail folded in at k = 5
  dd q = dd_mul({tail, 0.0}, z);
  for (int k = 4; k >= 0; --k) {
    q = dd_add(q, dd_recip((double)(2 * k + 3)));
    if (k > 0) q = dd_mul(q, z);
  }
  const dd s2 = {2.0 * s.hi, 2.0 * s.lo};
  const dd series = dd_add(s2, dd_mul(dd_mul(s2, z), q));
  // ln 2 = 0x1.62e42fefa39efp-1 + 0x1.abc9e3b39803fp-56 (to 2^-106)
  const dd ln2 = {0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};
  const double ed = (double)e;
  dd el = two_prod(ed, ln2.hi);
  el.lo += ed * ln2.lo;
  el = quick_two_sum(el.hi, el.lo);
  const dd r = dd_add(el, series);
  return r.hi + r.lo;
}

// (w0 + w1 2^32) / 2^64, clamped below 1 (generate_canonical<double, 53>).
__device__ inline double canonical53(uint32_t w0, uint32_t w1) {
#pragma clang fp contract(off)
  double sum = (double)w0;
  sum = sum + (double)w1 * 4294967296.0;
  double r = sum / 18446744073709551616.0;
  if (r >= 1.0) r = 0x1.fffffffffffffp-1;
  return r;
}

}  // namespace

// out[:, s] (len values, column stride ld) = the Gaussian stream of seeds[s].
__global__ __launch_bounds__(256) void k_gauss_mt(const uint32_t* __restrict__ seeds, int64_t len,
                                                  double* __restrict__ out, int64_t ld) {
#pragma clang fp contract(off)
  __shared__ uint32_t st[kMtN];
  __shared__ int wsum[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double* o = out + (int64_t)blockIdx.x * ld;
  if (t == 0) {
    uint32_t x = seeds[blockIdx.x];
    st[0] = x;
    for (int i = 1; i < kMtN; ++i) {
      x = 1812433253u * (x ^ (x >> 30)) + (uint32_t)i;
      st[i] = x;
    }
  }
  __syncthreads();
  int64_t produced = 0;
  while (produced < len) {
    // twist: phase A reads only old words
    uint32_t v = 0;
    if (t < 227) v = mt_twist_word(st[t], st[t + 1], st[t + 397]);
    __syncthreads();
    if (t < 227) st[t] = v;
    __syncthreads();
    // phase B: k in [227, 454), far word k - 227 is from phase A
    if (t < 227) v = mt_twist_word(st[227 + t], st[228 + t], st[t]);
    __syncthreads();
    if (t < 227) st[227 + t] = v;
    __syncthreads();
    // phase C: k in [454, 623) with far word k - 227 from phase B; the last
    // word pairs with the updated x[0] and x[396]
    if (t < 169) v = mt_twist_word(st[454 + t], st[455 + t], st[227 + t]);
    else if (t == 169) v = mt_twist_word(st[623], st[0], st[396]);
    __syncthreads();
    if (t < 169) st[454 + t] = v;
    else if (t == 169) st[623] = v;
    __syncthreads();
    // 156 polar attempts of four words each
    bool acc = false;
    double gx = 0.0, gy = 0.0;
    if (t < 156) {
      const uint32_t w0 = mt_temper(st[4 * t]), w1 = mt_temper(st[4 * t + 1]);
      const uint32_t w2 = mt_temper(st[4 * t + 2]), w3 = mt_temper(st[4 * t + 3]);
      const double x = 2.0 * canonical53(w0, w1) - 1.0;
      const double y = 2.0 * canonical53(w2, w3) - 1.0;
      const double xx = x * x, yy = y * y;
      const double r2 = xx + yy;
      acc = !(r2 > 1.0 || r2 == 0.0);
      if (acc) {
        const double q = (-2.0 * log_cr(r2)) / r2;
        const double mult = sqrt(q);
        gy = y * mult + 0.0;
        gx = x * mult + 0.0;
      }
    }
    const unsigned long long b = __ballot(acc);
    if (lane == 0) wsum[w] = __popcll(b);
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      before += q < w ? wsum[q] : 0;
      total += wsum[q];
    }
    if (acc) {
      const int rank = before + __popcll(b & ((1ull << lane) - 1ull));
      const int64_t idx = produced + 2 * (int64_t)rank;
      if (idx < len) o[idx] = gy;
      if (idx + 1 < len) o[idx + 1] = gx;
    }
    produced += 2 * (int64_t)total;
    __syncthreads();  // wsum and st are rewritten by the next block
  }
}

void gaussian_dev(Ctx* ctx, const uint32_t* seeds_dev, int n_streams, int64_t len, double* out,
                  int64_t ld) {
  if (n_streams <= 0 || len <= 0) return;
  FPLS_REQUIRE(ld >= len, FPLS_ERR_ARG, "gaussian_dev: ld < len");
  launch(ctx, "gauss_mt19937", k_gauss_mt, dim3((unsigned)n_streams), dim3(256), 0, seeds_dev,
         len, out, ld);
}

}  // namespace fpls

using namespace fpls;

extern "C" int fpls_gaussian_streams(fpls_ctx* ctx, const uint32_t* seeds, int32_t n_streams,
                                     int64_t len, double* out) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    FPLS_REQUIRE(n_streams >= 0 && len >= 0, FPLS_ERR_ARG, "negative size");
    if (n_streams == 0 || len == 0) return;
    FPLS_REQUIRE(seeds && out, FPLS_ERR_ARG, "NULL pointer");
    DBuf<uint32_t> sd(c, n_streams);
    h2d(c, sd.get(), seeds, n_streams * sizeof(uint32_t));
    DBuf<double> o(c, (size_t)n_streams * len);
    gaussian_dev(c, sd.get(), n_streams, len, o.get(), len);
    to_host(c, out, o.get(), (size_t)n_streams * len);
  });
}
