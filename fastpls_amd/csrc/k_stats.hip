// Streaming column statistics and label bookkeeping.
//
// Column-major X makes every column a contiguous run of n elements, so one
// workgroup reduces one column with coalesced loads and a fixed-order FP64
// tree (deterministic).  HBM-bound: 4 (or 8) bytes per element, one pass.
#include <algorithm>

#include "fpls_internal.h"

namespace fpls {

template <typename T>
__device__ inline double block_sum_256(double v, double* red) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (t == 0) s = red[0] + red[1] + red[2] + red[3];
  return s;
}

// out[j] = sum / div; post 1: sqrt(sum / div), post 2: the same with zero or
// non-finite mapped to 1; ones (optional) gets 1 in every column.
template <typename T>
__global__ __launch_bounds__(256) void k_col_sums(int64_t n, int64_t p, const T* __restrict__ X,
                                                  int64_t ld, const double* __restrict__ center,
                                                  int squared, double* __restrict__ out,
                                                  double div, int post,
                                                  double* __restrict__ ones) {
  __shared__ double red[4];
  const int64_t j = blockIdx.x;
  if (j >= p) return;
  const T* col = X + j * ld;
  const double c = center ? center[j] : 0.0;
  auto term = [&](T x) {
    const double v = static_cast<double>(x) - c;
    return squared ? v * v : v;
  };
  // four independent partial sums over 16-byte vectors keep several loads in
  // flight per thread
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  constexpr int VW = 16 / sizeof(T);
  int64_t head = 0;
  if ((reinterpret_cast<uintptr_t>(col) & 15) == 0) {
    typedef T vT __attribute__((ext_vector_type(VW)));
    const vT* cv = reinterpret_cast<const vT*>(col);
    const int64_t nv = n / VW;
#pragma unroll 4
    for (int64_t i = threadIdx.x; i < nv; i += 256) {
      const vT v = cv[i];
      if constexpr (VW == 4) {
        s0 += term(v[0]);
        s1 += term(v[1]);
        s2 += term(v[2]);
        s3 += term(v[3]);
      } else {
        s0 += term(v[0]);
        s1 += term(v[1]);
      }
    }
    head = nv * VW;
  }
  for (int64_t i = head + threadIdx.x; i < n; i += 256) s0 += term(col[i]);
  double s = (s0 + s1) + (s2 + s3);
  s = block_sum_256<T>(s, red);
  if (threadIdx.x == 0) {
    double r = s / div;
    if (post >= 1) {
      r = sqrt(r);
      if (post == 2 && (!isfinite(r) || r == 0.0)) r = 1.0;
    }
    out[j] = r;
    if (ones) ones[j] = 1.0;
  }
}

void col_sums(Ctx* ctx, const DMat& X, double* sums) { col_stat(ctx, X, nullptr, sums, 1.0, 0); }

void col_ssd(Ctx* ctx, const DMat& X, const double* center, double* ssd) {
  col_stat(ctx, X, center, ssd, 1.0, 0);
}

void col_stat(Ctx* ctx, const DMat& X, const double* center, double* out, double div, int post,
              double* ones) {
  if (X.cols <= 0) return;
  const int sq = center ? 1 : 0;
  if (X.dtype == FPLS_F32)
    launch(ctx, "col_sums", k_col_sums<float>, dim3((unsigned)X.cols), dim3(256), 0, X.rows,
           X.cols, X.f(), X.ld, center, sq, out, div, post, ones);
  else
    launch(ctx, "col_sums", k_col_sums<double>, dim3((unsigned)X.cols), dim3(256), 0, X.rows,
           X.cols, X.d(), X.ld, center, sq, out, div, post, ones);
}

// Integer class counts: per-block LDS histogram then integer atomics (exact,
// order-independent).  bad[0] counts codes outside 1..C.
__global__ __launch_bounds__(256) void k_label_counts(const int32_t* __restrict__ y, int64_t n,
                                                      int32_t C,
                                                      unsigned long long* __restrict__ counts,
                                                      int32_t* __restrict__ bad) {
  extern __shared__ unsigned int hist[];
  for (int c = threadIdx.x; c < C; c += blockDim.x) hist[c] = 0u;
  __syncthreads();
  int nbad = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = y[i];
    if (c >= 1 && c <= C) atomicAdd(&hist[c - 1], 1u);
    else ++nbad;
  }
  if (nbad) atomicAdd(bad, nbad);
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x)
    if (hist[c]) atomicAdd(&counts[c], (unsigned long long)hist[c]);
}

// Class counts and proportions in one launch: every block adds its LDS
// histogram into gcounts (a persistent zero buffer) with integer atomics
// (exact, order-independent); the last block to finish (a ticket counter,
// reset by that block) takes every count out of gcounts (leaving it zero for
// the next launch) and writes prop[c] = count / n and counts[c] (FP64).
__global__ __launch_bounds__(256) void k_label_counts_prop(const int32_t* __restrict__ y,
                                                           int64_t n, int32_t C,
                                                           unsigned long long* __restrict__ gcounts,
                                                           unsigned* __restrict__ ticket,
                                                           double* __restrict__ prop,
                                                           double* __restrict__ counts,
                                                           int32_t* __restrict__ bad) {
  extern __shared__ unsigned int hist[];
  __shared__ bool last;
  for (int c = threadIdx.x; c < C; c += blockDim.x) hist[c] = 0u;
  __syncthreads();
  int nbad = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = y[i];
    if (c >= 1 && c <= C) atomicAdd(&hist[c - 1], 1u);
    else ++nbad;
  }
  if (nbad) atomicAdd(bad, nbad);
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x)
    if (hist[c]) atomicAdd(&gcounts[c], (unsigned long long)hist[c]);
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    const double v = (double)atomicExch(&gcounts[c], 0ull);
    prop[c] = v / (double)n;
    if (counts) counts[c] = v;
  }
  if (threadIdx.x == 0) atomicExch(ticket, 0u);
}

__global__ void k_label_counts_global(const int32_t* __restrict__ y, int64_t n, int32_t C,
                                      unsigned long long* __restrict__ counts,
                                      int32_t* __restrict__ bad) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = y[i];
    if (c >= 1 && c <= C) atomicAdd(&counts[c - 1], 1ull);
    else atomicAdd(bad, 1);
  }
}

void label_counts(Ctx* ctx, const int32_t* y, int64_t n, int32_t C, int64_t* counts,
                  int32_t* bad, bool bad_zeroed) {
  fill_f64(ctx, reinterpret_cast<double*>(counts), 0.0, C);  // all-zero bit pattern
  if (!bad_zeroed) fill_i32(ctx, bad, 0, 1);
  if (n <= 0) return;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
  if (C <= 16384)
    launch(ctx, "label_counts", k_label_counts, dim3(blocks), dim3(256), C * sizeof(unsigned),
           y, n, C, reinterpret_cast<unsigned long long*>(counts), bad);
  else
    launch(ctx, "label_counts", k_label_counts_global, dim3(blocks), dim3(256), 0, y, n, C,
           reinterpret_cast<unsigned long long*>(counts), bad);
}

bool label_counts_prop(Ctx* ctx, const int32_t* y, int64_t n, int32_t C, double* prop,
                       double* counts, int32_t* bad) {
  if (n <= 0 || C <= 0 || C > Ctx::kZeroCounts) return false;
  // about 2048 rows per block keeps the global atomics (C per block) few
  const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 2047) / 2048, 256));
  launch(ctx, "label_counts", k_label_counts_prop, dim3(blocks), dim3(256), C * sizeof(unsigned),
         y, n, C, ctx->zero_counts(), ctx->ticket(), prop, counts, bad);
  return true;
}

}  // namespace fpls
