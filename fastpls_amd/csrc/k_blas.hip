// Dense contractions for gfx950.
//
//  * gemm_f64     -- LDS-tiled FP64 GEMM (64x64 tiles, split-K for long K)
//                    for the small/medium FP64 products of the factorisations.
//  (the FP32-data MFMA contractions live in k_gemm2.hip)
//  * gemv_n/_t    -- the SIMPLS matrix-vector pair t = Xc r, p = Xc^T t.
//
// Layouts are column-major throughout.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "fpls_internal.h"
#include "fpls_wave.h"

namespace fpls {


// ---------------------------------------------------------------- elementwise
__global__ void k_fill_f64(double* x, double v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}
__global__ void k_fill_f32(float* x, float v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}
__global__ void k_fill_i32(int32_t* x, int32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}
__global__ void k_copy_f64(double* d, const double* s, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}
__global__ void k_copy2d_f64(double* d, int64_t ldd, const double* s, int64_t lds, int64_t rows,
                             int64_t cols) {
  const int64_t n = rows * cols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % rows, j = t / rows;
    d[i + j * ldd] = s[i + j * lds];
  }
}
__global__ void k_f64_to_f32(float* d, const double* s, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = static_cast<float>(s[i]);
}
__global__ void k_f32_to_f64(double* d, const float* s, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = static_cast<double>(s[i]);
}

__global__ void k_transpose_f64(const double* __restrict__ A, int64_t lda, int64_t rows,
                                int64_t cols, double* __restrict__ B, int64_t ldb) {
  const int64_t n = rows * cols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % rows, j = t / rows;
    B[j + i * ldb] = A[i + j * lda];
  }
}

static inline unsigned grid1d(int64_t n, int threads = 256) {
  int64_t g = (n + threads - 1) / threads;
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(g, 65536)));
}

void fill_f64(Ctx* ctx, double* x, double v, int64_t n) {
  if (n > 0) launch(ctx, "fill", k_fill_f64, dim3(grid1d(n)), dim3(256), 0, x, v, n);
}
void fill_f32(Ctx* ctx, float* x, float v, int64_t n) {
  if (n > 0) launch(ctx, "fill", k_fill_f32, dim3(grid1d(n)), dim3(256), 0, x, v, n);
}
void fill_i32(Ctx* ctx, int32_t* x, int32_t v, int64_t n) {
  if (n > 0) launch(ctx, "fill", k_fill_i32, dim3(grid1d(n)), dim3(256), 0, x, v, n);
}
__global__ void k_copy_i32(int32_t* __restrict__ d, const int32_t* __restrict__ s, int64_t n) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    d[t] = s[t];
}
void copy_i32(Ctx* ctx, int32_t* d, const int32_t* s, int64_t n) {
  if (n > 0) launch(ctx, "copy", k_copy_i32, dim3(grid1d(n)), dim3(256), 0, d, s, n);
}
void copy_f64(Ctx* ctx, double* d, const double* s, int64_t n) {
  if (n > 0) launch(ctx, "copy", k_copy_f64, dim3(grid1d(n)), dim3(256), 0, d, s, n);
}
void copy2d_f64(Ctx* ctx, double* d, int64_t ldd, const double* s, int64_t lds, int64_t rows,
                int64_t cols) {
  if (rows > 0 && cols > 0)
    launch(ctx, "copy", k_copy2d_f64, dim3(grid1d(rows * cols)), dim3(256), 0, d, ldd, s, lds,
           rows, cols);
}
void transpose_f64(Ctx* ctx, const double* A, int64_t lda, int64_t rows, int64_t cols,
                   double* B, int64_t ldb) {
  if (rows > 0 && cols > 0)
    launch(ctx, "transpose", k_transpose_f64, dim3(grid1d(rows * cols)), dim3(256), 0, A, lda,
           rows, cols, B, ldb);
}
void f64_to_f32(Ctx* ctx, float* d, const double* s, int64_t n) {
  if (n > 0) launch(ctx, "convert", k_f64_to_f32, dim3(grid1d(n)), dim3(256), 0, d, s, n);
}
void f32_to_f64(Ctx* ctx, double* d, const float* s, int64_t n) {
  if (n > 0) launch(ctx, "convert", k_f32_to_f64, dim3(grid1d(n)), dim3(256), 0, d, s, n);
}

// ---------------------------------------------------------------- FP64 GEMM
constexpr int G64_BM = 64, G64_BN = 64, G64_BK = 16;

// grid (ceil(M/64), ceil(N/64), splits); block 256.
// splits == 1: C = alpha acc + beta C.  splits > 1: partial[z] = acc.
__global__ __launch_bounds__(256) void k_gemm_f64(bool transA, bool transB, int64_t M, int64_t N,
                                                  int64_t K, int64_t kchunk, double alpha,
                                                  const double* __restrict__ A, int64_t lda,
                                                  const double* __restrict__ shA,
                                                  const double* __restrict__ B, int64_t ldb,
                                                  const int32_t* __restrict__ labB,
                                                  const double* __restrict__ shB, double beta,
                                                  double* __restrict__ C, int64_t ldc,
                                                  double* __restrict__ partial) {
  __shared__ double As[G64_BK][G64_BM + 1];
  __shared__ double Bs[G64_BK][G64_BN + 1];
  const int t = threadIdx.x;
  const int tx = t % 16, ty = t / 16;
  const int64_t i0 = (int64_t)blockIdx.x * G64_BM;
  const int64_t j0 = (int64_t)blockIdx.y * G64_BN;
  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  const int64_t kend = std::min<int64_t>(K, kbeg + kchunk);
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;

  // register prefetch: the loads of tile k0 + BK are in flight during the
  // FMAs of tile k0
  double ra[4], rb[4];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int i, k;
      if (!transA) { i = t % 64; k = t / 64 + 4 * r; }
      else { k = t % 16; i = t / 16 + 16 * r; }
      const int64_t gi = i0 + i, gk = k0 + k;
      double v = 0.0;
      if (gi < M && gk < kend) {
        if (!transA) v = A[gi + gk * lda] - (shA ? shA[gk] : 0.0);
        else v = A[gk + gi * lda] - (shA ? shA[gi] : 0.0);
      }
      ra[r] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int j, k;
      if (!transB) { k = t % 16; j = t / 16 + 16 * r; }
      else { j = t % 64; k = t / 64 + 4 * r; }
      const int64_t gj = j0 + j, gk = k0 + k;
      double v = 0.0;
      if (gj < N && gk < kend) {
        if (labB) v = (labB[gk] == gj + 1 ? 1.0 : 0.0) - (shB ? shB[gj] : 0.0);
        else if (!transB) v = B[gk + gj * ldb] - (shB ? shB[gj] : 0.0);
        else v = B[gj + gk * ldb] - (shB ? shB[gk] : 0.0);
      }
      rb[r] = v;
    }
  };
  if (kbeg < kend) fetch(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += G64_BK) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (!transA) As[t / 64 + 4 * r][t % 64] = ra[r];
      else As[t % 16][t / 16 + 16 * r] = ra[r];
      if (!transB) Bs[t % 16][t / 16 + 16 * r] = rb[r];
      else Bs[t / 64 + 4 * r][t % 64] = rb[r];
    }
    __syncthreads();
    if (k0 + G64_BK < kend) fetch(k0 + G64_BK);
#pragma unroll
    for (int k = 0; k < G64_BK; ++k) {
      double a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[k][ty + 16 * r];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = Bs[k][tx + 16 * c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = fma(a[r], b[c], acc[r][c]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t gi = i0 + ty + 16 * r, gj = j0 + tx + 16 * c;
      if (gi < M && gj < N) {
        if (partial) {
          partial[(int64_t)blockIdx.z * M * N + gi + gj * M] = acc[r][c];
        } else {
          double v = alpha * acc[r][c];
          if (beta != 0.0) v += beta * C[gi + gj * ldc];
          C[gi + gj * ldc] = v;
        }
      }
    }
  }
}

// Large FP64 products: 128 x 128 tiles, 256 threads each holding an 8 x 8
// register block (rows ty*4 + {0..3} and 64 + ty*4 + {0..3}, columns likewise
// with tx), 16-deep K steps through double-buffered LDS (k-major, 2 doubles of
// padding per row) with the next step's global loads in registers during the
// FMAs; one barrier per step.  Operand fragments are 16-byte LDS reads; the
// A fragment is a broadcast across the 16 lanes of a row group.
constexpr int GB_M = 128, GB_N = 128, GB_K = 16, GB_LD = GB_M + 2;

template <bool TA, bool TB, bool LAB>
__device__ __forceinline__ void gb_fetch(int t, int64_t i0, int64_t j0, int64_t k0, int64_t kend,
                                         int64_t M, int64_t N, const double* __restrict__ A,
                                         int64_t lda, const double* __restrict__ shA,
                                         const double* __restrict__ B, int64_t ldb,
                                         const int32_t* __restrict__ labB,
                                         const double* __restrict__ shB, double (&ra)[8],
                                         double (&rb)[8]) {
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int i = TA ? (t >> 4) + 16 * r : (t & 127);
    const int k = TA ? (t & 15) : (t >> 7) + 2 * r;
    const int64_t gi = i0 + i, gk = k0 + k;
    double v = 0.0;
    if (gi < M && gk < kend) {
      v = TA ? A[gk + gi * lda] : A[gi + gk * lda];
      if (shA) v -= TA ? shA[gi] : shA[gk];
    }
    ra[r] = v;
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int j = TB ? (t & 127) : (t >> 4) + 16 * r;
    const int k = TB ? (t >> 7) + 2 * r : (t & 15);
    const int64_t gj = j0 + j, gk = k0 + k;
    double v = 0.0;
    if (gj < N && gk < kend) {
      // labels (non-transposed B only): B(k, j) = [lab_k == j + 1]
      if (LAB) v = labB[gk] == gj + 1 ? 1.0 : 0.0;
      else v = TB ? B[gj + gk * ldb] : B[gk + gj * ldb];
      if (shB) v -= TB ? shB[gk] : shB[gj];
    }
    rb[r] = v;
  }
}

template <bool TA, bool TB>
__device__ __forceinline__ void gb_store(int t, double (*As)[GB_LD], double (*Bs)[GB_LD],
                                         const double (&ra)[8], const double (&rb)[8]) {
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int i = TA ? (t >> 4) + 16 * r : (t & 127);
    const int k = TA ? (t & 15) : (t >> 7) + 2 * r;
    As[k][i] = ra[r];
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int j = TB ? (t & 127) : (t >> 4) + 16 * r;
    const int k = TB ? (t >> 7) + 2 * r : (t & 15);
    Bs[k][j] = rb[r];
  }
}

// grid (ceil(M/128), ceil(N/128), splits); splits == 1: C = alpha acc + beta C,
// else partial[z] = acc.  LAB: B is the one-hot of labB (not transposed).
template <bool TA, bool TB, bool LAB>
__global__ __launch_bounds__(256, 2) void k_gemm_f64_big(int64_t M, int64_t N, int64_t K,
                                                        int64_t kchunk, double alpha,
                                                        const double* __restrict__ A, int64_t lda,
                                                        const double* __restrict__ shA,
                                                        const double* __restrict__ B, int64_t ldb,
                                                        const int32_t* __restrict__ labB,
                                                        const double* __restrict__ shB,
                                                        double beta, double* __restrict__ C,
                                                        int64_t ldc, double* __restrict__ partial) {
  __shared__ __align__(16) double As[2][GB_K][GB_LD];
  __shared__ __align__(16) double Bs[2][GB_K][GB_LD];
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const int64_t i0 = (int64_t)blockIdx.x * GB_M, j0 = (int64_t)blockIdx.y * GB_N;
  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  const int64_t kend = std::min<int64_t>(K, kbeg + kchunk);
  double acc[8][8];
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[r][c] = 0.0;
  double ra[8], rb[8];
  int buf = 0;
  if (kbeg < kend) {
    gb_fetch<TA, TB, LAB>(t, i0, j0, kbeg, kend, M, N, A, lda, shA, B, ldb, labB, shB, ra, rb);
    gb_store<TA, TB>(t, As[0], Bs[0], ra, rb);
  }
  __syncthreads();
  for (int64_t k0 = kbeg; k0 < kend; k0 += GB_K) {
    const bool more = k0 + GB_K < kend;
    if (more)
      gb_fetch<TA, TB, LAB>(t, i0, j0, k0 + GB_K, kend, M, N, A, lda, shA, B, ldb, labB, shB, ra,
                            rb);
#pragma unroll 2
    for (int k = 0; k < GB_K; ++k) {
      const double2 a0 = *reinterpret_cast<const double2*>(&As[buf][k][ty * 4]);
      const double2 a1 = *reinterpret_cast<const double2*>(&As[buf][k][ty * 4 + 2]);
      const double2 a2 = *reinterpret_cast<const double2*>(&As[buf][k][64 + ty * 4]);
      const double2 a3 = *reinterpret_cast<const double2*>(&As[buf][k][64 + ty * 4 + 2]);
      const double2 b0 = *reinterpret_cast<const double2*>(&Bs[buf][k][tx * 4]);
      const double2 b1 = *reinterpret_cast<const double2*>(&Bs[buf][k][tx * 4 + 2]);
      const double2 b2 = *reinterpret_cast<const double2*>(&Bs[buf][k][64 + tx * 4]);
      const double2 b3 = *reinterpret_cast<const double2*>(&Bs[buf][k][64 + tx * 4 + 2]);
      const double a[8] = {a0.x, a0.y, a1.x, a1.y, a2.x, a2.y, a3.x, a3.y};
      const double b[8] = {b0.x, b0.y, b1.x, b1.y, b2.x, b2.y, b3.x, b3.y};
#pragma unroll
      for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[r][c] = fma(a[r], b[c], acc[r][c]);
    }
    if (more) gb_store<TA, TB>(t, As[buf ^ 1], Bs[buf ^ 1], ra, rb);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int64_t gi = i0 + (r < 4 ? ty * 4 + r : 64 + ty * 4 + (r - 4));
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int64_t gj = j0 + (c < 4 ? tx * 4 + c : 64 + tx * 4 + (c - 4));
      if (gi < M && gj < N) {
        if (partial) {
          partial[(int64_t)blockIdx.z * M * N + gi + gj * M] = acc[r][c];
        } else {
          double v = alpha * acc[r][c];
          if (beta != 0.0) v += beta * C[gi + gj * ldc];
          C[gi + gj * ldc] = v;
        }
      }
    }
  }
}

// Deep split-K reductions: one wave per output, lanes over the splits, then a
// fixed butterfly (deterministic order).
__global__ __launch_bounds__(256) void k_reduce_partials_wave(int64_t M, int64_t N, int splits,
                                                             const double* __restrict__ partial,
                                                             double alpha, double beta,
                                                             double* __restrict__ C,
                                                             int64_t ldc) {
  const int64_t n = M * N;
  const int lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= n) return;
  double s = 0.0;
  for (int z = lane; z < splits; z += 64) s += partial[(int64_t)z * n + t];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) {
    const int64_t i = t % M, j = t / M;
    double v = alpha * s;
    if (beta != 0.0) v += beta * C[i + j * ldc];
    C[i + j * ldc] = v;
  }
}

__global__ void k_reduce_partials_f64(int64_t M, int64_t N, int splits,
                                      const double* __restrict__ partial, double alpha,
                                      double beta, double* __restrict__ C, int64_t ldc) {
  const int64_t n = M * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < splits; ++z) s += partial[(int64_t)z * n + t];
    const int64_t i = t % M, j = t / M;
    double v = alpha * s;
    if (beta != 0.0) v += beta * C[i + j * ldc];
    C[i + j * ldc] = v;
  }
}

// C(i, j) = alpha sum_k A(i, k) op(B)(k, j) + beta C(i, j) for short K.  Block
// (., j): 64 rows i over the lanes, the four waves split k four ways
// (k = w + 4u), four accumulators each, and the wave sums are added in wave
// order through LDS.  Column j of op(B) is the same for the whole block
// (broadcast loads), A's column k is read coalesced.
__global__ __launch_bounds__(256) void k_gemm_small(int transB, int64_t M, int64_t K,
                                                    double alpha, const double* __restrict__ A,
                                                    int64_t lda, const double* __restrict__ B,
                                                    int64_t ldb, double beta,
                                                    double* __restrict__ C, int64_t ldc) {
  __shared__ double part[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t j = blockIdx.y, i = blockIdx.x * (int64_t)64 + lane;
  const int64_t bs = transB ? ldb : 1, b0 = transB ? j : j * ldb;
  const int64_t ir = i < M ? i : M - 1;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t k = w;
  for (; k + 12 < K; k += 16) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
      a[u] = fma(A[ir + (k + 4 * u) * lda], B[b0 + (k + 4 * u) * bs], a[u]);
  }
  for (; k < K; k += 4) a[0] = fma(A[ir + k * lda], B[b0 + k * bs], a[0]);
  part[w][lane] = (a[0] + a[1]) + (a[2] + a[3]);
  __syncthreads();
  if (w == 0 && i < M) {
    const double v = alpha * ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]));
    double* c = C + i + j * ldc;
    *c = beta == 0.0 ? v : fma(beta, *c, v);
  }
}

void gemm_f64(Ctx* ctx, bool transA, bool transB, int64_t M, int64_t N, int64_t K, double alpha,
              const double* A, int64_t lda, const double* shiftA, const double* B, int64_t ldb,
              const double* shiftB, double beta, double* C, int64_t ldc,
              const int32_t* labelsB, const char* name) {
  if (M <= 0 || N <= 0) return;
  const char* kname = name ? name : "gemm_f64";
  if (K <= 0) {
    // C = beta C
    DBuf<double> zero(ctx, 1);
    fill_f64(ctx, zero.get(), 0.0, 1);
    launch(ctx, "gemm_f64_reduce", k_reduce_partials_f64, dim3(grid1d(M * N)), dim3(256), 0, M,
           N, 0, zero.get(), alpha, beta, C, ldc);
    return;
  }
  // short products (K <= 256) of a small plain A (every block re-reads all of
  // A from L2, so A stays within 2 MiB and N within 512): one output per
  // thread in one launch, no split-K partials to reduce
  if (!transA && K <= 256 && M >= 32 && M * K <= (int64_t(1) << 18) && N <= 512 && !labelsB &&
      !shiftA && !shiftB) {
    launch(ctx, kname, k_gemm_small, dim3((unsigned)((M + 63) / 64), (unsigned)N), dim3(256), 0,
           transB ? 1 : 0, M, K, alpha, A, lda, B, ldb, beta, C, ldc);
    return;
  }
  // the large tile pays when there are enough of them (or a long K to split);
  // small factor products stay on the 64 x 64 tiles below
  const int64_t big_tiles = ((M + GB_M - 1) / GB_M) * ((N + GB_N - 1) / GB_N);
  if ((!labelsB || !transB) && M >= 128 && N >= 64 && K >= 64 &&
      (big_tiles >= ctx->num_cus / 2 || K >= 4096)) {
    const int64_t tiles = big_tiles;
    int64_t splits = 1;
    const int64_t want = 2 * ctx->num_cus;
    if (tiles < want && K >= 512) {
      splits = std::min<int64_t>((want + tiles - 1) / tiles, K / 256);
      const int64_t cap = std::max<int64_t>(
          1, (int64_t)(ctx->partial_budget(size_t(512) << 20) / 8) / std::max<int64_t>(1, M * N));
      splits = std::max<int64_t>(1, std::min(splits, cap));
    }
    int64_t kchunk = (K + splits - 1) / splits;
    kchunk = (kchunk + GB_K - 1) / GB_K * GB_K;
    splits = (K + kchunk - 1) / kchunk;
    dim3 grid((unsigned)((M + GB_M - 1) / GB_M), (unsigned)((N + GB_N - 1) / GB_N),
              (unsigned)splits);
    DBuf<double> part;
    if (splits > 1) part = DBuf<double>(ctx, (size_t)splits * M * N);
    double* pp = splits > 1 ? part.get() : nullptr;
    if (labelsB && !transA)
      launch(ctx, kname, k_gemm_f64_big<false, false, true>, grid, dim3(256), 0, M, N, K, kchunk,
             alpha, A, lda, shiftA, B, ldb, labelsB, shiftB, beta, C, ldc, pp);
    else if (labelsB)
      launch(ctx, kname, k_gemm_f64_big<true, false, true>, grid, dim3(256), 0, M, N, K, kchunk,
             alpha, A, lda, shiftA, B, ldb, labelsB, shiftB, beta, C, ldc, pp);
    else if (!transA && !transB)
      launch(ctx, kname, k_gemm_f64_big<false, false, false>, grid, dim3(256), 0, M, N, K,
             kchunk, alpha, A, lda, shiftA, B, ldb, labelsB, shiftB, beta, C, ldc, pp);
    else if (!transA && transB)
      launch(ctx, kname, k_gemm_f64_big<false, true, false>, grid, dim3(256), 0, M, N, K,
             kchunk, alpha, A, lda, shiftA, B, ldb, labelsB, shiftB, beta, C, ldc, pp);
    else if (transA && !transB)
      launch(ctx, kname, k_gemm_f64_big<true, false, false>, grid, dim3(256), 0, M, N, K,
             kchunk, alpha, A, lda, shiftA, B, ldb, labelsB, shiftB, beta, C, ldc, pp);
    else
      launch(ctx, kname, k_gemm_f64_big<true, true, false>, grid, dim3(256), 0, M, N, K, kchunk,
             alpha, A, lda, shiftA, B, ldb, labelsB, shiftB, beta, C, ldc, pp);
    if (splits > 1)
      launch(ctx, "gemm_f64_reduce", k_reduce_partials_f64, dim3(grid1d(M * N)), dim3(256), 0, M,
             N, (int)splits, (const double*)part.get(), alpha, beta, C, ldc);
    return;
  }
  const int64_t tiles = ((M + 63) / 64) * ((N + 63) / 64);
  int64_t splits = 1;
  const int64_t want = 2 * ctx->num_cus;
  if (tiles < want && K > 2 * G64_BK) {
    // small outputs: spread K over many workgroups (two tiles of K each at
    // least), latency rather than FLOPs bounds these products
    splits = std::min<int64_t>((want + tiles - 1) / tiles, (K + 2 * G64_BK - 1) / (2 * G64_BK));
    // keep the partial buffer modest
    const int64_t cap = std::max<int64_t>(
          1, (int64_t)(ctx->partial_budget(size_t(512) << 20) / 8) / std::max<int64_t>(1, M * N));
    splits = std::max<int64_t>(1, std::min(splits, cap));
  }
  int64_t kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + G64_BK - 1) / G64_BK * G64_BK;
  splits = (K + kchunk - 1) / kchunk;
  dim3 grid((unsigned)((M + 63) / 64), (unsigned)((N + 63) / 64), (unsigned)splits);
  if (splits == 1) {
    launch(ctx, kname, k_gemm_f64, grid, dim3(256), 0, transA, transB, M, N, K, kchunk,
           alpha, A, lda, shiftA, B, ldb, labelsB, shiftB, beta, C, ldc, (double*)nullptr);
  } else {
    DBuf<double> part(ctx, (size_t)splits * M * N);
    launch(ctx, kname, k_gemm_f64, grid, dim3(256), 0, transA, transB, M, N, K, kchunk,
           alpha, A, lda, shiftA, B, ldb, labelsB, shiftB, beta, C, ldc, part.get());
    if (splits >= 16)
      launch(ctx, "gemm_f64_reduce", k_reduce_partials_wave,
             dim3((unsigned)((M * N + 3) / 4)), dim3(256), 0, M, N, (int)splits,
             (const double*)part.get(), alpha, beta, C, ldc);
    else
      launch(ctx, "gemm_f64_reduce", k_reduce_partials_f64, dim3(grid1d(M * N)), dim3(256), 0,
             M, N, (int)splits, (const double*)part.get(), alpha, beta, C, ldc);
  }
}

// ---------------------------------------------------------------- GEMV pair
// y(M) partial over K-chunks: ypart[z][i] = sum_{k in chunk z} (A[i,k]-sh[k]) x[k]
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_n(int64_t M, int64_t K, int64_t kchunk,
                                                const T* __restrict__ A, int64_t lda,
                                                const double* __restrict__ sh,
                                                const double* __restrict__ x,
                                                double* __restrict__ ypart) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t kb = (int64_t)blockIdx.y * kchunk;
  const int64_t ke = std::min<int64_t>(K, kb + kchunk);
  if (i >= M) return;
  double s = 0.0;
  for (int64_t k = kb; k < ke; ++k) {
    const double a = static_cast<double>(A[i + k * lda]) - (sh ? sh[k] : 0.0);
    s = fma(a, x[k], s);
  }
  ypart[(int64_t)blockIdx.y * M + i] = s;
}

// 16-byte streaming load of a matrix that is read once per pass: the
// nontemporal hint keeps it from displacing the reused vectors in the caches
// (measured: 5.6 -> 6.3 TB/s for a 4 GB FP32 GEMV pass).
template <typename T>
struct StreamVec;
template <>
struct StreamVec<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct StreamVec<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <typename T, typename V>
__device__ __forceinline__ V load_stream(const T* p) {
  using S = typename StreamVec<T>::type;
  const S t = __builtin_nontemporal_load(reinterpret_cast<const S*>(p));
  V r;
#pragma unroll
  for (int i = 0; i < (int)(16 / sizeof(T)); ++i) r.v[i] = t[i];
  return r;
}

// max or sum over a 256-thread workgroup (every thread gets the result)
__device__ __forceinline__ double gn_block_reduce(double v, double* red, bool is_max) {
  v = is_max ? wave_max_d(v) : wave_sum_d(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int w = 1; w < 4; ++w) r = is_max ? fmax(r, red[w]) : r + red[w];
  return r;
}

// One lane's VW consecutive rows of the K slab [kb, ke) of (A - 1 sh^T) x.
template <typename T, int VW, int KU, typename Vec, bool AGENT_OUT = false>
__device__ __forceinline__ void gemv_nv_part(int64_t kb, int64_t ke,
                                             const T* __restrict__ a0, int64_t lda,
                                             const double* __restrict__ sh,
                                             const double* __restrict__ x,
                                             double* __restrict__ out) {
  double s[VW];
#pragma unroll
  for (int r = 0; r < VW; ++r) s[r] = 0.0;
  int64_t k = kb;
  for (; k + KU <= ke; k += KU) {
    Vec a[KU];
#pragma unroll
    for (int u = 0; u < KU; ++u) a[u] = load_stream<T, Vec>(a0 + (k + u) * lda);
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const double xv = x[k + u], sv = sh ? sh[k + u] : 0.0;
#pragma unroll
      for (int r = 0; r < VW; ++r) s[r] = fma(static_cast<double>(a[u].v[r]) - sv, xv, s[r]);
    }
  }
  for (; k < ke; ++k) {
    const Vec a = *reinterpret_cast<const Vec*>(a0 + k * lda);
    const double xv = x[k], sv = sh ? sh[k] : 0.0;
#pragma unroll
    for (int r = 0; r < VW; ++r) s[r] = fma(static_cast<double>(a.v[r]) - sv, xv, s[r]);
  }
#pragma unroll
  for (int r = 0; r < VW; ++r) {
    if (AGENT_OUT)
      __hip_atomic_store(&out[r], s[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      out[r] = s[r];
  }
}

// The same partials with VW = 16 / sizeof(T) consecutive rows per lane (one
// 16-byte load per column) and 8 columns loaded before any is used, so each
// lane keeps 128 bytes of X in flight; needs M and lda multiples of VW and a
// 16-byte aligned A.
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_nv(int64_t M, int64_t K, int64_t kchunk,
                                                 const T* __restrict__ A, int64_t lda,
                                                 const double* __restrict__ sh,
                                                 const double* __restrict__ x,
                                                 double* __restrict__ ypart) {
  constexpr int VW = 16 / sizeof(T);
  constexpr int KU = 8;
  struct alignas(16) Vec {
    T v[VW];
  };
  const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * VW;
  if (i0 >= M) return;
  const int64_t kb = (int64_t)blockIdx.y * kchunk;
  const int64_t ke = std::min<int64_t>(K, kb + kchunk);
  gemv_nv_part<T, VW, KU, Vec>(kb, ke, A + i0, lda, sh, x, ypart + (int64_t)blockIdx.y * M + i0);
}

// y = sum of the partials, four independent sums so that the loads overlap
__global__ void k_sum_parts(int64_t M, int parts, const double* __restrict__ part,
                            double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    int z = 0;
    for (; z + 4 <= parts; z += 4)
#pragma unroll
      for (int u = 0; u < 4; ++u) s[u] += part[(int64_t)(z + u) * M + i];
    for (; z < parts; ++z) s[0] += part[(int64_t)z * M + i];
    y[i] = (s[0] + s[1]) + (s[2] + s[3]);
  }
}

// y = (A - 1 sh^T) x over K slabs as k_gemv_nv + k_sum_parts, then, in the
// last workgroup to finish (atomic ticket; partials written and read with
// agent-scope atomics), y /= |y| and also /= |y| with the scaled norm of a
// one-partial normalisation (a zero or non-finite norm sets flags[0] and
// leaves y unnormalised; nothing is normalised once flags[0] is set).  One
// launch for M <= GN_MMAX rows.
constexpr int GN_MMAX = 2048;
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_nv_norm(int64_t M, int64_t K, int64_t kchunk,
                                                      const T* __restrict__ A, int64_t lda,
                                                      const double* __restrict__ sh,
                                                      const double* x, double* ypart, double* y,
                                                      double* also, int64_t n2,
                                                      int32_t* flags, unsigned* counter) {
  constexpr int VW = 16 / sizeof(T);
  constexpr int KU = 8;
  struct alignas(16) Vec {
    T v[VW];
  };
  __shared__ unsigned ticket;
  __shared__ double red[4];
  const int t = threadIdx.x;
  const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + t) * VW;
  if (i0 < M) {
    const int64_t kb = (int64_t)blockIdx.y * kchunk;
    const int64_t ke = std::min<int64_t>(K, kb + kchunk);
    gemv_nv_part<T, VW, KU, Vec, true>(kb, ke, A + i0, lda, sh, x,
                                       ypart + (int64_t)blockIdx.y * M + i0);
  }
  __builtin_amdgcn_s_waitcnt(0);  // this thread's partial stores have completed
  __syncthreads();
  if (t == 0)
    ticket = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (ticket != gridDim.x * gridDim.y - 1) return;
  if (t == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int parts = (int)gridDim.y;
  constexpr int NE = GN_MMAX / 256;
  double v[NE], amax = 0.0, bad = 0.0;
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t i = t + (int64_t)e * 256;
    v[e] = 0.0;
    if (i < M) {
      double s4[4] = {0.0, 0.0, 0.0, 0.0};
      int z = 0;
      for (; z + 4 <= parts; z += 4)
#pragma unroll
        for (int u = 0; u < 4; ++u)
          s4[u] += __hip_atomic_load(&ypart[(int64_t)(z + u) * M + i], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
      for (; z < parts; ++z)
        s4[0] += __hip_atomic_load(&ypart[(int64_t)z * M + i], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
      v[e] = (s4[0] + s4[1]) + (s4[2] + s4[3]);
      y[i] = v[e];
    }
    const double a = fabs(v[e]);
    if (!(a <= DBL_MAX)) bad = 1.0;
    else amax = fmax(amax, a);
  }
  if (flags[0]) return;  // uniform over the workgroup
  bad = gn_block_reduce(bad, red, true);
  amax = gn_block_reduce(amax, red, true);
  double ssq = 0.0;
  if (bad == 0.0 && amax > 0.0) {
    const double inv = 1.0 / amax;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      const double u = v[e] * inv;
      ssq = fma(u, u, ssq);
    }
  }
  ssq = gn_block_reduce(ssq, red, false);
  const double nr = bad != 0.0 ? __builtin_nan("") : amax * sqrt(ssq);
  if (!(nr > 0.0 && nr <= DBL_MAX)) {
    if (t == 0) flags[0] = 1;
    return;
  }
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t i = t + (int64_t)e * 256;
    if (i < M) y[i] = v[e] / nr;
  }
  if (also)
    for (int64_t i = t; i < n2; i += 256) also[i] /= nr;
}

bool gemv_n_normalize(Ctx* ctx, const DMat& A, const double* shift, const double* x, double* y,
                      double* also, int64_t n2, int32_t* flags) {
  const int64_t M = A.rows, K = A.cols;
  const int vw = A.dtype == FPLS_F32 ? 4 : 2;
  if (M <= 0 || M > GN_MMAX || K <= 0 || M % vw != 0 || A.ld % vw != 0 ||
      (reinterpret_cast<uintptr_t>(A.p) & 15) != 0 || ctx->arena_base)
    return false;
  // K slabs of at least 128 columns: the last workgroup's sum over the
  // partials is a chain of dependent L2 round trips, so fewer, longer slabs
  // (the matrices that fit this form are at most 2048 rows) finish sooner
  const int64_t rb = (M / vw + 255) / 256;
  int64_t parts = std::max<int64_t>(1, std::min<int64_t>((8 * ctx->num_cus + rb - 1) / rb,
                                                         (K + 127) / 128));
  const int64_t kchunk = (K + parts - 1) / parts;
  parts = std::max<int64_t>(1, (K + kchunk - 1) / kchunk);
  DBuf<double> part(ctx, (size_t)parts * M);
  dim3 grid((unsigned)rb, (unsigned)parts);
  if (A.dtype == FPLS_F32)
    launch(ctx, "gemv_n_norm", k_gemv_nv_norm<float>, grid, dim3(256), 0, M, K, kchunk, A.f(),
           A.ld, shift, x, part.get(), y, also, n2, flags, ctx->ticket() + 1);
  else
    launch(ctx, "gemv_n_norm", k_gemv_nv_norm<double>, grid, dim3(256), 0, M, K, kchunk, A.d(),
           A.ld, shift, x, part.get(), y, also, n2, flags, ctx->ticket() + 1);
  return true;
}

// y(K): one wavefront per column, lanes stride the rows.
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_t(int64_t M, int64_t K, const T* __restrict__ A,
                                                int64_t lda, const double* __restrict__ sh,
                                                const double* __restrict__ x,
                                                double* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= K) return;
  const double s0 = sh ? sh[k] : 0.0;
  const T* col = A + k * lda;
  double s = 0.0;
  for (int64_t i = lane; i < M; i += 64) s = fma(static_cast<double>(col[i]) - s0, x[i], s);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[k] = s;
}

// One wave's CW columns k0 .. k0 + CW - 1 over the rows of chunk blockIdx.y.
template <typename T, int CW>
__device__ __forceinline__ void gemv_tw_part(int64_t M, int64_t K, int64_t mchunk, int64_t k0,
                                             const T* __restrict__ A, int64_t lda,
                                             const double* __restrict__ sh,
                                             const double* __restrict__ x,
                                             double* __restrict__ o) {
  constexpr int VW = 16 / sizeof(T);
  constexpr int U = 2;
  constexpr int64_t STEP = 64 * VW;
  struct alignas(16) Vec {
    T v[VW];
  };
  struct alignas(16) XV {
    double v[VW];
  };
  const int lane = threadIdx.x & 63;
  const int64_t mb = (int64_t)blockIdx.y * mchunk;
  const int64_t me = std::min<int64_t>(M, mb + mchunk);
  const T* col[CW];
  double s0[CW], s[CW];
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    const int64_t kc = std::min<int64_t>(k0 + c, K - 1);  // columns past K: loaded, not stored
    col[c] = A + kc * lda;
    s0[c] = sh ? sh[kc] : 0.0;
    s[c] = 0.0;
  }
  int64_t i = mb + (int64_t)lane * VW;
  // i and me multiples of VW: i + (U - 1) STEP < me keeps all U steps in range
  for (; i + (U - 1) * STEP < me; i += U * STEP) {
    XV xv[U];
    Vec a[U][CW];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = *reinterpret_cast<const XV*>(x + i + u * STEP);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int c = 0; c < CW; ++c) a[u][c] = load_stream<T, Vec>(col[c] + i + u * STEP);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int c = 0; c < CW; ++c)
#pragma unroll
        for (int r = 0; r < VW; ++r)
          s[c] = fma(static_cast<double>(a[u][c].v[r]) - s0[c], xv[u].v[r], s[c]);
  }
  for (; i < me; i += STEP) {
    const XV xv = *reinterpret_cast<const XV*>(x + i);
#pragma unroll
    for (int c = 0; c < CW; ++c) {
      const Vec a = *reinterpret_cast<const Vec*>(col[c] + i);
#pragma unroll
      for (int r = 0; r < VW; ++r) s[c] = fma(static_cast<double>(a.v[r]) - s0[c], xv.v[r], s[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    const double t = wave_sum_d(s[c]);
    if (lane == 0 && k0 + c < K) o[k0 + c] = t;
  }
}

// y(K) = (A - 1 sh^T)^T x with CW columns per wave: each step a lane loads
// VW rows of x once and the same rows of the wave's CW columns, so x is read
// once per CW columns, and U steps of loads are issued before any is used.
// gridDim.y row chunks of mchunk rows (a multiple of VW) write the partial
// sums out[chunk * K + k] (y itself with one chunk).  Needs M and lda
// multiples of VW and a 16-byte aligned A.
template <typename T, int CW>
__global__ __launch_bounds__(256) void k_gemv_tw(int64_t M, int64_t K, int64_t mchunk,
                                                 const T* __restrict__ A, int64_t lda,
                                                 const double* __restrict__ sh,
                                                 const double* __restrict__ x,
                                                 double* __restrict__ out) {
  const int64_t k0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * CW;
  if (k0 < K) gemv_tw_part<T, CW>(M, K, mchunk, k0, A, lda, sh, x, out + (int64_t)blockIdx.y * K);
}

void gemv_n(Ctx* ctx, const DMat& A, const double* shift, const double* x, double* y) {
  const int64_t M = A.rows, K = A.cols;
  if (M <= 0) return;
  const int vw = A.dtype == FPLS_F32 ? 4 : 2;
  if (K > 0 && M % vw == 0 && A.ld % vw == 0 && (reinterpret_cast<uintptr_t>(A.p) & 15) == 0) {
    // vector path: about eight 256-thread workgroups per CU, K slabs of >= 16
    const int64_t rb = (M / vw + 255) / 256;
    int64_t parts = std::max<int64_t>(1, std::min<int64_t>((8 * ctx->num_cus + rb - 1) / rb,
                                                           (K + 15) / 16));
    const int64_t kchunk = (K + parts - 1) / parts;
    parts = std::max<int64_t>(1, (K + kchunk - 1) / kchunk);
    // one K slab writes y itself
    DBuf<double> part;
    double* out = y;
    if (parts > 1) {
      part = DBuf<double>(ctx, (size_t)parts * M);
      out = part.get();
    }
    dim3 grid((unsigned)rb, (unsigned)parts);
    if (A.dtype == FPLS_F32)
      launch(ctx, "gemv_n", k_gemv_nv<float>, grid, dim3(256), 0, M, K, kchunk, A.f(), A.ld,
             shift, x, out);
    else
      launch(ctx, "gemv_n", k_gemv_nv<double>, grid, dim3(256), 0, M, K, kchunk, A.d(), A.ld,
             shift, x, out);
    if (parts > 1)
      launch(ctx, "gemv_n_reduce", k_sum_parts, dim3(grid1d(M)), dim3(256), 0, M, (int)parts,
             (const double*)part.get(), y);
    return;
  }
  const int64_t rb = (M + 255) / 256;
  int64_t parts = std::max<int64_t>(1, std::min<int64_t>((2 * ctx->num_cus + rb - 1) / rb,
                                                         (K + 63) / 64));
  const int64_t kchunk = (K + parts - 1) / parts;
  parts = std::max<int64_t>(1, (K + kchunk - 1) / kchunk);
  DBuf<double> part(ctx, (size_t)parts * M);
  dim3 grid((unsigned)rb, (unsigned)parts);
  if (A.dtype == FPLS_F32)
    launch(ctx, "gemv_n", k_gemv_n<float>, grid, dim3(256), 0, M, K, kchunk, A.f(), A.ld, shift,
           x, part.get());
  else
    launch(ctx, "gemv_n", k_gemv_n<double>, grid, dim3(256), 0, M, K, kchunk, A.d(), A.ld,
           shift, x, part.get());
  launch(ctx, "gemv_n_reduce", k_sum_parts, dim3(grid1d(M)), dim3(256), 0, M, (int)parts,
         part.get(), y);
}

template <int CW>
static void gemv_t_wide(Ctx* ctx, const DMat& A, const double* shift, const double* x,
                        double* y) {
  const int64_t M = A.rows, K = A.cols;
  const int vw = A.dtype == FPLS_F32 ? 4 : 2;
  const int64_t step = 64 * vw;
  const int64_t wblocks = (K + 4 * CW - 1) / (4 * CW);
  // row chunks of >= 4 steps until about four workgroups per CU
  int64_t splits = 1;
  if (wblocks < 4 * ctx->num_cus)
    splits = std::max<int64_t>(1, std::min<int64_t>((4 * ctx->num_cus + wblocks - 1) / wblocks,
                                                    M / (4 * step)));
  int64_t mchunk = (M + splits - 1) / splits;
  mchunk = (mchunk + step - 1) / step * step;
  splits = (M + mchunk - 1) / mchunk;
  DBuf<double> part;
  double* out = y;
  if (splits > 1) {
    part = DBuf<double>(ctx, (size_t)splits * K);
    out = part.get();
  }
  dim3 grid((unsigned)wblocks, (unsigned)splits);
  if (A.dtype == FPLS_F32)
    launch(ctx, "gemv_t", k_gemv_tw<float, CW>, grid, dim3(256), 0, M, K, mchunk, A.f(), A.ld,
           shift, x, out);
  else
    launch(ctx, "gemv_t", k_gemv_tw<double, CW>, grid, dim3(256), 0, M, K, mchunk, A.d(), A.ld,
           shift, x, out);
  if (splits > 1)
    launch(ctx, "gemv_t_reduce", k_sum_parts, dim3(grid1d(K)), dim3(256), 0, K, (int)splits,
           (const double*)part.get(), y);
}

void gemv_t(Ctx* ctx, const DMat& A, const double* shift, const double* x, double* y) {
  const int64_t M = A.rows, K = A.cols;
  if (K <= 0) return;
  dim3 grid((unsigned)((K + 3) / 4));
  const int vw = A.dtype == FPLS_F32 ? 4 : 2;
  if (M > 0 && M % vw == 0 && A.ld % vw == 0 && (reinterpret_cast<uintptr_t>(A.p) & 15) == 0 &&
      (reinterpret_cast<uintptr_t>(x) & 15) == 0) {
    if (K >= 16)
      gemv_t_wide<8>(ctx, A, shift, x, y);
    else
      gemv_t_wide<1>(ctx, A, shift, x, y);
    return;
  }
  if (A.dtype == FPLS_F32)
    launch(ctx, "gemv_t", k_gemv_t<float>, grid, dim3(256), 0, M, K, A.f(), A.ld, shift, x, y);
  else
    launch(ctx, "gemv_t", k_gemv_t<double>, grid, dim3(256), 0, M, K, A.d(), A.ld, shift, x, y);
}

}  // namespace fpls
