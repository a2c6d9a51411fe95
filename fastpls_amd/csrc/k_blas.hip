// Dense contractions for gfx950.
//
//  * gemm_f64     -- LDS-tiled FP64 GEMM (64x64 tiles, split-K for long K)
//                    for the small/medium FP64 products of the factorisations.
//  * gemm_nn_f32  -- X (n x p, FP32) times a skinny FP64 factor on MFMA
//                    (v_mfma_f32_16x16x4f32), centring applied on load and an
//                    offset epilogue: T = (X - 1 mX^T) U - 1 o^T.
//  * gemm_tn_f32  -- cross products reduced over the long row dimension
//                    (X^T Y, T^T T) on MFMA with FP32 K-slab partials reduced
//                    in FP64 in a fixed order (deterministic).
//  * gemv_n/_t    -- the SIMPLS matrix-vector pair t = Xc r, p = Xc^T t.
//
// Layouts are column-major throughout.  MFMA 16x16x4 f32 lane layout
// (CDNA3/CDNA4): lane l supplies A[l%16][l/16] and B[l/16][l%16]; the four
// accumulator registers hold D[4*(l/16)+r][l%16], r = 0..3.
#include <algorithm>
#include <cmath>

#include "fpls_internal.h"

namespace fpls {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

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

  for (int64_t k0 = kbeg; k0 < kend; k0 += G64_BK) {
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
      As[k][i] = v;
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
      Bs[k][j] = v;
    }
    __syncthreads();
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

void gemm_f64(Ctx* ctx, bool transA, bool transB, int64_t M, int64_t N, int64_t K, double alpha,
              const double* A, int64_t lda, const double* shiftA, const double* B, int64_t ldb,
              const double* shiftB, double beta, double* C, int64_t ldc,
              const int32_t* labelsB) {
  if (M <= 0 || N <= 0) return;
  if (K <= 0) {
    // C = beta C
    DBuf<double> zero(ctx, 1);
    fill_f64(ctx, zero.get(), 0.0, 1);
    launch(ctx, "gemm_f64_reduce", k_reduce_partials_f64, dim3(grid1d(M * N)), dim3(256), 0, M,
           N, 0, zero.get(), alpha, beta, C, ldc);
    return;
  }
  const int64_t tiles = ((M + 63) / 64) * ((N + 63) / 64);
  int64_t splits = 1;
  const int64_t want = 2 * ctx->num_cus;
  if (tiles < want && K > 256) {
    splits = std::min<int64_t>((want + tiles - 1) / tiles, (K + 255) / 256);
    // keep the partial buffer modest
    const int64_t cap = std::max<int64_t>(1, (int64_t(64) << 20) / std::max<int64_t>(1, M * N));
    splits = std::max<int64_t>(1, std::min(splits, cap));
  }
  int64_t kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + G64_BK - 1) / G64_BK * G64_BK;
  splits = (K + kchunk - 1) / kchunk;
  dim3 grid((unsigned)((M + 63) / 64), (unsigned)((N + 63) / 64), (unsigned)splits);
  if (splits == 1) {
    launch(ctx, "gemm_f64", k_gemm_f64, grid, dim3(256), 0, transA, transB, M, N, K, kchunk,
           alpha, A, lda, shiftA, B, ldb, labelsB, shiftB, beta, C, ldc, (double*)nullptr);
  } else {
    DBuf<double> part(ctx, (size_t)splits * M * N);
    launch(ctx, "gemm_f64", k_gemm_f64, grid, dim3(256), 0, transA, transB, M, N, K, kchunk,
           alpha, A, lda, shiftA, B, ldb, labelsB, shiftB, beta, C, ldc, part.get());
    launch(ctx, "gemm_f64_reduce", k_reduce_partials_f64, dim3(grid1d(M * N)), dim3(256), 0, M,
           N, (int)splits, part.get(), alpha, beta, C, ldc);
  }
}

// ---------------------------------------------------------------- MFMA helpers
__device__ inline v4f mfma4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------- NN on MFMA
// C(M x N) = (A - 1 shA^T) B - 1 off^T.  BM = 64, BK = 16, BN in {64, 128}.
// 4 waves: wave w covers rows [32*(w%2), +32) and cols [(BN/2)*(w/2), +BN/2).
template <int BN>
__global__ __launch_bounds__(256) void k_gemm_nn_f32(int64_t M, int64_t N, int64_t K,
                                                     const float* __restrict__ A, int64_t lda,
                                                     const double* __restrict__ shA,
                                                     const float* __restrict__ B, int64_t ldb,
                                                     const double* __restrict__ off,
                                                     float* __restrict__ Cf,
                                                     double* __restrict__ Cd, int64_t ldc,
                                                     int crow) {
  constexpr int BM = 64, BK = 16;
  constexpr int NB = BN / 32;  // 16-col blocks per wave
  __shared__ float As[BK][BM];
  __shared__ float Bs[BK][BN];
  __shared__ float Cs[BN][BM + 1];
  __shared__ float shs[BK];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int64_t row0 = (int64_t)blockIdx.x * BM;
  const int64_t col0 = (int64_t)blockIdx.y * BN;
  v4f acc[2][NB];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};

  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    if (t < BK) shs[t] = (shA && k0 + t < K) ? static_cast<float>(shA[k0 + t]) : 0.f;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = t % 64, k = t / 64 + 4 * r;
      const int64_t gi = row0 + i, gk = k0 + k;
      float v = 0.f;
      if (gi < M && gk < K) v = A[gi + gk * lda] - shs[k];
      As[k][i] = v;
    }
#pragma unroll
    for (int r = 0; r < BN / 16; ++r) {
      const int k = t % 16, j = t / 16 + 16 * r;
      const int64_t gj = col0 + j, gk = k0 + k;
      Bs[k][j] = (gj < N && gk < K) ? B[gk + gj * ldb] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float af[2], bf[NB];
#pragma unroll
      for (int a = 0; a < 2; ++a) af[a] = As[kk + lane / 16][wm * 32 + a * 16 + lane % 16];
#pragma unroll
      for (int b = 0; b < NB; ++b) bf[b] = Bs[kk + lane / 16][wn * (BN / 2) + b * 16 + lane % 16];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = mfma4(af[a], bf[b], acc[a][b]);
    }
    __syncthreads();
  }
  // stage the tile in LDS (column-major) and write coalesced columns
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = wm * 32 + a * 16 + 4 * (lane / 16) + r;
        const int j = wn * (BN / 2) + b * 16 + lane % 16;
        Cs[j][i] = acc[a][b][r];
      }
  __syncthreads();
  for (int idx = t; idx < BM * BN; idx += 256) {
    // column-major: rows fastest; row-major (crow): columns fastest
    const int i = crow ? idx / BN : idx % BM, j = crow ? idx % BN : idx / BM;
    const int64_t gi = row0 + i, gj = col0 + j;
    if (gi < M && gj < N) {
      double v = Cs[j][i];
      if (off) v -= off[gj];
      const int64_t o = crow ? gi * ldc + gj : gi + gj * ldc;
      if (Cd) Cd[o] = v;
      else Cf[o] = static_cast<float>(v);
    }
  }
}

// Split-BF16x3 variant: a = ah + al, b = bh + bl (bf16 parts), and
// a b ~= ah bh + ah bl + al bh on v_mfma_f32_16x16x32_bf16.
// Operands are staged as FP32 in LDS and split in registers.
__device__ inline void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = static_cast<__bf16>(x);
  lo = static_cast<__bf16>(x - static_cast<float>(hi));
}

template <int BN>
__global__ __launch_bounds__(256) void k_gemm_nn_bf16x3(int64_t M, int64_t N, int64_t K,
                                                        const float* __restrict__ A, int64_t lda,
                                                        const double* __restrict__ shA,
                                                        const float* __restrict__ B,
                                                        int64_t ldb,
                                                        const double* __restrict__ off,
                                                        float* __restrict__ Cf,
                                                        double* __restrict__ Cd, int64_t ldc,
                                                        int crow) {
  constexpr int BM = 64, BK = 32;
  constexpr int NB = BN / 32;
  __shared__ float As[BM][BK + 1];   // row-major per tile row: lane reads 8 consecutive k
  __shared__ float Bs[BN][BK + 1];
  __shared__ float Cs[BN][BM + 1];
  __shared__ float shs[BK];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int64_t row0 = (int64_t)blockIdx.x * BM;
  const int64_t col0 = (int64_t)blockIdx.y * BN;
  v4f acc[2][NB];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};

  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    if (t < BK) shs[t] = (shA && k0 + t < K) ? static_cast<float>(shA[k0 + t]) : 0.f;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int i = t % 64, k = t / 64 + 4 * r;
      const int64_t gi = row0 + i, gk = k0 + k;
      As[i][k] = (gi < M && gk < K) ? A[gi + gk * lda] - shs[k] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < BN / 8; ++r) {
      const int k = t % 32, j = t / 32 + 8 * r;
      const int64_t gj = col0 + j, gk = k0 + k;
      Bs[j][k] = (gj < N && gk < K) ? B[gk + gj * ldb] : 0.f;
    }
    __syncthreads();
    v8bf ah[2], al[2], bh[NB], bl[NB];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int i = wm * 32 + a * 16 + lane % 16;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        __bf16 h, l;
        split_bf16(As[i][8 * (lane / 16) + e], h, l);
        ah[a][e] = h;
        al[a][e] = l;
      }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = wn * (BN / 2) + b * 16 + lane % 16;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        __bf16 h, l;
        split_bf16(Bs[j][8 * (lane / 16) + e], h, l);
        bh[b][e] = h;
        bl[b][e] = l;
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
      }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = wm * 32 + a * 16 + 4 * (lane / 16) + r;
        const int j = wn * (BN / 2) + b * 16 + lane % 16;
        Cs[j][i] = acc[a][b][r];
      }
  __syncthreads();
  for (int idx = t; idx < BM * BN; idx += 256) {
    // column-major: rows fastest; row-major (crow): columns fastest
    const int i = crow ? idx / BN : idx % BM, j = crow ? idx % BN : idx / BM;
    const int64_t gi = row0 + i, gj = col0 + j;
    if (gi < M && gj < N) {
      double v = Cs[j][i];
      if (off) v -= off[gj];
      const int64_t o = crow ? gi * ldc + gj : gi + gj * ldc;
      if (Cd) Cd[o] = v;
      else Cf[o] = static_cast<float>(v);
    }
  }
}

void gemm_nn_f32(Ctx* ctx, int precision, int64_t M, int64_t N, int64_t K, const float* A,
                 int64_t lda, const double* shiftA, const float* B, int64_t ldb,
                 const double* offset, float* Cf, double* Cd, int64_t ldc, bool c_rowmajor) {
  const int crow = c_rowmajor ? 1 : 0;
  if (M <= 0 || N <= 0) return;
  FPLS_REQUIRE((Cf != nullptr) != (Cd != nullptr), FPLS_ERR_ARG, "gemm_nn_f32: one output");
  const bool wide = N > 64;
  const int BN = wide ? 128 : 64;
  dim3 grid((unsigned)((M + 63) / 64), (unsigned)((N + BN - 1) / BN));
  if (precision == FPLS_PREC_BF16X3) {
    if (wide)
      launch(ctx, "gemm_nn_bf16x3", k_gemm_nn_bf16x3<128>, grid, dim3(256), 0, M, N, K, A, lda,
             shiftA, B, ldb, offset, Cf, Cd, ldc, crow);
    else
      launch(ctx, "gemm_nn_bf16x3", k_gemm_nn_bf16x3<64>, grid, dim3(256), 0, M, N, K, A, lda,
             shiftA, B, ldb, offset, Cf, Cd, ldc, crow);
  } else {
    if (wide)
      launch(ctx, "gemm_nn_f32", k_gemm_nn_f32<128>, grid, dim3(256), 0, M, N, K, A, lda, shiftA,
             B, ldb, offset, Cf, Cd, ldc, crow);
    else
      launch(ctx, "gemm_nn_f32", k_gemm_nn_f32<64>, grid, dim3(256), 0, M, N, K, A, lda, shiftA,
             B, ldb, offset, Cf, Cd, ldc, crow);
  }
}

// ---------------------------------------------------------------- TN on MFMA
// partial[z](M x N) = sum over rows k in slab z of (A[k,m]-shA[m]) (B[k,n]-shB[n]).
// A: K x M (lda), B: K x N (ldb).  BM = BN = 64, BK = 32.
__global__ __launch_bounds__(256) void k_gemm_tn_f32(int64_t M, int64_t N, int64_t K,
                                                     int64_t kchunk, const float* __restrict__ A,
                                                     int64_t lda, const double* __restrict__ shA,
                                                     const float* __restrict__ B, int64_t ldb,
                                                     const int32_t* __restrict__ labB,
                                                     const double* __restrict__ shB,
                                                     float* __restrict__ partial) {
  constexpr int BM = 64, BN = 64, BK = 32;
  __shared__ float As[BM][BK + 1];
  __shared__ float Bs[BN][BK + 1];
  __shared__ float Cs[BN][BM + 1];
  __shared__ float sha[BM], shb[BN];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int64_t m0 = (int64_t)blockIdx.x * BM;
  const int64_t n0 = (int64_t)blockIdx.y * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  const int64_t kend = std::min<int64_t>(K, kbeg + kchunk);
  if (t < BM) sha[t] = (shA && m0 + t < M) ? static_cast<float>(shA[m0 + t]) : 0.f;
  else if (t < BM + BN) {
    const int j = t - BM;
    shb[j] = (shB && n0 + j < N) ? static_cast<float>(shB[n0 + j]) : 0.f;
  }
  v4f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  __syncthreads();
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int k = t % 32, m = t / 32 + 8 * r;
      const int64_t gk = k0 + k, gm = m0 + m;
      As[m][k] = (gk < kend && gm < M) ? A[gk + gm * lda] - sha[m] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int k = t % 32, n = t / 32 + 8 * r;
      const int64_t gk = k0 + k, gn = n0 + n;
      float bv = 0.f;
      if (gk < kend && gn < N) {
        const float raw = labB ? (labB[gk] == gn + 1 ? 1.f : 0.f) : B[gk + gn * ldb];
        bv = raw - shb[n];
      }
      Bs[n][k] = bv;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float af[2], bf[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) af[a] = As[wm * 32 + a * 16 + lane % 16][kk + lane / 16];
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = Bs[wn * 32 + b * 16 + lane % 16][kk + lane / 16];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = mfma4(af[a], bf[b], acc[a][b]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = wm * 32 + a * 16 + 4 * (lane / 16) + r;
        const int j = wn * 32 + b * 16 + lane % 16;
        Cs[j][i] = acc[a][b][r];
      }
  __syncthreads();
  float* P = partial + (int64_t)blockIdx.z * M * N;
  for (int idx = t; idx < BM * BN; idx += 256) {
    const int i = idx % BM, j = idx / BM;
    const int64_t gm = m0 + i, gn = n0 + j;
    if (gm < M && gn < N) P[gm + gn * M] = Cs[j][i];
  }
}

__global__ __launch_bounds__(256) void k_gemm_tn_bf16x3(int64_t M, int64_t N, int64_t K,
                                                        int64_t kchunk,
                                                        const float* __restrict__ A, int64_t lda,
                                                        const double* __restrict__ shA,
                                                        const float* __restrict__ B, int64_t ldb,
                                                        const int32_t* __restrict__ labB,
                                                        const double* __restrict__ shB,
                                                        float* __restrict__ partial) {
  constexpr int BM = 64, BN = 64, BK = 32;
  __shared__ float As[BM][BK + 1];
  __shared__ float Bs[BN][BK + 1];
  __shared__ float Cs[BN][BM + 1];
  __shared__ float sha[BM], shb[BN];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int64_t m0 = (int64_t)blockIdx.x * BM;
  const int64_t n0 = (int64_t)blockIdx.y * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  const int64_t kend = std::min<int64_t>(K, kbeg + kchunk);
  if (t < BM) sha[t] = (shA && m0 + t < M) ? static_cast<float>(shA[m0 + t]) : 0.f;
  else if (t < BM + BN) {
    const int j = t - BM;
    shb[j] = (shB && n0 + j < N) ? static_cast<float>(shB[n0 + j]) : 0.f;
  }
  v4f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  __syncthreads();
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int k = t % 32, m = t / 32 + 8 * r;
      const int64_t gk = k0 + k, gm = m0 + m;
      As[m][k] = (gk < kend && gm < M) ? A[gk + gm * lda] - sha[m] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int k = t % 32, n = t / 32 + 8 * r;
      const int64_t gk = k0 + k, gn = n0 + n;
      float bv = 0.f;
      if (gk < kend && gn < N) {
        const float raw = labB ? (labB[gk] == gn + 1 ? 1.f : 0.f) : B[gk + gn * ldb];
        bv = raw - shb[n];
      }
      Bs[n][k] = bv;
    }
    __syncthreads();
    v8bf ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int i = wm * 32 + a * 16 + lane % 16;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        __bf16 h, l;
        split_bf16(As[i][8 * (lane / 16) + e], h, l);
        ah[a][e] = h;
        al[a][e] = l;
      }
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = wn * 32 + b * 16 + lane % 16;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        __bf16 h, l;
        split_bf16(Bs[j][8 * (lane / 16) + e], h, l);
        bh[b][e] = h;
        bl[b][e] = l;
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
      }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = wm * 32 + a * 16 + 4 * (lane / 16) + r;
        const int j = wn * 32 + b * 16 + lane % 16;
        Cs[j][i] = acc[a][b][r];
      }
  __syncthreads();
  float* P = partial + (int64_t)blockIdx.z * M * N;
  for (int idx = t; idx < BM * BN; idx += 256) {
    const int i = idx % BM, j = idx / BM;
    const int64_t gm = m0 + i, gn = n0 + j;
    if (gm < M && gn < N) P[gm + gn * M] = Cs[j][i];
  }
}

__global__ void k_reduce_partials_f32(int64_t M, int64_t N, int splits,
                                      const float* __restrict__ partial,
                                      double* __restrict__ C, int64_t ldc) {
  const int64_t n = M * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < splits; ++z) s += static_cast<double>(partial[(int64_t)z * n + t]);
    const int64_t i = t % M, j = t / M;
    C[i + j * ldc] = s;
  }
}

void gemm_tn_f32(Ctx* ctx, int precision, int64_t M, int64_t N, int64_t K, const float* A,
                 int64_t lda, const double* shiftA, const float* B, int64_t ldb,
                 const int32_t* labelsB, const double* shiftB, double* C, int64_t ldc) {
  if (M <= 0 || N <= 0) return;
  if (K <= 0) {
    for (int64_t j = 0; j < N; ++j) fill_f64(ctx, C + j * ldc, 0.0, M);
    return;
  }
  const int64_t tiles = ((M + 63) / 64) * ((N + 63) / 64);
  const int64_t want = 4 * ctx->num_cus;
  int64_t splits = std::max<int64_t>(1, (want + tiles - 1) / tiles);
  splits = std::min<int64_t>(splits, (K + 255) / 256);  // at least 256 rows per slab
  const int64_t cap = std::max<int64_t>(1, (int64_t(256) << 20) / (4 * M * N));
  splits = std::max<int64_t>(1, std::min(splits, cap));
  int64_t kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + 31) / 32 * 32;
  splits = (K + kchunk - 1) / kchunk;
  DBuf<float> part(ctx, (size_t)splits * M * N);
  dim3 grid((unsigned)((M + 63) / 64), (unsigned)((N + 63) / 64), (unsigned)splits);
  if (precision == FPLS_PREC_BF16X3)
    launch(ctx, "gemm_tn_bf16x3", k_gemm_tn_bf16x3, grid, dim3(256), 0, M, N, K, kchunk, A, lda,
           shiftA, B, ldb, labelsB, shiftB, part.get());
  else
    launch(ctx, "gemm_tn_f32", k_gemm_tn_f32, grid, dim3(256), 0, M, N, K, kchunk, A, lda,
           shiftA, B, ldb, labelsB, shiftB, part.get());
  launch(ctx, "gemm_tn_reduce", k_reduce_partials_f32, dim3(grid1d(M * N)), dim3(256), 0, M, N,
         (int)splits, part.get(), C, ldc);
}

// C (M x N, column-major FP32, ld M) = A B^T with A (M x K) and B (N x K) both
// row-major (K contiguous): the TN kernel with a single K slab writes C directly.
void gemm_rr_f32(Ctx* ctx, int precision, int64_t M, int64_t N, int64_t K, const float* A,
                 int64_t lda, const float* B, int64_t ldb, float* C) {
  if (M <= 0 || N <= 0) return;
  const int64_t kchunk = ((std::max<int64_t>(K, 1) + 31) / 32) * 32;
  dim3 grid((unsigned)((M + 63) / 64), (unsigned)((N + 63) / 64), 1);
  if (precision == FPLS_PREC_BF16X3)
    launch(ctx, "gemm_rr_bf16x3", k_gemm_tn_bf16x3, grid, dim3(256), 0, M, N, K, kchunk, A, lda,
           (const double*)nullptr, B, ldb, (const int32_t*)nullptr, (const double*)nullptr, C);
  else
    launch(ctx, "gemm_rr_f32", k_gemm_tn_f32, grid, dim3(256), 0, M, N, K, kchunk, A, lda,
           (const double*)nullptr, B, ldb, (const int32_t*)nullptr, (const double*)nullptr, C);
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

__global__ void k_sum_parts(int64_t M, int parts, const double* __restrict__ part,
                            double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < parts; ++z) s += part[(int64_t)z * M + i];
    y[i] = s;
  }
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

void gemv_n(Ctx* ctx, const DMat& A, const double* shift, const double* x, double* y) {
  const int64_t M = A.rows, K = A.cols;
  if (M <= 0) return;
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

void gemv_t(Ctx* ctx, const DMat& A, const double* shift, const double* x, double* y) {
  const int64_t M = A.rows, K = A.cols;
  if (K <= 0) return;
  dim3 grid((unsigned)((K + 3) / 4));
  if (A.dtype == FPLS_F32)
    launch(ctx, "gemv_t", k_gemv_t<float>, grid, dim3(256), 0, M, K, A.f(), A.ld, shift, x, y);
  else
    launch(ctx, "gemv_t", k_gemv_t<double>, grid, dim3(256), 0, M, K, A.d(), A.ld, shift, x, y);
}

}  // namespace fpls
