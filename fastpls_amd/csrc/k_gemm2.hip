// Pipelined MFMA contractions for FP32 data (gfx950).
//
//   NN:  C (M x N) = (A - 1 shA^T) B - 1 off^T      A: M x K col-major (X), B: K x N
//   TN:  P_z (M x N) = sum_{k in slab z} (A[k,:] - shA)^T (B[k,:] - shB)
//        A: K x M col-major (X), B: K x N col-major or the one-hot of int32 labels
//
// 128 x 128 output tiles, BK = 32, 256 threads = 4 waves in a 2 x 2 grid, each
// wave 64 x 64 = 4 x 4 blocks of 16 x 16.  The next K tile is fetched into
// registers (float4 loads) while the MFMAs of the current tile run; one barrier
// per tile.  FP32 path: v_mfma_f32_16x16x4f32.  Split-BF16x3 path: each element
// is split once, at LDS-store time, into bf16 hi + lo stored k-contiguous, and
// three v_mfma_f32_16x16x32_bf16 (lo*hi + hi*lo + hi*hi) replace eight FP32
// MFMAs.  Lane layouts: 16x16x4 f32 -- lane l supplies A[l%16][l/16],
// B[l/16][l%16]; 16x16x32 bf16 -- lane l supplies A[l%16][8(l/16)..+7],
// B[8(l/16)..+7][l%16]; both accumulate D[4(l/16)+r][l%16].
#include <algorithm>
#include <cstdio>
#include <type_traits>
#include <utility>
#include <vector>

#include "fpls_internal.h"

namespace fpls {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int G2_BM = 128, G2_BN = 128, G2_BK = 32, G2_NT = 256;
constexpr int G2_PADF = 4;   // FP32 LDS row padding (floats)
constexpr int G2_PADH = 8;   // bf16 LDS row padding (elements)

__device__ inline void split_bf(float x, __bf16& hi, __bf16& lo) {
  hi = static_cast<__bf16>(x);
  lo = static_cast<__bf16>(x - static_cast<float>(hi));
}

// 4 consecutive floats p[0..3] with bounds (count valid), vectorised when aligned
__device__ inline float4 load4(const float* __restrict__ p, int valid, bool aligned) {
  if (valid >= 4 && aligned) return *reinterpret_cast<const float4*>(p);
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  if (valid > 0) v.x = p[0];
  if (valid > 1) v.y = p[1];
  if (valid > 2) v.z = p[2];
  if (valid > 3) v.w = p[3];
  return v;
}

// LDS layout shared by both kernels, as raw floats.
//  FP32:  As[2][BK][BM + PADF] (k-major, m contiguous), Bs[2][BK][BN + PADF]
//  BF16:  Ah/Al[2][BM][BK + PADH] (m-major, k contiguous), Bh/Bl[2][BN][BK + PADH]
constexpr int G2_F32_TILE = G2_BK * (G2_BM + G2_PADF);              // floats per buffer
constexpr int G2_BF_TILE = G2_BM * (G2_BK + G2_PADH);               // bf16 per buffer
constexpr int G2_SMEM_F32 = 2 * 2 * G2_F32_TILE;                    // floats
constexpr int G2_SMEM_BF = (4 * 2 * G2_BF_TILE) / 2;                // floats
constexpr int G2_SMEM_C = G2_BN * (G2_BM + 1);                      // epilogue tile
template <bool BF>
struct G2Smem {
  static constexpr int tiles = BF ? G2_SMEM_BF : G2_SMEM_F32;
  static constexpr int value = tiles > G2_SMEM_C ? tiles : G2_SMEM_C;
};

// The register-staged tile of one thread: A part (4 float4) and B part (4 float4)
struct G2Regs {
  float4 a[4];
  float4 b[4];
  double sa[4];   // NN: centring shift of the four A columns
  int4 lab[4];    // TN with labels: codes of the four rows
};

// Store the staged registers into LDS buffer `buf`.
//  amaj: for NN, A regs hold 4 rows (m..m+3) of one column k; for TN, 4 k of one m.
template <bool BF, bool NN>
__device__ inline void g2_store(float* smem, int buf, const G2Regs& r, int t) {
  if (!BF) {
    float* As = smem + buf * G2_F32_TILE;
    float* Bs = smem + 2 * G2_F32_TILE + buf * G2_F32_TILE;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (NN) {  // A: rows m4..m4+3 of column k
        const int m4 = (t % 32) * 4, k = t / 32 + 8 * j;
        float* d = As + k * (G2_BM + G2_PADF) + m4;
        *reinterpret_cast<float4*>(d) = r.a[j];
      } else {   // A: k4..k4+3 of column m
        const int k4 = (t % 8) * 4, m = t / 8 + 32 * j;
        As[(k4 + 0) * (G2_BM + G2_PADF) + m] = r.a[j].x;
        As[(k4 + 1) * (G2_BM + G2_PADF) + m] = r.a[j].y;
        As[(k4 + 2) * (G2_BM + G2_PADF) + m] = r.a[j].z;
        As[(k4 + 3) * (G2_BM + G2_PADF) + m] = r.a[j].w;
      }
      const int k4 = (t % 8) * 4, n = t / 8 + 32 * j;  // B: k4..k4+3 of column n
      Bs[(k4 + 0) * (G2_BN + G2_PADF) + n] = r.b[j].x;
      Bs[(k4 + 1) * (G2_BN + G2_PADF) + n] = r.b[j].y;
      Bs[(k4 + 2) * (G2_BN + G2_PADF) + n] = r.b[j].z;
      Bs[(k4 + 3) * (G2_BN + G2_PADF) + n] = r.b[j].w;
    }
  } else {
    __bf16* base = reinterpret_cast<__bf16*>(smem);
    __bf16* Ah = base + (0 * 2 + buf) * G2_BF_TILE;
    __bf16* Al = base + (1 * 2 + buf) * G2_BF_TILE;
    __bf16* Bh = base + (2 * 2 + buf) * G2_BF_TILE;
    __bf16* Bl = base + (3 * 2 + buf) * G2_BF_TILE;
    constexpr int LD = G2_BK + G2_PADH;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float av[4] = {r.a[j].x, r.a[j].y, r.a[j].z, r.a[j].w};
      if (NN) {
        const int m4 = (t % 32) * 4, k = t / 32 + 8 * j;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          __bf16 h, l;
          split_bf(av[e], h, l);
          Ah[(m4 + e) * LD + k] = h;
          Al[(m4 + e) * LD + k] = l;
        }
      } else {
        const int k4 = (t % 8) * 4, m = t / 8 + 32 * j;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          __bf16 h, l;
          split_bf(av[e], h, l);
          Ah[m * LD + k4 + e] = h;
          Al[m * LD + k4 + e] = l;
        }
      }
      const float bv[4] = {r.b[j].x, r.b[j].y, r.b[j].z, r.b[j].w};
      const int k4 = (t % 8) * 4, n = t / 8 + 32 * j;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        __bf16 h, l;
        split_bf(bv[e], h, l);
        Bh[n * LD + k4 + e] = h;
        Bl[n * LD + k4 + e] = l;
      }
    }
  }
}

// MFMAs of one K tile from LDS buffer `buf` into the 4 x 4 accumulators.
template <bool BF>
__device__ inline void g2_compute(const float* smem, int buf, v4f (&acc)[4][4], int lane, int wm,
                                  int wn) {
  if (!BF) {
    const float* As = smem + buf * G2_F32_TILE;
    const float* Bs = smem + 2 * G2_F32_TILE + buf * G2_F32_TILE;
#pragma unroll
    for (int kk = 0; kk < G2_BK; kk += 4) {
      float af[4], bf[4];
      const int kr = kk + lane / 16;
#pragma unroll
      for (int a = 0; a < 4; ++a) af[a] = As[kr * (G2_BM + G2_PADF) + wm * 64 + a * 16 + lane % 16];
#pragma unroll
      for (int b = 0; b < 4; ++b) bf[b] = Bs[kr * (G2_BN + G2_PADF) + wn * 64 + b * 16 + lane % 16];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
  } else {
    const __bf16* base = reinterpret_cast<const __bf16*>(smem);
    const __bf16* Ah = base + (0 * 2 + buf) * G2_BF_TILE;
    const __bf16* Al = base + (1 * 2 + buf) * G2_BF_TILE;
    const __bf16* Bh = base + (2 * 2 + buf) * G2_BF_TILE;
    const __bf16* Bl = base + (3 * 2 + buf) * G2_BF_TILE;
    constexpr int LD = G2_BK + G2_PADH;
    const int ko = 8 * (lane / 16);
    v8bf ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int m = wm * 64 + a * 16 + lane % 16;
      ah[a] = *reinterpret_cast<const v8bf*>(Ah + m * LD + ko);
      al[a] = *reinterpret_cast<const v8bf*>(Al + m * LD + ko);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int n = wn * 64 + b * 16 + lane % 16;
      bh[b] = *reinterpret_cast<const v8bf*>(Bh + n * LD + ko);
      bl[b] = *reinterpret_cast<const v8bf*>(Bl + n * LD + ko);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
      }
  }
}

// Accumulators -> LDS tile Cs[n][m] (ld BM + 1).
__device__ inline void g2_stage_c(float* smem, const v4f (&acc)[4][4], int lane, int wm, int wn) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = wm * 64 + a * 16 + 4 * (lane / 16) + r;
        const int j = wn * 64 + b * 16 + lane % 16;
        smem[j * (G2_BM + 1) + i] = acc[a][b][r];
      }
}

// ---------------------------------------------------------------- NN
// B: zero-padded to Kp = ceil(K/32)*32 rows and Np = ceil(N/128)*128 columns,
// ldb = Kp (multiple of 4), 16-byte aligned.  CEN: A is centred on load by shA
// (the latent projections of raw X); otherwise A is used as stored.
template <bool BF, bool CEN>
__global__ __launch_bounds__(G2_NT) void k_g2_nn(int64_t M, int64_t N, int64_t K,
                                                 const float* __restrict__ A, int64_t lda,
                                                 const double* __restrict__ shA,
                                                 const float* __restrict__ B, int64_t ldb,
                                                 const double* __restrict__ off,
                                                 float* __restrict__ Cf, double* __restrict__ Cd,
                                                 int64_t ldc, int crow) {
  __shared__ __align__(16) float smem[G2Smem<BF>::value];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int64_t row0 = (int64_t)blockIdx.x * G2_BM, col0 = (int64_t)blockIdx.y * G2_BN;
  v4f acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  // Addresses are clamped into the matrix (rows >= M and columns >= K read valid
  // memory); B is zero-padded to whole tiles, so clamped columns contribute 0 and
  // clamped rows are never written.  No per-element branches in the loop.
  const bool vecA = ((lda & 3) == 0) && ((M & 3) == 0) &&
                    ((reinterpret_cast<uintptr_t>(A) & 15) == 0);
  auto fetch = [&](int64_t k0, G2Regs& r) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m4 = (t % 32) * 4, k = t / 32 + 8 * j;
      const int64_t gk = std::min<int64_t>(k0 + k, K - 1);
      const float* col = A + gk * lda;
      float4 a;
      if (vecA) {
        a = *reinterpret_cast<const float4*>(col + std::min<int64_t>(row0 + m4, M - 4));
      } else {
        a.x = col[std::min<int64_t>(row0 + m4 + 0, M - 1)];
        a.y = col[std::min<int64_t>(row0 + m4 + 1, M - 1)];
        a.z = col[std::min<int64_t>(row0 + m4 + 2, M - 1)];
        a.w = col[std::min<int64_t>(row0 + m4 + 3, M - 1)];
      }
      if (CEN) r.sa[j] = shA[gk];
      r.a[j] = a;
      const int kq = (t % 8) * 4, n = t / 8 + 32 * j;
      r.b[j] = *reinterpret_cast<const float4*>(B + (k0 + kq) + (col0 + n) * ldb);
    }
  };
  // arithmetic on the fetched registers runs after the MFMAs of the current tile
  auto finish = [&](G2Regs& r) {
    if (!CEN) return;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float sv = static_cast<float>(r.sa[j]);
      r.a[j].x -= sv;
      r.a[j].y -= sv;
      r.a[j].z -= sv;
      r.a[j].w -= sv;
    }
  };
  const int nk = (int)((K + G2_BK - 1) / G2_BK);
  G2Regs r;
  fetch(0, r);
  finish(r);
  g2_store<BF, true>(smem, 0, r, t);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) fetch((int64_t)(kt + 1) * G2_BK, r);
    g2_compute<BF>(smem, cur, acc, lane, wm, wn);
    if (more) {
      finish(r);
      g2_store<BF, true>(smem, cur ^ 1, r, t);
    }
    __syncthreads();
  }
  g2_stage_c(smem, acc, lane, wm, wn);
  __syncthreads();
  for (int idx = t; idx < G2_BM * G2_BN; idx += G2_NT) {
    const int i = crow ? idx / G2_BN : idx % G2_BM;
    const int j = crow ? idx % G2_BN : idx / G2_BM;
    const int64_t gi = row0 + i, gj = col0 + j;
    if (gi < M && gj < N) {
      double v = smem[j * (G2_BM + 1) + i];
      if (off) v -= off[gj];
      const int64_t o = crow ? gi * ldc + gj : gi + gj * ldc;
      if (Cd) Cd[o] = v;
      else Cf[o] = static_cast<float>(v);
    }
  }
}

// ---------------------------------------------------------------- TN (split-K partials)
template <bool BF>
__global__ __launch_bounds__(G2_NT) void k_g2_tn(int64_t M, int64_t N, int64_t K, int64_t kchunk,
                                                 const float* __restrict__ A, int64_t lda,
                                                 const double* __restrict__ shA,
                                                 const float* __restrict__ B, int64_t ldb,
                                                 const int32_t* __restrict__ labB,
                                                 const double* __restrict__ shB,
                                                 float* __restrict__ partial, int sym) {
  // sym: C is symmetric (A == B); tiles below the diagonal are left to the reduction
  if (sym && blockIdx.y < blockIdx.x) return;
  __shared__ __align__(16) float smem[G2Smem<BF>::value];
  __shared__ float sha[G2_BM], shb[G2_BN];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int64_t m0 = (int64_t)blockIdx.x * G2_BM, n0 = (int64_t)blockIdx.y * G2_BN;
  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  const int64_t kend = std::min<int64_t>(K, kbeg + kchunk);
  const bool a_al = ((lda & 3) == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0) &&
                    ((kbeg & 3) == 0);
  const bool b_al = B && ((ldb & 3) == 0) && ((reinterpret_cast<uintptr_t>(B) & 15) == 0) &&
                    ((kbeg & 3) == 0);
  if (t < G2_BM) sha[t] = (shA && m0 + t < M) ? static_cast<float>(shA[m0 + t]) : 0.f;
  else shb[t - G2_BM] = (shB && n0 + t - G2_BM < N) ? static_cast<float>(shB[n0 + t - G2_BM]) : 0.f;
  __syncthreads();
  v4f acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  // Clamped addresses, then masks (selects, not branches) zero everything outside
  // the slab / the matrix.
  auto ld4 = [&](const float* base, int64_t gk, bool vec) -> float4 {
    if (vec && gk + 4 <= K) return *reinterpret_cast<const float4*>(base + gk);
    float4 v;
    v.x = base[std::min<int64_t>(gk + 0, K - 1)];
    v.y = base[std::min<int64_t>(gk + 1, K - 1)];
    v.z = base[std::min<int64_t>(gk + 2, K - 1)];
    v.w = base[std::min<int64_t>(gk + 3, K - 1)];
    return v;
  };
  auto fetch = [&](int64_t k0, G2Regs& r) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kq = (t % 8) * 4, c = t / 8 + 32 * j;
      const int64_t gk = k0 + kq;
      r.a[j] = ld4(A + std::min<int64_t>(m0 + c, M - 1) * lda, gk, a_al);
      if (labB) {
        r.lab[j].x = labB[std::min<int64_t>(gk + 0, K - 1)];
        r.lab[j].y = labB[std::min<int64_t>(gk + 1, K - 1)];
        r.lab[j].z = labB[std::min<int64_t>(gk + 2, K - 1)];
        r.lab[j].w = labB[std::min<int64_t>(gk + 3, K - 1)];
      } else {
        r.b[j] = ld4(B + std::min<int64_t>(n0 + c, N - 1) * ldb, gk, b_al);
      }
    }
  };
  // centring, one-hot generation and slab masks, applied after the MFMAs
  auto finish = [&](int64_t k0, G2Regs& r) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kq = (t % 8) * 4, c = t / 8 + 32 * j;
      const int64_t gk = k0 + kq;
      const bool in0 = gk + 0 < kend, in1 = gk + 1 < kend, in2 = gk + 2 < kend, in3 = gk + 3 < kend;
      const bool mok = m0 + c < M, nok = n0 + c < N;
      const float s = sha[c];
      float4 a = r.a[j];
      a.x = (in0 && mok) ? a.x - s : 0.f;
      a.y = (in1 && mok) ? a.y - s : 0.f;
      a.z = (in2 && mok) ? a.z - s : 0.f;
      a.w = (in3 && mok) ? a.w - s : 0.f;
      r.a[j] = a;
      float4 b;
      if (labB) {
        const int lab_n = (int)(n0 + c) + 1;
        b.x = (r.lab[j].x == lab_n ? 1.f : 0.f);
        b.y = (r.lab[j].y == lab_n ? 1.f : 0.f);
        b.z = (r.lab[j].z == lab_n ? 1.f : 0.f);
        b.w = (r.lab[j].w == lab_n ? 1.f : 0.f);
      } else {
        b = r.b[j];
      }
      const float sb = shb[c];
      b.x = (in0 && nok) ? b.x - sb : 0.f;
      b.y = (in1 && nok) ? b.y - sb : 0.f;
      b.z = (in2 && nok) ? b.z - sb : 0.f;
      b.w = (in3 && nok) ? b.w - sb : 0.f;
      r.b[j] = b;
    }
  };
  const int nk = (int)((kend - kbeg + G2_BK - 1) / G2_BK);
  G2Regs r;
  if (nk > 0) {
    fetch(kbeg, r);
    finish(kbeg, r);
    g2_store<BF, false>(smem, 0, r, t);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    const int64_t knext = kbeg + (int64_t)(kt + 1) * G2_BK;
    if (more) fetch(knext, r);
    g2_compute<BF>(smem, cur, acc, lane, wm, wn);
    if (more) {
      finish(knext, r);
      g2_store<BF, false>(smem, cur ^ 1, r, t);
    }
    __syncthreads();
  }
  g2_stage_c(smem, acc, lane, wm, wn);
  __syncthreads();
  float* P = partial + (int64_t)blockIdx.z * M * N;
  for (int idx = t; idx < G2_BM * G2_BN; idx += G2_NT) {
    const int i = idx % G2_BM, j = idx / G2_BM;
    const int64_t gm = m0 + i, gn = n0 + j;
    if (gm < M && gn < N) P[gm + gn * M] = smem[j * (G2_BM + 1) + i];
  }
}

static unsigned g1(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

__global__ void k_g2_pad(const float* __restrict__ B, int64_t ldb, int64_t K, int64_t N,
                         float* __restrict__ out, int64_t Kp, int64_t Np) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < Kp * Np;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = t % Kp, n = t / Kp;
    out[t] = (k < K && n < N) ? B[k + n * ldb] : 0.f;
  }
}

// C(m, n) = sum_z P_z(m, n) in FP64, fixed order.  Block: 32 outputs x 8
// split groups; group g sums splits g, g + 8, ... and the 8 group sums are
// added in order, so many loads are in flight per output.
// sym (M == N): the partials hold only the 128 x 128 tiles on and above the
// diagonal; outputs below it read their mirror.
__global__ __launch_bounds__(256) void k_g2_reduce(int64_t M, int64_t N, int splits,
                                                   const float* __restrict__ partial,
                                                   double* __restrict__ C, int64_t ldc,
                                                   int sym) {
  __shared__ double red[8][32];
  const int64_t n = M * N;
  const int o = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int64_t t = (int64_t)blockIdx.x * 32 + o;
  double s = 0.0;
  if (t < n) {
    int64_t src = t;
    if (sym) {
      const int64_t i = t % M, j = t / M;
      if (j / G2_BN < i / G2_BM) src = j + i * M;
    }
#pragma unroll 4
    for (int z = g; z < splits; z += 8) s += static_cast<double>(partial[(int64_t)z * n + src]);
  }
  red[g][o] = s;
  __syncthreads();
  if (g == 0 && t < n) {
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) r += red[q][o];
    C[(t % M) + (t / M) * ldc] = r;
  }
}

// The split-K sum with the epilogue of k_proj: C = sum_z P_z - 1 offset^T
// into the FP32 or FP64 output, column- or row-major; one output per thread,
// the slabs in order (coalesced reads of every partial).
__global__ __launch_bounds__(256) void k_g2_reduce_out(int64_t M, int64_t N, int splits,
                                                       const float* __restrict__ partial,
                                                       const double* __restrict__ off,
                                                       float* __restrict__ Cf,
                                                       double* __restrict__ Cd, int64_t ldc,
                                                       int crow) {
  const int64_t n = M * N, t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double r = 0.0;
  int z = 0;
  for (; z + 4 <= splits; z += 4) {
    const float p0 = partial[(int64_t)z * n + t], p1 = partial[(int64_t)(z + 1) * n + t];
    const float p2 = partial[(int64_t)(z + 2) * n + t], p3 = partial[(int64_t)(z + 3) * n + t];
    r += static_cast<double>(p0);
    r += static_cast<double>(p1);
    r += static_cast<double>(p2);
    r += static_cast<double>(p3);
  }
  for (; z < splits; ++z) r += static_cast<double>(partial[(int64_t)z * n + t]);
  const int64_t row = t % M, col = t / M;
  if (off) r -= off[col];
  const int64_t idx = crow ? row * ldc + col : row + col * ldc;
  if (Cd) Cd[idx] = r;
  else Cf[idx] = static_cast<float>(r);
}

// ---------------------------------------------------------------- skinny NN
// T (M x N, N <= 128) = (X - 1 shA^T) B with X streamed straight from HBM into
// MFMA operand registers (no LDS for the big operand).  Workgroup = 4 waves x
// 16 MB rows; each wave holds MB x NB accumulator blocks (NB 16-column blocks
// of B staged: 4 for N <= 64, else 8; MB, the tile height, chosen per shape by
// the autotuner below).  B (K x N, the zero-padded copy) goes
// through LDS one 32-deep K chunk at a time (double buffered, one barrier per
// chunk); the X operand of chunk c+1 is loaded into registers while the MFMAs
// of chunk c run.  With split-K (part != NULL) each workgroup writes its
// slab's FP32 partial.  The kernel is latency bound: loading two consecutive
// rows per lane as one vector costs registers and, at the lower occupancy,
// time (measured), so rows go one per lane.
constexpr int P_BK = 32, P_NT = 256;
constexpr int p_bm(int mb) { return 64 * mb; }  // rows of a workgroup tile
constexpr int P_LDB = P_BK + 8;  // bf16 row length of the [n][k] B tiles
template <bool BF, int NB>
struct PSmem {
  static constexpr int value = BF ? (2 * 2 * 16 * NB * P_LDB) / 2 : 2 * P_BK * (16 * NB + 4);
};

// NB: 16-column blocks of B staged and multiplied (N <= 16 NB); MB: 16-row
// blocks per wave
template <bool BF, bool CEN, int NB, int MB>
__global__ __launch_bounds__(P_NT) void k_proj(int64_t M, int N, int64_t K,
                                               const float* A, int64_t lda,
                                               const float* shA,
                                               const float* B, int64_t ldb,
                                               const double* __restrict__ off,
                                               float* __restrict__ Cf, double* __restrict__ Cd,
                                               int64_t ldc, int crow, int64_t kchunk,
                                               float* __restrict__ part) {
  __shared__ __align__(16) float smem[PSmem<BF, NB>::value];
  constexpr int BW = 16 * NB;              // staged B columns
  constexpr int TPC = P_NT / BW;           // staging threads per column
  constexpr int KPT = P_BK / TPC;          // k values per staging thread (8 or 16)
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t row0 = (int64_t)blockIdx.x * p_bm(MB) + w * 16 * MB;
  // K slab of this workgroup (kchunk is a multiple of P_BK; B's zero rows pad
  // the last slab up to a whole chunk)
  const int64_t kbeg = (int64_t)blockIdx.y * kchunk;
  const int64_t kend = std::min<int64_t>(K, kbeg + kchunk);
  const int nb = (N + 15) >> 4;
  v4f acc[MB][NB];
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  const float* arow[MB];
#pragma unroll
  for (int a = 0; a < MB; ++a) arow[a] = A + std::min<int64_t>(row0 + 16 * a + lane % 16, M - 1);
  // k offset (within a chunk) of operand element j of this lane
  auto kof = [&](int j) { return BF ? 8 * (lane / 16) + j : 4 * j + lane / 16; };
  // the centring shifts arrive as FP32 (converted once per call): any
  // arithmetic on a value at load time would wait on every outstanding load
  // three register sets (chunks c, c + 1, c + 2) used in a fixed rotation by
  // the loop unrolled three times: loads run two chunks ahead of the MFMAs
  // (a workgroup is bound by the bytes it has in flight, not by arithmetic),
  // and no register that is still the target of a load is ever copied
  struct ASet {
    float a[MB][8], m[8];
    float4 q[KPT / 4];
  };
  ASet R0, R1, R2;
  auto fetchA = [&](int64_t k0, float (&x)[MB][8], float (&mx)[8]) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int64_t gk = std::min<int64_t>(k0 + kof(j), K - 1);
#pragma unroll
      for (int a = 0; a < MB; ++a) x[a][j] = arow[a][gk * lda];
      mx[j] = CEN ? shA[gk] : 0.f;
    }
  };
  // B chunk staging: thread -> column n = t / TPC, KPT consecutive k
  const int bn = t / TPC, bk = (t % TPC) * KPT;
  auto fetchB = [&](int64_t k0, float4 (&dst)[KPT / 4]) {
    const float* src = B + (k0 + bk) + (int64_t)bn * ldb;
#pragma unroll
    for (int q = 0; q < KPT / 4; ++q) dst[q] = *reinterpret_cast<const float4*>(src + 4 * q);
  };
  auto storeB = [&](int buf, const float4 (&src)[KPT / 4]) {
    if (BF) {
      __bf16* base = reinterpret_cast<__bf16*>(smem);
      __bf16* Bh = base + (0 * 2 + buf) * BW * P_LDB;
      __bf16* Bl = base + (1 * 2 + buf) * BW * P_LDB;
#pragma unroll
      for (int h = 0; h < KPT / 8; ++h) {
        const float v[8] = {src[2 * h].x, src[2 * h].y, src[2 * h].z, src[2 * h].w,
                            src[2 * h + 1].x, src[2 * h + 1].y, src[2 * h + 1].z,
                            src[2 * h + 1].w};
        v8bf hv, lv;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          __bf16 hh, ll;
          split_bf(v[e], hh, ll);
          hv[e] = hh;
          lv[e] = ll;
        }
        *reinterpret_cast<v8bf*>(Bh + bn * P_LDB + bk + 8 * h) = hv;
        *reinterpret_cast<v8bf*>(Bl + bn * P_LDB + bk + 8 * h) = lv;
      }
    } else {
      float* Bs = smem + buf * P_BK * (BW + 4);
#pragma unroll
      for (int q = 0; q < KPT / 4; ++q) {
        Bs[(bk + 4 * q + 0) * (BW + 4) + bn] = src[q].x;
        Bs[(bk + 4 * q + 1) * (BW + 4) + bn] = src[q].y;
        Bs[(bk + 4 * q + 2) * (BW + 4) + bn] = src[q].z;
        Bs[(bk + 4 * q + 3) * (BW + 4) + bn] = src[q].w;
      }
    }
  };
  auto compute = [&](int buf, const float (&ac)[MB][8], const float (&mc)[8]) {
    if (BF) {
      const __bf16* base = reinterpret_cast<const __bf16*>(smem);
      const __bf16* Bh = base + (0 * 2 + buf) * BW * P_LDB;
      const __bf16* Bl = base + (1 * 2 + buf) * BW * P_LDB;
      v8bf ah[MB], al[MB];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float me = mc[e];
#pragma unroll
        for (int a = 0; a < MB; ++a) {
          __bf16 hh, ll;
          split_bf(ac[a][e] - me, hh, ll);
          ah[a][e] = hh;
          al[a][e] = ll;
        }
      }
      const int ko = 8 * (lane / 16);
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (b < nb) {
          const int n = b * 16 + lane % 16;
          const v8bf bh = *reinterpret_cast<const v8bf*>(Bh + n * P_LDB + ko);
          const v8bf bl = *reinterpret_cast<const v8bf*>(Bl + n * P_LDB + ko);
#pragma unroll
          for (int a = 0; a < MB; ++a) {
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh, acc[a][b], 0, 0, 0);
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl, acc[a][b], 0, 0, 0);
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh, acc[a][b], 0, 0, 0);
          }
        }
      }
    } else {
      const float* Bs = smem + buf * P_BK * (BW + 4);
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const float ms = mc[s];
        float x[MB];
#pragma unroll
        for (int a = 0; a < MB; ++a) x[a] = ac[a][s] - ms;
        const int kr = 4 * s + lane / 16;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (b < nb) {
            const float bv = Bs[kr * (BW + 4) + b * 16 + lane % 16];
#pragma unroll
            for (int a = 0; a < MB; ++a)
              acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[a], bv, acc[a][b], 0, 0, 0);
          }
        }
      }
    }
  };
  const int nk = (int)((kend - kbeg + P_BK - 1) / P_BK);
  // chunk c uses register set c % 3 and LDS buffer c & 1
  auto step = [&](int c, ASet& cur, ASet& nxt, ASet& far) {
    if (c + 2 < nk) {
      fetchA(kbeg + (int64_t)(c + 2) * P_BK, far.a, far.m);
      fetchB(kbeg + (int64_t)(c + 2) * P_BK, far.q);
    }
    __builtin_amdgcn_sched_barrier(0);
    compute(c & 1, cur.a, cur.m);
    if (c + 1 < nk) storeB((c + 1) & 1, nxt.q);
    __syncthreads();
  };
  fetchA(kbeg, R0.a, R0.m);
  fetchB(kbeg, R0.q);
  if (nk > 1) {
    fetchA(kbeg + P_BK, R1.a, R1.m);
    fetchB(kbeg + P_BK, R1.q);
  }
  storeB(0, R0.q);
  __syncthreads();
  for (int c = 0; c < nk; c += 3) {
    step(c, R0, R1, R2);
    if (c + 1 < nk) step(c + 1, R1, R2, R0);
    if (c + 2 < nk) step(c + 2, R2, R0, R1);
  }
  // epilogue: lane holds D[4(l/16) + r][l%16] of each 16 x 16 block
  if (part) {  // split-K: raw FP32 slab partial, M x N column-major
    float* P = part + (int64_t)blockIdx.y * M * N;
#pragma unroll
    for (int a = 0; a < MB; ++a) {
      const int64_t rbase = row0 + 16 * a + 4 * (lane / 16);
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int col = b * 16 + lane % 16;
        if (b >= nb || col >= N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (rbase + r < M) P[rbase + r + (int64_t)col * M] = acc[a][b][r];
      }
    }
    return;
  }
#pragma unroll
  for (int a = 0; a < MB; ++a) {
    const int64_t rbase = row0 + 16 * a + 4 * (lane / 16);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int col = b * 16 + lane % 16;
      if (b >= nb || col >= N) continue;
      const double o = off ? off[col] : 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = rbase + r;
        if (row >= M) continue;
        const double v = static_cast<double>(acc[a][b][r]) - o;
        const int64_t idx = crow ? row * ldc + col : row + (int64_t)col * ldc;
        if (Cd) Cd[idx] = v;
        else Cf[idx] = static_cast<float>(v);
      }
    }
  }
}

// k_proj with the column-block count picked from N and the tile height mb.
template <bool BF, bool CEN, typename... Args>
static void launch_proj(Ctx* ctx, const char* name, int N, int mb, dim3 grid, Args... args) {
  if (N <= 64) {
    if (mb == 1) launch(ctx, name, k_proj<BF, CEN, 4, 1>, grid, dim3(P_NT), 0, args...);
    else if (mb == 4) launch(ctx, name, k_proj<BF, CEN, 4, 4>, grid, dim3(P_NT), 0, args...);
    else launch(ctx, name, k_proj<BF, CEN, 4, 2>, grid, dim3(P_NT), 0, args...);
  } else {
    if (mb == 1) launch(ctx, name, k_proj<BF, CEN, 8, 1>, grid, dim3(P_NT), 0, args...);
    else launch(ctx, name, k_proj<BF, CEN, 8, 2>, grid, dim3(P_NT), 0, args...);
  }
}

// B (K x N, FP64, ld ldb; or its transpose, B(k, n) = src[n + k * ldb]) as the
// zero-padded FP32 operand of gemm_nn_f32_padded, in one pass.
__global__ void k_g2_pad_f64(const double* __restrict__ B, int64_t ldb, int trans, int64_t K,
                             int64_t N, float* __restrict__ out, int64_t Kp, int64_t Np) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < Kp * Np;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = t % Kp, n = t / Kp;
    out[t] = (k < K && n < N) ? (float)(trans ? B[n + k * ldb] : B[k + n * ldb]) : 0.f;
  }
}

__global__ void k_g2_shift_f32(const double* __restrict__ s, int64_t K, float* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < K;
       t += (int64_t)gridDim.x * blockDim.x)
    out[t] = (float)s[t];
}

int64_t nn_pad_k(int64_t K) { return (K + G2_BK - 1) / G2_BK * G2_BK; }
int64_t nn_pad_n(int64_t N) { return (N + G2_BN - 1) / G2_BN * G2_BN; }

DBuf<float> nn_pad_operand(Ctx* ctx, const double* B, int64_t ldb, bool trans, int64_t K,
                           int64_t N) {
  const int64_t Kp = nn_pad_k(K), Np = nn_pad_n(N);
  DBuf<float> out(ctx, (size_t)Kp * Np);
  launch(ctx, "pad_b", k_g2_pad_f64, dim3(g1(Kp * Np)), dim3(256), 0, B, ldb, trans ? 1 : 0, K,
         N, out.get(), Kp, Np);
  return out;
}

DBuf<float> nn_shift_operand(Ctx* ctx, const double* shift, int64_t K) {
  DBuf<float> out(ctx, (size_t)K);
  launch(ctx, "shift_f32", k_g2_shift_f32, dim3(g1(K)), dim3(256), 0, shift, K, out.get());
  return out;
}

// ---------------------------------------------------------------- projection autotuner
// One skinny projection (N <= 128) and its launch configurations: tile height
// mb (64 mb rows per workgroup) and split-K count (FP32 slab partials summed
// in FP64 by the epilogue kernel).
struct ProjShape {
  bool bf, cen;
  int64_t M;
  int N;
  int64_t K;
  const float* A;
  int64_t lda;
  const float* sh32;
  const float* B;
  int64_t ldb;
  const double* offset;
  float* Cf;
  double* Cd;
  int64_t ldc;
  int crow;
};

// The built-in split rule: fewer row tiles than CUs splits K into slabs so
// that about two workgroups run per CU, each slab at least 128 deep.
static int proj_rule_splits(const Ctx* ctx, const ProjShape& s, int mb) {
  const int64_t tiles = (s.M + p_bm(mb) - 1) / p_bm(mb);
  if (!(tiles < (int64_t)ctx->num_cus && s.K >= 256)) return 1;
  int64_t splits = std::max<int64_t>(1, (2 * (int64_t)ctx->num_cus + tiles - 1) / tiles);
  splits = std::min<int64_t>(splits, s.K / 128);
  splits = std::min<int64_t>(
      splits, std::max<int64_t>(1, (int64_t)ctx->partial_budget() / (4 * s.M * s.N)));
  return (int)std::max<int64_t>(1, splits);
}

static void proj_run(Ctx* ctx, const ProjShape& s, int mb, int splits_req) {
  const int64_t tiles = (s.M + p_bm(mb) - 1) / p_bm(mb);
  const char* name = s.cen ? (s.bf ? "project_bf16x3" : "project_f32")
                           : (s.bf ? "gemm_nn_bf16x3" : "gemm_nn_f32");
  if (splits_req > 1) {
    int64_t kchunk = (s.K + splits_req - 1) / splits_req;
    kchunk = (kchunk + P_BK - 1) / P_BK * P_BK;
    const int64_t splits = (s.K + kchunk - 1) / kchunk;
    DBuf<float> part(ctx, (size_t)splits * s.M * s.N);
    const dim3 g((unsigned)tiles, (unsigned)splits);
    const double* noff = nullptr;
    float* ncf = nullptr;
    double* ncd = nullptr;
    if (s.cen) {
      if (s.bf)
        launch_proj<true, true>(ctx, name, s.N, mb, g, s.M, s.N, s.K, s.A, s.lda, s.sh32, s.B,
                                s.ldb, noff, ncf, ncd, (int64_t)0, 0, kchunk, part.get());
      else
        launch_proj<false, true>(ctx, name, s.N, mb, g, s.M, s.N, s.K, s.A, s.lda, s.sh32, s.B,
                                 s.ldb, noff, ncf, ncd, (int64_t)0, 0, kchunk, part.get());
    } else {
      if (s.bf)
        launch_proj<true, false>(ctx, name, s.N, mb, g, s.M, s.N, s.K, s.A, s.lda, s.sh32, s.B,
                                 s.ldb, noff, ncf, ncd, (int64_t)0, 0, kchunk, part.get());
      else
        launch_proj<false, false>(ctx, name, s.N, mb, g, s.M, s.N, s.K, s.A, s.lda, s.sh32, s.B,
                                  s.ldb, noff, ncf, ncd, (int64_t)0, 0, kchunk, part.get());
    }
    launch(ctx, "gemm_nn_reduce", k_g2_reduce_out, dim3((unsigned)((s.M * s.N + 255) / 256)),
           dim3(256), 0, s.M, (int64_t)s.N, (int)splits, (const float*)part.get(), s.offset, s.Cf,
           s.Cd, s.ldc, s.crow);
    return;
  }
  const dim3 g((unsigned)tiles);
  if (s.cen) {
    if (s.bf)
      launch_proj<true, true>(ctx, name, s.N, mb, g, s.M, s.N, s.K, s.A, s.lda, s.sh32, s.B,
                              s.ldb, s.offset, s.Cf, s.Cd, s.ldc, s.crow, s.K, (float*)nullptr);
    else
      launch_proj<false, true>(ctx, name, s.N, mb, g, s.M, s.N, s.K, s.A, s.lda, s.sh32, s.B,
                               s.ldb, s.offset, s.Cf, s.Cd, s.ldc, s.crow, s.K, (float*)nullptr);
  } else {
    if (s.bf)
      launch_proj<true, false>(ctx, name, s.N, mb, g, s.M, s.N, s.K, s.A, s.lda,
                               (const float*)nullptr, s.B, s.ldb, s.offset, s.Cf, s.Cd, s.ldc,
                               s.crow, s.K, (float*)nullptr);
    else
      launch_proj<false, false>(ctx, name, s.N, mb, g, s.M, s.N, s.K, s.A, s.lda,
                                (const float*)nullptr, s.B, s.ldb, s.offset, s.Cf, s.Cd, s.ldc,
                                s.crow, s.K, (float*)nullptr);
  }
}

// The configuration of a shape: pinned, cached, timed now, or the rule.
// Timing first runs the rule's configuration three times (clocks and caches
// settle), then times every candidate three times, each run between its own
// HIP events on the context stream, and keeps its fastest run; a candidate
// replaces the rule only when it is at least 3 % faster.  Every candidate
// writes the same output, and the caller's launch with the winner writes it
// last.
static Ctx::TuneChoice proj_choose(Ctx* ctx, const ProjShape& s) {
  Ctx::TuneChoice ch;
  if (ctx->force_mb > 0 || ctx->force_splits > 0) {
    ch.mb = ctx->force_mb > 0 ? ctx->force_mb : 2;
    if (s.N > 64 && ch.mb == 4) ch.mb = 2;  // NB = 8 tiles exist for mb 1 and 2
    ch.splits = ctx->force_splits > 0 ? ctx->force_splits : proj_rule_splits(ctx, s, ch.mb);
    return ch;
  }
  char key[160];
  std::snprintf(key, sizeof(key), "proj:%s%s:M%lld:N%d:K%lld", s.bf ? "bf16x3" : "f32",
                s.cen ? ":centred" : "", (long long)s.M, s.N, (long long)s.K);
  auto it = ctx->tune_table.find(key);
  if (it != ctx->tune_table.end()) return it->second;
  // the rule: 64-row tiles for the split-BF16x3 form (its per-chunk split
  // work favours more workgroups), 128-row tiles for FP32 (measured on the
  // 50,000 x 768 projection)
  ch.mb = s.bf ? 1 : 2;
  ch.splits = proj_rule_splits(ctx, s, ch.mb);
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  const bool capturing = hipStreamIsCapturing(ctx->stream, &cap) == hipSuccess &&
                         cap != hipStreamCaptureStatusNone;
  if (ctx->tune_mode == 0 || capturing || ctx->arena_base || s.M * (int64_t)s.N < 4096) {
    if (!capturing && !ctx->arena_base) ctx->tune_table[key] = ch;
    return ch;
  }
  std::vector<std::pair<int, int>> cand;
  for (int mb : {1, 2, 4}) {
    if (mb == 4 && s.N > 64) continue;
    const int r = proj_rule_splits(ctx, s, mb);
    for (int sp : {1, r, 2 * r}) {
      if (sp > 1 && (sp > s.K / 128 || (int64_t)sp * 4 * s.M * s.N > (int64_t)ctx->partial_budget()))
        continue;
      if (std::find(cand.begin(), cand.end(), std::make_pair(mb, sp)) == cand.end())
        cand.emplace_back(mb, sp);
    }
  }
  hipEvent_t e0, e1;
  FPLS_HIP(hipEventCreate(&e0));
  FPLS_HIP(hipEventCreate(&e1));
  auto time_one = [&](int mb, int sp) {
    double us = 1e30;
    for (int r = 0; r < 3; ++r) {
      FPLS_HIP(hipEventRecord(e0, ctx->stream));
      proj_run(ctx, s, mb, sp);
      FPLS_HIP(hipEventRecord(e1, ctx->stream));
      FPLS_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      FPLS_HIP(hipEventElapsedTime(&ms, e0, e1));
      us = std::min(us, 1e3 * (double)ms);
    }
    return us;
  };
  for (int r = 0; r < 3; ++r) proj_run(ctx, s, ch.mb, ch.splits);
  const double rule_us = time_one(ch.mb, ch.splits);
  double best = rule_us;
  int bmb = ch.mb, bsp = ch.splits;
  for (const auto& c : cand) {
    const double us = (c.first == ch.mb && c.second == ch.splits) ? rule_us
                                                                    : time_one(c.first, c.second);
    if (ctx->tune_log)
      std::fprintf(stderr, "fpls autotune %s: rows %d splits %d  %.2f us\n", key, 64 * c.first,
                   c.second, us);
    if (us < best) {
      best = us;
      bmb = c.first;
      bsp = c.second;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best < 0.97 * rule_us) {
    ch.mb = bmb;
    ch.splits = bsp;
  } else {
    best = rule_us;
  }
  ch.us = best;
  ch.tuned = true;
  ctx->tune_table[key] = ch;
  return ch;
}

void gemm_nn_f32(Ctx* ctx, int precision, int64_t M, int64_t N, int64_t K, const float* A,
                 int64_t lda, const double* shiftA, const float* B, int64_t ldb,
                 const double* offset, float* Cf, double* Cd, int64_t ldc, bool c_rowmajor) {
  if (M <= 0 || N <= 0) return;
  FPLS_REQUIRE(K >= 1, FPLS_ERR_ARG, "gemm_nn_f32: inner dimension must be positive");
  // zero-padded copy of the (small) B operand: Kp x Np, ld Kp
  const int64_t Kp = nn_pad_k(K), Np = nn_pad_n(N);
  DBuf<float> Bp(ctx, (size_t)Kp * Np);
  launch(ctx, "pad_b", k_g2_pad, dim3(g1(Kp * Np)), dim3(256), 0, B, ldb, K, N, Bp.get(), Kp, Np);
  DBuf<float> shf;
  if (shiftA) shf = nn_shift_operand(ctx, shiftA, K);
  gemm_nn_f32_padded(ctx, precision, M, N, K, A, lda, shiftA, shf.get(), Bp.get(), offset, Cf,
                     Cd, ldc, c_rowmajor);
}

constexpr int64_t kProjSliceRows = 32768, kProjSliceMaxN = 1024;

void gemm_nn_f32_padded(Ctx* ctx, int precision, int64_t M, int64_t N, int64_t K,
                        const float* A, int64_t lda, const double* shiftA,
                        const float* shift32, const float* B, const double* offset, float* Cf,
                        double* Cd, int64_t ldc, bool c_rowmajor) {
  if (M <= 0 || N <= 0) return;
  FPLS_REQUIRE((Cf != nullptr) != (Cd != nullptr), FPLS_ERR_ARG, "gemm_nn_f32: one output");
  FPLS_REQUIRE(K >= 1, FPLS_ERR_ARG, "gemm_nn_f32: inner dimension must be positive");
  FPLS_REQUIRE(!shiftA || shift32, FPLS_ERR_ARG, "gemm_nn_f32: the FP32 shift is missing");
  dim3 grid((unsigned)((M + G2_BM - 1) / G2_BM), (unsigned)((N + G2_BN - 1) / G2_BN));
  const int crow = c_rowmajor ? 1 : 0;
  const int64_t ldb = nn_pad_k(K);
  // centred contractions (projections of X) are timed as "project_*"
  const bool bf = precision == FPLS_PREC_BF16X3;
  if (N <= 128) {
    const ProjShape sh{bf, shiftA != nullptr, M, (int)N, K, A, lda, shift32, B, ldb, offset,
                       Cf, Cd, ldc, crow};
    const Ctx::TuneChoice ch = proj_choose(ctx, sh);
    proj_run(ctx, sh, ch.mb, ch.splits);
    return;
  }
  if (M >= kProjSliceRows && N <= kProjSliceMaxN) {
    // tall A, a few hundred columns: one streaming pass of the skinny
    // projection per 128 columns of B (B's padded columns make every slice
    // whole) reads A once per slice at the streaming kernel's rate; the
    // 128 x 128 tiles would read it as often, far slower
    for (int64_t j0 = 0; j0 < N; j0 += 128) {
      const int nj = (int)std::min<int64_t>(128, N - j0);
      const int64_t co = crow ? j0 : j0 * ldc;
      const ProjShape sh{bf, shiftA != nullptr, M, nj, K, A, lda, shift32, B + j0 * ldb, ldb,
                         offset ? offset + j0 : nullptr, Cf ? Cf + co : nullptr,
                         Cd ? Cd + co : nullptr, ldc, crow};
      const Ctx::TuneChoice ch = proj_choose(ctx, sh);
      proj_run(ctx, sh, ch.mb, ch.splits);
    }
    return;
  }
  if (shiftA) {
    if (bf)
      launch(ctx, "project_bf16x3", k_g2_nn<true, true>, grid, dim3(G2_NT), 0, M, N, K, A, lda,
             shiftA, B, ldb, offset, Cf, Cd, ldc, crow);
    else
      launch(ctx, "project_f32", k_g2_nn<false, true>, grid, dim3(G2_NT), 0, M, N, K, A, lda,
             shiftA, B, ldb, offset, Cf, Cd, ldc, crow);
  } else {
    if (bf)
      launch(ctx, "gemm_nn_bf16x3", k_g2_nn<true, false>, grid, dim3(G2_NT), 0, M, N, K, A, lda,
             shiftA, B, ldb, offset, Cf, Cd, ldc, crow);
    else
      launch(ctx, "gemm_nn_f32", k_g2_nn<false, false>, grid, dim3(G2_NT), 0, M, N, K, A, lda,
             shiftA, B, ldb, offset, Cf, Cd, ldc, crow);
  }
}

// ---------------------------------------------------------------- TN, one-hot B
// P_z (M x N, N <= 128) = sum_{k in slab z} (A[k, m] - shA[m]) [y_k == n + 1]:
// the class sums of the centred columns of A (K x M col-major) from int32 labels.
// For column-major A the MFMA A operand (8 consecutive k of one m per lane) is
// two contiguous float4 loads, and the one-hot B operand is generated in
// registers from 8 labels, so the kernel needs neither LDS staging nor barriers
// in its main loop.  Labels outside 1..N (and rows past the slab, read as 0)
// fall in no class.  Workgroup: 32 columns m (2 blocks of 16) x all N; its 4
// waves take interleaved 32-deep K steps of the slab and are summed through LDS
// at the end.  Split-BF16x3 splits A three ways (hi + mid + lo), which with the
// exact 0/1 operand reproduces the FP32 product; FP32 mode uses
// v_mfma_f32_16x16x4f32 with lane-private k (a summation index may be assigned
// to lanes in any order shared by both operands).
constexpr int TL_BM = 32, TL_BK = 32, TL_NT = 256;

// C(m, n) = sum_z P_z(m, n) - shB(n) sum_z P_z(m, N)   (second term with shB)
__global__ __launch_bounds__(256) void k_tl_reduce(int64_t M, int64_t N, int splits,
                                                   const float* __restrict__ partial,
                                                   const double* __restrict__ shB,
                                                   double* __restrict__ C, int64_t ldc) {
  __shared__ double red[8][32][2];
  const int64_t NC = N + (shB ? 1 : 0);
  const int64_t stride = M * NC;
  const int o = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int64_t t = (int64_t)blockIdx.x * 32 + o;
  const int64_t m = t % M, n = t / M;
  double s = 0.0, u = 0.0;
  if (t < M * N) {
#pragma unroll 4
    for (int z = g; z < splits; z += 8) {
      s += static_cast<double>(partial[(int64_t)z * stride + m + n * M]);
      if (shB) u += static_cast<double>(partial[(int64_t)z * stride + m + N * M]);
    }
  }
  red[g][o][0] = s;
  red[g][o][1] = u;
  __syncthreads();
  if (g == 0 && t < M * N) {
    double r = 0.0, v = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      r += red[q][o][0];
      v += red[q][o][1];
    }
    if (shB) r -= v * shB[n];
    C[m + n * ldc] = r;
  }
}

// k_tl_reduce (with shB) that also finishes the column statistics of the
// shifted A: threads past M * N write mean[m] = shA[m] + sum_z P_z(m, N) / n
// and ones[m] = 1 (either optional).
__global__ __launch_bounds__(256) void k_tl_reduce_stats(int64_t M, int64_t N, int splits,
                                                         const float* __restrict__ partial,
                                                         const double* __restrict__ shB,
                                                         double* __restrict__ C, int64_t ldc,
                                                         const double* __restrict__ shA,
                                                         int64_t nrows, double* __restrict__ mean,
                                                         double* __restrict__ ones) {
  const int64_t NC = N + 1, stride = M * NC;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  auto slab_sum = [&](int64_t off) {
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int z = 0;
    for (; z + 8 <= splits; z += 8) {
#pragma unroll
      for (int q = 0; q < 8; ++q) a[q] += (double)partial[(int64_t)(z + q) * stride + off];
    }
    for (int q = 0; z + q < splits; ++q) a[q] += (double)partial[(int64_t)(z + q) * stride + off];
    return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  };
  if (t < M * N) {
    const int64_t m = t % M, n = t / M;
    C[m + n * ldc] = slab_sum(m + n * M) - shB[n] * slab_sum(m + N * M);
  } else if (t < M * N + M) {
    const int64_t m = t - M * N;
    if (mean) mean[m] = shA[m] + slab_sum(m + N * M) / (double)nrows;
    if (ones) ones[m] = 1.0;
  }
}

template <bool BF>
__global__ __launch_bounds__(TL_NT) void k_tns_lab(int64_t M, int N, int64_t K, int64_t kchunk,
                                                   const float* A, int64_t lda,
                                                   const double* shA, const int32_t* lab,
                                                   int ones, float* partial,
                                                   double* __restrict__ self_shift) {
  __shared__ float red[4][TL_BM][128 + 1];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * TL_BM;
  const int64_t kbeg = (int64_t)blockIdx.y * kchunk;
  const int64_t kend = std::min<int64_t>(K, kbeg + kchunk);
  // with `ones`, column N is the all-ones column: P(:, N) = column sums of
  // the centred A over the slab (for the one-hot centring correction)
  const int NC = N + (ones ? 1 : 0);
  const int nb = (NC + 15) >> 4;
  const int64_t mr0 = std::min<int64_t>(m0 + lane % 16, M - 1);
  const int64_t mr1 = std::min<int64_t>(m0 + 16 + lane % 16, M - 1);
  const float* col0 = A + mr0 * lda;
  const float* col1 = A + mr1 * lda;
  // self_shift: each lane shifts its columns by their first row (a sample of
  // the column, the classic shifted-data rule; the same in every workgroup,
  // recorded by slab 0)
  double d0 = 0.0, d1 = 0.0;
  if (self_shift) {
    d0 = (double)col0[0];
    d1 = (double)col1[0];
    if (blockIdx.y == 0 && lane < 16) {
      if (m0 + lane < M) self_shift[m0 + lane] = d0;
      if (m0 + 16 + lane < M) self_shift[m0 + 16 + lane] = d1;
    }
  }
  const float s0 = shA ? static_cast<float>(shA[mr0]) : static_cast<float>(d0);
  const float s1 = shA ? static_cast<float>(shA[mr1]) : static_cast<float>(d1);
  const bool vec = ((lda & 3) == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0) &&
                   ((reinterpret_cast<uintptr_t>(lab) & 15) == 0) && ((kbeg & 3) == 0);
  v4f acc[2][8];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  const int ko = 8 * (lane / 16);
  auto fetch = [&](int64_t k0, float (&a0)[8], float (&a1)[8], int (&l8)[8]) {
    const int64_t kb = k0 + ko;
    if (vec && kb + 8 <= kend) {
      const float4 p = *reinterpret_cast<const float4*>(col0 + kb);
      const float4 q = *reinterpret_cast<const float4*>(col0 + kb + 4);
      const float4 u = *reinterpret_cast<const float4*>(col1 + kb);
      const float4 v = *reinterpret_cast<const float4*>(col1 + kb + 4);
      const int4 la = *reinterpret_cast<const int4*>(lab + kb);
      const int4 lz = *reinterpret_cast<const int4*>(lab + kb + 4);
      a0[0] = p.x; a0[1] = p.y; a0[2] = p.z; a0[3] = p.w;
      a0[4] = q.x; a0[5] = q.y; a0[6] = q.z; a0[7] = q.w;
      a1[0] = u.x; a1[1] = u.y; a1[2] = u.z; a1[3] = u.w;
      a1[4] = v.x; a1[5] = v.y; a1[6] = v.z; a1[7] = v.w;
      l8[0] = la.x; l8[1] = la.y; l8[2] = la.z; l8[3] = la.w;
      l8[4] = lz.x; l8[5] = lz.y; l8[6] = lz.z; l8[7] = lz.w;
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int64_t k = kb + j;
        const int64_t kc = std::min<int64_t>(k, K - 1);
        a0[j] = col0[kc];
        a1[j] = col1[kc];
        l8[j] = k < kend ? lab[kc] : 0;  // label 0 matches no class
      }
    }
  };
  auto compute = [&](const float (&a0)[8], const float (&a1)[8], const int (&l8)[8]) {
    if (BF) {
      v8bf h0, m0v, l0, h1, m1v, l1;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float c0 = a0[j] - s0, c1 = a1[j] - s1;
        const __bf16 hh0 = static_cast<__bf16>(c0), hh1 = static_cast<__bf16>(c1);
        const float r0 = c0 - static_cast<float>(hh0), r1 = c1 - static_cast<float>(hh1);
        const __bf16 mm0 = static_cast<__bf16>(r0), mm1 = static_cast<__bf16>(r1);
        h0[j] = hh0;
        m0v[j] = mm0;
        l0[j] = static_cast<__bf16>(r0 - static_cast<float>(mm0));
        h1[j] = hh1;
        m1v[j] = mm1;
        l1[j] = static_cast<__bf16>(r1 - static_cast<float>(mm1));
      }
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        if (b < nb) {
          const int cls = b * 16 + lane % 16 + 1;
          const bool all = ones && cls == N + 1;
          v8bf oh;
#pragma unroll
          for (int j = 0; j < 8; ++j)
            oh[j] = static_cast<__bf16>((l8[j] == cls || (all && l8[j] > 0)) ? 1.f : 0.f);
          acc[0][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l0, oh, acc[0][b], 0, 0, 0);
          acc[0][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m0v, oh, acc[0][b], 0, 0, 0);
          acc[0][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h0, oh, acc[0][b], 0, 0, 0);
          acc[1][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l1, oh, acc[1][b], 0, 0, 0);
          acc[1][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m1v, oh, acc[1][b], 0, 0, 0);
          acc[1][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h1, oh, acc[1][b], 0, 0, 0);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float c0 = a0[j] - s0, c1 = a1[j] - s1;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          if (b < nb) {
            const int cls = b * 16 + lane % 16 + 1;
            const bool all = ones && cls == N + 1;
            const float oh = (l8[j] == cls || (all && l8[j] > 0)) ? 1.f : 0.f;
            acc[0][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(c0, oh, acc[0][b], 0, 0, 0);
            acc[1][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(c1, oh, acc[1][b], 0, 0, 0);
          }
        }
      }
    }
  };
  // wave w takes K steps w, w + 4, w + 8, ... of the slab; three register
  // sets in a fixed rotation (the step loop unrolled three times) keep the
  // loads of steps i + 1 and i + 2 in flight during the MFMAs of step i, and
  // no register still waiting for its load is copied
  struct Set {
    float a[8], b[8];
    int l[8];
  };
  Set S0, S1, S2;
  const int64_t kst = kbeg + (int64_t)w * TL_BK, kstr = 4 * TL_BK;
  const int ns = kst < kend ? (int)((kend - kst + kstr - 1) / kstr) : 0;
  auto step = [&](int i, Set& cur, Set& far) {
    if (i + 2 < ns) fetch(kst + (int64_t)(i + 2) * kstr, far.a, far.b, far.l);
    __builtin_amdgcn_sched_barrier(0);
    compute(cur.a, cur.b, cur.l);
  };
  if (ns > 0) fetch(kst, S0.a, S0.b, S0.l);
  if (ns > 1) fetch(kst + kstr, S1.a, S1.b, S1.l);
  for (int i = 0; i < ns; i += 3) {
    step(i, S0, S2);
    if (i + 1 < ns) step(i + 1, S1, S0);
    if (i + 2 < ns) step(i + 2, S2, S1);
  }
  // sum the four waves: lane holds D[4(l/16) + r][l%16] of each 16 x 16 block
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[w][a * 16 + 4 * (lane / 16) + r][b * 16 + lane % 16] = acc[a][b][r];
  __syncthreads();
  float* P = partial + (int64_t)blockIdx.y * M * NC;
  for (int idx = t; idx < TL_BM * 128; idx += TL_NT) {
    const int mi = idx % TL_BM, n = idx / TL_BM;
    const int64_t m = m0 + mi;
    if (m < M && n < NC)
      P[m + (int64_t)n * M] = (red[0][mi][n] + red[1][mi][n]) + (red[2][mi][n] + red[3][mi][n]);
  }
}

// ---------------------------------------------------------------- TN, skinny dense B
// P_z (M x N, N <= 16 NB) = sum_{k in slab z} (A[k, m] - shA[m]) B[k, n] for a
// dense, zero-padded B (K x N): the batched fits' X^T T.  Workgroup = 4 waves
// x 32 columns m (128 per workgroup), all on the same 32-deep K step: A goes
// straight from HBM into operand registers (two float4 per lane and column,
// next step prefetched), and the step's B tile is loaded once per workgroup
// into LDS (double buffered, one barrier per step), so the L2 traffic of B
// is a quarter of one tile per wave.  The B tile is stored [n][k] (k
// contiguous), from which each lane reads its operand with 16-byte LDS loads.
// Split-BF16x3 takes B pre-split into bf16 hi / lo planes ([n][k], ld ldb >=
// K rounded up to 32, zero rows past K), FP32 takes B itself (column-major,
// same padding).  kchunk is a multiple of 32, so only the last slab runs past
// K, into B's zero rows.  Each wave owns its columns: no cross-wave sum.
constexpr int TS_LDK = TL_BK + 8;  // LDS row length (elements) of the [n][k] B tile

template <bool BF, int NB>
__global__ __launch_bounds__(TL_NT) __attribute__((amdgpu_waves_per_eu(BF && NB == 8 ? 2 : 3))) void k_tns_dense(int64_t M, int N, int64_t K, int64_t kchunk,
                                                     const float* A, int64_t lda,
                                                     const double* shA, const float* Bf,
                                                     const __bf16* Bh, const __bf16* Bl,
                                                     int64_t ldb, float* partial) {
  // per buffer: BF -> hi and lo [16 NB][TS_LDK] bf16; FP32 -> [16 NB][TS_LDK] float
  constexpr int TILE_F = 16 * NB * TS_LDK;  // floats per buffer (FP32) = bf16 pairs (BF)
  __shared__ __align__(16) float smem[2 * TILE_F];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * (4 * TL_BM) + w * TL_BM;
  const int64_t kbeg = (int64_t)blockIdx.y * kchunk;
  const int64_t kend = std::min<int64_t>(K, kbeg + kchunk);
  const int nb = (N + 15) >> 4;
  const int64_t mr0 = std::min<int64_t>(m0 + lane % 16, M - 1);
  const int64_t mr1 = std::min<int64_t>(m0 + 16 + lane % 16, M - 1);
  const float* col0 = A + mr0 * lda;
  const float* col1 = A + mr1 * lda;
  const float s0 = shA ? static_cast<float>(shA[mr0]) : 0.f;
  const float s1 = shA ? static_cast<float>(shA[mr1]) : 0.f;
  const bool vec = ((lda & 3) == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0) &&
                   ((kbeg & 3) == 0);
  v4f acc[2][NB];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  const int ko = 8 * (lane / 16);
  float x0[8], x1[8], y0[8], y1[8];
  auto fetchA = [&](int64_t k0, float (&a0)[8], float (&a1)[8]) {
    const int64_t kb = k0 + ko;
    if (vec && kb + 8 <= K) {
      const float4 p = *reinterpret_cast<const float4*>(col0 + kb);
      const float4 q = *reinterpret_cast<const float4*>(col0 + kb + 4);
      const float4 u = *reinterpret_cast<const float4*>(col1 + kb);
      const float4 v = *reinterpret_cast<const float4*>(col1 + kb + 4);
      a0[0] = p.x; a0[1] = p.y; a0[2] = p.z; a0[3] = p.w;
      a0[4] = q.x; a0[5] = q.y; a0[6] = q.z; a0[7] = q.w;
      a1[0] = u.x; a1[1] = u.y; a1[2] = u.z; a1[3] = u.w;
      a1[4] = v.x; a1[5] = v.y; a1[6] = v.z; a1[7] = v.w;
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int64_t kc = std::min<int64_t>(kb + j, K - 1);
        a0[j] = col0[kc];
        a1[j] = col1[kc];
      }
    }
  };
  // B staging: thread t -> column bn = t / 4 + 64 pass (NB / 4 passes; clamped
  // to N - 1: columns >= N only feed outputs that are never stored), 8
  // consecutive k at 8 (t % 4)
  // (named registers, not arrays: arrays captured by the lambdas below were
  // left in scratch memory)
  constexpr int NPASS = NB / 4;
  static_assert(NPASS == 1 || NPASS == 2, "k_tns_dense: NB must be 4 or 8");
  const int bk = 8 * (t & 3);
  const int64_t bsA = (int64_t)std::min(t >> 2, N - 1) * ldb + bk;
  const int64_t bsB = (int64_t)std::min((t >> 2) + 64, N - 1) * ldb + bk;
  float4 bA0, bA1, bB0, bB1;  // FP32: 8 floats; BF: hi (x0) / lo (x1), 8 bf16 each
  auto fetch1 = [&](int64_t src, float4& q0, float4& q1) {
    if constexpr (BF) {
      q0 = *reinterpret_cast<const float4*>(Bh + src);
      q1 = *reinterpret_cast<const float4*>(Bl + src);
    } else {
      q0 = *reinterpret_cast<const float4*>(Bf + src);
      q1 = *reinterpret_cast<const float4*>(Bf + src + 4);
    }
  };
  auto fetchB = [&](int64_t k0) {
    fetch1(bsA + k0, bA0, bA1);
    if constexpr (NPASS == 2) fetch1(bsB + k0, bB0, bB1);
  };
  auto store1 = [&](float* base, int bn, const float4& q0, const float4& q1) {
    if constexpr (BF) {
      __bf16* hb = reinterpret_cast<__bf16*>(base);
      __bf16* lb = hb + 16 * NB * TS_LDK;
      *reinterpret_cast<float4*>(hb + bn * TS_LDK + bk) = q0;
      *reinterpret_cast<float4*>(lb + bn * TS_LDK + bk) = q1;
    } else {
      *reinterpret_cast<float4*>(base + bn * TS_LDK + bk) = q0;
      *reinterpret_cast<float4*>(base + bn * TS_LDK + bk + 4) = q1;
    }
  };
  auto storeB = [&](int buf) {
    float* base = smem + buf * TILE_F;
    store1(base, t >> 2, bA0, bA1);
    if constexpr (NPASS == 2) store1(base, (t >> 2) + 64, bB0, bB1);
  };
  auto compute = [&](int buf, const float (&a0)[8], const float (&a1)[8]) {
    const float* base = smem + buf * TILE_F;
    if constexpr (BF) {
      const __bf16* hb = reinterpret_cast<const __bf16*>(base);
      const __bf16* lb = hb + 16 * NB * TS_LDK;
      v8bf h0, l0, h1, l1;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        __bf16 hh, ll;
        split_bf(a0[j] - s0, hh, ll);
        h0[j] = hh;
        l0[j] = ll;
        split_bf(a1[j] - s1, hh, ll);
        h1[j] = hh;
        l1[j] = ll;
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (b < nb) {
          const int o = (b * 16 + lane % 16) * TS_LDK + ko;
          const v8bf bh = *reinterpret_cast<const v8bf*>(hb + o);
          const v8bf bl = *reinterpret_cast<const v8bf*>(lb + o);
          acc[0][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l0, bh, acc[0][b], 0, 0, 0);
          acc[0][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h0, bl, acc[0][b], 0, 0, 0);
          acc[0][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h0, bh, acc[0][b], 0, 0, 0);
          acc[1][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l1, bh, acc[1][b], 0, 0, 0);
          acc[1][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h1, bl, acc[1][b], 0, 0, 0);
          acc[1][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h1, bh, acc[1][b], 0, 0, 0);
        }
      }
    } else {
      float c0[8], c1[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        c0[j] = a0[j] - s0;
        c1[j] = a1[j] - s1;
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (b < nb) {
          const float* bc = base + (b * 16 + lane % 16) * TS_LDK + ko;
          const float4 p = *reinterpret_cast<const float4*>(bc);
          const float4 q = *reinterpret_cast<const float4*>(bc + 4);
          const float bv[8] = {p.x, p.y, p.z, p.w, q.x, q.y, q.z, q.w};
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            acc[0][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(c0[j], bv[j], acc[0][b], 0, 0, 0);
            acc[1][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(c1[j], bv[j], acc[1][b], 0, 0, 0);
          }
        }
      }
    }
  };
  // A runs two steps ahead of the MFMAs in three register sets (x, y, z)
  // whose roles rotate through a 3-way unrolled loop, so no set is copied.
  // Each step fetches the next B tile before the A operand two steps ahead:
  // loads complete in issue order, so the wait for B before its LDS store
  // leaves the younger A loads in flight across the barrier.
  const int nk = (int)((kend - kbeg + TL_BK - 1) / TL_BK);
  float z0[8], z1[8];
  if (nk > 0) {
    fetchB(kbeg);
    fetchA(kbeg, x0, x1);
    if (nk > 1) fetchA(kbeg + TL_BK, y0, y1);
    storeB(0);
  }
  __syncthreads();
  auto step = [&](int c, const float (&c0)[8], const float (&c1)[8], float (&d0)[8],
                  float (&d1)[8]) {
    if (c + 1 < nk) fetchB(kbeg + (int64_t)(c + 1) * TL_BK);
    if (c + 2 < nk) fetchA(kbeg + (int64_t)(c + 2) * TL_BK, d0, d1);
    __builtin_amdgcn_sched_barrier(0);
    compute(c & 1, c0, c1);
    if (c + 1 < nk) storeB((c + 1) & 1);
    __syncthreads();
  };
  for (int c = 0; c < nk; c += 3) {
    step(c, x0, x1, z0, z1);
    if (c + 1 < nk) step(c + 1, y0, y1, x0, x1);
    if (c + 2 < nk) step(c + 2, z0, z1, y0, y1);
  }
  // lane holds D[4(l/16) + r][l%16] of each 16 x 16 block: rows are columns m
  // of A, columns are n
  float* P = partial + (int64_t)blockIdx.y * M * N;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int64_t mb = m0 + 16 * a + 4 * (lane / 16);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int n = b * 16 + lane % 16;
      if (b >= nb || n >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (mb + r < M) P[mb + r + (int64_t)n * M] = acc[a][b][r];
    }
  }
}

// C (M x N FP64, ldc) = (A - 1 shA^T)^T B for a skinny dense B (N <= 128)
// given as FP32 (Bf) or as bf16 hi / lo planes (Bh, Bl), see k_tns_dense.
void tn_skinny(Ctx* ctx, int precision, int64_t M, int N, int64_t K, const float* A, int64_t lda,
               const double* shA, const float* Bf, const __bf16* Bh, const __bf16* Bl,
               int64_t ldb, int64_t splits, double* C, int64_t ldc) {
  const bool bf = precision == FPLS_PREC_BF16X3;
  FPLS_REQUIRE(N >= 1 && N <= 128 && ldb >= (K + TL_BK - 1) / TL_BK * TL_BK &&
                   (bf ? (Bh && Bl) : Bf != nullptr),
               FPLS_ERR_ARG, "tn_skinny: bad operands");
  splits = std::max<int64_t>(1, std::min<int64_t>(splits, (K + TL_BK - 1) / TL_BK));
  int64_t kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + TL_BK - 1) / TL_BK * TL_BK;
  splits = (K + kchunk - 1) / kchunk;
  DBuf<float> part(ctx, (size_t)splits * M * N);
  dim3 grid((unsigned)((M + 4 * TL_BM - 1) / (4 * TL_BM)), (unsigned)splits);
  if (N <= 64) {
    if (bf)
      launch(ctx, "tn_skinny_bf16x3", k_tns_dense<true, 4>, grid, dim3(TL_NT), 0, M, N, K,
             kchunk, A, lda, shA, Bf, Bh, Bl, ldb, part.get());
    else
      launch(ctx, "tn_skinny_f32", k_tns_dense<false, 4>, grid, dim3(TL_NT), 0, M, N, K, kchunk,
             A, lda, shA, Bf, Bh, Bl, ldb, part.get());
  } else {
    if (bf)
      launch(ctx, "tn_skinny_bf16x3", k_tns_dense<true, 8>, grid, dim3(TL_NT), 0, M, N, K,
             kchunk, A, lda, shA, Bf, Bh, Bl, ldb, part.get());
    else
      launch(ctx, "tn_skinny_f32", k_tns_dense<false, 8>, grid, dim3(TL_NT), 0, M, N, K, kchunk,
             A, lda, shA, Bf, Bh, Bl, ldb, part.get());
  }
  launch(ctx, "tn_skinny_reduce", k_g2_reduce, dim3((unsigned)((M * N + 31) / 32)), dim3(256), 0,
         M, (int64_t)N, (int)splits, (const float*)part.get(), C, ldc, 0);
}

// The B operand of k_tns_dense from a column-major B (K x N, ld ldb): [n][k]
// with ld Kp (K rounded up to 32) and zero rows K .. Kp - 1, as FP32 (Bf) or
// as bf16 hi / lo planes (Bh, Bl).
__global__ void k_g2_bsplit(const float* __restrict__ B, int64_t ldb, int64_t K, int N,
                            int64_t Kp, const double* __restrict__ shift, float* __restrict__ Bf,
                            __bf16* __restrict__ Bh, __bf16* __restrict__ Bl) {
  const int64_t tot = Kp * N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e % Kp, n = e / Kp;
    const float v = k >= K ? 0.f
                    : shift ? (float)((double)B[k + n * ldb] - shift[n])
                            : B[k + n * ldb];
    if (Bf) {
      Bf[e] = v;
    } else {
      __bf16 hi, lo;
      split_bf(v, hi, lo);
      Bh[e] = hi;
      Bl[e] = lo;
    }
  }
}

// ---------------------------------------------------------------- Gram, M <= 128
// P_z (M x M) = sum_{k in slab z} (A[k, :] - sh)^T (A[k, :] - sh) for a tall
// A (K x M col-major).  One 32-deep K step of A is staged once in LDS in the
// [m][k] layout that serves both MFMA operands (hi/lo bf16 for split-BF16x3,
// FP32 [k][m] for the plain path), double buffered with one barrier per step;
// wave w owns output rows 32w .. 32w + 31.
template <bool BF>
struct GSmem {
  static constexpr int value = BF ? (2 * 2 * 128 * P_LDB) / 2 : 2 * P_BK * (128 + 4);
};

template <bool BF>
__global__ __launch_bounds__(P_NT) void k_gram(int64_t M, int64_t K, int64_t kchunk,
                                               const float* A, int64_t lda, const double* sh,
                                               float* partial) {
  __shared__ __align__(16) float smem[GSmem<BF>::value];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t kbeg = (int64_t)blockIdx.x * kchunk;
  const int64_t kend = std::min<int64_t>(K, kbeg + kchunk);
  const int nb = (int)((M + 15) >> 4);
  v4f acc[2][8];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  // staging: thread -> column m = t / 2, 16 consecutive k
  const int sm = t >> 1, sk = (t & 1) * 16;
  const bool mok = sm < M;
  const float* scol = A + std::min<int64_t>(sm, M - 1) * lda;
  const float shv = (sh && mok) ? static_cast<float>(sh[sm]) : 0.f;
  const bool vec = ((lda & 3) == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0) &&
                   ((kbeg & 3) == 0);
  // three register sets (chunks c, c + 1, c + 2) in a fixed rotation of the
  // loop unrolled three times: the loads run two chunks ahead and no register
  // still waiting for its load is copied
  struct GSet {
    float v[16];
  };
  GSet V0, V1, V2;
  auto fetch = [&](int64_t k0, float (&v)[16]) {
    const int64_t kb = k0 + sk;
    if (vec && kb + 16 <= kend) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 f = *reinterpret_cast<const float4*>(scol + kb + 4 * q);
        v[4 * q + 0] = f.x;
        v[4 * q + 1] = f.y;
        v[4 * q + 2] = f.z;
        v[4 * q + 3] = f.w;
      }
    } else {
#pragma unroll
      for (int e = 0; e < 16; ++e) v[e] = scol[std::min<int64_t>(kb + e, K - 1)];
    }
  };
  auto store = [&](int64_t k0, int buf, const float (&v)[16]) {
    const int64_t kb = k0 + sk;
    float c[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) c[e] = (mok && kb + e < kend) ? v[e] - shv : 0.f;
    if (BF) {
      __bf16* base = reinterpret_cast<__bf16*>(smem);
      __bf16* Th = base + (0 * 2 + buf) * 128 * P_LDB;
      __bf16* Tl = base + (1 * 2 + buf) * 128 * P_LDB;
      v8bf h0, l0, h1, l1;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        __bf16 hh, ll;
        split_bf(c[e], hh, ll);
        h0[e] = hh;
        l0[e] = ll;
        split_bf(c[8 + e], hh, ll);
        h1[e] = hh;
        l1[e] = ll;
      }
      *reinterpret_cast<v8bf*>(Th + sm * P_LDB + sk) = h0;
      *reinterpret_cast<v8bf*>(Th + sm * P_LDB + sk + 8) = h1;
      *reinterpret_cast<v8bf*>(Tl + sm * P_LDB + sk) = l0;
      *reinterpret_cast<v8bf*>(Tl + sm * P_LDB + sk + 8) = l1;
    } else {
      float* Ts = smem + buf * P_BK * (128 + 4);
#pragma unroll
      for (int e = 0; e < 16; ++e) Ts[(sk + e) * (128 + 4) + sm] = c[e];
    }
  };
  auto compute = [&](int buf) {
    if (BF) {
      const __bf16* base = reinterpret_cast<const __bf16*>(smem);
      const __bf16* Th = base + (0 * 2 + buf) * 128 * P_LDB;
      const __bf16* Tl = base + (1 * 2 + buf) * 128 * P_LDB;
      const int ko = 8 * (lane / 16);
      v8bf ah[2], al[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int m = w * 32 + a * 16 + lane % 16;
        ah[a] = *reinterpret_cast<const v8bf*>(Th + m * P_LDB + ko);
        al[a] = *reinterpret_cast<const v8bf*>(Tl + m * P_LDB + ko);
      }
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        if (b < nb) {
          const int n = b * 16 + lane % 16;
          const v8bf bh = *reinterpret_cast<const v8bf*>(Th + n * P_LDB + ko);
          const v8bf bl = *reinterpret_cast<const v8bf*>(Tl + n * P_LDB + ko);
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh, acc[a][b], 0, 0, 0);
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl, acc[a][b], 0, 0, 0);
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh, acc[a][b], 0, 0, 0);
          }
        }
      }
    } else {
      const float* Ts = smem + buf * P_BK * (128 + 4);
#pragma unroll
      for (int s4 = 0; s4 < 8; ++s4) {
        const int kr = 4 * s4 + lane / 16;
        float af[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) af[a] = Ts[kr * (128 + 4) + w * 32 + a * 16 + lane % 16];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          if (b < nb) {
            const float bv = Ts[kr * (128 + 4) + b * 16 + lane % 16];
#pragma unroll
            for (int a = 0; a < 2; ++a)
              acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bv, acc[a][b], 0, 0, 0);
          }
        }
      }
    }
  };
  const int nk = (int)((kend - kbeg + P_BK - 1) / P_BK);
  auto step = [&](int c, GSet& nxt, GSet& far) {
    if (c + 2 < nk) fetch(kbeg + (int64_t)(c + 2) * P_BK, far.v);
    __builtin_amdgcn_sched_barrier(0);
    compute(c & 1);
    if (c + 1 < nk) store(kbeg + (int64_t)(c + 1) * P_BK, (c + 1) & 1, nxt.v);
    __syncthreads();
  };
  if (nk > 0) {
    fetch(kbeg, V0.v);
    if (nk > 1) fetch(kbeg + P_BK, V1.v);
    store(kbeg, 0, V0.v);
  }
  __syncthreads();
  for (int c = 0; c < nk; c += 3) {
    step(c, V1, V2);
    if (c + 1 < nk) step(c + 1, V2, V0);
    if (c + 2 < nk) step(c + 2, V0, V1);
  }
  float* P = partial + (int64_t)blockIdx.x * M * M;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int col = b * 16 + lane % 16;
      if (b >= nb || col >= M) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = w * 32 + a * 16 + 4 * (lane / 16) + r;
        if (row < M) P[row + (int64_t)col * M] = acc[a][b][r];
      }
    }
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
  if (!labelsB && A == B && lda == ldb && shiftA == shiftB && M == N && M <= 128) {
    // Gram matrix of a tall A
    int64_t splits = std::max<int64_t>(1, std::min<int64_t>(2 * ctx->num_cus, (K + 255) / 256));
    const int64_t cap = std::max<int64_t>(1, (int64_t(32) << 20) / (4 * M * M));
    splits = std::min(splits, cap);
    int64_t kchunk = (K + splits - 1) / splits;
    kchunk = (kchunk + P_BK - 1) / P_BK * P_BK;
    splits = (K + kchunk - 1) / kchunk;
    DBuf<float> part(ctx, (size_t)splits * M * M);
    if (precision == FPLS_PREC_BF16X3)
      launch(ctx, "gemm_tn_bf16x3", k_gram<true>, dim3((unsigned)splits), dim3(P_NT), 0, M, K,
             kchunk, A, lda, shiftA, part.get());
    else
      launch(ctx, "gemm_tn_f32", k_gram<false>, dim3((unsigned)splits), dim3(P_NT), 0, M, K,
             kchunk, A, lda, shiftA, part.get());
    launch(ctx, "gemm_tn_reduce", k_g2_reduce, dim3((unsigned)((M * N + 31) / 32)), dim3(256), 0, M, N, (int)splits,
           (const float*)part.get(), C, ldc, 0);
    return;
  }
  // more classes than one 128-column one-hot tile: class sums in one pass over
  // A (the centred product, shiftB set, does not depend on the centring of A,
  // and raw class sums take none)
  if (labelsB && N + (shiftB ? 1 : 0) > 128 && (shiftB || !shiftA) &&
      class_sums_f32(ctx, K, M, A, lda, labelsB, (int)N, shiftB, C, ldc, nullptr, nullptr))
    return;
  if (labelsB && N + (shiftB ? 1 : 0) <= 128) {
    // class sums from the labels; the centring of the one-hot by shiftB uses
    // an extra all-ones column (the column sums of the centred A)
    const int64_t mt = (M + TL_BM - 1) / TL_BM;
    const int64_t NC = N + (shiftB ? 1 : 0);
    int64_t splits = std::max<int64_t>(1, (2 * ctx->num_cus + mt - 1) / mt);
    splits = std::min<int64_t>(splits, (K + 4 * TL_BK * 4 - 1) / (4 * TL_BK * 4));
    int64_t kchunk = (K + splits - 1) / splits;
    kchunk = (kchunk + TL_BK - 1) / TL_BK * TL_BK;
    splits = (K + kchunk - 1) / kchunk;
    DBuf<float> part(ctx, (size_t)splits * M * NC);
    dim3 grid((unsigned)mt, (unsigned)splits);
    const int n32 = (int)N, ones = shiftB ? 1 : 0;
    if (precision == FPLS_PREC_BF16X3)
      launch(ctx, "gemm_tn_bf16x3", k_tns_lab<true>, grid, dim3(TL_NT), 0, M, n32, K, kchunk, A,
             lda, shiftA, labelsB, ones, part.get(), (double*)nullptr);
    else
      launch(ctx, "gemm_tn_f32", k_tns_lab<false>, grid, dim3(TL_NT), 0, M, n32, K, kchunk, A,
             lda, shiftA, labelsB, ones, part.get(), (double*)nullptr);
    launch(ctx, "gemm_tn_reduce", k_tl_reduce, dim3((unsigned)((M * N + 31) / 32)), dim3(256), 0, M, N, (int)splits,
           (const float*)part.get(), shiftB, C, ldc);
    return;
  }
  if (!labelsB && N <= 128) {
    // skinny dense B: the B tile of each K step is shared through LDS by the
    // workgroup's four waves (k_tns_dense); split K so that the grid holds
    // about five workgroups per CU.  B's shift is applied as it is padded.
    const bool bf = precision == FPLS_PREC_BF16X3;
    const int64_t Kp = (K + TL_BK - 1) / TL_BK * TL_BK;
    DBuf<float> Bf;
    DBuf<__bf16> Bh, Bl;
    if (bf) {
      Bh = DBuf<__bf16>(ctx, (size_t)Kp * N);
      Bl = DBuf<__bf16>(ctx, (size_t)Kp * N);
    } else {
      Bf = DBuf<float>(ctx, (size_t)Kp * N);
    }
    launch(ctx, "pad_b", k_g2_bsplit, dim3(g1(Kp * N)), dim3(256), 0, B, ldb, K, (int)N, Kp,
           shiftB, Bf.get(), Bh.get(), Bl.get());
    const int64_t tiles = (M + 4 * TL_BM - 1) / (4 * TL_BM);
    int64_t splits = std::max<int64_t>(1, (5 * (int64_t)ctx->num_cus + tiles - 1) / tiles);
    splits = std::min<int64_t>(splits, std::max<int64_t>(1, K / 256));
    splits = std::min<int64_t>(splits, std::max<int64_t>(1, (int64_t)ctx->partial_budget() / (4 * M * N)));
    tn_skinny(ctx, precision, M, (int)N, K, A, lda, shiftA, Bf.get(), Bh.get(), Bl.get(), Kp,
              splits, C, ldc);
    return;
  }
  const int64_t splits = tn_default_splits(ctx, M, N, K);
  gemm_tn_dense(ctx, precision, M, N, K, A, lda, shiftA, B, ldb, labelsB, shiftB, C, ldc,
                splits);
}

bool class_cross_cov_f32(Ctx* ctx, int precision, int64_t p, int64_t C, int64_t n,
                         const float* X, int64_t ld, const int32_t* y, const double* mY,
                         double* S, int64_t lds, double* mean, double* ones) {
  if (C + 1 > 128 || n <= 0 || p <= 0) return false;
  // the shift: any value gives the same centred product; a sample of each
  // column (its first row) keeps the FP32 partials of data far from zero
  // accurate
  DBuf<double> sh(ctx, (size_t)p);
  const int64_t mt = (p + TL_BM - 1) / TL_BM, NC = C + 1;
  int64_t splits = std::max<int64_t>(1, (2 * ctx->num_cus + mt - 1) / mt);
  splits = std::min<int64_t>(splits, (n + 4 * TL_BK * 4 - 1) / (4 * TL_BK * 4));
  int64_t kchunk = (n + splits - 1) / splits;
  kchunk = (kchunk + TL_BK - 1) / TL_BK * TL_BK;
  splits = (n + kchunk - 1) / kchunk;
  DBuf<float> part(ctx, (size_t)splits * p * NC);
  dim3 grid((unsigned)mt, (unsigned)splits);
  if (precision == FPLS_PREC_BF16X3)
    launch(ctx, "gemm_tn_bf16x3", k_tns_lab<true>, grid, dim3(TL_NT), 0, p, (int)C, n, kchunk, X,
           ld, (const double*)nullptr, y, 1, part.get(), sh.get());
  else
    launch(ctx, "gemm_tn_f32", k_tns_lab<false>, grid, dim3(TL_NT), 0, p, (int)C, n, kchunk, X,
           ld, (const double*)nullptr, y, 1, part.get(), sh.get());
  const int64_t outs = p * C + p;
  launch(ctx, "gemm_tn_reduce", k_tl_reduce_stats, dim3((unsigned)((outs + 255) / 256)),
         dim3(256), 0, p, C, (int)splits, (const float*)part.get(), mY, S, lds,
         (const double*)sh.get(), n, mean, ones);
  return true;
}

int64_t tn_default_splits(const Ctx* ctx, int64_t M, int64_t N, int64_t K) {
  const int64_t tiles = ((M + G2_BM - 1) / G2_BM) * ((N + G2_BN - 1) / G2_BN);
  const int64_t want = 2 * ctx->num_cus;
  int64_t splits = std::max<int64_t>(1, (want + tiles - 1) / tiles);
  return std::min<int64_t>(splits, (K + 511) / 512);  // at least 512 rows per slab
}

void gemm_tn_dense(Ctx* ctx, int precision, int64_t M, int64_t N, int64_t K, const float* A,
                   int64_t lda, const double* shiftA, const float* B, int64_t ldb,
                   const int32_t* labelsB, const double* shiftB, double* C, int64_t ldc,
                   int64_t splits) {
  // the Gram matrix of A: only the tiles on and above the diagonal
  const int sym = (!labelsB && A == B && lda == ldb && shiftA == shiftB && M == N) ? 1 : 0;
  if (M <= 0 || N <= 0) return;
  const int64_t cap = std::max<int64_t>(1, (int64_t)ctx->partial_budget() / (4 * M * N));
  splits = std::max<int64_t>(1, std::min(splits, cap));
  int64_t kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + G2_BK - 1) / G2_BK * G2_BK;
  splits = (K + kchunk - 1) / kchunk;
  DBuf<float> part(ctx, (size_t)splits * M * N);
  dim3 grid((unsigned)((M + G2_BM - 1) / G2_BM), (unsigned)((N + G2_BN - 1) / G2_BN),
            (unsigned)splits);
  if (precision == FPLS_PREC_BF16X3)
    launch(ctx, "gemm_tn_bf16x3", k_g2_tn<true>, grid, dim3(G2_NT), 0, M, N, K, kchunk, A, lda,
           shiftA, B, ldb, labelsB, shiftB, part.get(), sym);
  else
    launch(ctx, "gemm_tn_f32", k_g2_tn<false>, grid, dim3(G2_NT), 0, M, N, K, kchunk, A, lda,
           shiftA, B, ldb, labelsB, shiftB, part.get(), sym);
  launch(ctx, "gemm_tn_reduce", k_g2_reduce, dim3((unsigned)((M * N + 31) / 32)), dim3(256), 0, M, N, (int)splits,
         (const float*)part.get(), C, ldc, sym);
}

// Split-K skinny NN (N <= 128): part[z] (M x N FP32, column-major) is slab z
// (rows z kchunk ... of B) of (A - 1 shA^T) B.  B is the zero-padded FP32
// operand (at least ceil(K / 32) 32 rows and 128 columns, ld ldb), so the
// last slab needs no masking.
void proj_splitk(Ctx* ctx, int precision, int64_t M, int N, int64_t K, const float* A,
                 int64_t lda, const float* shA32, const float* Bpad, int64_t ldb, int64_t kchunk,
                 int splits, float* part) {
  FPLS_REQUIRE(N >= 1 && N <= 128 && kchunk % P_BK == 0 && splits >= 1 &&
                   (int64_t)(splits - 1) * kchunk < K && (int64_t)splits * kchunk >= K,
               FPLS_ERR_ARG, "proj_splitk: bad shape");
  dim3 grid((unsigned)((M + p_bm(2) - 1) / p_bm(2)), (unsigned)splits);
  const bool bf = precision == FPLS_PREC_BF16X3;
  if (shA32) {
    if (bf)
      launch_proj<true, true>(ctx, "cv_project_bf16x3", N, 2, grid, M, N, K, A, lda,
             shA32, Bpad, ldb, (const double*)nullptr, (float*)nullptr, (double*)nullptr,
             (int64_t)0, 0, kchunk, part);
    else
      launch_proj<false, true>(ctx, "cv_project_f32", N, 2, grid, M, N, K, A, lda,
             shA32, Bpad, ldb, (const double*)nullptr, (float*)nullptr, (double*)nullptr,
             (int64_t)0, 0, kchunk, part);
  } else {
    if (bf)
      launch_proj<true, false>(ctx, "cv_project_bf16x3", N, 2, grid, M, N, K, A,
             lda, (const float*)nullptr, Bpad, ldb, (const double*)nullptr, (float*)nullptr,
             (double*)nullptr, (int64_t)0, 0, kchunk, part);
    else
      launch_proj<false, false>(ctx, "cv_project_f32", N, 2, grid, M, N, K, A, lda,
             (const float*)nullptr, Bpad, ldb, (const double*)nullptr, (float*)nullptr,
             (double*)nullptr, (int64_t)0, 0, kchunk, part);
  }
}

// C (M x N col-major FP32, ld M) = A B^T, A: M x K and B: N x K row-major.
void gemm_rr_f32(Ctx* ctx, int precision, int64_t M, int64_t N, int64_t K, const float* A,
                 int64_t lda, const float* B, int64_t ldb, float* C) {
  if (M <= 0 || N <= 0) return;
  const int64_t kchunk = ((std::max<int64_t>(K, 1) + G2_BK - 1) / G2_BK) * G2_BK;
  dim3 grid((unsigned)((M + G2_BM - 1) / G2_BM), (unsigned)((N + G2_BN - 1) / G2_BN), 1);
  if (precision == FPLS_PREC_BF16X3)
    launch(ctx, "gemm_rr_bf16x3", k_g2_tn<true>, grid, dim3(G2_NT), 0, M, N, K, kchunk, A, lda,
           (const double*)nullptr, B, ldb, (const int32_t*)nullptr, (const double*)nullptr, C, 0);
  else
    launch(ctx, "gemm_rr_f32", k_g2_tn<false>, grid, dim3(G2_NT), 0, M, N, K, kchunk, A, lda,
           (const double*)nullptr, B, ldb, (const int32_t*)nullptr, (const double*)nullptr, C, 0);
}

}  // namespace fpls
