// One-pass class sums of an FP32 matrix from int32 labels: the segmented
// reduction behind S = X^T (onehot(y) - 1 mY^T) of the label-aware
// cross-product (the rowsum() accumulation of the label-aware stream model,
// R/main.R:563-588) and the class sums of the LDA head.
//
// X (n x p, column-major) is read once, whatever the number of classes.  Each
// wave owns COLS columns (64 / COLS row phases share a wave instruction) and
// streams tiles of 2048 / COLS rows through LDS (8 KiB of X a tile, loads two
// tiles ahead); every row is added into its class bin with an LDS float atomic
// that only this wave's lanes touch, so the order of the additions is fixed:
// the result is deterministic.  Rows are shifted by u = x - s with s the mean
// of the first 64 rows, which keeps the FP32 slab sums accurate for data far
// from zero; the shift cancels from the centred product,
//   S = sum_k u_k (onehot_k - mY)^T = U - (sum_k u_k) mY^T   (labels in 1..C),
// and the column means come out of the same pass: mX = s + (sum_k u_k) / n.
// Slab partials (p x (C + 1) per slab, FP32) are summed in FP64 in slab order.
#include <algorithm>

#include "fpls_internal.h"
#include "fpls_wave.h"

namespace fpls {

namespace {

constexpr int CS_ACC = 6656;     // class-bin floats per wave ((C + 1) * COLS, bin C spare)
constexpr int CS_TILE = 2560;    // tile floats per wave (2048 / COLS rows x COLS + 1)
constexpr int CS_LAB = 512;      // labels per tile (2048 / 4 at COLS = 4)
constexpr int CS_WAVE = CS_ACC + CS_TILE + CS_LAB;

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <int COLS>
__global__ __launch_bounds__(256) void k_class_sums(int64_t n, int64_t p,
                                                    const float* __restrict__ X, int64_t ld,
                                                    const int32_t* __restrict__ y, int C,
                                                    int use_shift, int64_t kchunk,
                                                    float* __restrict__ part,
                                                    float* __restrict__ shift_out) {
  constexpr int PH = 64 / COLS;        // row phases of a wave instruction
  constexpr int TR = 2048 / COLS;      // tile rows (8 KiB of X)
  constexpr int RPL = TR / PH;         // rows of a tile per lane (32)
  constexpr int LDT = COLS + 1;        // tile row stride
  constexpr int NLD = 8;               // float4 loads a lane per tile (512 a tile)
  constexpr int QPC = TR / 4;          // row quads of a tile column
  constexpr int NLB = TR >= 64 ? TR / 64 : 1;  // labels a lane per tile
  __shared__ float lds[4 * CS_WAVE];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t j0 = ((int64_t)blockIdx.x * 4 + w) * COLS;
  if (j0 >= p) return;  // the whole wave (no workgroup barriers below)
  float* acc = lds + w * CS_WAVE;
  float* tile = acc + CS_ACC;
  int* labs = reinterpret_cast<int*>(tile + CS_TILE);
  const int c = lane % COLS, ph = lane / COLS;
  const int64_t j = j0 + c;
  const bool jok = j < p;
  const int64_t jc = jok ? j : p - 1;
  const int64_t kb = (int64_t)blockIdx.y * kchunk, ke = std::min<int64_t>(n, kb + kchunk);

  // shift: the mean of the first 64 rows (every workgroup the same)
  float s = 0.f;
  if (use_shift) {
    const int64_t r0 = std::min<int64_t>(64, n);
    float a = 0.f;
    for (int64_t r = ph; r < r0; r += PH) a += X[r + jc * ld];
    tile[ph * LDT + c] = a;
    wave_fence();
    float t = 0.f;
    for (int q = 0; q < PH; ++q) t += tile[q * LDT + c];
    s = t / (float)r0;
    wave_fence();
    if (blockIdx.y == 0 && ph == 0 && jok) shift_out[j] = s;
  }
  for (int i = lane; i < (C + 1) * COLS; i += 64) acc[i] = 0.f;

  // tile loads: lane -> (column cc, row quad rq) of one float4, NLD a lane
  // (consecutive lanes read consecutive quads of a column); labels NLB a lane
  const bool vec = (ld % 4 == 0) && ((reinterpret_cast<uintptr_t>(X) & 15) == 0) && (kb % 4 == 0);
  struct TSet {
    float4 v[NLD];
    int lab[NLB];
  };
  auto fetch = [&](int64_t k0, TSet& d) {
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      const int t4 = lane + 64 * i, cc = t4 / QPC, rq = t4 % QPC;
      const int64_t col = std::min<int64_t>(j0 + cc, p - 1);
      const int64_t r = k0 + 4 * rq;
      if (vec && r + 4 <= ke) {
        d.v[i] = *reinterpret_cast<const float4*>(X + r + col * ld);
      } else {
        float e[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) e[q] = r + q < ke ? X[r + q + col * ld] : 0.f;
        d.v[i] = make_float4(e[0], e[1], e[2], e[3]);
      }
    }
#pragma unroll
    for (int i = 0; i < NLB; ++i) {
      const int r = lane + 64 * i;
      d.lab[i] = (r < TR && k0 + r < ke) ? y[k0 + r] : 0;
    }
  };
  auto stash = [&](const TSet& d) {
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      const int t4 = lane + 64 * i, cc = t4 / QPC, rq = t4 % QPC;
      tile[(4 * rq + 0) * LDT + cc] = d.v[i].x;
      tile[(4 * rq + 1) * LDT + cc] = d.v[i].y;
      tile[(4 * rq + 2) * LDT + cc] = d.v[i].z;
      tile[(4 * rq + 3) * LDT + cc] = d.v[i].w;
    }
#pragma unroll
    for (int i = 0; i < NLB; ++i) {
      const int r = lane + 64 * i;
      if (r < TR) labs[r] = d.lab[i];
    }
  };
  float ones = 0.f;
  // branch-free: the tile's values and labels are read in one batch, rows past
  // the slab and labels outside 1..C go to the spare bin C
  auto accumulate = [&](int64_t k0) {
    const int nr = (int)std::min<int64_t>(TR, ke - k0);
    float uv[RPL];
    int lv[RPL];
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int r = ph + PH * q;
      uv[q] = tile[r * LDT + c];
      lv[q] = labs[r];
    }
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const bool ok = ph + PH * q < nr;
      const float u = uv[q] - s;
      ones += ok ? u : 0.f;
      const int lb = lv[q];
      const int bin = (ok && lb >= 1 && lb <= C) ? lb - 1 : C;
      atomicAdd(&acc[bin * COLS + c], u);
    }
  };
  const int nt = ke > kb ? (int)((ke - kb + TR - 1) / TR) : 0;
  TSet T0, T1, T2;
  auto step = [&](int i, TSet& cur, TSet& far) {
    const int64_t k0 = kb + (int64_t)i * TR;
    if (i + 2 < nt) fetch(k0 + 2 * TR, far);
    stash(cur);
    wave_fence();
    accumulate(k0);
    wave_fence();
  };
  if (nt > 0) fetch(kb, T0);
  if (nt > 1) fetch(kb + TR, T1);
  for (int i = 0; i < nt; i += 3) {
    step(i, T0, T2);
    if (i + 1 < nt) step(i + 1, T1, T0);
    if (i + 2 < nt) step(i + 2, T2, T1);
  }
  // the ones column: the phases of each column summed in phase order
  tile[ph * LDT + c] = ones;
  wave_fence();
  float* P = part + (int64_t)blockIdx.y * p * (C + 1);
  if (ph == 0 && jok) {
    float t = 0.f;
    for (int q = 0; q < PH; ++q) t += tile[q * LDT + c];
    P[j + (int64_t)C * p] = t;
  }
  for (int i = lane; i < C * COLS; i += 64) {
    const int cls = i / COLS, cc = i % COLS;
    if (j0 + cc < p) P[(j0 + cc) + (int64_t)cls * p] = acc[i];
  }
}

// S(m, c) = sum_z P_z(m, c) - shB[c] sum_z P_z(m, C) (shB optional), FP64 in
// slab order; threads past p * C finish the column statistics: mean[m] =
// s[m] + sum_z P_z(m, C) / n and ones[m] = 1.
__global__ __launch_bounds__(256) void k_cs_reduce(int64_t p, int C, int slabs,
                                                   const float* __restrict__ part,
                                                   const double* __restrict__ shB,
                                                   double* __restrict__ S, int64_t lds,
                                                   const float* __restrict__ shift, int64_t n,
                                                   double* __restrict__ mean,
                                                   double* __restrict__ ones) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = p * (C + 1);
  // eight slab sums in flight (slab z into sum z % 8), combined in order
  auto slab_sum = [&](int64_t off) {
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int z = 0;
    for (; z + 8 <= slabs; z += 8) {
#pragma unroll
      for (int q = 0; q < 8; ++q) a[q] += (double)part[(int64_t)(z + q) * stride + off];
    }
    for (int q = 0; z + q < slabs; ++q) a[q] += (double)part[(int64_t)(z + q) * stride + off];
    return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  };
  if (t < p * C) {
    const int64_t m = t % p, cls = t / p;
    const double a = slab_sum(m + cls * p);
    S[m + cls * lds] = shB ? a - shB[cls] * slab_sum(m + (int64_t)C * p) : a;
  } else if (t < p * C + p) {
    const int64_t m = t - p * C;
    if (mean) mean[m] = (shift ? (double)shift[m] : 0.0) + slab_sum(m + (int64_t)C * p) / (double)n;
    if (ones) ones[m] = 1.0;
  }
}

template <int COLS>
void launch_cs(Ctx* ctx, int64_t n, int64_t p, const float* X, int64_t ld, const int32_t* y,
               int C, bool use_shift, const double* shB, double* S, int64_t lds, double* mean,
               double* ones) {
  const int64_t cb = (p + 4 * COLS - 1) / (4 * COLS);
  constexpr int TR = 2048 / COLS;
  int64_t slabs = std::max<int64_t>(1, (ctx->num_cus + cb - 1) / cb);
  slabs = std::min<int64_t>(slabs, (n + TR - 1) / TR);
  int64_t kchunk = (n + slabs - 1) / slabs;
  kchunk = (kchunk + TR - 1) / TR * TR;
  slabs = (n + kchunk - 1) / kchunk;
  DBuf<float> part(ctx, (size_t)slabs * p * (C + 1));
  DBuf<float> sh(ctx, use_shift ? (size_t)p : 1);
  launch(ctx, "class_sums", k_class_sums<COLS>, dim3((unsigned)cb, (unsigned)slabs), dim3(256), 0,
         n, p, X, ld, y, C, use_shift ? 1 : 0, kchunk, part.get(), sh.get());
  const int64_t outs = p * C + p;
  launch(ctx, "class_sums_reduce", k_cs_reduce, dim3((unsigned)((outs + 255) / 256)), dim3(256), 0,
         p, C, (int)slabs, (const float*)part.get(), shB, S, lds,
         use_shift ? (const float*)sh.get() : (const float*)nullptr, n, mean, ones);
}

}  // namespace

bool class_sums_f32(Ctx* ctx, int64_t n, int64_t p, const float* X, int64_t ld, const int32_t* y,
                    int C, const double* shiftB, double* S, int64_t lds, double* mean,
                    double* ones) {
  if (n <= 0 || p <= 0 || C < 1) return false;
  // the centred product is shift-invariant and takes the shift; raw class
  // sums are taken unshifted
  const bool use_shift = shiftB != nullptr;
  const int64_t bins = (int64_t)C + 1;
  if (bins * 64 <= CS_ACC) {
    launch_cs<64>(ctx, n, p, X, ld, y, C, use_shift, shiftB, S, lds, mean, ones);
  } else if (bins * 32 <= CS_ACC) {
    launch_cs<32>(ctx, n, p, X, ld, y, C, use_shift, shiftB, S, lds, mean, ones);
  } else if (bins * 16 <= CS_ACC) {
    launch_cs<16>(ctx, n, p, X, ld, y, C, use_shift, shiftB, S, lds, mean, ones);
  } else if (bins * 8 <= CS_ACC) {
    launch_cs<8>(ctx, n, p, X, ld, y, C, use_shift, shiftB, S, lds, mean, ones);
  } else if (bins * 4 <= CS_ACC) {
    launch_cs<4>(ctx, n, p, X, ld, y, C, use_shift, shiftB, S, lds, mean, ones);
  } else {
    return false;
  }
  return true;
}

}  // namespace fpls
