// Candidate kNN classifier head (the reference's cknn classifier):
//   fit      .fit_candidate_knn            R/main.R:795-873 (score space :700-730,
//            row_l2 :657-662, centroids :686-698)
//   predict  candidate_knn_predict_native_impl  src/fastPLS.cpp:1322-1467
//            (candidate_insert_top :1249-1267, smoothed class kNN :1287-1320)
//
// Layout: training scores in the score space are row-normalised, stored
// row-major (k contiguous) and sorted by class (stable: ascending row id inside
// a class), so every class is one contiguous block of rows.  Prediction:
//   base = Ttest_n C^T (MFMA), top-m candidate classes per row (wave argmax,
//   ties -> lower class), (row, candidate) pairs bucketed by class, then one
//   work item per (class, 128 pairs, 4096-row chunk) computes the 128 x chunk
//   similarity block as an MFMA GEMM and keeps running top-k lists; partial lists are
//   merged per pair, the temperature-smoothed score is formed in FP64 and the
//   candidates are re-ranked with the reference's strict '>' insertion.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <set>
#include <vector>

#include "model.h"

namespace fpls {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int CK_PT = 128;      // pairs per work item (CKG_P)
constexpr int CK_CHUNK = 4096;  // training rows per work item
constexpr int CK_KNN_MAX = 32;  // largest knn_k (same clamp as the CUDA kernel of the reference)
constexpr int SORT_CHUNK = 4096;

static unsigned g1(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

// ---------------------------------------------------------------- stable class sort
__global__ __launch_bounds__(256) void k_chunk_counts(const int32_t* __restrict__ y, int64_t n,
                                                      int C, int* __restrict__ counts) {
  extern __shared__ int hist[];
  const int64_t b = blockIdx.x;
  for (int c = threadIdx.x; c < C; c += blockDim.x) hist[c] = 0;
  __syncthreads();
  const int64_t i0 = b * SORT_CHUNK, i1 = std::min<int64_t>(n, i0 + SORT_CHUNK);
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) atomicAdd(&hist[y[i] - 1], 1);
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) counts[b * C + c] = hist[c];
}

// per class: totals over chunks
__global__ void k_class_totals(const int* __restrict__ counts, int B, int C, int* __restrict__ tot) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  int s = 0;
  for (int b = 0; b < B; ++b) s += counts[(int64_t)b * C + c];
  tot[c] = s;
}

// per class: running base per chunk (in place over counts)
__global__ void k_chunk_base(int* __restrict__ counts, int B, int C, const int* __restrict__ off) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  int run = off[c];
  for (int b = 0; b < B; ++b) {
    const int v = counts[(int64_t)b * C + c];
    counts[(int64_t)b * C + c] = run;
    run += v;
  }
}

// one wavefront per chunk: stable ranks of equal labels inside each batch of 64
__global__ __launch_bounds__(64) void k_chunk_scatter(const int32_t* __restrict__ y, int64_t n,
                                                      int C, const int* __restrict__ base,
                                                      int* __restrict__ rows) {
  extern __shared__ int sb[];
  const int64_t b = blockIdx.x;
  const int lane = threadIdx.x;
  for (int c = lane; c < C; c += 64) sb[c] = base[b * C + c];
  __syncthreads();
  const int64_t i0 = b * SORT_CHUNK, i1 = std::min<int64_t>(n, i0 + SORT_CHUNK);
  for (int64_t s = i0; s < i1; s += 64) {
    const int64_t i = s + lane;
    const int lab = (i < i1) ? y[i] : -1;
    int rank = 0, tot = 0;
    for (int j = 0; j < 64; ++j) {
      const int lj = __shfl(lab, j, 64);
      if (lj == lab) {
        ++tot;
        if (j < lane) ++rank;
      }
    }
    if (lab > 0) {
      rows[sb[lab - 1] + rank] = (int)i;
    }
    __syncthreads();
    if (lab > 0 && rank == tot - 1) sb[lab - 1] += tot;
    __syncthreads();
  }
}

// ---------------------------------------------------------------- row normalisation
// Tn[s, :] = S[rows[s], :] / |S[rows[s], :]|, row-major (ld k); zero or
// non-finite norms map to 1 (R/main.R:657-662).  One wavefront per row.
__global__ __launch_bounds__(256) void k_rownorm_permute(const float* __restrict__ S, int64_t n,
                                                         int k, const int* __restrict__ rows,
                                                         float* __restrict__ Tn) {
  const int lane = threadIdx.x & 63;
  const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= n) return;
  const int64_t r = rows ? rows[s] : s;
  const float* src = S + r * k;
  double acc = 0.0;
  for (int j = lane; j < k; j += 64) acc = fma((double)src[j], (double)src[j], acc);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  double nr = sqrt(acc);
  if (!isfinite(nr) || nr == 0.0) nr = 1.0;
  float* dst = Tn + s * k;
  for (int j = lane; j < k; j += 64) dst[j] = (float)((double)src[j] / nr);
}

// centroid partial sums: one block per (class, chunk) work item
__global__ __launch_bounds__(256) void k_class_partial_sums(const float* __restrict__ Tn, int k,
                                                            const int4* __restrict__ items,
                                                            double* __restrict__ part) {
  const int4 it = items[blockIdx.x];  // (class, row begin, row end, unused)
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    double s = 0.0;
    for (int r = it.y; r < it.z; ++r) s += (double)Tn[(int64_t)r * k + j];
    part[(int64_t)blockIdx.x * k + j] = s;
  }
}

// centroid[c] = row_l2(mean of the class rows); empty classes use count 1
__global__ __launch_bounds__(256) void k_centroids(const double* __restrict__ part, int k,
                                                   const int* __restrict__ item_begin,
                                                   const int* __restrict__ off,
                                                   double* __restrict__ cen,
                                                   float* __restrict__ cen_f) {
  __shared__ double red[4];
  __shared__ double nrm;
  const int c = blockIdx.x;
  const int cnt = off[c + 1] - off[c];
  const double denom = cnt > 0 ? (double)cnt : 1.0;
  double ss = 0.0;
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    double s = 0.0;
    for (int it = item_begin[c]; it < item_begin[c + 1]; ++it) s += part[(int64_t)it * k + j];
    s /= denom;
    cen[(int64_t)c * k + j] = s;
    ss += s * s;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) ss += __shfl_down(ss, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) {
    double nr = sqrt(red[0] + red[1] + red[2] + red[3]);
    if (!isfinite(nr) || nr == 0.0) nr = 1.0;
    nrm = nr;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    const double v = cen[(int64_t)c * k + j] / nrm;
    cen[(int64_t)c * k + j] = v;
    cen_f[(int64_t)c * k + j] = (float)v;
  }
}

// ---------------------------------------------------------------- top-m candidates
// One wavefront per test row: m rounds of a wave argmax with the ordering
// (value descending, class ascending), i.e. repeated strict-'>' insertion in
// ascending class order.  base: w x C row-major.  With boundary set, one more
// round gives the best class left out: boundary[i] = the m-th candidate's base
// score minus that class's (+inf when no class is left out), the tie-log entry
// of the candidate set.
__global__ __launch_bounds__(256) void k_topm(const float* __restrict__ base, int64_t w, int C,
                                              int m, int32_t* __restrict__ cand,
                                              double* __restrict__ cand_base,
                                              double* __restrict__ boundary) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= w) return;
  const float* row = base + i * C;
  unsigned long long taken = 0ull;  // bit b: class lane + 64 b already selected
  float last = -INFINITY;           // base score of the latest selected candidate
  const int rounds = boundary ? m + 1 : m;
  for (int r = 0; r < rounds; ++r) {
    float bv = -INFINITY;
    int bc = -1;
    for (int b = 0, c = lane; c < C; ++b, c += 64) {
      if (taken & (1ull << b)) continue;
      const float v = row[c];
      if (!(v > -INFINITY)) continue;  // NaN / -inf never pass the strict '>' insertion
      if (bc < 0 || v > bv) {
        bv = v;
        bc = c;
      }
    }
    // wave reduction: larger value wins, equal values -> lower class
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float ov = __shfl_xor(bv, o, 64);
      const int oc = __shfl_xor(bc, o, 64);
      const bool take = (oc >= 0) && (bc < 0 || ov > bv || (ov == bv && oc < bc));
      if (take) {
        bv = ov;
        bc = oc;
      }
    }
    if (r == m) {  // the boundary round (boundary is set)
      if (lane == 0)
        boundary[i] = (bc >= 0 && last > -INFINITY) ? (double)last - (double)bv : INFINITY;
      break;
    }
    if (bc >= 0 && (bc % 64) == lane) taken |= 1ull << (bc / 64);
    last = bc >= 0 ? bv : -INFINITY;
    if (lane == 0) {
      cand[i * m + r] = bc + 1;
      cand_base[i * m + r] = (double)bv;
    }
  }
}

// ---------------------------------------------------------------- pair buckets
__global__ void k_pair_hist(const int32_t* __restrict__ cand, int64_t npairs, int* __restrict__ cnt) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < npairs;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int c = cand[t];
    if (c >= 1) atomicAdd(&cnt[c - 1], 1);
  }
}

__global__ void k_pair_scatter(const int32_t* __restrict__ cand, int64_t npairs,
                               int* __restrict__ cursor, int* __restrict__ bucket) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < npairs;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int c = cand[t];
    if (c >= 1) bucket[atomicAdd(&cursor[c - 1], 1)] = (int)t;
  }
}

// ---------------------------------------------------------------- candidate kNN partials
struct CkItem {
  int cls;        // 0-based class
  int pair_beg;   // first bucket slot of the item's pairs
  int pair_end;
  int row_beg;    // sorted training rows of this chunk
  int row_end;
  int chunk;      // chunk index inside the class
  int nchunks;
  int part_base;  // float offset of this class's partial lists
};

// GEMM form of the same partial lists: work item = class c x up to CKG_P
// pairs (test rows whose candidates include c) x one chunk of c's training
// rows.  Sim = A B^T for A the pairs' normalised test rows and B a tile of
// CKG_R normalised training rows, K streamed through LDS in CKG_KS-deep
// steps (double buffered, one barrier per step).  4 waves, each a 64 x 64
// block of Sim: 4 x 4 v_mfma_f32_16x16x4f32 accumulators; within a step lane
// l takes k = 4 (l / 16) + s in sub-step s (any bijection of k shared by both
// operands gives the same sum), so each operand fragment is one 16-byte LDS
// read.  After each tile Sim goes through LDS (aliasing the operand buffers)
// and thread t keeps a register top-knn_k list of pair t / 2 over half the
// tile's columns; lists are written in the layout k_cand_merge reads.
constexpr int CKG_P = CK_PT, CKG_R = 128, CKG_KS = 16, CKG_LD = CKG_KS + 4;
constexpr int CKG_OPS = 2 * (CKG_P + CKG_R) * CKG_LD;  // floats, both buffers
constexpr int CKG_SIM = CKG_P * (CKG_R + 1);
constexpr int CKG_SMEM = CKG_OPS > CKG_SIM ? CKG_OPS : CKG_SIM;

__global__ __launch_bounds__(256) void k_cand_gemm(const float* __restrict__ Ttest, int k,
                                                   const float* __restrict__ Tn,
                                                   const int* __restrict__ bucket, int m,
                                                   const CkItem* __restrict__ items,
                                                   int knn_k, float* __restrict__ part) {
  __shared__ __align__(16) float smem[CKG_SMEM];
  const CkItem it = items[blockIdx.x];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int np = it.pair_end - it.pair_beg;
  const int ph = w & 1, th = w >> 1;  // 64-row block of pairs, 64-column block of rows
  // staging: thread -> operand row t / 2, 8 consecutive k at 8 (t % 2)
  const int srow = t >> 1, sk = (t & 1) * 8;
  const float* arow = nullptr;
  if (srow < np) arow = Ttest + (int64_t)(bucket[it.pair_beg + srow] / m) * k;
  const bool vec = (k & 3) == 0 && ((reinterpret_cast<uintptr_t>(Ttest) |
                                     reinterpret_cast<uintptr_t>(Tn)) & 15) == 0;
  float qa[8], qb[8];
  auto load8 = [&](const float* src, int k0, float (&q)[8]) {
    const int j0 = k0 + sk;
    if (src && vec && j0 + 8 <= k) {
      const float4 u = *reinterpret_cast<const float4*>(src + j0);
      const float4 v = *reinterpret_cast<const float4*>(src + j0 + 4);
      q[0] = u.x; q[1] = u.y; q[2] = u.z; q[3] = u.w;
      q[4] = v.x; q[5] = v.y; q[6] = v.z; q[7] = v.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) q[e] = (src && j0 + e < k) ? src[j0 + e] : 0.f;
    }
  };
  auto store8 = [&](float* dst, const float (&q)[8]) {
    *reinterpret_cast<float4*>(dst) = make_float4(q[0], q[1], q[2], q[3]);
    *reinterpret_cast<float4*>(dst + 4) = make_float4(q[4], q[5], q[6], q[7]);
  };
  auto opA = [&](int buf) { return smem + buf * (CKG_P + CKG_R) * CKG_LD; };
  auto opB = [&](int buf) { return smem + buf * (CKG_P + CKG_R) * CKG_LD + CKG_P * CKG_LD; };
  // list entries are only ever indexed by compile-time constants (a runtime
  // index would put the array in scratch memory); thr = L[knn_k - 1]
  float L[CK_KNN_MAX];
#pragma unroll
  for (int j = 0; j < CK_KNN_MAX; ++j) L[j] = -INFINITY;
  float thr = -INFINITY;
  const int nks = (k + CKG_KS - 1) / CKG_KS;
  for (int r0 = it.row_beg; r0 < it.row_end; r0 += CKG_R) {
    const float* brow = (r0 + srow < it.row_end) ? Tn + (int64_t)(r0 + srow) * k : nullptr;
    v4f acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    load8(arow, 0, qa);
    load8(brow, 0, qb);
    __syncthreads();  // the previous tile's Sim reads are done
    store8(opA(0) + srow * CKG_LD + sk, qa);
    store8(opB(0) + srow * CKG_LD + sk, qb);
    __syncthreads();
    const int q4 = 4 * (lane >> 4), li = lane & 15;
    for (int st = 0; st < nks; ++st) {
      const bool more = st + 1 < nks;
      if (more) {
        load8(arow, (st + 1) * CKG_KS, qa);
        load8(brow, (st + 1) * CKG_KS, qb);
      }
      const float* A = opA(st & 1);
      const float* B = opB(st & 1);
      float4 af[4], bf[4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
        af[a] = *reinterpret_cast<const float4*>(A + (ph * 64 + a * 16 + li) * CKG_LD + q4);
#pragma unroll
      for (int b = 0; b < 4; ++b)
        bf[b] = *reinterpret_cast<const float4*>(B + (th * 64 + b * 16 + li) * CKG_LD + q4);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a].x, bf[b].x, acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a].y, bf[b].y, acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a].z, bf[b].z, acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a].w, bf[b].w, acc[a][b], 0, 0, 0);
        }
      if (more) {
        store8(opA((st + 1) & 1) + srow * CKG_LD + sk, qa);
        store8(opB((st + 1) & 1) + srow * CKG_LD + sk, qb);
      }
      __syncthreads();
    }
    // Sim tile through LDS: lane holds D[4 (l / 16) + r][l % 16] of each block
    float* Sim = smem;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          Sim[(ph * 64 + a * 16 + q4 + r) * (CKG_R + 1) + th * 64 + b * 16 + li] = acc[a][b][r];
    __syncthreads();
    if (srow < np) {
      const int cend = min(CKG_R, it.row_end - r0);
      const float* srow_p = Sim + srow * (CKG_R + 1);
      for (int c = (t & 1) * 64; c < (t & 1) * 64 + 64 && c < cend; ++c) {
        float v = srow_p[c];
        if (v > thr) {  // a value the list keeps: bubble it into place
#pragma unroll
          for (int j = 0; j < CK_KNN_MAX; ++j) {
            if (j < knn_k && v > L[j]) {
              const float o = L[j];
              L[j] = v;
              v = o;
            }
          }
#pragma unroll
          for (int j = 0; j < CK_KNN_MAX; ++j)
            if (j == knn_k - 1) thr = L[j];
        }
      }
    }
  }
  if (srow < np) {
    const int slot = it.pair_beg + srow;
    float* dst = part + (int64_t)it.part_base +
                 ((int64_t)slot * it.nchunks + it.chunk) * 2 * knn_k + (t & 1) * knn_k;
#pragma unroll
    for (int j = 0; j < CK_KNN_MAX; ++j)
      if (j < knn_k) dst[j] = L[j];
  }
}

// Split-FP16 form of k_cand_gemm.  Both operands are unit-normalised rows
// (|x| <= 1), so x 2^12 splits exactly enough into an FP16 pair
// x 2^12 = h + l + r with h = fp16(x 2^12), l = fp16(x 2^12 - h),
// |r| <= 2^-22 |x 2^12| (h, l normal for every entry above 2^-13), and
// x . y = 2^-24 (h_x h_y + h_x l_y + l_x h_y) up to 3 2^-22 sum |x_i y_i|:
// FP32-level similarities from three FP16 MFMA products (16x16x32) at the
// FP16 matrix-core rate.  Operands arrive pre-split, row-major with ld kp (a
// multiple of 32, zero padded); tiles and lists as in k_cand_gemm, K streamed
// in 32-deep steps (hi and lo of both operands, double buffered).
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
constexpr int CKH_KS = 32;                      // halves per K step
constexpr int CKH_LD = CKH_KS + 8;              // LDS row stride (80 bytes)
constexpr int CKH_ROWS = CKG_P + CKG_R;         // operand rows per buffer (A then B)
constexpr int CKH_BUF = 2 * CKH_ROWS * CKH_LD;  // halves per buffer (hi, lo)
constexpr int CKH_OPS = CKH_BUF * 2;            // bytes, one buffer
constexpr int CKH_SIMH = (CKG_P / 2) * (CKG_R + 1) * 4;  // bytes, half a Sim tile
constexpr int CKH_SMEM = CKH_OPS > CKH_SIMH ? CKH_OPS : CKH_SIMH;
constexpr float CKH_SCALE = 4096.0f;

__global__ void k_split_f16(const float* __restrict__ X, int64_t rows, int k, int kp,
                            _Float16* __restrict__ Xh, _Float16* __restrict__ Xl) {
  const int cpr = kp / 8;
  const int64_t nchunk = rows * cpr;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nchunk;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = q / cpr;
    const int c0 = (int)(q % cpr) * 8;
    h8v hv, lv;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int c = c0 + e;
      const float x = c < k ? X[r * k + c] * CKH_SCALE : 0.0f;
      const _Float16 h = (_Float16)x;
      hv[e] = h;
      lv[e] = (_Float16)(x - (float)h);
    }
    *reinterpret_cast<h8v*>(Xh + r * kp + c0) = hv;
    *reinterpret_cast<h8v*>(Xl + r * kp + c0) = lv;
  }
}

template <int KM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_cand_gemm_h(
    const _Float16* __restrict__ Tth, const _Float16* __restrict__ Ttl, int kp,
    const _Float16* __restrict__ Tnh, const _Float16* __restrict__ Tnl,
    const int* __restrict__ bucket, int m, const CkItem* __restrict__ items, int knn_k,
    float* __restrict__ part) {
  __shared__ __align__(16) unsigned char smem_raw[CKH_SMEM];
  _Float16* hs = reinterpret_cast<_Float16*>(smem_raw);
  const CkItem it = items[blockIdx.x];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int np = it.pair_end - it.pair_beg;
  const int ph = w & 1, th = w >> 1;
  // staging: thread -> operand row t / 2, 16 consecutive k at 16 (t % 2)
  const int srow = t >> 1, sk = (t & 1) * 16;
  // rows past the item's pairs (training rows) stage a valid row whose
  // similarities are never read, so the staging loads need no branches
  const int64_t tr = srow < np ? bucket[it.pair_beg + srow] / m : 0;
  const _Float16* ah = Tth + tr * kp;
  const _Float16* al = Ttl + tr * kp;
  // hi at [buf][0][row][k], lo at [buf][1][row][k]; A rows 0..127, B rows 128..255
  auto lds = [&](int lo, int row) { return hs + lo * CKH_ROWS * CKH_LD + row * CKH_LD; };
  float L[KM];
#pragma unroll
  for (int j = 0; j < KM; ++j) L[j] = -INFINITY;
  float thr = -INFINITY;
  const int nks = kp / CKH_KS;
  for (int r0 = it.row_beg; r0 < it.row_end; r0 += CKG_R) {
    const int64_t brow = r0 + srow < it.row_end ? r0 + srow : r0;
    const _Float16* bh = Tnh + brow * kp;
    const _Float16* bl = Tnl + brow * kp;
    v4f acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    // staged 16-byte pieces: A hi, A lo, B hi, B lo (two each)
    uint4 q0, q1, q2, q3, q4, q5, q6, q7;
#define CKH_LOAD(k0)                                                                   \
  do {                                                                                 \
    const int o_ = (k0) + sk;                                                          \
    q0 = *reinterpret_cast<const uint4*>(ah + o_);                                 \
    q1 = *reinterpret_cast<const uint4*>(ah + o_ + 8);                         \
    q2 = *reinterpret_cast<const uint4*>(al + o_);                                 \
    q3 = *reinterpret_cast<const uint4*>(al + o_ + 8);                         \
    q4 = *reinterpret_cast<const uint4*>(bh + o_);                                 \
    q5 = *reinterpret_cast<const uint4*>(bh + o_ + 8);                         \
    q6 = *reinterpret_cast<const uint4*>(bl + o_);                                 \
    q7 = *reinterpret_cast<const uint4*>(bl + o_ + 8);                         \
  } while (0)
#define CKH_STORE()                                                                    \
  do {                                                                                 \
    *reinterpret_cast<uint4*>(lds(0, srow) + sk) = q0;                          \
    *reinterpret_cast<uint4*>(lds(0, srow) + sk + 8) = q1;                      \
    *reinterpret_cast<uint4*>(lds(1, srow) + sk) = q2;                          \
    *reinterpret_cast<uint4*>(lds(1, srow) + sk + 8) = q3;                      \
    *reinterpret_cast<uint4*>(lds(0, CKG_P + srow) + sk) = q4;                  \
    *reinterpret_cast<uint4*>(lds(0, CKG_P + srow) + sk + 8) = q5;              \
    *reinterpret_cast<uint4*>(lds(1, CKG_P + srow) + sk) = q6;                  \
    *reinterpret_cast<uint4*>(lds(1, CKG_P + srow) + sk + 8) = q7;              \
  } while (0)
    CKH_LOAD(0);
    __syncthreads();  // the previous tile's Sim reads are done
    CKH_STORE();
    __syncthreads();
    const int g8 = 8 * (lane >> 4), li = lane & 15;
    for (int st = 0; st < nks; ++st) {
      const bool more = st + 1 < nks;
      if (more) CKH_LOAD((st + 1) * CKH_KS);
      h8v ahf[4], alf[4], bhf[4], blf[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int row = ph * 64 + a * 16 + li;
        ahf[a] = *reinterpret_cast<const h8v*>(lds(0, row) + g8);
        alf[a] = *reinterpret_cast<const h8v*>(lds(1, row) + g8);
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int row = CKG_P + th * 64 + b * 16 + li;
        bhf[b] = *reinterpret_cast<const h8v*>(lds(0, row) + g8);
        blf[b] = *reinterpret_cast<const h8v*>(lds(1, row) + g8);
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahf[a], blf[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(alf[a], bhf[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahf[a], bhf[b], acc[a][b], 0, 0, 0);
        }
      __syncthreads();  // every wave is done reading this step's operands
      if (more) {
        CKH_STORE();
        __syncthreads();
      }
    }
#undef CKH_LOAD
#undef CKH_STORE
    // Sim through LDS (aliasing the operands) in two halves of 64 pairs, so
    // the workgroup needs half the LDS: lane holds D[4 (l / 16) + r][l % 16]
    // of each block, scaled by 2^24
    float* Sim = reinterpret_cast<float*>(smem_raw);
    const int r4 = 4 * (lane >> 4);
    constexpr float unscale = 1.0f / (CKH_SCALE * CKH_SCALE);
    for (int h = 0; h < 2; ++h) {
      if (ph == h) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              Sim[(a * 16 + r4 + r) * (CKG_R + 1) + th * 64 + b * 16 + li] =
                  acc[a][b][r] * unscale;
      }
      __syncthreads();
      if (srow < np && (srow >> 6) == h) {
        const int cend = min(CKG_R, it.row_end - r0);
        const float* srow_p = Sim + (srow - 64 * h) * (CKG_R + 1);
        for (int c = (t & 1) * 64; c < (t & 1) * 64 + 64 && c < cend; ++c) {
          float v = srow_p[c];
          if (v > thr) {
#pragma unroll
            for (int j = 0; j < KM; ++j) {
              if (j < knn_k && v > L[j]) {
                const float o = L[j];
                L[j] = v;
                v = o;
              }
            }
#pragma unroll
            for (int j = 0; j < KM; ++j)
              if (j == knn_k - 1) thr = L[j];
          }
        }
      }
      __syncthreads();  // the half's reads are done before the next write
    }
  }
  if (srow < np) {
    const int slot = it.pair_beg + srow;
    float* dst = part + (int64_t)it.part_base +
                 ((int64_t)slot * it.nchunks + it.chunk) * 2 * knn_k + (t & 1) * knn_k;
#pragma unroll
    for (int j = 0; j < KM; ++j)
      if (j < knn_k) dst[j] = L[j];
  }
}

// merge partial lists of one pair and form the smoothed score (FP64)
__global__ void k_cand_merge(const int* __restrict__ bucket, int nslots,
                             const int* __restrict__ slot_class, const int* __restrict__ class_nchunks,
                             const int* __restrict__ class_part_base, const int* __restrict__ class_slot0,
                             const int* __restrict__ class_size, const float* __restrict__ part,
                             int knn_k, double tau, double alpha, const double* __restrict__ bias,
                             const double* __restrict__ cand_base, double* __restrict__ score) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  const int c = slot_class[slot];
  const int pair = bucket[slot];
  const int nsz = class_size[c];
  double local;
  if (nsz <= 0) {
    local = -INFINITY;
  } else {
    const int use_k = knn_k < nsz ? knn_k : nsz;
    const int nl = class_nchunks[c] * 2;
    const float* L = part + (int64_t)class_part_base[c] +
                     (int64_t)(slot - class_slot0[c]) * nl * knn_k;
    int head[2 * 64];  // list heads (nchunks <= 64 enforced on the host)
    for (int q = 0; q < nl; ++q) head[q] = 0;
    double vals[CK_KNN_MAX];
    for (int j = 0; j < use_k; ++j) {
      int bq = -1;
      float bv = -INFINITY;
      for (int q = 0; q < nl; ++q) {
        if (head[q] < knn_k) {
          const float v = L[q * knn_k + head[q]];
          if (bq < 0 || v > bv) {
            bv = v;
            bq = q;
          }
        }
      }
      vals[j] = (double)bv;
      if (bq >= 0) ++head[bq];
    }
    const double mx = vals[0];
    if (!isfinite(mx)) {
      local = mx;
    } else if (!(tau > 0.0)) {
      double s = 0.0;
      for (int j = 0; j < use_k; ++j) s += vals[j];
      local = s / use_k;
    } else {
      double acc = 0.0;
      for (int j = 0; j < use_k; ++j) acc += exp((vals[j] - mx) / tau);
      local = mx + tau * log(acc / (double)use_k);
    }
  }
  score[pair] = local + alpha * cand_base[pair] + (bias ? bias[c] : 0.0);
}

// rerank: strict '>' insertion of the m candidates in candidate order
// (src/fastPLS.cpp:1428-1457); one thread per test row.
// With margin set, margin[i] = rank-1 score - rank-2 score (+inf with a single
// ranked class): the tie-log entry of the winning class.
__global__ void k_rerank(const int32_t* __restrict__ cand, const double* __restrict__ score,
                         int64_t w, int m, int top, int C, int32_t* __restrict__ top_index,
                         double* __restrict__ top_score, int64_t ld_out,
                         double* __restrict__ margin) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= w) return;
  double sv[64];
  int si[64];
  for (int j = 0; j < m; ++j) {
    sv[j] = -INFINITY;
    si[j] = -1;
  }
  for (int j = 0; j < m; ++j) {
    const int cls = cand[i * m + j];
    if (cls < 1 || cls > C) continue;
    const double v = score[i * m + j];
    for (int q = 0; q < m; ++q) {
      if (v > sv[q]) {
        for (int h = m - 1; h > q; --h) {
          sv[h] = sv[h - 1];
          si[h] = si[h - 1];
        }
        sv[q] = v;
        si[q] = cls;
        break;
      }
    }
  }
  for (int j = 0; j < top; ++j) {
    top_index[i + j * ld_out] = si[j];
    if (top_score) top_score[i + j * ld_out] = sv[j];
  }
  if (margin) margin[i] = (m > 1 && sv[1] > -INFINITY) ? sv[0] - sv[1] : INFINITY;
}

// ---------------------------------------------------------------- host core
void stable_class_sort(Ctx* ctx, const int32_t* y, int64_t n, int C, CkClassIndex& ci) {
  const int B = (int)((n + SORT_CHUNK - 1) / SORT_CHUNK);
  FPLS_REQUIRE(C <= 16384, FPLS_ERR_ARG, "candidate kNN supports up to 16384 classes");
  {
    DBuf<int64_t> cnt(ctx, C);
    DBuf<int32_t> bad(ctx, 1);
    label_counts(ctx, y, n, C, cnt.get(), bad.get());
    int32_t hbad = 0;
    to_host_i32(ctx, &hbad, bad.get(), 1);
    FPLS_REQUIRE(hbad == 0, FPLS_ERR_LABEL_RANGE, "class codes must be in 1..n_classes");
  }
  DBuf<int> counts(ctx, (size_t)std::max(B, 1) * C), tot(ctx, C);
  launch(ctx, "class_sort", k_chunk_counts, dim3(B), dim3(256), C * sizeof(int), y, n, C,
         counts.get());
  launch(ctx, "class_sort", k_class_totals, dim3((C + 255) / 256), dim3(256), 0, counts.get(), B,
         C, tot.get());
  std::vector<int> ht(C);
  FPLS_HIP(hipMemcpyAsync(ht.data(), tot.get(), C * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  sync(ctx);
  ci.off.assign(C + 1, 0);
  for (int c = 0; c < C; ++c) ci.off[c + 1] = ci.off[c] + ht[c];
  FPLS_REQUIRE(ci.off[C] == n, FPLS_ERR_LABEL_RANGE, "labels must be encoded as 1..n_classes");
  ci.d_off = DBuf<int>(ctx, C + 1);
  h2d(ctx, ci.d_off.get(), ci.off.data(), (C + 1) * sizeof(int));
  launch(ctx, "class_sort", k_chunk_base, dim3((C + 255) / 256), dim3(256), 0, counts.get(), B, C,
         ci.d_off.get());
  ci.d_rows = DBuf<int>(ctx, std::max<int64_t>(n, 1));
  launch(ctx, "class_sort", k_chunk_scatter, dim3(B), dim3(64), C * sizeof(int), y, n, C,
         counts.get(), ci.d_rows.get());
}

// centroids (C x k row-major FP64 + FP32) of the class-sorted normalised rows
void class_centroids(Ctx* ctx, const float* Tn, int k, const CkClassIndex& ci, int C,
                     DBuf<double>& cen, DBuf<float>& cen_f) {
  std::vector<int4> items;
  std::vector<int> ib(C + 1, 0);
  for (int c = 0; c < C; ++c) {
    ib[c] = (int)items.size();
    for (int r = ci.off[c]; r < ci.off[c + 1]; r += CK_CHUNK)
      items.push_back(make_int4(c, r, std::min(ci.off[c + 1], r + CK_CHUNK), 0));
  }
  ib[C] = (int)items.size();
  DBuf<int4> d_items(ctx, std::max<size_t>(items.size(), 1));
  DBuf<int> d_ib(ctx, C + 1);
  if (!items.empty())
    h2d(ctx, d_items.get(), items.data(), items.size() * sizeof(int4));
  h2d(ctx, d_ib.get(), ib.data(), (C + 1) * sizeof(int));
  DBuf<double> part(ctx, std::max<size_t>(items.size(), 1) * k);
  if (!items.empty())
    launch(ctx, "centroid_sums", k_class_partial_sums, dim3((unsigned)items.size()), dim3(256), 0,
           Tn, k, (const int4*)d_items.get(), part.get());
  cen = DBuf<double>(ctx, (size_t)C * k);
  cen_f = DBuf<float>(ctx, (size_t)C * k);
  launch(ctx, "centroids", k_centroids, dim3(C), dim3(256), 0, (const double*)part.get(), k,
         (const int*)d_ib.get(), (const int*)ci.d_off.get(), cen.get(), cen_f.get());
  sync(ctx);  // host vectors above must outlive the async copies
}

// Predict for one score space: Ttest_n (w x k row-major, normalised), Tn
// (class-sorted, normalised), centroids (C x k row-major FP32).
void cknn_predict_core(Ctx* ctx, int precision, const float* Ttn, int64_t w, int k, const float* Tn,
                       const std::vector<int>& off, const float* cen_f, int C, const double* bias,
                       int top, int top_m, int knn_k, double tau, double alpha,
                       int32_t* top_index, double* top_score, int64_t ld_out,
                       double* margins) {
  FPLS_REQUIRE(k >= 1 && k <= 1024, FPLS_ERR_ARG, "candidate kNN supports up to 1024 components");
  FPLS_REQUIRE(C <= 4096, FPLS_ERR_ARG, "candidate selection supports up to 4096 classes");
  const int top_k = std::max(1, std::min(top, C));
  const int m = std::max(top_k, std::min(std::max(1, top_m), C));
  FPLS_REQUIRE(m <= 64, FPLS_ERR_ARG, "candidate_top_m must be <= 64");
  knn_k = std::max(1, std::min(knn_k, CK_KNN_MAX));
  if (!std::isfinite(tau) || tau <= 0.0) tau = 0.2;
  if (!std::isfinite(alpha)) alpha = 0.5;
  // base scores, row-major w x C: (centroids x Ttn^T) as C x w column-major
  DBuf<float> base(ctx, (size_t)w * C);
  gemm_rr_f32(ctx, precision, C, w, k, cen_f, k, Ttn, k, base.get());
  const int64_t npairs = w * m;
  DBuf<int32_t> cand(ctx, npairs);
  DBuf<double> cbase(ctx, npairs), score(ctx, npairs);
  launch(ctx, "cknn_topm", k_topm, dim3((unsigned)((w + 3) / 4)), dim3(256), 0,
         (const float*)base.get(), w, C, m, cand.get(), cbase.get(),
         margins ? margins + w : (double*)nullptr);
  // bucket the pairs by class
  DBuf<int> cnt(ctx, C), cursor(ctx, C), bucket(ctx, npairs);
  fill_i32(ctx, cnt.get(), 0, C);
  launch(ctx, "cknn_pairs", k_pair_hist, dim3(g1(npairs)), dim3(256), 0,
         (const int32_t*)cand.get(), npairs, cnt.get());
  std::vector<int> hc(C);
  FPLS_HIP(hipMemcpyAsync(hc.data(), cnt.get(), C * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  sync(ctx);
  std::vector<int> slot0(C + 1, 0), nch(C, 0), pbase(C, 0), csize(C, 0);
  std::vector<CkItem> items;
  int64_t ptotal = 0;
  for (int c = 0; c < C; ++c) {
    slot0[c + 1] = slot0[c] + hc[c];
    csize[c] = off[c + 1] - off[c];
    nch[c] = (csize[c] + CK_CHUNK - 1) / CK_CHUNK;
    FPLS_REQUIRE(nch[c] <= 64, FPLS_ERR_ARG, "a class exceeds 262144 training rows");
    pbase[c] = (int)ptotal;
    ptotal += (int64_t)hc[c] * nch[c] * 2 * knn_k;
    // chunk-major order: the pair groups of one chunk run together, so the
    // chunk's training rows are fetched from HBM once and then served by the
    // caches to the other groups
    for (int ch = 0; ch < nch[c]; ++ch)
      for (int pb = 0; pb < hc[c]; pb += CK_PT) {
        CkItem it;
        it.cls = c;
        it.pair_beg = slot0[c] + pb;
        it.pair_end = slot0[c] + std::min(hc[c], pb + CK_PT);
        it.row_beg = off[c] + ch * CK_CHUNK;
        it.row_end = std::min(off[c + 1], it.row_beg + CK_CHUNK);
        it.chunk = ch;
        it.nchunks = nch[c];
        it.part_base = pbase[c] - slot0[c] * nch[c] * 2 * knn_k;
        items.push_back(it);
      }
  }
  FPLS_REQUIRE(ptotal < (int64_t(1) << 31), FPLS_ERR_ARG, "candidate kNN partial buffer too large");
  h2d(ctx, cursor.get(), slot0.data(), C * sizeof(int));
  launch(ctx, "cknn_pairs", k_pair_scatter, dim3(g1(npairs)), dim3(256), 0,
         (const int32_t*)cand.get(), npairs, cursor.get(), bucket.get());
  std::vector<int> slot_class(std::max<int64_t>(npairs, 1), 0);
  for (int c = 0; c < C; ++c)
    for (int s = slot0[c]; s < slot0[c + 1]; ++s) slot_class[s] = c;
  const int nslots = slot0[C];
  DBuf<int> d_slot_class(ctx, slot_class.size()), d_nch(ctx, C), d_pbase(ctx, C), d_slot0(ctx, C),
      d_csize(ctx, C);
  DBuf<CkItem> d_items(ctx, std::max<size_t>(items.size(), 1));
  DBuf<float> part(ctx, std::max<int64_t>(ptotal, 1));
  h2d(ctx, d_slot_class.get(), slot_class.data(), slot_class.size() * sizeof(int));
  h2d(ctx, d_nch.get(), nch.data(), C * sizeof(int));
  h2d(ctx, d_pbase.get(), pbase.data(), C * sizeof(int));
  h2d(ctx, d_slot0.get(), slot0.data(), C * sizeof(int));
  h2d(ctx, d_csize.get(), csize.data(), C * sizeof(int));
  if (!items.empty()) {
    h2d(ctx, d_items.get(), items.data(), items.size() * sizeof(CkItem));
    if (precision == FPLS_PREC_BF16X3) {
      // split form: FP16 pairs of both operands, three FP16 MFMA products
      const int kp = (k + CKH_KS - 1) / CKH_KS * CKH_KS;
      const int64_t ntr = off[C];
      DBuf<_Float16> th(ctx, (size_t)w * kp), tl(ctx, (size_t)w * kp);
      DBuf<_Float16> nh(ctx, (size_t)ntr * kp), nl(ctx, (size_t)ntr * kp);
      launch(ctx, "cknn_split", k_split_f16, dim3(g1(w * (kp / 8))), dim3(256), 0, Ttn, w, k, kp,
             th.get(), tl.get());
      launch(ctx, "cknn_split", k_split_f16, dim3(g1(ntr * (kp / 8))), dim3(256), 0, Tn, ntr, k,
             kp, nh.get(), nl.get());
      // list capacity: 16 registers for knn_k <= 16, else 32
      auto kern = knn_k <= 16 ? k_cand_gemm_h<16> : k_cand_gemm_h<CK_KNN_MAX>;
      launch(ctx, "cknn_partial", kern, dim3((unsigned)items.size()), dim3(256), 0,
             (const _Float16*)th.get(), (const _Float16*)tl.get(), kp, (const _Float16*)nh.get(),
             (const _Float16*)nl.get(), (const int*)bucket.get(), m,
             (const CkItem*)d_items.get(), knn_k, part.get());
    } else {
      launch(ctx, "cknn_partial", k_cand_gemm, dim3((unsigned)items.size()), dim3(256), 0, Ttn,
             k, Tn, (const int*)bucket.get(), m, (const CkItem*)d_items.get(), knn_k,
             part.get());
    }
  }
  if (nslots > 0)
    launch(ctx, "cknn_merge", k_cand_merge, dim3((nslots + 255) / 256), dim3(256), 0,
           (const int*)bucket.get(), nslots, (const int*)d_slot_class.get(),
           (const int*)d_nch.get(), (const int*)d_pbase.get(), (const int*)d_slot0.get(),
           (const int*)d_csize.get(), (const float*)part.get(), knn_k, tau, alpha, bias,
           (const double*)cbase.get(), score.get());
  launch(ctx, "cknn_rerank", k_rerank, dim3((unsigned)((w + 255) / 256)), dim3(256), 0,
         (const int32_t*)cand.get(), (const double*)score.get(), w, m, top_k, C, top_index,
         top_score, ld_out, margins);
  sync(ctx);  // host vectors above must outlive the async copies
}

void rownorm_rows(Ctx* ctx, const float* S, int64_t n, int k, const int* rows, float* Tn) {
  if (n > 0)
    launch(ctx, "cknn_rownorm", k_rownorm_permute, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, S, n,
           k, rows, Tn);
}

// Score space of a model for k components (R/main.R:700-730): PLSSVD scores are
// multiplied by the C_latent slice (exact ncomp match, else the first >= k,
// else the last).  Output row-major n x k FP32.
void score_space(Ctx* ctx, const Model* M, const DMat& T, int k, float* out) {
  DBuf<float> Tf;
  const float* A = nullptr;
  int64_t lda = T.ld;
  if (T.dtype == FPLS_F32) {
    A = T.f();
  } else {
    Tf = DBuf<float>(ctx, (size_t)T.rows * k);
    for (int j = 0; j < k; ++j) f64_to_f32(ctx, Tf.get() + (int64_t)j * T.rows, T.d() + j * T.ld, T.rows);
    A = Tf.get();
    lda = T.rows;
  }
  const double* Ck = nullptr;
  if (M->method == FPLS_METHOD_PLSSVD && M->C_latent.get()) {
    const int L = (int)M->ncomp.size();
    int sl = -1;
    for (int a = 0; a < L; ++a)
      if (M->ncomp[a] == k) { sl = a; break; }
    if (sl < 0)
      for (int a = 0; a < L; ++a)
        if (M->ncomp[a] >= k) { sl = a; break; }
    if (sl < 0) sl = L - 1;
    Ck = M->C_latent.get() + (size_t)sl * M->kmax * M->kmax;
  }
  DBuf<float> ident;
  DBuf<float> Cf(ctx, (size_t)k * k);
  if (Ck) {
    DBuf<double> Cd(ctx, (size_t)k * k);
    copy2d_f64(ctx, Cd.get(), k, Ck, M->kmax, k, k);
    f64_to_f32(ctx, Cf.get(), Cd.get(), (int64_t)k * k);
  } else {
    std::vector<float> I((size_t)k * k, 0.f);
    for (int j = 0; j < k; ++j) I[j + (size_t)j * k] = 1.f;
    h2d(ctx, Cf.get(), I.data(), I.size() * sizeof(float));
    sync(ctx);
  }
  // out = T_k C_k in row-major layout
  gemm_nn_f32(ctx, FPLS_PREC_FP32, T.rows, k, k, A, lda, nullptr, Cf.get(), k, nullptr, out,
              nullptr, k, true);
}

// The candidate-kNN head for every k of ks on training scores T (rows x >=
// max(ks)) with checked labels y (device, 1..C): the class-sorted,
// row-normalised score space per k (R/main.R:845-848, 911-914) and the class
// centroids -- computed from it (R/main.R:686-698), or, when cen_host is set,
// taken from it (C x k column-major per k, packed in the order of ks).
static void cknn_attach(Ctx* c, Model* M, const DMat& T, const int32_t* y, int C,
                        const std::vector<int>& ks, const double* cen_host) {
  const int64_t n = T.rows;
  CkClassIndex ci;
  stable_class_sort(c, y, n, C, ci);
  std::vector<CknnSlice> slices;
  size_t oc = 0;
  for (int k : ks) {
    CknnSlice sl;
    sl.k = k;
    DBuf<float> S(c, (size_t)n * k);
    score_space(c, M, T, k, S.get());
    sl.Tn = DBuf<float>(c, (size_t)n * k);
    rownorm_rows(c, S.get(), n, k, ci.d_rows.get(), sl.Tn.get());
    if (cen_host) {
      // column-major C x k -> the row-major layout of the device head
      std::vector<double> rm((size_t)C * k);
      for (int j = 0; j < k; ++j)
        for (int r = 0; r < C; ++r) rm[(size_t)r * k + j] = cen_host[oc + (size_t)j * C + r];
      sl.centroids = DBuf<double>(c, rm.size());
      to_device(c, sl.centroids.get(), rm.data(), rm.size());
      sl.centroids_f = DBuf<float>(c, rm.size());
      f64_to_f32(c, sl.centroids_f.get(), sl.centroids.get(), (int64_t)rm.size());
      oc += rm.size();
    } else {
      CkClassIndex cix;
      cix.off = ci.off;
      cix.d_off = DBuf<int>(c, C + 1);
      h2d(c, cix.d_off.get(), ci.off.data(), (C + 1) * sizeof(int));
      class_centroids(c, sl.Tn.get(), k, cix, C, sl.centroids, sl.centroids_f);
    }
    slices.push_back(std::move(sl));
  }
  sync(c);
  M->cknn = std::move(slices);
  M->class_offsets_host.assign(ci.off.begin(), ci.off.end());
  M->class_offsets = std::move(ci.d_off);
  M->class_rows = std::move(ci.d_rows);
  M->n_classes = C;
}

}  // namespace fpls

using namespace fpls;

extern "C" int fpls_cknn_fit(fpls_ctx* ctx, fpls_model* model, const fpls_mat* Xm,
                             const fpls_ivec* yv, int32_t n_classes, int32_t knn_k, double tau,
                             double alpha, int32_t top_m) {
  if (!model) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    Model* M = &model->m;
    FPLS_REQUIRE(std::isfinite(tau) && tau > 0.0, FPLS_ERR_ARG, "candidate_tau must be positive");
    FPLS_REQUIRE(std::isfinite(alpha), FPLS_ERR_ARG, "candidate_alpha must be finite");
    FPLS_REQUIRE(n_classes >= 1, FPLS_ERR_ARG, "candidate kNN requires at least one class");
    DBuf<int32_t> hl;
    const int32_t* y = stage_labels(c, yv, M->n, hl, "cknn: y");
    DBuf<int64_t> cnt(c, n_classes);
    DBuf<int32_t> bad(c, 1);
    label_counts(c, y, M->n, n_classes, cnt.get(), bad.get());
    int32_t hbad = 0;
    to_host_i32(c, &hbad, bad.get(), 1);
    FPLS_REQUIRE(hbad == 0, FPLS_ERR_LABEL_RANGE,
                 "candidate-kNN received labels outside the training levels");
    DBuf<unsigned char> hx;
    DMat T;
    DBuf<float> Tproj;
    std::set<int> ks;
    for (int v : M->ncomp) ks.insert(std::max(1, std::min(v, M->kmax)));
    const int kmax = *ks.rbegin();
    if (Xm) {
      DMat X = stage(c, Xm, hx, "cknn: Xtrain");
      FPLS_REQUIRE(X.rows == M->n, FPLS_ERR_ARG, "cknn: one label per training row");
      Tproj = DBuf<float>(c, (size_t)X.rows * kmax);
      if (X.dtype == FPLS_F32) {
        project_scores(c, M, X, kmax, Tproj.get(), nullptr, X.rows);
      } else {
        DBuf<double> Td(c, (size_t)X.rows * kmax);
        project_scores(c, M, X, kmax, nullptr, Td.get(), X.rows);
        f64_to_f32(c, Tproj.get(), Td.get(), X.rows * (int64_t)kmax);
      }
      T.p = Tproj.get();
      T.dtype = FPLS_F32;
      T.rows = X.rows;
      T.cols = kmax;
      T.ld = X.rows;
    } else {
      FPLS_REQUIRE(M->has_T(), FPLS_ERR_STATE, "model holds no training scores; pass Xtrain");
      T = M->T();
    }
    FPLS_REQUIRE(M->lda.empty() || M->n_classes == n_classes, FPLS_ERR_ARG,
                 "the LDA and candidate-kNN heads disagree on the number of classes");
    cknn_attach(c, M, T, y, n_classes, std::vector<int>(ks.begin(), ks.end()), nullptr);
    M->knn_k = std::max(1, (int)knn_k);
    M->tau = tau;
    M->alpha = alpha;
    M->top_m = std::max(1, (int)top_m);
    sync(c);
  });
}

static int cknn_predict_entry(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xm,
                              int32_t top, int32_t* top_index, double* top_score,
                              double* margins_out) {
  if (!model) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    const Model* M = &model->m;
    FPLS_REQUIRE(!M->cknn.empty(), FPLS_ERR_STATE,
                 "this model does not contain fitted candidate-kNN parameters");
    DBuf<unsigned char> hx;
    DMat X = stage(c, Xm, hx, "cknn_predict: X");
    FPLS_REQUIRE(X.cols == M->p, FPLS_ERR_ARG, "cknn_predict: X has the wrong number of columns");
    const int64_t w = X.rows;
    const int C = M->n_classes;
    const int kmax = M->cknn.back().k;
    DMat T;
    DBuf<float> Tf;
    DBuf<double> Td;
    T.rows = w;
    T.cols = kmax;
    T.ld = w;
    if (X.dtype == FPLS_F32) {
      Tf = DBuf<float>(c, (size_t)w * kmax);
      project_scores(c, M, X, kmax, Tf.get(), nullptr, w);
      T.p = Tf.get();
      T.dtype = FPLS_F32;
    } else {
      Td = DBuf<double>(c, (size_t)w * kmax);
      project_scores(c, M, X, kmax, nullptr, Td.get(), w);
      T.p = Td.get();
      T.dtype = FPLS_F64;
    }
    const int topc = std::max(1, std::min((int)top, C));
    const int top_m_eff = std::min(std::max(topc, M->top_m), C);  // R/main.R:746-747
    const int L = (int)M->ncomp.size();
    DBuf<int32_t> ti(c, (size_t)w * topc * L);
    DBuf<double> ts(c, (size_t)w * topc * L);
    DBuf<double> mg;
    if (margins_out) mg = DBuf<double>(c, (size_t)w * 2 * L);
    for (int a = 0; a < L; ++a) {
      const int kk = std::max(1, std::min(M->ncomp[a], kmax));
      const CknnSlice* sl = nullptr;
      for (const auto& s : M->cknn)
        if (s.k == kk) sl = &s;
      FPLS_REQUIRE(sl != nullptr, FPLS_ERR_STATE, "no fitted candidate-kNN classifier for ncomp");
      DBuf<float> S(c, (size_t)w * kk), Ttn(c, (size_t)w * kk);
      score_space(c, M, T, kk, S.get());
      launch(c, "cknn_rownorm", k_rownorm_permute, dim3((unsigned)((w + 3) / 4)), dim3(256), 0,
             (const float*)S.get(), w, kk, (const int*)nullptr, Ttn.get());
      cknn_predict_core(c, M->precision, Ttn.get(), w, kk, sl->Tn.get(), M->class_offsets_host,
                        sl->centroids_f.get(), C, nullptr, topc, top_m_eff, M->knn_k, M->tau,
                        M->alpha, ti.get() + (size_t)a * w * topc, ts.get() + (size_t)a * w * topc,
                        w, margins_out ? mg.get() + (size_t)a * w * 2 : nullptr);
    }
    to_host_i32(c, top_index, ti.get(), (size_t)w * topc * L);
    if (margins_out) to_host(c, margins_out, mg.get(), (size_t)w * 2 * L);
    if (top_score) to_host(c, top_score, ts.get(), (size_t)w * topc * L);
  });
}

extern "C" int fpls_cknn_predict(fpls_ctx* ctx, const fpls_model* model, const fpls_mat* Xm,
                                 int32_t top, int32_t* top_index, double* top_score) {
  return cknn_predict_entry(ctx, model, Xm, top, top_index, top_score, nullptr);
}

extern "C" int fpls_cknn_predict_margins(fpls_ctx* ctx, const fpls_model* model,
                                         const fpls_mat* Xm, int32_t top, int32_t* top_index,
                                         double* top_score, double* margins) {
  if (!margins) return FPLS_ERR_ARG;
  return cknn_predict_entry(ctx, model, Xm, top, top_index, top_score, margins);
}

static int cknn_scores_entry(fpls_ctx* ctx, const fpls_mat* Ttest_m, const fpls_mat* Ttrain_m,
                             const fpls_ivec* yv, const fpls_mat* cen_m,
                             const double* candidate_bias, int32_t top_k, int32_t top_m,
                             int32_t knn_k, double tau, double alpha, int32_t* top_index,
                             double* top_score, double* margins_out) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    DBuf<unsigned char> h1, h2, h3;
    DMat Te = stage(c, Ttest_m, h1, "cknn: Ttest_norm");
    DMat Tr = stage(c, Ttrain_m, h2, "cknn: Ttrain_norm");
    DMat Ce = stage(c, cen_m, h3, "cknn: centroids");
    FPLS_REQUIRE(Te.cols == Tr.cols && Te.cols == Ce.cols, FPLS_ERR_ARG,
                 "candidate-kNN score dimensions do not match");
    const int k = (int)Te.cols, C = (int)Ce.rows;
    FPLS_REQUIRE(C >= 1, FPLS_ERR_ARG, "candidate-kNN requires at least one class");
    DBuf<int32_t> hl;
    const int32_t* y = stage_labels(c, yv, Tr.rows, hl, "cknn: y_codes");
    // row-major FP32 copies (the inputs are column-major)
    auto to_rowmajor = [&](const DMat& A, DBuf<float>& out) {
      out = DBuf<float>(c, (size_t)A.rows * A.cols);
      DBuf<double> A64(c, (size_t)A.rows * A.cols), At(c, (size_t)A.rows * A.cols);
      if (A.dtype == FPLS_F64) copy2d_f64(c, A64.get(), A.rows, A.d(), A.ld, A.rows, A.cols);
      else
        for (int64_t j = 0; j < A.cols; ++j)
          f32_to_f64(c, A64.get() + j * A.rows, A.f() + j * A.ld, A.rows);
      transpose_f64(c, A64.get(), A.rows, A.rows, A.cols, At.get(), A.cols);
      f64_to_f32(c, out.get(), At.get(), A.rows * A.cols);
    };
    DBuf<float> te, tr, ce;
    to_rowmajor(Te, te);
    to_rowmajor(Tr, tr);
    to_rowmajor(Ce, ce);
    CkClassIndex ci;
    // labels outside 1..C are skipped by the reference (src/fastPLS.cpp:1379-1384); the device
    // sort requires them to be valid
    stable_class_sort(c, y, Tr.rows, C, ci);
    DBuf<float> sorted(c, (size_t)Tr.rows * k);
    launch(c, "cknn_rownorm", k_rownorm_permute, dim3((unsigned)((Tr.rows + 3) / 4)), dim3(256), 0,
           (const float*)tr.get(), Tr.rows, k, (const int*)ci.d_rows.get(), sorted.get());
    // the inputs are normalised already; re-normalising is idempotent up to rounding
    DBuf<double> bias;
    if (candidate_bias) {
      bias = DBuf<double>(c, C);
      to_device(c, bias.get(), candidate_bias, C);
    }
    const int topc = std::max(1, std::min((int)top_k, C));
    DBuf<int32_t> ti(c, (size_t)Te.rows * topc);
    DBuf<double> ts(c, (size_t)Te.rows * topc);
    DBuf<double> mg;
    if (margins_out) mg = DBuf<double>(c, (size_t)Te.rows * 2);
    cknn_predict_core(c, FPLS_PREC_FP32, te.get(), Te.rows, k, sorted.get(), ci.off, ce.get(), C,
                      bias.get(), topc, top_m, knn_k, tau, alpha, ti.get(), ts.get(), Te.rows,
                      mg.get());
    to_host_i32(c, top_index, ti.get(), (size_t)Te.rows * topc);
    if (top_score) to_host(c, top_score, ts.get(), (size_t)Te.rows * topc);
    if (margins_out) to_host(c, margins_out, mg.get(), (size_t)Te.rows * 2);
  });
}

extern "C" int fpls_cknn_predict_scores(fpls_ctx* ctx, const fpls_mat* Ttest_m,
                                        const fpls_mat* Ttrain_m, const fpls_ivec* yv,
                                        const fpls_mat* cen_m, const double* candidate_bias,
                                        int32_t top_k, int32_t top_m, int32_t knn_k, double tau,
                                        double alpha, int32_t* top_index, double* top_score) {
  return cknn_scores_entry(ctx, Ttest_m, Ttrain_m, yv, cen_m, candidate_bias, top_k, top_m,
                           knn_k, tau, alpha, top_index, top_score, nullptr);
}

extern "C" int fpls_cknn_predict_scores_margins(
    fpls_ctx* ctx, const fpls_mat* Ttest_m, const fpls_mat* Ttrain_m, const fpls_ivec* yv,
    const fpls_mat* cen_m, const double* candidate_bias, int32_t top_k, int32_t top_m,
    int32_t knn_k, double tau, double alpha, int32_t* top_index, double* top_score,
    double* margins) {
  if (!margins) return FPLS_ERR_ARG;
  return cknn_scores_entry(ctx, Ttest_m, Ttrain_m, yv, cen_m, candidate_bias, top_k, top_m,
                           knn_k, tau, alpha, top_index, top_score, margins);
}

// The reference's model$candidate_knn list (R/main.R:861-872) as the device
// candidate-kNN head: Ttrain, y_codes, the per-model centroids and the
// parameters.  The score spaces are formed from Ttrain with the device model's
// C_latent, as .candidate_knn_predictions does at every call (R/main.R:903-918).
extern "C" int fpls_model_import_cknn(fpls_ctx* ctx, fpls_model* model,
                                      const fpls_cknn_fields* f) {
  if (!model || !f) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    Model* M = &model->m;
    const int C = f->n_classes, L = f->n_models;
    const int64_t n = f->n_train;
    FPLS_REQUIRE(C >= 1, FPLS_ERR_ARG, "candidate kNN requires at least one class");
    FPLS_REQUIRE(L >= 1 && f->ncomp, FPLS_ERR_ARG,
                 "candidate-kNN import requires at least one model");
    FPLS_REQUIRE(f->centroids && f->Ttrain && f->y_codes && n >= 1, FPLS_ERR_ARG,
                 "candidate-kNN import requires centroids, Ttrain and y_codes");
    FPLS_REQUIRE(std::isfinite(f->tau) && f->tau > 0.0, FPLS_ERR_ARG,
                 "candidate_tau must be positive");
    FPLS_REQUIRE(std::isfinite(f->alpha), FPLS_ERR_ARG, "candidate_alpha must be finite");
    FPLS_REQUIRE(M->lda.empty() || M->n_classes == C, FPLS_ERR_ARG,
                 "the LDA and candidate-kNN heads disagree on the number of classes");
    std::vector<int> ks(f->ncomp, f->ncomp + L);
    for (int a = 0; a < L; ++a) {
      FPLS_REQUIRE(ks[a] >= 1 && ks[a] <= M->kmax, FPLS_ERR_ARG,
                   "candidate-kNN score dimension does not match the model");
      FPLS_REQUIRE(a == 0 || ks[a] > ks[a - 1], FPLS_ERR_ARG,
                   "candidate-kNN ncomp must be strictly ascending");
    }
    const int kmax = ks.back();
    FPLS_REQUIRE(f->ttrain_cols >= kmax, FPLS_ERR_ARG,
                 "candidate-kNN Ttrain has fewer columns than max(ncomp)");
    // labels: the range check of .fit_candidate_knn (R/main.R:840-843)
    DBuf<int32_t> y(c, n);
    FPLS_HIP(hipMemcpyAsync(y.get(), f->y_codes, sizeof(int32_t) * n, hipMemcpyHostToDevice,
                            c->stream));
    DBuf<int64_t> cnt(c, C);
    DBuf<int32_t> bad(c, 1);
    label_counts(c, y.get(), n, C, cnt.get(), bad.get());
    int32_t hbad = 0;
    to_host_i32(c, &hbad, bad.get(), 1);
    FPLS_REQUIRE(hbad == 0, FPLS_ERR_LABEL_RANGE,
                 "candidate-kNN received labels outside the training levels");
    // Ttrain[:, :kmax] in FP32, the precision of the device head's scores
    DBuf<double> T64(c, (size_t)n * kmax);
    to_device(c, T64.get(), f->Ttrain, (size_t)n * kmax);
    DBuf<float> T32(c, (size_t)n * kmax);
    f64_to_f32(c, T32.get(), T64.get(), n * (int64_t)kmax);
    DMat T;
    T.p = T32.get();
    T.dtype = FPLS_F32;
    T.rows = n;
    T.cols = kmax;
    T.ld = n;
    cknn_attach(c, M, T, y.get(), C, ks, f->centroids);
    M->knn_k = std::max(1, (int)f->knn_k);
    M->tau = f->tau;
    M->alpha = f->alpha;
    M->top_m = std::max(1, (int)f->top_m);
  });
}
