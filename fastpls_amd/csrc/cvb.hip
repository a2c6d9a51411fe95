// Batched cross-validated SIMPLS: F fits over one resident X, advanced
// together component by component, replayed as one captured HIP graph.
//
// The reference's CV driver (pls_cv_predict_compiled, src/fastPLS.cpp:6615-7151)
// fits every fold on its own: it gathers the fold's rows into a fresh copy,
// fits pls_model2_fast (src/fastPLS.cpp:2454-2778) with seed + f and predicts
// the held-out rows for every requested ncomp; pls.double.cv
// (R/main.R:6586-6894) repeats that for every outer fold.  Each SIMPLS
// component streams the fold's X twice (t = Xc r, p = Xc^T t), so F folds
// stream X 2 F times per component.
//
// Here a fit is a row mask over the one resident X (0 unused, 1 train,
// 2 test) and the F fits share every pass over X:
//   T = (X - 1 g^T) [r_1 .. r_F]          one skinny NN product (N = F)
//   P = (X - 1 g^T)^T [w_1 t_1 .. w_F t_F] one TN product (N = F)
// with g the column means over all rows.  Per-fit centring is a rank-1
// correction: with d_f = mX_f - g,
//   Xc_f r = (X - 1 g^T) r - 1 (d_f . r),  Xc_f^T t = (X - g)^T (w t) - d_f (1^T w t)
// and the fold's cross-covariance is
//   S_f = (X - g)^T (w_f o Y) - n_f d_f mY_f^T
// from one TN product against the masked responses of all fits.  Everything
// per fit (power steps on S_f, norms, Gram-Schmidt against V_f, deflation,
// prediction of the held-out rows) is FP64 on p- or n-long vectors, spread
// over row chunks x fits so the whole chip works on it.  Held-out predictions
// accumulate t q^T per component (Ypred = mY + sum_a t_a q_a^T, the compact
// prediction of src/fastPLS.cpp:4115-4168 without forming B), and every
// requested ncomp is scored when its component is done.
//
// Fits stop individually where the reference's append_component stops
// (|t| or |v| non-finite or zero, src/fastPLS.cpp:2580-2620): later
// components add nothing, as the reference's zero columns of R and Q do.
// Fold status follows the sequential driver: 1 ok, 2 no test rows, 3 no
// training rows (predictions = observed responses), 4 fewer than two classes
// among the training rows (predictions = that class).
//
// Gaussian starts are drawn on the device (k_rng.hip), one stream per fit.
// The launch sequence depends only on the shapes, so it is captured once into
// a hipGraph and replayed; per-call inputs (masks, labels, seeds, counts) are
// copied into the plan's buffers before the replay.  The split-K counts of the
// two big products are chosen by timing candidates on the first call.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <vector>

#include "fpls_wave.h"
#include "model.h"

namespace fpls {

constexpr int CV_NT = 256;                 // threads of the chunked kernels
constexpr int CV_CH = 1024;                // rows of p per chunk (4 per thread)
constexpr int CV_MAX_M = 512;              // responses / classes per fit
constexpr int CV_MAX_K = 512;              // components
constexpr int CV_MAX_F = 128;              // fits per batch (skinny NN limit)

struct CvArgs {
  int64_t n, p;
  int F, m, K, L, nch;
  int64_t Kp;                              // padded rows of RR32
  int classification, m1, power, scaling3;
  int nsplit;                              // split-K slabs of the NN product
  const float* X;
  int64_t ldx;
  const float* Yf;
  const double* Yd;
  int64_t ldy;
  const int8_t* mask;                      // n x F
  const int32_t* y;                        // n labels (classification)
  const int32_t* stat;                     // F
  const double* nf;                        // F train counts
  double* mY;                              // F x m train means of Y
  double* gy;                              // m overall means of Y (regression)
  const double* C0;                        // p x (F m [+ F])
  double* S;                               // p x F m
  double* d;                               // p x F
  const double* omega;                     // p x F Gaussian starts
  double* ybuf;                            // p x F
  double* rr;                              // p x F
  float* RR32;                             // Kp x 128 (zero padded)
  double* zpart;                           // 2 x F x nch x m
  double* npart;                           // F x nch
  double* cpart;                           // F x nch
  const float* projpart;                   // nsplit x n x F
  double* T64;                             // n x F (t before division by |t|)
  int bf;                                  // split-BF16x3 operands for the TN product
  int64_t npad;                            // rows of the TN operand (n rounded up to 32)
  float* Tw32;                             // npad x F: masked t (FP32 mode)
  __bf16* Twh;                             // npad x F: masked t, bf16 hi (split mode)
  __bf16* Twl;                             // npad x F: masked t, bf16 lo
  int nblk;                                // row blocks of k_cvb_treduce
  double* tpart;                           // F x nblk x (m + 2) partial sums
  double* q;                               // F x m
  double* stv;                             // F
  double* tnv;                             // F
  int32_t* act;                            // F
  const double* P;                         // p x F
  double* vbuf;                            // p x F
  double* vpart;                           // F x nch x K
  double* v2part;                          // F x nch
  double* vspart;                          // F x nch x m
  double* V;                               // p x K x F
  double* Yacc;                            // n x m x F
  int32_t* cls;                            // n x L x F
  int32_t* ncorrect;                       // L x F
  double* ypred;                           // n x m x L x F (optional)
};

// Fixed-order sum over the block (any multiple of 64 threads up to 1024);
// every thread gets the result.
__device__ inline double cv_block_sum(double v, double* red) {
  v = wave_sum_d(v);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < nw; ++q) s += red[q];
  return s;
}

// out[c * ostride] = sum_{r < rows} M[r + c ldm] vec[r], c < ncols: wave w
// takes columns w, w + 4, ... and its lanes stride the (contiguous) rows.
__device__ inline void cv_chunk_colsums(const double* M, int64_t ldm, int rows, int ncols,
                                        const double* vec, double* out, int64_t ostride) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int c = w; c < ncols; c += CV_NT / 64) {
    const double* col = M + (int64_t)c * ldm;
    double s = 0.0;
    for (int r = lane; r < rows; r += 64) s = fma(col[r], vec[r], s);
    s = wave_sum_d(s);
    if (lane == 0) out[(int64_t)c * ostride] = s;
  }
}

__device__ inline double cv_yval(const CvArgs& A, int64_t i, int c) {
  return A.Yd ? A.Yd[i + (int64_t)c * A.ldy] : (double)A.Yf[i + (int64_t)c * A.ldy];
}

// ---------------------------------------------------------------- setup
__global__ void k_cvb_init(CvArgs A) {
  const int64_t nt = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t i = t0; i < A.F; i += nt) A.act[i] = A.stat[i] == 1 ? 1 : 0;
  for (int64_t i = t0; i < (int64_t)A.n * A.m * A.F; i += nt) A.Yacc[i] = 0.0;
  for (int64_t i = t0; i < (int64_t)A.n * A.L * A.F; i += nt) A.cls[i] = 0;
  for (int64_t i = t0; i < (int64_t)A.L * A.F; i += nt) A.ncorrect[i] = 0;
  if (A.ypred)
    for (int64_t i = t0; i < (int64_t)A.n * A.m * A.L * A.F; i += nt) A.ypred[i] = 0.0;
}

// g = column sums / n (FP64 and FP32 copies); zero for scaling "none".
__global__ void k_cvb_gscale(int64_t p, double inv_n, int zero, double* g, float* g32) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < p;
       j += (int64_t)gridDim.x * blockDim.x) {
    const double v = zero ? 0.0 : g[j] * inv_n;
    g[j] = v;
    g32[j] = (float)v;
  }
}

// Regression means: block f < F -> train mean of fit f, block F -> all rows.
__global__ __launch_bounds__(CV_NT) void k_cvb_ymeans(CvArgs A) {
  __shared__ double red[CV_NT / 64];
  const int f = blockIdx.x;
  for (int c = 0; c < A.m; ++c) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < A.n; i += CV_NT)
      if (f == A.F || A.mask[i + (int64_t)f * A.n] == 1) s += cv_yval(A, i, c);
    s = cv_block_sum(s, red);
    if (threadIdx.x == 0) {
      if (f == A.F) A.gy[c] = s / (double)A.n;
      else A.mY[(int64_t)f * A.m + c] = A.nf[f] > 0.0 ? s / A.nf[f] : 0.0;
    }
  }
}

// Masked responses of all fits: column f m + c is w_f o onehot_c (labels) or
// w_f o (Y_c - gy_c); regression adds the F columns w_f.
__global__ void k_cvb_zbuild(CvArgs A, float* Z, int64_t ncols) {
  const int64_t tot = A.n * ncols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % A.n, col = e / A.n;
    float z = 0.f;
    if (col < (int64_t)A.F * A.m) {
      const int f = (int)(col / A.m), c = (int)(col % A.m);
      if (A.mask[i + (int64_t)f * A.n] == 1) {
        if (A.classification) z = A.y[i] == c + 1 ? 1.f : 0.f;
        else z = (float)(cv_yval(A, i, c) - A.gy[c]);
      }
    } else {
      const int f = (int)(col - (int64_t)A.F * A.m);
      z = A.mask[i + (int64_t)f * A.n] == 1 ? 1.f : 0.f;
    }
    Z[e] = z;
  }
}

// S_f and d_f from C0 = (X - g)^T Z (grid: chunks x fits).
__global__ __launch_bounds__(CV_NT) void k_cvb_sinit(CvArgs A) {
  const int f = blockIdx.y;
  const int64_t j0 = (int64_t)blockIdx.x * CV_CH;
  const int64_t p = A.p, m = A.m;
  const double nf = A.nf[f];
  const bool ok = A.stat[f] == 1 && nf > 0.0;
  for (int64_t j = j0 + threadIdx.x; j < std::min<int64_t>(p, j0 + CV_CH); j += CV_NT) {
    double w;  // (X - g)^T w_f at row j
    if (A.classification) {
      w = 0.0;
      for (int c = 0; c < m; ++c) w += A.C0[j + ((int64_t)f * m + c) * p];
    } else {
      w = A.C0[j + ((int64_t)A.F * m + f) * p];
    }
    for (int c = 0; c < m; ++c) {
      const int64_t col = (int64_t)f * m + c;
      const double mu = A.classification ? A.mY[col] : A.mY[col] - A.gy[c];
      A.S[j + col * p] = ok ? A.C0[j + col * p] - w * mu : 0.0;
    }
    A.d[j + (int64_t)f * p] = (ok && !A.scaling3) ? w / nf : 0.0;
  }
}

// ---------------------------------------------------------------- refresh
// Deflation by the previous component (S -= v (v^T S)), the start vector
// (Gaussian stream, previous r, or for one response the deflated S itself)
// and the partial sums of the first power step (or of |y|^2 when there is
// no power step).
__global__ __launch_bounds__(CV_NT) void k_cvb_start(CvArgs A, int a) {
  __shared__ double vec[CV_CH];
  __shared__ double vs[CV_MAX_M];
  __shared__ double red[CV_NT / 64];
  const int f = blockIdx.y, ch = blockIdx.x, t = threadIdx.x;
  if (!A.act[f]) return;
  const int64_t p = A.p, j0 = (int64_t)ch * CV_CH;
  const int rows = (int)std::min<int64_t>(CV_CH, p - j0);
  const int m = A.m;
  double* Sf = A.S + (int64_t)f * m * p + j0;
  if (a > 0) {
    for (int c = t; c < m; c += CV_NT) {
      double s = 0.0;
      for (int q = 0; q < A.nch; ++q) s += A.vspart[((int64_t)f * A.nch + q) * m + c];
      vs[c] = s;
    }
    __syncthreads();
    const double* v = A.V + ((int64_t)f * A.K + (a - 1)) * p + j0;
    for (int r = t; r < rows; r += CV_NT) {
      const double vr = v[r];
      for (int c = 0; c < m; ++c) Sf[r + (int64_t)c * p] -= vr * vs[c];
    }
    __syncthreads();
  }
  double s2 = 0.0;
  for (int r = t; r < rows; r += CV_NT) {
    double y0;
    if (A.m1) y0 = Sf[r];
    else if (a == 0) y0 = A.omega[(int64_t)f * p + j0 + r];
    else y0 = A.rr[(int64_t)f * p + j0 + r];
    vec[r] = y0;
    if (!A.power) {
      A.ybuf[(int64_t)f * p + j0 + r] = y0;
      s2 = fma(y0, y0, s2);
    }
  }
  __syncthreads();
  if (A.power) {
    cv_chunk_colsums(Sf, p, rows, m, vec, A.zpart + ((int64_t)f * A.nch + ch) * m, 1);
  } else {
    s2 = cv_block_sum(s2, red);
    if (t == 0) A.npart[(int64_t)f * A.nch + ch] = s2;
  }
}

// One power step y <- S (S^T y): z from the partials of buffer it & 1, then
// y = S z for the chunk; then the next step's partials (other buffer) or,
// after the last step, the partials of |y|^2.
__global__ __launch_bounds__(CV_NT) void k_cvb_power(CvArgs A, int it, int last) {
  __shared__ double vec[CV_CH];
  __shared__ double zs[CV_MAX_M];
  __shared__ double red[CV_NT / 64];
  const int f = blockIdx.y, ch = blockIdx.x, t = threadIdx.x;
  if (!A.act[f]) return;
  const int64_t p = A.p, j0 = (int64_t)ch * CV_CH;
  const int rows = (int)std::min<int64_t>(CV_CH, p - j0);
  const int m = A.m, F = A.F;
  const double* zin = A.zpart + (int64_t)(it & 1) * F * A.nch * m;
  double* zout = A.zpart + (int64_t)((it + 1) & 1) * F * A.nch * m;
  for (int c = t; c < m; c += CV_NT) {
    double s = 0.0;
    for (int q = 0; q < A.nch; ++q) s += zin[((int64_t)f * A.nch + q) * m + c];
    zs[c] = s;
  }
  __syncthreads();
  const double* Sf = A.S + (int64_t)f * m * p + j0;
  double s2 = 0.0;
  for (int r = t; r < rows; r += CV_NT) {
    double y = 0.0;
    for (int c = 0; c < m; ++c) y = fma(Sf[r + (int64_t)c * p], zs[c], y);
    vec[r] = y;
    if (last) {
      A.ybuf[(int64_t)f * p + j0 + r] = y;
      s2 = fma(y, y, s2);
    }
  }
  __syncthreads();
  if (!last) {
    cv_chunk_colsums(Sf, p, rows, m, vec, zout + ((int64_t)f * A.nch + ch) * m, 1);
  } else {
    s2 = cv_block_sum(s2, red);
    if (t == 0) A.npart[(int64_t)f * A.nch + ch] = s2;
  }
}

// r = y / |y| (FP64), its FP32 copy as column f of the NN operand, and the
// partial sums of d_f . r32 (the centring correction of t).
__global__ __launch_bounds__(CV_NT) void k_cvb_rr(CvArgs A) {
  __shared__ double red[CV_NT / 64];
  const int f = blockIdx.y, ch = blockIdx.x, t = threadIdx.x;
  const int64_t p = A.p, j0 = (int64_t)ch * CV_CH;
  const int rows = (int)std::min<int64_t>(CV_CH, p - j0);
  float* col = A.RR32 + (int64_t)f * A.Kp + j0;
  if (!A.act[f]) {
    for (int r = t; r < rows; r += CV_NT) col[r] = 0.f;
    if (t == 0) A.cpart[(int64_t)f * A.nch + ch] = 0.0;
    return;
  }
  double nrm2 = 0.0;
  for (int q = 0; q < A.nch; ++q) nrm2 += A.npart[(int64_t)f * A.nch + q];
  const double nrm = sqrt(nrm2);
  double cd = 0.0;
  for (int r = t; r < rows; r += CV_NT) {
    const int64_t j = (int64_t)f * p + j0 + r;
    const double rv = A.ybuf[j] / nrm;
    A.rr[j] = rv;
    const float r32 = (float)rv;
    col[r] = r32;
    cd = fma(A.d[j], (double)r32, cd);
  }
  cd = cv_block_sum(cd, red);
  if (t == 0) A.cpart[(int64_t)f * A.nch + ch] = cd;
}

// ---------------------------------------------------------------- append
// t before its division by |t| (grid: row blocks x fits): slab sums of the NN
// product minus d.r, the masked operand of the TN product (FP32, or split into
// bf16 hi / lo), and per block the training-row sums of each class's t (or of
// t (Y_c - gy_c)), of t^2 and of t.  The TN product and q are linear in t, so
// the division by |t| is applied to their results.
__device__ inline void cv_split_bf(float x, __bf16& hi, __bf16& lo) {
  hi = static_cast<__bf16>(x);
  lo = static_cast<__bf16>(x - static_cast<float>(hi));
}

__global__ __launch_bounds__(CV_NT) void k_cvb_treduce(CvArgs A) {
  __shared__ double red[CV_NT / 64][CV_MAX_M + 2];
  const int f = blockIdx.y, t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t n = A.n, F = A.F;
  const int64_t i = (int64_t)blockIdx.x * CV_NT + t;
  const int64_t oi = i + (int64_t)f * A.npad;
  if (!A.act[f]) {
    if (i < n) {
      if (A.bf) {
        A.Twh[oi] = static_cast<__bf16>(0.f);
        A.Twl[oi] = static_cast<__bf16>(0.f);
      } else {
        A.Tw32[oi] = 0.f;
      }
    }
    return;
  }
  double cf = 0.0;
  for (int q = 0; q < A.nch; ++q) cf += A.cpart[(int64_t)f * A.nch + q];
  double v = 0.0;
  bool tr = false;
  if (i < n) {
    for (int z = 0; z < A.nsplit; ++z) v += (double)A.projpart[(int64_t)z * n * F + i + (int64_t)f * n];
    v -= cf;
    A.T64[i + (int64_t)f * n] = v;
    tr = A.mask[i + (int64_t)f * n] == 1;
    const float tw = tr ? (float)v : 0.f;
    if (A.bf) {
      __bf16 hh, ll;
      cv_split_bf(tw, hh, ll);
      A.Twh[oi] = hh;
      A.Twl[oi] = ll;
    } else {
      A.Tw32[oi] = tw;
    }
  }
  const int m = A.m;
  const double vt = tr ? v : 0.0;
  for (int c = 0; c < m; ++c) {
    double x;
    if (A.classification) x = (tr && A.y[i] == c + 1) ? v : 0.0;
    else x = tr ? v * (cv_yval(A, i, c) - A.gy[c]) : 0.0;
    x = wave_sum_d(x);
    if (lane == 0) red[w][c] = x;
  }
  const double s2 = wave_sum_d(vt * vt), s1 = wave_sum_d(vt);
  if (lane == 0) {
    red[w][m] = s2;
    red[w][m + 1] = s1;
  }
  __syncthreads();
  for (int q = t; q < m + 2; q += CV_NT)
    A.tpart[((int64_t)f * A.nblk + blockIdx.x) * (m + 2) + q] =
        (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
}

// |t| from the row-block sums, the stop test (src/fastPLS.cpp:2580-2583), and
// (first chunk) q = Yc^T t, 1^T w t and |t| for later kernels; then
// p_f = (P_f - d_f 1^T w T) / |t|, r /= |t|, and the partial sums of V_f^T p_f
// (or of |p|^2 for the first component).
__global__ __launch_bounds__(CV_NT) void k_cvb_vproj(CvArgs A, int a) {
  __shared__ double vec[CV_CH];
  __shared__ double red[CV_NT / 64];
  const int f = blockIdx.y, ch = blockIdx.x, t = threadIdx.x;
  if (!A.act[f]) return;
  const int64_t p = A.p, j0 = (int64_t)ch * CV_CH;
  const int rows = (int)std::min<int64_t>(CV_CH, p - j0);
  const int m = A.m, m2 = A.m + 2;
  const double* tp = A.tpart + (int64_t)f * A.nblk * m2;
  double s2 = 0.0, s1 = 0.0;
  for (int b = 0; b < A.nblk; ++b) {
    s2 += tp[(int64_t)b * m2 + m];
    s1 += tp[(int64_t)b * m2 + m + 1];
  }
  const double tn = sqrt(s2);
  if (!(isfinite(tn) && tn > 0.0)) {
    // every workgroup of the fit computes the same |t| and stops here
    if (ch == 0 && t == 0) A.act[f] = 0;
    return;
  }
  const double st = s1 / tn;
  if (ch == 0) {
    for (int c = t; c < m; c += CV_NT) {
      double cs = 0.0;
      for (int b = 0; b < A.nblk; ++b) cs += tp[(int64_t)b * m2 + c];
      const double mu = A.classification ? A.mY[(int64_t)f * m + c]
                                         : A.mY[(int64_t)f * m + c] - A.gy[c];
      A.q[(int64_t)f * m + c] = (cs - mu * s1) / tn;
    }
    if (t == 0) {
      A.stv[f] = st;
      A.tnv[f] = tn;
    }
  }
  double pp = 0.0;
  for (int r = t; r < rows; r += CV_NT) {
    const int64_t j = (int64_t)f * p + j0 + r;
    A.rr[j] /= tn;
    const double pj = A.P[j] / tn - A.d[j] * st;
    vec[r] = pj;
    A.vbuf[j] = pj;
    pp = fma(pj, pj, pp);
  }
  __syncthreads();
  if (a == 0) {
    pp = cv_block_sum(pp, red);
    if (t == 0) A.v2part[(int64_t)f * A.nch + ch] = pp;
  } else {
    cv_chunk_colsums(A.V + (int64_t)f * A.K * p + j0, p, rows, a, vec,
                     A.vpart + ((int64_t)f * A.nch + ch) * A.K, 1);
  }
}

// v = p - V (V^T p) and the partial sums of |v|^2.
__global__ __launch_bounds__(CV_NT) void k_cvb_vorth(CvArgs A, int a) {
  __shared__ double cb[CV_MAX_K];
  __shared__ double red[CV_NT / 64];
  const int f = blockIdx.y, ch = blockIdx.x, t = threadIdx.x;
  if (!A.act[f]) return;
  const int64_t p = A.p, j0 = (int64_t)ch * CV_CH;
  const int rows = (int)std::min<int64_t>(CV_CH, p - j0);
  for (int b = t; b < a; b += CV_NT) {
    double s = 0.0;
    for (int q = 0; q < A.nch; ++q) s += A.vpart[((int64_t)f * A.nch + q) * A.K + b];
    cb[b] = s;
  }
  __syncthreads();
  const double* Vf = A.V + (int64_t)f * A.K * p + j0;
  double s2 = 0.0;
  for (int r = t; r < rows; r += CV_NT) {
    double acc = 0.0;
    for (int b = 0; b < a; ++b) acc = fma(Vf[r + (int64_t)b * p], cb[b], acc);
    const int64_t j = (int64_t)f * p + j0 + r;
    const double v = A.vbuf[j] - acc;
    A.vbuf[j] = v;
    s2 = fma(v, v, s2);
  }
  s2 = cv_block_sum(s2, red);
  if (t == 0) A.v2part[(int64_t)f * A.nch + ch] = s2;
}

// |v|, the stop test, V[:, a] = v / |v| and the partial sums of v^T S_f
// (applied by the next component's k_cvb_start).
__global__ __launch_bounds__(CV_NT) void k_cvb_vnorm(CvArgs A, int a) {
  __shared__ double vec[CV_CH];
  const int f = blockIdx.y, ch = blockIdx.x, t = threadIdx.x;
  if (!A.act[f]) return;
  const int64_t p = A.p, j0 = (int64_t)ch * CV_CH;
  const int rows = (int)std::min<int64_t>(CV_CH, p - j0);
  double v2 = 0.0;
  for (int q = 0; q < A.nch; ++q) v2 += A.v2part[(int64_t)f * A.nch + q];
  const double vn = sqrt(v2);
  if (!(isfinite(vn) && vn > 0.0)) {
    // every workgroup of the fit computes the same vn and stops here
    if (ch == 0 && t == 0) A.act[f] = 0;
    return;
  }
  double* Vcol = A.V + ((int64_t)f * A.K + a) * p + j0;
  for (int r = t; r < rows; r += CV_NT) {
    const double v = A.vbuf[(int64_t)f * p + j0 + r] / vn;
    Vcol[r] = v;
    vec[r] = v;
  }
  __syncthreads();
  cv_chunk_colsums(A.S + (int64_t)f * A.m * p + j0, p, rows, A.m, vec,
                   A.vspart + ((int64_t)f * A.nch + ch) * A.m, 1);
}

// Held-out rows: Yacc += t q^T for fits still running, then every ncomp slice
// that ends at this component is scored (first maximum wins, as index_max).
__global__ __launch_bounds__(CV_NT) void k_cvb_pred(CvArgs A, const int* sl, int ns) {
  const int f = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * CV_NT + threadIdx.x;
  if (i >= A.n || A.stat[f] != 1 || A.mask[i + (int64_t)f * A.n] != 2) return;
  const int64_t n = A.n;
  const int m = A.m;
  double* Y = A.Yacc + (int64_t)f * m * n + i;
  const double* mu = A.mY + (int64_t)f * m;
  if (A.act[f]) {
    const double tv = A.T64[i + (int64_t)f * n] / A.tnv[f];
    const double* qf = A.q + (int64_t)f * m;
    for (int c = 0; c < m; ++c) Y[(int64_t)c * n] = fma(tv, qf[c], Y[(int64_t)c * n]);
  }
  for (int k = 0; k < ns; ++k) {
    const int s = sl[k];
    const int64_t slab = (int64_t)s + (int64_t)A.L * f;
    int best = 0;
    double bv = Y[0] + mu[0];
    if (A.ypred) A.ypred[i + n * (m * slab)] = bv;
    for (int c = 1; c < m; ++c) {
      const double v = Y[(int64_t)c * n] + mu[c];
      if (A.ypred) A.ypred[i + n * (c + m * slab)] = v;
      if (v > bv) {
        bv = v;
        best = c;
      }
    }
    if (A.classification) {
      A.cls[i + n * slab] = best + 1;
      if (best + 1 == A.y[i]) atomicAdd(&A.ncorrect[slab], 1);
    }
  }
}

// ---------------------------------------------------------------- plan
struct CvPlan {
  // shape key
  const void* X = nullptr;
  int64_t n = 0, p = 0, ldx = 0;
  int F = 0, m = 0, K = 0, L = 0;
  std::vector<int> nc;
  int precision = 0, scaling = 0, inc_iters = 0, classification = 0, want_ypred = 0;
  const void* Y = nullptr;
  int ydtype = 0;
  int64_t ldy = 0;
  // derived
  int nch = 0;
  int64_t Kp = 0, ncols = 0;
  int nsplit = 1;
  int64_t nn_kchunk = 0, tn_splits = 1;
  std::vector<int> slice_off;   // K + 1 offsets into the slice list
  Pool pool;                    // temporaries of the recorded sequence
  std::vector<void*> owned;
  CvArgs A{};
  float* g32 = nullptr;
  double* g = nullptr;
  float* Z = nullptr;
  double* C0w = nullptr;
  double* Pw = nullptr;
  float* projw = nullptr;
  uint32_t* seeds = nullptr;
  int* slices = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  bool captured = false;

  template <typename T>
  T* dalloc(size_t count) {
    void* ptr = nullptr;
    FPLS_HIP(hipMalloc(&ptr, std::max<size_t>(count, 1) * sizeof(T)));
    owned.push_back(ptr);
    return static_cast<T*>(ptr);
  }
  ~CvPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    pool.release_all();
    for (void* ptr : owned) (void)hipFree(ptr);
  }
};

// Swaps the context pool for the plan's while the sequence is recorded, so the
// temporaries baked into the graph are never handed to anything else.
struct PoolSwap {
  Ctx* c;
  CvPlan* pl;
  PoolSwap(Ctx* c_, CvPlan* p_) : c(c_), pl(p_) { std::swap(c->pool, pl->pool); }
  ~PoolSwap() { std::swap(c->pool, pl->pool); }
};

static unsigned cv_g1(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
}

// The whole launch sequence of one batched call.
static void cv_record(Ctx* c, CvPlan& pl) {
  CvArgs& A = pl.A;
  const dim3 chunks((unsigned)pl.nch, (unsigned)pl.F);
  launch(c, "cv_init", k_cvb_init, dim3(cv_g1(pl.n * pl.m * pl.F)), dim3(256), 0, A);
  // overall column means (the common shift of every product)
  DMat Xd;
  Xd.p = pl.X;
  Xd.dtype = FPLS_F32;
  Xd.rows = pl.n;
  Xd.cols = pl.p;
  Xd.ld = pl.ldx;
  if (pl.scaling != 3) col_sums(c, Xd, pl.g);
  launch(c, "cv_gscale", k_cvb_gscale, dim3(cv_g1(pl.p)), dim3(256), 0, pl.p,
         1.0 / (double)pl.n, pl.scaling == 3 ? 1 : 0, pl.g, pl.g32);
  if (!pl.classification)
    launch(c, "cv_ymeans", k_cvb_ymeans, dim3((unsigned)pl.F + 1), dim3(CV_NT), 0, A);
  launch(c, "cv_zbuild", k_cvb_zbuild, dim3(cv_g1(pl.n * pl.ncols)), dim3(256), 0, A, pl.Z,
         pl.ncols);
  // cross-covariances of every fit from one TN product (FP32 MFMA: the mask /
  // one-hot operand is exact and S is computed once)
  gemm_tn_dense(c, FPLS_PREC_FP32, pl.p, pl.ncols, pl.n, (const float*)pl.X, pl.ldx, pl.g,
                pl.Z, pl.n, nullptr, nullptr, pl.C0w, pl.p,
                tn_default_splits(c, pl.p, pl.ncols, pl.n));
  launch(c, "cv_sinit", k_cvb_sinit, chunks, dim3(CV_NT), 0, A);
  if (!A.m1) gaussian_dev(c, pl.seeds, pl.F, pl.p, const_cast<double*>(A.omega), pl.p);
  for (int a = 0; a < pl.K; ++a) {
    launch(c, "cv_start", k_cvb_start, chunks, dim3(CV_NT), 0, A, a);
    for (int it = 0; it < A.power; ++it)
      launch(c, "cv_power", k_cvb_power, chunks, dim3(CV_NT), 0, A, it,
             it == A.power - 1 ? 1 : 0);
    launch(c, "cv_rr", k_cvb_rr, chunks, dim3(CV_NT), 0, A);
    proj_splitk(c, pl.precision, pl.n, pl.F, pl.p, (const float*)pl.X, pl.ldx,
                pl.scaling == 3 ? nullptr : pl.g32, pl.A.RR32, pl.Kp, pl.nn_kchunk, pl.nsplit,
                pl.projw);
    launch(c, "cv_treduce", k_cvb_treduce, dim3((unsigned)A.nblk, (unsigned)pl.F), dim3(CV_NT),
           0, A);
    tn_skinny(c, pl.precision, pl.p, pl.F, pl.n, (const float*)pl.X, pl.ldx,
              pl.scaling == 3 ? nullptr : pl.g, A.Tw32, A.Twh, A.Twl, A.npad, pl.tn_splits, pl.Pw,
              pl.p);
    launch(c, "cv_vproj", k_cvb_vproj, chunks, dim3(CV_NT), 0, A, a);
    if (a > 0) launch(c, "cv_vorth", k_cvb_vorth, chunks, dim3(CV_NT), 0, A, a);
    launch(c, "cv_vnorm", k_cvb_vnorm, chunks, dim3(CV_NT), 0, A, a);
    const int ns = pl.slice_off[a + 1] - pl.slice_off[a];
    launch(c, "cv_pred", k_cvb_pred, dim3((unsigned)((pl.n + CV_NT - 1) / CV_NT), (unsigned)pl.F),
           dim3(CV_NT), 0, A, (const int*)(pl.slices + pl.slice_off[a]), ns);
  }
}

// Mean time (ms) of `run` over a few launches, after one warm-up.
template <typename Fn>
static double cv_time(Ctx* c, Fn&& run) {
  hipEvent_t e0 = c->event(), e1 = c->event();
  run();
  FPLS_HIP(hipEventRecord(e0, c->stream));
  for (int r = 0; r < 3; ++r) run();
  FPLS_HIP(hipEventRecord(e1, c->stream));
  FPLS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  FPLS_HIP(hipEventElapsedTime(&ms, e0, e1));
  c->spare_events.push_back(e0);
  c->spare_events.push_back(e1);
  return ms / 3.0;
}

// Split-K choices of the two per-component products, by timing candidates on
// the plan's own buffers (contents irrelevant: both are pure products).
static void cv_autotune(Ctx* c, CvPlan& pl) {
  const bool timing = c->timing;
  c->timing = false;
  // FPLS_TUNE_LOG=1 prints every candidate's time to stderr
  const char* lg = getenv("FPLS_TUNE_LOG");
  const bool log = lg && lg[0] == '1';
  // NN: slabs of at least 256 rows of p, at most 32 slabs
  double best = 1e30;
  int best_s = 1;
  for (int s : {1, 2, 4, 8, 16, 32}) {
    int64_t kchunk = (pl.p + s - 1) / s;
    kchunk = (kchunk + 31) / 32 * 32;
    const int ss = (int)((pl.p + kchunk - 1) / kchunk);
    if (ss != s || (s > 1 && kchunk < 256)) continue;
    const double ms = cv_time(c, [&] {
      proj_splitk(c, pl.precision, pl.n, pl.F, pl.p, (const float*)pl.X, pl.ldx, pl.g32,
                  pl.A.RR32, pl.Kp, kchunk, s, pl.projw);
    });
    if (log) fprintf(stderr, "cv autotune: NN splits %d  %.1f us\n", s, 1e3 * ms);
    if (ms < best) {
      best = ms;
      best_s = s;
    }
  }
  pl.nsplit = best_s;
  pl.nn_kchunk = ((pl.p + best_s - 1) / best_s + 31) / 32 * 32;
  // TN: slabs of at least 256 rows of n
  best = 1e30;
  int64_t best_t = 1;
  for (int64_t s : {1, 2, 4, 8, 16, 32}) {
    if (s > 1 && pl.n / s < 256) continue;
    const double ms = cv_time(c, [&] {
      tn_skinny(c, pl.precision, pl.p, pl.F, pl.n, (const float*)pl.X, pl.ldx, pl.g, pl.A.Tw32,
                pl.A.Twh, pl.A.Twl, pl.A.npad, s, pl.Pw, pl.p);
    });
    if (log) fprintf(stderr, "cv autotune: TN splits %ld  %.1f us\n", (long)s, 1e3 * ms);
    if (ms < best) {
      best = ms;
      best_t = s;
    }
  }
  pl.tn_splits = best_t;
  c->timing = timing;
}

static std::shared_ptr<CvPlan> cv_build_plan(Ctx* c, const DMat& X, const DMat* Y, int F, int m,
                                             const std::vector<int>& nc, const fpls_opts& o,
                                             bool classification, bool want_ypred) {
  auto pl = std::make_shared<CvPlan>();
  pl->X = X.p;
  pl->n = X.rows;
  pl->p = X.cols;
  pl->ldx = X.ld;
  pl->F = F;
  pl->m = m;
  pl->nc = nc;
  pl->K = *std::max_element(nc.begin(), nc.end());
  pl->L = (int)nc.size();
  pl->precision = o.precision;
  pl->scaling = o.scaling;
  pl->inc_iters = std::max(o.inc_iters, 0);
  pl->classification = classification ? 1 : 0;
  pl->want_ypred = want_ypred ? 1 : 0;
  if (Y) {
    pl->Y = Y->p;
    pl->ydtype = Y->dtype;
    pl->ldy = Y->ld;
  }
  const int64_t n = pl->n, p = pl->p;
  const int K = pl->K, L = pl->L;
  pl->nch = (int)((p + CV_CH - 1) / CV_CH);
  pl->Kp = (p + 31) / 32 * 32;
  pl->ncols = (int64_t)F * m + (classification ? 0 : F);
  // ncomp slices grouped by the component that completes them
  std::vector<std::vector<int>> by(K);
  for (int s = 0; s < L; ++s) by[nc[s] - 1].push_back(s);
  std::vector<int> flat;
  pl->slice_off.assign(K + 1, 0);
  for (int a = 0; a < K; ++a) {
    pl->slice_off[a] = (int)flat.size();
    flat.insert(flat.end(), by[a].begin(), by[a].end());
  }
  pl->slice_off[K] = (int)flat.size();
  CvArgs& A = pl->A;
  A.n = n;
  A.p = p;
  A.F = F;
  A.m = m;
  A.K = K;
  A.L = L;
  A.nch = pl->nch;
  A.Kp = pl->Kp;
  A.classification = pl->classification;
  A.m1 = m == 1 ? 1 : 0;
  A.power = m > 1 ? pl->inc_iters : 0;
  A.scaling3 = o.scaling == 3 ? 1 : 0;
  A.X = (const float*)X.p;
  A.ldx = X.ld;
  if (Y) {
    A.Yf = Y->dtype == FPLS_F32 ? Y->f() : nullptr;
    A.Yd = Y->dtype == FPLS_F64 ? Y->d() : nullptr;
    A.ldy = Y->ld;
  }
  CvPlan& P = *pl;
  A.mask = P.dalloc<int8_t>((size_t)n * F);
  A.y = P.dalloc<int32_t>(n);
  A.stat = P.dalloc<int32_t>(F);
  A.nf = P.dalloc<double>(F);
  A.mY = P.dalloc<double>((size_t)F * m);
  A.gy = P.dalloc<double>(m);
  P.g = P.dalloc<double>(p);
  P.g32 = P.dalloc<float>(p);
  P.Z = P.dalloc<float>((size_t)n * pl->ncols);
  P.C0w = P.dalloc<double>((size_t)p * pl->ncols);
  A.C0 = P.C0w;
  A.S = P.dalloc<double>((size_t)p * F * m);
  A.d = P.dalloc<double>((size_t)p * F);
  A.omega = P.dalloc<double>((size_t)p * F);
  A.ybuf = P.dalloc<double>((size_t)p * F);
  A.rr = P.dalloc<double>((size_t)p * F);
  A.RR32 = P.dalloc<float>((size_t)pl->Kp * 128);
  A.zpart = P.dalloc<double>((size_t)2 * F * pl->nch * m);
  A.npart = P.dalloc<double>((size_t)F * pl->nch);
  A.cpart = P.dalloc<double>((size_t)F * pl->nch);
  P.projw = P.dalloc<float>((size_t)32 * n * F);
  A.projpart = P.projw;
  A.T64 = P.dalloc<double>((size_t)n * F);
  A.bf = o.precision == FPLS_PREC_BF16X3 ? 1 : 0;
  A.npad = (n + 31) / 32 * 32;
  if (A.bf) {
    A.Twh = P.dalloc<__bf16>((size_t)A.npad * F);
    A.Twl = P.dalloc<__bf16>((size_t)A.npad * F);
  } else {
    A.Tw32 = P.dalloc<float>((size_t)A.npad * F);
  }
  A.nblk = (int)((n + CV_NT - 1) / CV_NT);
  A.tpart = P.dalloc<double>((size_t)F * A.nblk * (m + 2));
  A.q = P.dalloc<double>((size_t)F * m);
  A.stv = P.dalloc<double>(F);
  A.tnv = P.dalloc<double>(F);
  A.act = P.dalloc<int32_t>(F);
  P.Pw = P.dalloc<double>((size_t)p * F);
  A.P = P.Pw;
  A.vbuf = P.dalloc<double>((size_t)p * F);
  A.vpart = P.dalloc<double>((size_t)F * pl->nch * K);
  A.v2part = P.dalloc<double>((size_t)F * pl->nch);
  A.vspart = P.dalloc<double>((size_t)F * pl->nch * m);
  A.V = P.dalloc<double>((size_t)p * K * F);
  A.Yacc = P.dalloc<double>((size_t)n * m * F);
  A.cls = P.dalloc<int32_t>((size_t)n * L * F);
  A.ncorrect = P.dalloc<int32_t>((size_t)L * F);
  A.ypred = want_ypred ? P.dalloc<double>((size_t)n * m * L * F) : nullptr;
  P.seeds = P.dalloc<uint32_t>(F);
  P.slices = P.dalloc<int>(std::max<size_t>(flat.size(), 1));
  h2d(c, P.slices, flat.data(), flat.size() * sizeof(int));
  // the NN operand's padding rows and columns stay zero
  fill_f32(c, A.RR32, 0.f, pl->Kp * 128);
  fill_f64(c, P.g, 0.0, p);
  fill_f32(c, P.g32, 0.f, p);
  // the TN operand's padding rows stay zero (bf16 zero is the all-zero pattern;
  // npad F is even)
  if (A.bf) {
    fill_i32(c, reinterpret_cast<int32_t*>(A.Twh), 0, A.npad * F / 2);
    fill_i32(c, reinterpret_cast<int32_t*>(A.Twl), 0, A.npad * F / 2);
  } else {
    fill_f32(c, A.Tw32, 0.f, A.npad * F);
  }
  {
    PoolSwap sw(c, pl.get());
    cv_autotune(c, P);
  }
  fill_f32(c, A.RR32, 0.f, pl->Kp * 128);
  A.nsplit = P.nsplit;
  return pl;
}

static bool cv_plan_matches(const CvPlan& pl, const DMat& X, const DMat* Y, int F, int m,
                            const std::vector<int>& nc, const fpls_opts& o, bool classification,
                            bool want_ypred) {
  return pl.X == X.p && pl.n == X.rows && pl.p == X.cols && pl.ldx == X.ld && pl.F == F &&
         pl.m == m && pl.nc == nc && pl.precision == o.precision && pl.scaling == o.scaling &&
         pl.inc_iters == std::max(o.inc_iters, 0) && pl.classification == (classification ? 1 : 0) &&
         pl.want_ypred == (want_ypred ? 1 : 0) && pl.Y == (Y ? Y->p : nullptr) &&
         (!Y || (pl.ydtype == Y->dtype && pl.ldy == Y->ld));
}

// Runs the recorded sequence: eagerly the first time (and whenever kernel
// timing is on), then from the captured graph.
static void cv_execute(Ctx* c, CvPlan& pl) {
  PoolSwap sw(c, &pl);
  if (pl.captured && !c->timing) {
    FPLS_HIP(hipGraphLaunch(pl.exec, c->stream));
    return;
  }
  cv_record(c, pl);
  if (!pl.captured && !c->timing) {
    // every temporary of the sequence now sits in the plan's pool, so the
    // capture allocates nothing
    FPLS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    try {
      cv_record(c, pl);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(c->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    FPLS_HIP(hipStreamEndCapture(c->stream, &pl.graph));
    FPLS_HIP(hipGraphInstantiate(&pl.exec, pl.graph, nullptr, nullptr, 0));
    pl.captured = true;
  }
}

}  // namespace fpls

using namespace fpls;

extern "C" int fpls_cv_batched(fpls_ctx* ctx, const fpls_mat* Xm, const fpls_mat* Ym,
                               const fpls_ivec* yv, int32_t n_classes, const int8_t* masks,
                               int32_t n_fits, const uint32_t* seeds, const int32_t* ncomp,
                               int32_t n_ncomp, const fpls_opts* opts, int32_t* status_out,
                               int32_t* class_pred, int32_t* n_correct, double* ypred) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    fpls_opts o;
    fpls_opts_default(&o);
    if (opts) o = *opts;
    DBuf<unsigned char> hx, hy;
    DMat X = stage(c, Xm, hx, "cv_batched: X");
    FPLS_REQUIRE(X.dtype == FPLS_F32, FPLS_ERR_ARG,
                 "cv_batched: X must be FP32 (the batched products run on FP32 / BF16x3 MFMA)");
    const int64_t n = X.rows, p = X.cols;
    const int F = n_fits;
    FPLS_REQUIRE(n >= 2, FPLS_ERR_ARG, "Xdata must contain at least two samples");
    FPLS_REQUIRE(p > 5, FPLS_ERR_ARG,
                 "cv_batched: needs more than 5 columns (the incremental SIMPLS refresh)");
    FPLS_REQUIRE(F >= 1 && F <= CV_MAX_F, FPLS_ERR_ARG, "cv_batched: n_fits must be 1..128");
    FPLS_REQUIRE(masks != nullptr && status_out != nullptr, FPLS_ERR_ARG,
                 "cv_batched: masks and status_out are required");
    FPLS_REQUIRE(ncomp && n_ncomp >= 1, FPLS_ERR_ARG, "ncomp must contain at least one value");
    FPLS_REQUIRE(o.scaling == 1 || o.scaling == 3, FPLS_ERR_ARG,
                 "cv_batched: scaling must be 1 (centering) or 3 (none)");
    const bool classification = yv && yv->ptr;
    std::vector<int32_t> yh;
    DMat Y;
    int m = 0;
    if (classification) {
      FPLS_REQUIRE(yv->n == n, FPLS_ERR_ARG, "one label per sample is required");
      FPLS_REQUIRE(n_classes >= 1, FPLS_ERR_ARG, "n_classes must be >= 1");
      yh.resize(n);
      if (yv->loc == FPLS_DEVICE) {
        FPLS_HIP(hipMemcpyAsync(yh.data(), yv->ptr, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                                c->stream));
        sync(c);
      } else {
        std::copy(yv->ptr, yv->ptr + n, yh.begin());
      }
      for (int32_t v : yh)
        FPLS_REQUIRE(v >= 1 && v <= n_classes, FPLS_ERR_LABEL_RANGE,
                     "labels must be encoded as 1..n_classes");
      m = n_classes;
    } else {
      Y = stage(c, Ym, hy, "cv_batched: Y");
      FPLS_REQUIRE(Y.rows == n, FPLS_ERR_ARG, "Ydata must have the same number of rows as Xdata");
      m = (int)Y.cols;
      FPLS_REQUIRE(m >= 1, FPLS_ERR_ARG, "Ydata has no columns");
    }
    FPLS_REQUIRE(m <= CV_MAX_M, FPLS_ERR_ARG, "cv_batched: at most 512 responses / classes");
    std::vector<int> nc(ncomp, ncomp + n_ncomp);
    for (int& v : nc) v = std::max(v, 1);
    FPLS_REQUIRE(*std::max_element(nc.begin(), nc.end()) <= CV_MAX_K, FPLS_ERR_ARG,
                 "cv_batched: at most 512 components");
    const int L = n_ncomp;
    // fold status, training counts and class counts on the host
    std::vector<int32_t> stat(F, 1);
    std::vector<double> nfh(F, 0.0), mYh((size_t)F * m, 0.0);
    std::vector<int> single(F, 0);
    for (int f = 0; f < F; ++f) {
      int64_t ntr = 0, nte = 0;
      std::vector<int64_t> cnt(classification ? m : 0, 0);
      for (int64_t i = 0; i < n; ++i) {
        const int8_t v = masks[i + (int64_t)f * n];
        FPLS_REQUIRE(v >= 0 && v <= 2, FPLS_ERR_ARG, "masks hold 0 (unused), 1 (train), 2 (test)");
        if (v == 1) {
          ++ntr;
          if (classification) ++cnt[yh[i] - 1];
        } else if (v == 2) {
          ++nte;
        }
      }
      nfh[f] = (double)ntr;
      if (nte == 0) {
        stat[f] = 2;
      } else if (ntr == 0) {
        stat[f] = 3;
      } else if (classification) {
        int active = 0, last = -1;
        for (int j = 0; j < m; ++j)
          if (cnt[j] > 0) {
            ++active;
            last = j;
          }
        if (active <= 1) {
          stat[f] = 4;
          single[f] = last + 1;
        }
        for (int j = 0; j < m; ++j) mYh[(size_t)f * m + j] = (double)cnt[j] / (double)ntr;
      }
    }
    std::vector<uint32_t> sd(F);
    for (int f = 0; f < F; ++f) sd[f] = seeds ? seeds[f] : o.seed + (uint32_t)f;
    // plan (reused while the shapes, X and options stay the same)
    const bool want_ypred = ypred != nullptr;
    auto pl = std::static_pointer_cast<CvPlan>(c->cv_plan);
    if (!pl || !cv_plan_matches(*pl, X, classification ? nullptr : &Y, F, m, nc, o,
                                classification, want_ypred)) {
      c->cv_plan.reset();
      pl = cv_build_plan(c, X, classification ? nullptr : &Y, F, m, nc, o, classification,
                         want_ypred);
      c->cv_plan = pl;
    }
    CvPlan& P = *pl;
    h2d(c, const_cast<int8_t*>(P.A.mask), masks, (size_t)n * F);
    if (classification) h2d(c, const_cast<int32_t*>(P.A.y), yh.data(), n * sizeof(int32_t));
    h2d(c, const_cast<int32_t*>(P.A.stat), stat.data(), F * sizeof(int32_t));
    h2d(c, const_cast<double*>(P.A.nf), nfh.data(), F * sizeof(double));
    if (classification) h2d(c, P.A.mY, mYh.data(), (size_t)F * m * sizeof(double));
    h2d(c, P.seeds, sd.data(), F * sizeof(uint32_t));
    cv_execute(c, P);
    // results
    std::copy(stat.begin(), stat.end(), status_out);
    if (classification && (class_pred || n_correct)) {
      if (class_pred) to_host_i32(c, class_pred, P.A.cls, (size_t)n * L * F);
      if (n_correct) to_host_i32(c, n_correct, P.A.ncorrect, (size_t)L * F);
    }
    if (ypred) to_host(c, ypred, P.A.ypred, (size_t)n * m * L * F);
    sync(c);
    // degenerate folds: observed responses (3) or the single training class (4)
    std::vector<double> yrow;
    for (int f = 0; f < F; ++f) {
      if (stat[f] != 3 && stat[f] != 4) continue;
      if (!classification) {  // status 3: the observed rows of Y
        if (!ypred) continue;
        std::vector<double> ycol(n);
        for (int j = 0; j < m; ++j) {
          if (Y.dtype == FPLS_F64) {
            to_host(c, ycol.data(), Y.d() + (int64_t)j * Y.ld, n);
          } else {
            std::vector<float> tmp(n);
            FPLS_HIP(hipMemcpyAsync(tmp.data(), Y.f() + (int64_t)j * Y.ld, n * sizeof(float),
                                    hipMemcpyDeviceToHost, c->stream));
            sync(c);
            for (int64_t i = 0; i < n; ++i) ycol[i] = tmp[i];
          }
          for (int s = 0; s < L; ++s)
            for (int64_t i = 0; i < n; ++i)
              if (masks[i + (int64_t)f * n] == 2)
                ypred[i + n * (j + (int64_t)m * (s + (int64_t)L * f))] = ycol[i];
        }
        continue;
      }
      for (int s = 0; s < L; ++s) {
        int good = 0;
        for (int64_t i = 0; i < n; ++i) {
          if (masks[i + (int64_t)f * n] != 2) continue;
          const int lab = stat[f] == 3 ? yh[i] : single[f];
          const int64_t slab = s + (int64_t)L * f;
          if (class_pred) class_pred[i + n * slab] = lab;
          good += lab == yh[i];
          if (ypred)
            for (int j = 0; j < m; ++j) ypred[i + n * (j + (int64_t)m * slab)] = (j + 1 == lab);
        }
        if (n_correct) n_correct[s + (int64_t)L * f] = good;
      }
    }
  });
}
