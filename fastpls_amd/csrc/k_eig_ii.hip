// Symmetric eigenpairs of orders <= 128 (Gram matrices of the cross-covariance
// and of rSVD sketches, Jacobi pair blocks), batched, in three launches:
//
//  1. k_trd_reg: Householder reduction to tridiagonal form T = Q^T A Q
//     (Golub & Van Loan, Algorithm 8.3.1), one workgroup of 256 threads per
//     matrix with the whole matrix in registers: thread (tr, tc) of a 16 x 16
//     grid holds the cyclic 8 x 8 set A[tr + 16 i][tc + 16 q], so the live
//     trailing block stays spread over every thread as it shrinks.  Per column
//     the owners publish the column to LDS, every wave forms the reflector,
//     p = tau A v is reduced over the 16 lanes of a thread row by DPP moves
//     (reduce-scatter: each lane ends with one row's total), and the rank-2
//     update A -= v w^T + w v^T runs on the registers: two barriers a column.
//  2. k_trd_pairs: one workgroup per eigenvalue k of T: multisection on the
//     Sturm count (number of eigenvalues below x from the LDL^T recurrence
//     q_i = d_i - x - e_{i-1}^2 / q_{i-1}; Barth, Martin & Wilkinson 1967,
//     with the pivmin guard and Gershgorin start of LAPACK's dstebz), 512
//     points per step, the division by the hardware reciprocal and one Newton
//     step (about 11 ulp, so the counts are those of a matrix perturbed by a
//     few ulp); then the eigenvector from the twisted factorisation
//     of T - lambda I (top-down L D L^T and bottom-up U D U^T pivots on two
//     lanes at once, twist index r = argmin |gamma_r|, z_r = 1 and the two
//     product recurrences outward; Fernando 1997, Parlett & Dhillon 2000),
//     normalised with its largest entry positive, checked by its residual,
//     and back-transformed on one wave by blocks of four reflectors in
//     compact-WY form (their T factors built by the other three waves).
//  3. k_trd_clorth: eigenvalues closer than 1e-3 |T| form a cluster whose
//     vectors are made orthogonal by modified Gram-Schmidt (one wave per
//     cluster); a cluster wider than CL_MAX or tighter than 1e-9 |T|, a
//     non-finite vector or a residual above the bound sets the problem's flag,
//     and the divide-and-conquer solver (k_eig.hip) then recomputes that
//     problem.
#include <cfloat>
#include <cmath>

#include "fpls_internal.h"
#include "fpls_wave.h"

// Timing hook for stage-level measurements; compiled out unless a probe build
// defines it (it then records a clock value at each call site).
#ifndef FPLS_TRD_STAMP
#define FPLS_TRD_STAMP(slot)
#endif

namespace fpls {

namespace {

constexpr int II_N = 128;   // largest order
constexpr int CL_MAX = 8;   // largest cluster the Gram-Schmidt pass takes
constexpr double II_EPS = DBL_EPSILON * 0.5;  // unit roundoff (dlamch('E'))

constexpr int TD_T = 256;  // tridiagonalisation: 16 x 16 thread grid
constexpr int TD_R = 8;    // 8 x 8 cyclic register set per thread

// DPP lane moves inside a 16-lane row (row_ror:8 pairs lane i with i ^ 8)
constexpr int DPP_ROR8 = 0x128;

__device__ __forceinline__ double wmin(double v) { return -wave_max_d(-v); }

// 1/d: hardware reciprocal and two Newton steps (within an ulp or two; the
// sequential recurrences below are latency bound on their divisions)
__device__ __forceinline__ double rcp2(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = fma(-d, y, 1.0);
  y = fma(y, e, y);
  e = fma(-d, y, 1.0);
  return fma(y, e, y);
}

// 1/d from the hardware reciprocal and one Newton step (within about 11 ulp)
__device__ __forceinline__ double rcp1(double d) {
  const double y = __builtin_amdgcn_rcp(d);
  return fma(y, fma(-d, y, 1.0), y);
}

// Gershgorin interval, |T| bounds and pivmin (every lane of a wave gets the
// same values).
struct TriBounds {
  double gl, gu, tnorm, pivmin, onenrm;
};
__device__ TriBounds tri_bounds(const double* d, const double* e, const double* ee, int n) {
  const int lane = threadIdx.x & 63;
  double lo = DBL_MAX, hi = -DBL_MAX, emax = 0.0, one = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
    one = fmax(one, fabs(d[i]) + r);
    if (i + 1 < n) emax = fmax(emax, ee[i]);
  }
  TriBounds b;
  b.gl = wmin(lo);
  b.gu = wave_max_d(hi);
  b.onenrm = wave_max_d(one);
  b.tnorm = fmax(fabs(b.gl), fabs(b.gu));
  b.pivmin = DBL_MIN * fmax(1.0, wave_max_d(emax));
  const double fudge = 2.1 * (b.tnorm * II_EPS * n + 2.0 * b.pivmin);
  b.gl -= fudge;
  b.gu += fudge;
  return b;
}

// ---------------------------------------------------------------- stage 1
// Shared state of the column loop (LDS buffers and the thread's roles).
struct TrdState {
  double (*xs)[II_N];  // published column, by parity of j
  double* ps;          // tau A v
  double* kred;        // per-wave shares of v^T A v
  double* tri;
  double* refl;
  int n, t, tr, tc, lane, wv;
  int hb3, hb2, hb1, hb0, slot;
};

// Columns j in [16 jb, jend) with L = 8 - jb live register rows / columns:
// register (i, q) holds A[tr + 16 (i + jb)][tc + 16 (q + jb)].  Rows and
// columns at or past n hold zeros throughout (their v and w are zero), so
// only the diagonal position relative to j needs a test.  The rank-2 update
// of column j is applied lazily: register column 0 (which holds the next
// column's owners' entries) before the first barrier of column j + 1, the
// other columns after it, where the FMAs overlap the latency of the next
// reflector's scalars.
template <int L>
__device__ __forceinline__ void trd_block(double (&a)[TD_R][TD_R], int jb, int jend,
                                          const TrdState& S) {
  const int n = S.n, tr = S.tr, tc = S.tc, lane = S.lane;
  double vr[L], vc[L], wr[L], wc[L];  // the pending update A -= v w^T + w v^T
#pragma unroll
  for (int i = 0; i < L; ++i) vr[i] = vc[i] = wr[i] = wc[i] = 0.0;
  for (int j = 16 * jb; j < jend; ++j) {
    double* __restrict__ x = S.xs[j & 1];
    if (j == 20) FPLS_TRD_STAMP(2);
#pragma unroll
    for (int i = 0; i < L; ++i) a[i][0] = fma(-vr[i], wc[0], fma(-wr[i], vc[0], a[i][0]));
    // column j from its owners (column register 0 of the threads with
    // tc == j % 16); rows below j are never read back
    if (tc == (j & 15)) {
#pragma unroll
      for (int i = 0; i < L; ++i) x[tr + 16 * (i + jb)] = a[i][0];
    }
    __syncthreads();
    if (j == 20) FPLS_TRD_STAMP(3);
    const double alpha = x[j + 1];
    double s = 0.0;
    {
      const int r0 = j + 2 + lane, r1 = r0 + 64;
      if (r0 < n) s = x[r0] * x[r0];
      if (r1 < n) s = fma(x[r1], x[r1], s);
    }
    double xr[L], xc[L];
#pragma unroll
    for (int i = 0; i < L; ++i) {
      xr[i] = x[tr + 16 * (i + jb)];
      xc[i] = x[tc + 16 * (i + jb)];
    }
#pragma unroll
    for (int i = 0; i < L; ++i)
#pragma unroll
      for (int q = 1; q < L; ++q) a[i][q] = fma(-vr[i], wc[q], fma(-wr[i], vc[q], a[i][q]));
    // reflector: beta = -sign(alpha) sqrt(alpha^2 + |x|^2), tau = 1 + |alpha| / |beta|,
    // v = x / (alpha - beta); r = 1 / sqrt(.) from rsq and two Newton steps
    const double xn2 = wave_sum_d(s);
    double beta = alpha, tau = 0.0, sv = 0.0;
    if (xn2 > 0.0) {
      const double q2 = fma(alpha, alpha, xn2);
      double r = __builtin_amdgcn_rsq(q2);
      double h = fma(-q2 * r, r, 1.0);
      r = fma(0.5 * r, h, r);
      h = fma(-q2 * r, r, 1.0);
      r = fma(0.5 * r, h, r);
      const double sq = q2 * r;
      beta = -copysign(sq, alpha);
      tau = fma(fabs(alpha), r, 1.0);
      sv = copysign(rcp2(fabs(alpha) + sq), alpha);
    }
    if (j == 20) FPLS_TRD_STAMP(4);
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const int r = tr + 16 * (i + jb), c = tc + 16 * (i + jb);
      vr[i] = (r == j + 1) ? 1.0 : (r > j + 1 ? xr[i] * sv : 0.0);
      vc[i] = (c == j + 1) ? 1.0 : (c > j + 1 ? xc[i] * sv : 0.0);
    }
    if (tau != 0.0) {  // uniform: H_j = I leaves A unchanged
      double pr[TD_R];
      double kp = 0.0;  // this thread's share of v^T A v
#pragma unroll
      for (int i = 0; i < TD_R; ++i) {
        double acc = 0.0;
        if (i < L) {
#pragma unroll
          for (int q = 0; q < L; ++q) acc = fma(a[i][q], vc[q], acc);
          kp = fma(vr[i], acc, kp);
        }
        pr[i] = acc;
      }
      // reduce-scatter over the 16 lanes of the thread row: partners by
      // row_ror:8 (bit 3), row_half_mirror (bit 2), quad_perm xor 2 (bit 1),
      // quad_perm xor 1 (bit 0); each exchange sends the half the partner keeps
      double h4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double keep = S.hb3 ? pr[4 + u] : pr[u];
        const double send = S.hb3 ? pr[u] : pr[4 + u];
        h4[u] = keep + dpp_mov<DPP_ROR8>(send);
      }
      double h2[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const double keep = S.hb2 ? h4[2 + u] : h4[u];
        const double send = S.hb2 ? h4[u] : h4[2 + u];
        h2[u] = keep + dpp_mov<dpp::HALF_MIRROR>(send);
      }
      double h1;
      {
        const double keep = S.hb1 ? h2[1] : h2[0];
        const double send = S.hb1 ? h2[0] : h2[1];
        h1 = keep + dpp_mov<dpp::QUAD_XOR2>(send);
      }
      h1 += dpp_mov<dpp::QUAD_XOR1>(h1);
      const double kw = wave_sum_d(kp);
      if (j == 20) FPLS_TRD_STAMP(5);
      if (S.hb0 == 0 && S.slot < L) S.ps[tr + 16 * (S.slot + jb)] = tau * h1;
      if (lane == 0) S.kred[S.wv] = kw;
      __syncthreads();
      if (j == 20) FPLS_TRD_STAMP(6);
      // p = tau A v (in ps), K = tau / 2 p^T v = tau^2 / 2 v^T A v, w = p - K v
      const double K =
          0.5 * tau * tau * ((S.kred[0] + S.kred[1]) + (S.kred[2] + S.kred[3]));
#pragma unroll
      for (int i = 0; i < L; ++i) {
        const int r = tr + 16 * (i + jb), c = tc + 16 * (i + jb);
        const double pr_ = S.ps[r], pc_ = S.ps[c];
        wr[i] = r > j ? fma(-K, vr[i], pr_) : 0.0;
        wc[i] = c > j ? fma(-K, vc[i], pc_) : 0.0;
      }
    } else {
#pragma unroll
      for (int i = 0; i < L; ++i) vr[i] = vc[i] = wr[i] = wc[i] = 0.0;
    }
    if (S.t == 0) {
      S.tri[j] = x[j];
      S.tri[n + j] = beta;
      S.tri[2 * n + j] = tau;
    }
    if (S.wv == 0) {
      for (int r = lane; r < n; r += 64)
        S.refl[r + (int64_t)j * n] = (r == j + 1) ? 1.0 : (r > j + 1 ? x[r] * sv : 0.0);
    }
    if (j == 20) FPLS_TRD_STAMP(7);
  }
  // the last pending update before the caller drops a block
#pragma unroll
  for (int i = 0; i < L; ++i)
#pragma unroll
    for (int q = 0; q < L; ++q) a[i][q] = fma(-vr[i], wc[q], fma(-wr[i], vc[q], a[i][q]));
}

// tri (3 n + 1 doubles): d, e (e[n-1] = 0), tau, then the scale (max |A|) the
// matrix was divided by.  refl (n x n, ld n): column j is reflector j's vector
// v_j (zero above row j + 1, v_j(j + 1) = 1); columns n - 2, n - 1 are zero.
// flags[b] is cleared for the later stages.
__global__ __launch_bounds__(TD_T) void k_trd_reg(const double* __restrict__ A0, int64_t lda,
                                                  int64_t strideA, const int* __restrict__ ns,
                                                  int n0, double* __restrict__ tri0,
                                                  int64_t strideTri, double* __restrict__ refl0,
                                                  int64_t strideRefl, int* __restrict__ flags) {
  const int64_t bi = blockIdx.x;
  const int n = ns ? ns[bi] : n0;
  const double* __restrict__ A = A0 + bi * strideA;
  double* __restrict__ tri = tri0 + bi * strideTri;
  double* __restrict__ refl = refl0 + bi * strideRefl;
  __shared__ double xs[2][II_N];
  __shared__ double ps[II_N];
  __shared__ double red[TD_T / 64];
  __shared__ double kred[TD_T / 64];
  __shared__ double tail[3];
  const int t = threadIdx.x, tr = t >> 4, tc = t & 15, lane = t & 63, wv = t >> 6;
  if (t == 0 && flags) flags[bi] = 0;
  if (n < 1 || n > II_N) return;
  FPLS_TRD_STAMP(0);

  // symmetric part of A, cyclic register set, and its largest entry
  double a[TD_R][TD_R];
  double mx = 0.0;
#pragma unroll
  for (int i = 0; i < TD_R; ++i)
#pragma unroll
    for (int q = 0; q < TD_R; ++q) {
      const int r = tr + 16 * i, c = tc + 16 * q;
      double x = 0.0;
      if (r < n && c < n) x = 0.5 * (A[r + (int64_t)c * lda] + A[c + (int64_t)r * lda]);
      a[i][q] = x;
      mx = fmax(mx, fabs(x));
    }
  mx = wave_max_d(mx);
  if (lane == 0) red[wv] = mx;
  __syncthreads();
  const double mall = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  const double scale = (mall > 0.0 && isfinite(mall)) ? mall : 1.0;
  const double iscale = 1.0 / scale;
#pragma unroll
  for (int i = 0; i < TD_R; ++i)
#pragma unroll
    for (int q = 0; q < TD_R; ++q) a[i][q] *= iscale;
  if (t == 0) tri[3 * n] = scale;

  TrdState S;
  S.xs = xs;
  S.ps = ps;
  S.kred = kred;
  S.tri = tri;
  S.refl = refl;
  S.n = n;
  S.t = t;
  S.tr = tr;
  S.tc = tc;
  S.lane = lane;
  S.wv = wv;
  S.hb3 = (tc >> 3) & 1;
  S.hb2 = (tc >> 2) & 1;
  S.hb1 = (tc >> 1) & 1;
  S.hb0 = tc & 1;
  S.slot = S.hb3 * 4 + S.hb2 * 2 + S.hb1;

  FPLS_TRD_STAMP(1);
  // reflectors j = 0 .. n - 3, sixteen columns per register block
  int jb = 0;
  for (;;) {
    const int jend = min(16 * jb + 16, n - 2);
    if (16 * jb < jend) {
      switch (jb) {
        case 0: trd_block<8>(a, jb, jend, S); break;
        case 1: trd_block<7>(a, jb, jend, S); break;
        case 2: trd_block<6>(a, jb, jend, S); break;
        case 3: trd_block<5>(a, jb, jend, S); break;
        case 4: trd_block<4>(a, jb, jend, S); break;
        case 5: trd_block<3>(a, jb, jend, S); break;
        case 6: trd_block<2>(a, jb, jend, S); break;
        default: trd_block<1>(a, jb, jend, S); break;
      }
    }
    if (16 * (jb + 1) >= n - 2) break;
    // block jb of rows and columns is finished: drop it
#pragma unroll
    for (int i = 0; i < TD_R - 1; ++i)
#pragma unroll
      for (int q = 0; q < TD_R - 1; ++q) a[i][q] = a[i + 1][q + 1];
#pragma unroll
    for (int i = 0; i < TD_R; ++i) {
      a[i][TD_R - 1] = 0.0;
      a[TD_R - 1][i] = 0.0;
    }
    ++jb;
  }
  // the last 2 x 2 block (or the 1 x 1 matrix): d[n-2], e[n-2], d[n-1]
  {
    const int live = TD_R - jb;
#pragma unroll
    for (int i = 0; i < TD_R; ++i)
#pragma unroll
      for (int q = 0; q < TD_R; ++q) {
        if (i < live && q < live) {
          const int r = tr + 16 * (i + jb), c = tc + 16 * (q + jb);
          if (n >= 2 && r == n - 2 && c == n - 2) tail[0] = a[i][q];
          if (n >= 2 && r == n - 1 && c == n - 2) tail[1] = a[i][q];
          if (r == n - 1 && c == n - 1) tail[2] = a[i][q];
        }
      }
  }
  __syncthreads();
  if (t == 0) {
    if (n >= 2) {
      tri[n - 2] = tail[0];
      tri[n + n - 2] = tail[1];
      tri[2 * n + n - 2] = 0.0;
    }
    tri[n - 1] = tail[2];
    tri[n + n - 1] = 0.0;
    tri[2 * n + n - 1] = 0.0;
  }
  for (int idx = t; idx < 2 * n; idx += TD_T) {
    const int j = n - 2 + idx / n, r = idx % n;
    if (j >= 0) refl[r + (int64_t)j * n] = 0.0;
  }
  FPLS_TRD_STAMP(8);
}

// ---------------------------------------------------------------- stage 2
constexpr int PR_T = 256;              // threads per eigenvalue
constexpr int PR_PTS = 2 * PR_T;       // multisection points per step
constexpr int BT_B = 4;                // reflectors per back-transformation block
constexpr int BT_TS = 10;              // packed upper-triangular T of one block

// Number of eigenvalues of T below x0 and below x1 (two interleaved chains).
__device__ __forceinline__ void sturm_count2(const double* d, const double* ee, int n, double x0,
                                             double x1, double pivmin, int& c0, int& c1) {
  double q0 = d[0] - x0, q1 = d[0] - x1;
  if (fabs(q0) < pivmin) q0 = -pivmin;
  if (fabs(q1) < pivmin) q1 = -pivmin;
  int n0 = q0 < 0.0 ? 1 : 0, n1 = q1 < 0.0 ? 1 : 0;
  for (int i = 1; i < n; ++i) {
    const double di = d[i], ei = ee[i - 1];
    q0 = (di - x0) - ei * rcp1(q0);
    q1 = (di - x1) - ei * rcp1(q1);
    if (fabs(q0) < pivmin) q0 = -pivmin;
    if (fabs(q1) < pivmin) q1 = -pivmin;
    n0 += q0 < 0.0 ? 1 : 0;
    n1 += q1 < 0.0 ? 1 : 0;
  }
  c0 = n0;
  c1 = n1;
}

// lam: ascending eigenvalues of the scaled T (stride nmax per problem);
// w / V: descending eigenvalues times the scale and the back-transformed
// eigenvectors (ld n); flags[b] set when a vector fails its checks.
__global__ __launch_bounds__(PR_T) void k_trd_pairs(
    const int* __restrict__ ns, int n0, const double* __restrict__ tri0, int64_t strideTri,
    const double* __restrict__ refl0, int64_t strideRefl, double* __restrict__ w0,
    int64_t strideW, double* __restrict__ V0, int64_t strideV, double* __restrict__ lam0,
    int64_t strideLam, int kwant, int* __restrict__ flags) {
  const int64_t bi = blockIdx.y;
  const int n = ns ? ns[bi] : n0;
  const int k = blockIdx.x;
  if (k >= n || n > II_N) return;
  const double* __restrict__ tri = tri0 + bi * strideTri;
  const double* __restrict__ refl = refl0 + bi * strideRefl;
  double* __restrict__ w = w0 + bi * strideW;
  double* __restrict__ V = V0 + bi * strideV;
  double* __restrict__ lam = lam0 + bi * strideLam;
  __shared__ double d[II_N], e[II_N], ee[II_N];
  __shared__ double fac[2][II_N];
  __shared__ double zv[II_N];
  __shared__ double R[II_N * II_N];  // the reflectors, staged for the back-transformation
  __shared__ double Tm[BT_TS * ((II_N + BT_B - 1) / BT_B)];
  __shared__ unsigned long long above_w[PR_T / 64][2];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int i = t; i < n; i += PR_T) {
    d[i] = tri[i];
    e[i] = tri[n + i];
    ee[i] = tri[n + i] * tri[n + i];
  }
  if (n >= 3) {
    const int nr = n * (n - 2);  // columns 0 .. n - 3
    for (int base = t; base < nr; base += 4 * PR_T) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int idx = base + u * PR_T;
        v[u] = idx < nr ? refl[idx] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int idx = base + u * PR_T;
        if (idx < nr) R[idx] = v[u];
      }
    }
  }
  __syncthreads();
  FPLS_TRD_STAMP(16);
  const double scale = tri[3 * n];
  const TriBounds b = tri_bounds(d, e, ee, n);

  // ---- eigenvalue k (ascending): count(lo) <= k < count(hi)
  double lo = b.gl, hi = b.gu;
  for (int it = 0; it < 12; ++it) {
    const double width = hi - lo;
    const double tol = fmax(II_EPS * b.tnorm,
                            fmax(b.pivmin, 2.0 * II_EPS * fmax(fabs(lo), fabs(hi))));
    if (!(width > tol)) break;  // uniform over the workgroup
    const double step = width / (double)(PR_PTS + 1);
    // thread t evaluates points t + 1 and t + 1 + PR_T
    const double xa = lo + step * (double)(t + 1);
    const double xb = lo + step * (double)(t + 1 + PR_T);
    int ca, cb;
    sturm_count2(d, ee, n, xa, xb, b.pivmin, ca, cb);
    const unsigned long long ma = __ballot(ca > k), mb = __ballot(cb > k);
    if (lane == 0) {
      above_w[wv][0] = ma;
      above_w[wv][1] = mb;
    }
    __syncthreads();
    // first point (in point order) with more than k eigenvalues below it
    int f = PR_PTS;
    for (int h = 0; h < 2 && f == PR_PTS; ++h)
      for (int q = 0; q < PR_T / 64; ++q) {
        const unsigned long long m = above_w[q][h];
        if (m) {
          f = h * PR_T + 64 * q + __ffsll((long long)m) - 1;
          break;
        }
      }
    __syncthreads();
    const double xf = f < PR_PTS ? lo + step * (double)(f + 1) : hi;
    const double xb0 = f > 0 ? lo + step * (double)f : lo;
    lo = xb0;
    hi = xf;
  }
  const double lambda = 0.5 * (lo + hi);
  FPLS_TRD_STAMP(17);
  const int m = n - 2;                          // reflectors 0 .. n - 3
  const int nblk = m > 0 ? (m + BT_B - 1) / BT_B : 0;
  const double* __restrict__ tauv = tri + 2 * n;
  bool bad = false;
  double z0 = 0.0, z1 = 0.0;  // wave 0: this lane's entries lane, lane + 64
  if (wv != 0) {
    // compact-WY factors of the blocks of BT_B reflectors (dlarft, forward,
    // columnwise): H_j0 ... H_j0+3 = I - V T V^T, T upper triangular with
    // T[:a, a] = -tau_a T[:a, :a] (V[:, :a]^T v_a); blocks shared round-robin
    // by waves 1 .. 3 while wave 0 forms the eigenvector
    for (int bk = wv - 1; bk < nblk; bk += PR_T / 64 - 1) {
      const int j0 = BT_B * bk;
      double va0[BT_B], va1[BT_B];
#pragma unroll
      for (int u = 0; u < BT_B; ++u) {
        const bool on = j0 + u < m;
        va0[u] = (on && lane < n) ? R[lane + (j0 + u) * n] : 0.0;
        va1[u] = (on && lane + 64 < n) ? R[lane + 64 + (j0 + u) * n] : 0.0;
      }
      // G01 G02 G12 G03 G13 G23
      const double g01 = wave_sum_d(fma(va0[0], va0[1], va1[0] * va1[1]));
      const double g02 = wave_sum_d(fma(va0[0], va0[2], va1[0] * va1[2]));
      const double g12 = wave_sum_d(fma(va0[1], va0[2], va1[1] * va1[2]));
      const double g03 = wave_sum_d(fma(va0[0], va0[3], va1[0] * va1[3]));
      const double g13 = wave_sum_d(fma(va0[1], va0[3], va1[1] * va1[3]));
      const double g23 = wave_sum_d(fma(va0[2], va0[3], va1[2] * va1[3]));
      if (lane == 0) {
        const double t0 = j0 < m ? tauv[j0] : 0.0, t1 = j0 + 1 < m ? tauv[j0 + 1] : 0.0;
        const double t2 = j0 + 2 < m ? tauv[j0 + 2] : 0.0, t3 = j0 + 3 < m ? tauv[j0 + 3] : 0.0;
        const double T00 = t0, T11 = t1, T22 = t2, T33 = t3;
        const double T01 = -t1 * (T00 * g01);
        const double T02 = -t2 * (T00 * g02 + T01 * g12), T12 = -t2 * (T11 * g12);
        const double T03 = -t3 * (T00 * g03 + T01 * g13 + T02 * g23);
        const double T13 = -t3 * (T11 * g13 + T12 * g23), T23 = -t3 * (T22 * g23);
        double* Tb = Tm + BT_TS * bk;
        Tb[0] = T00; Tb[1] = T01; Tb[2] = T11; Tb[3] = T02; Tb[4] = T12;
        Tb[5] = T22; Tb[6] = T03; Tb[7] = T13; Tb[8] = T23; Tb[9] = T33;
      }
    }
  } else {
    if (lane == 0) lam[k] = lambda;
    if (n == 1) {
      z0 = lane == 0 ? 1.0 : 0.0;
    } else {
      // top-down (lane 0) and bottom-up (lane 1) pivots of T - lambda I
      if (lane < 2) {
        const bool up = lane == 1;
        int ii = up ? n - 1 : 0;
        double q = d[ii] - lambda;
        if (fabs(q) < b.pivmin) q = -b.pivmin;
        fac[lane][ii] = q;
        for (int s = 1; s < n; ++s) {
          ii = up ? n - 1 - s : s;
          const double e2 = up ? ee[ii] : ee[ii - 1];
          q = (d[ii] - lambda) - e2 * rcp1(q);
          if (fabs(q) < b.pivmin) q = -b.pivmin;
          fac[lane][ii] = q;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      // twist index: smallest |gamma_r|, lowest index on ties; then the walk
      // ratios e_i / D+_i (up) and e_{i-1} / D-_i (down) in place
      double gbest = DBL_MAX;
      int rbest = n;
      for (int r = lane; r < n; r += 64) {
        const double g0 = fac[0][r], g1 = fac[1][r];
        const double g = fabs(g0 + g1 - (d[r] - lambda));
        if (g < gbest) {
          gbest = g;
          rbest = r;
        }
        fac[0][r] = r + 1 < n ? e[r] * rcp1(g0) : 0.0;
        fac[1][r] = r >= 1 ? e[r - 1] * rcp1(g1) : 0.0;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double og = __shfl_xor(gbest, o, 64);
        const int orr = __shfl_xor(rbest, o, 64);
        if (og < gbest || (og == gbest && orr < rbest)) {
          gbest = og;
          rbest = orr;
        }
      }
      const int rt = rbest < n ? rbest : 0;  // a NaN gamma everywhere: any index
      FPLS_TRD_STAMP(18);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      // z_rt = 1; lane 0 walks up (z_i = -(e_i / D+_i) z_{i+1}), lane 1 down
      // (z_i = -(e_{i-1} / D-_i) z_{i-1}); when the neighbour is exactly zero
      // the equation of the row between gives z_i from the entry two steps back
      if (lane < 2) {
        const bool down = lane == 1;
        const double* rat = fac[lane];
        if (lane == 0) zv[rt] = 1.0;
        const int steps = down ? n - 1 - rt : rt;
        double zp = 1.0, zpp = 0.0;
        int s = 1;
        for (; s + 3 <= steps; s += 4) {
          int ii[4];
          double rr[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            ii[u] = down ? rt + s + u : rt - s - u;
            rr[u] = rat[ii[u]];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            double z = -rr[u] * zp;
            if (zp == 0.0 && s + u >= 2) {
              const int i = ii[u];
              const double ei = down ? e[i - 1] : e[i], efar = down ? e[i - 2] : e[i + 1];
              z = ei != 0.0 ? -(efar / ei) * zpp : 0.0;
            }
            zv[ii[u]] = z;
            zpp = zp;
            zp = z;
          }
        }
        for (; s <= steps; ++s) {
          const int i = down ? rt + s : rt - s;
          double z = -rat[i] * zp;
          if (zp == 0.0 && s >= 2) {
            const double ei = down ? e[i - 1] : e[i], efar = down ? e[i - 2] : e[i + 1];
            z = ei != 0.0 ? -(efar / ei) * zpp : 0.0;
          }
          zv[i] = z;
          zpp = zp;
          zp = z;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      z0 = lane < n ? zv[lane] : 0.0;
      z1 = lane + 64 < n ? zv[lane + 64] : 0.0;
      // unit length, largest entry positive (lowest index on ties)
      double amax = -1.0, av = 0.0;
      int ai = n;
      if (lane < n && fabs(z0) > amax) {
        amax = fabs(z0);
        av = z0;
        ai = lane;
      }
      if (lane + 64 < n && fabs(z1) > amax) {
        amax = fabs(z1);
        av = z1;
        ai = lane + 64;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ou = __shfl_xor(amax, o, 64), ov = __shfl_xor(av, o, 64);
        const int oi = __shfl_xor(ai, o, 64);
        if (ou > amax || (ou == amax && oi < ai)) {
          amax = ou;
          av = ov;
          ai = oi;
        }
      }
      const double nrm2 = wave_sum_d(fma(z0, z0, z1 * z1));
      if (!(nrm2 > 0.0) || !isfinite(nrm2)) bad = true;
      const double inv = (av < 0.0 ? -1.0 : 1.0) / sqrt(nrm2);
      z0 *= inv;
      z1 *= inv;
      // residual |(T - lambda I) z|_inf against 1000 n eps |T|
      zv[lane] = z0;
      if (lane + 64 < II_N) zv[lane + 64] = z1;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      double res = 0.0;
      for (int i = lane; i < n; i += 64) {
        double ri = (d[i] - lambda) * zv[i];
        if (i > 0) ri = fma(e[i - 1], zv[i - 1], ri);
        if (i + 1 < n) ri = fma(e[i], zv[i + 1], ri);
        res = fmax(res, fabs(ri));
      }
      res = wave_max_d(res);
      if (!(res <= 1000.0 * (double)n * II_EPS * fmax(b.tnorm, DBL_MIN))) bad = true;
    }
  }
  __syncthreads();  // the T factors are in LDS
  if (wv != 0) return;
  FPLS_TRD_STAMP(19);
  // back-transformation z <- H_0 ... H_{n-3} z, block by block from the last:
  // z <- z - V (T (V^T z))
  for (int bk = nblk - 1; bk >= 0; --bk) {
    const int j0 = BT_B * bk;
    double va0[BT_B], va1[BT_B], sa[BT_B];
#pragma unroll
    for (int u = 0; u < BT_B; ++u) {
      const bool on = j0 + u < m;
      va0[u] = (on && lane < n) ? R[lane + (j0 + u) * n] : 0.0;
      va1[u] = (on && lane + 64 < n) ? R[lane + 64 + (j0 + u) * n] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < BT_B; ++u) sa[u] = wave_sum_d(fma(va0[u], z0, va1[u] * z1));
    const double* Tb = Tm + BT_TS * bk;
    const double u0 = fma(Tb[0], sa[0], fma(Tb[1], sa[1], fma(Tb[3], sa[2], Tb[6] * sa[3])));
    const double u1 = fma(Tb[2], sa[1], fma(Tb[4], sa[2], Tb[7] * sa[3]));
    const double u2 = fma(Tb[5], sa[2], Tb[8] * sa[3]);
    const double u3 = Tb[9] * sa[3];
    z0 = fma(-va0[0], u0, fma(-va0[1], u1, fma(-va0[2], u2, fma(-va0[3], u3, z0))));
    z1 = fma(-va1[0], u0, fma(-va1[1], u1, fma(-va1[2], u2, fma(-va1[3], u3, z1))));
  }
  FPLS_TRD_STAMP(20);
  const int oc = n - 1 - k;
  if (lane < n) V[lane + (int64_t)oc * n] = z0;
  if (lane + 64 < n) V[lane + 64 + (int64_t)oc * n] = z1;
  if (lane == 0) {
    w[oc] = lambda * scale;
    // only the wanted (largest kwant) eigenpairs can send the problem to the
    // fallback solver
    const bool wanted = kwant <= 0 || oc < kwant;
    if (wanted && (bad || !isfinite(lambda))) flags[bi] = 1;
  }
}

// ---------------------------------------------------------------- stage 3
// Modified Gram-Schmidt inside each cluster of eigenvalues closer than
// 1e-3 |T| (one wave per cluster, led by its first member), in ascending
// eigenvalue order: the members were computed independently from their own
// eigenvalues, each accurate to eps |T| / gap, and are made orthogonal here.
// Orthogonality between clusters is then within eps |T| / gap <= 2e-13.
constexpr int CL_W = 4;  // waves per workgroup
__global__ __launch_bounds__(64 * CL_W) void k_trd_clorth(
    const int* __restrict__ ns, int n0, const double* __restrict__ tri0, int64_t strideTri,
    const double* __restrict__ lam0, int64_t strideLam, double* __restrict__ V0,
    int64_t strideV, int kwant, int* __restrict__ flags) {
  const int64_t bi = blockIdx.y;
  const int n = ns ? ns[bi] : n0;
  if (n > II_N) return;
  const double* __restrict__ tri = tri0 + bi * strideTri;
  const double* __restrict__ lam = lam0 + bi * strideLam;
  double* __restrict__ V = V0 + bi * strideV;
  __shared__ double d[II_N], e[II_N], ee[II_N];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    d[i] = tri[i];
    e[i] = tri[n + i];
    ee[i] = tri[n + i] * tri[n + i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int k0 = blockIdx.x * CL_W + (threadIdx.x >> 6);
  if (k0 >= n) return;
  const TriBounds b = tri_bounds(d, e, ee, n);
  const double ortol = 1e-3 * b.onenrm, tight = 1e-9 * b.onenrm;
  if (k0 > 0 && lam[k0] - lam[k0 - 1] <= ortol) return;
  int cl = 1;
  while (k0 + cl < n && lam[k0 + cl] - lam[k0 + cl - 1] <= ortol) ++cl;
  if (cl == 1) return;
  bool badc = cl > CL_MAX;
  for (int q = k0; q + 1 < k0 + cl && !badc; ++q) badc = lam[q + 1] - lam[q] <= tight;
  if (badc) {
    // a cluster among the unwanted (smallest) eigenvalues keeps its vectors
    const bool wanted = kwant <= 0 || k0 + cl - 1 >= n - kwant;
    if (wanted && lane == 0) flags[bi] = 1;
    return;
  }
  for (int m = 1; m < cl; ++m) {
    double* zm = V + (int64_t)(n - 1 - (k0 + m)) * n;
    double u0 = lane < n ? zm[lane] : 0.0, u1 = lane + 64 < n ? zm[lane + 64] : 0.0;
    for (int r = 0; r < m; ++r) {
      const double* zr = V + (int64_t)(n - 1 - (k0 + r)) * n;
      const double v0 = lane < n ? zr[lane] : 0.0, v1 = lane + 64 < n ? zr[lane + 64] : 0.0;
      const double sr = wave_sum_d(fma(u0, v0, u1 * v1));
      u0 = fma(-sr, v0, u0);
      u1 = fma(-sr, v1, u1);
    }
    const double inv = 1.0 / sqrt(wave_sum_d(fma(u0, u0, u1 * u1)));
    if (lane < n) zm[lane] = u0 * inv;
    if (lane + 64 < n) zm[lane + 64] = u1 * inv;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

}  // namespace

void sym_eig_tri(Ctx* ctx, const double* A, int64_t lda, int64_t strideA, const int* ns_dev,
                 int n0, int nmax, int count, double* w, int64_t strideW, double* V,
                 int64_t strideV, int kwant, int* flags) {
  FPLS_REQUIRE(nmax >= 1 && nmax <= II_N && count >= 1, FPLS_ERR_ARG,
               "sym_eig_tri: orders must be in 1..128");
  const int64_t sTri = 3 * (int64_t)nmax + 1, sRefl = (int64_t)nmax * nmax, sLam = nmax;
  DBuf<double> tri(ctx, (size_t)sTri * count), refl(ctx, (size_t)sRefl * count),
      lam(ctx, (size_t)sLam * count);
  launch(ctx, "sym_eig_trd", k_trd_reg, dim3((unsigned)count), dim3(TD_T), 0, A, lda, strideA,
         ns_dev, n0, tri.get(), sTri, refl.get(), sRefl, flags);
  launch(ctx, "sym_eig_pairs", k_trd_pairs, dim3((unsigned)nmax, (unsigned)count), dim3(PR_T),
         0, ns_dev, n0, (const double*)tri.get(), sTri, (const double*)refl.get(), sRefl, w,
         strideW, V, strideV, lam.get(), sLam, kwant, flags);
  launch(ctx, "sym_eig_clorth", k_trd_clorth,
         dim3((unsigned)((nmax + CL_W - 1) / CL_W), (unsigned)count), dim3(64 * CL_W), 0, ns_dev,
         n0, (const double*)tri.get(), sTri, (const double*)lam.get(), sLam, V, strideV, kwant,
         flags);
}

}  // namespace fpls
