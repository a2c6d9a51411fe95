// Symmetric eigensolver for the small projected problems (Gram matrices of the
// cross-covariance and of rSVD sketches), n <= 128, in ONE workgroup of 1024
// threads with the matrix in LDS.  Three stages in one launch:
//
//  1. Householder reduction to tridiagonal form T = Q^T A Q (lower storage,
//     reflectors H_j = I - tau_j v_j v_j^T, v_j(j+1) = 1); Golub & Van Loan,
//     Algorithm 8.3.1.  Three barriers per column.
//  2. Divide and conquer on T from 1x1 leaves (Cuppen 1981): every merge of two
//     sorted halves is the rank-one problem D + rho z z^T.  Deflation of tiny
//     z_i and of close pairs by Givens rotations (the tests of LAPACK's
//     dlaed2), roots of the secular equation 1 + rho sum z_i^2/(d_i - lam)
//     with a two-pole rational model (Bunch, Nielsen & Sorensen 1978) kept in a
//     bisection bracket and measured from the nearer pole, then z recomputed
//     from the roots (Gu & Eisenstat 1994) so the eigenvectors come out
//     orthogonal.  All merges of a level run at once; 8 lanes per root.
//  3. Back-transformation V = H_0 ... H_{n-3} Z with each eigenvector column
//     held in the registers of 8 lanes (no barriers).
//
// Eigenvalues are returned in descending order with their eigenvectors (ld n).
#include <cfloat>
#include <cmath>

#include "fpls_internal.h"
#include "fpls_wave.h"

// Timing hook for stage-level measurements; compiled out unless a probe build
// defines it (it then records a clock value from thread 0 at each call site).
#ifndef FPLS_EIG_STAMP
#define FPLS_EIG_STAMP(slot)
#endif

namespace fpls {

namespace {

constexpr int EN = 128;           // largest order
constexpr int ELD = EN + 1;       // LDS leading dimension of A in stage 1
constexpr int ENT = 1024;         // threads
constexpr double EEPS = DBL_EPSILON;

__device__ __forceinline__ double sum8(double v) { return grp8_sum(v); }
// 1/d from the hardware reciprocal and two Newton steps (a few ulp; used only
// where d is bounded away from zero: secular-equation terms and their ratios)
__device__ __forceinline__ double rcp_nr(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = fma(-d, y, 1.0);
  y = fma(y, e, y);
  e = fma(-d, y, 1.0);
  return fma(y, e, y);
}
__device__ __forceinline__ double prod8(double v) { return grp8_prod(v); }
__device__ __forceinline__ int isum8(int v) { return grp8_isum(v); }
__device__ __forceinline__ double wave_sum(double v) { return wave_sum_d(v); }
__device__ __forceinline__ double wave_max(double v) { return wave_max_d(v); }

// first index in sorted x[lo, hi) with x >= v (strict = false) or x > v (strict = true)
__device__ __forceinline__ int bound(const double* x, int lo, int hi, double v, bool strict) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const bool go_right = strict ? (x[mid] <= v) : (x[mid] < v);
    if (go_right) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

}  // namespace

// Qa, Qb, Vr: n x n FP64 workspace (ld n).
// Batched: workgroup b solves problem b (A + b strideA, order ns[b] or n0,
// results at w + b strideW and V + b strideV with ld = its order, workspace of
// 3 nmax^2 doubles at ws + b strideWs).
// tri_out (when set): stop after stage 1 and write d (n), e (n), tau (n) and
// the scale of the problem at tri_out; the reflectors stay in the workspace
// (Vr, ld n).  run_flag (when set): return at once unless run_flag[b] != 0.
// zero_flag (when set) is cleared (the fallback flag the later kernels set).
__global__ __launch_bounds__(ENT) void k_sym_eig_dc(const double* __restrict__ Ain0, int64_t lda,
                                                    int64_t strideA, const int* __restrict__ ns,
                                                    int n0, int nmax, double* __restrict__ w0,
                                                    int64_t strideW, double* __restrict__ Vout0,
                                                    int64_t strideV, double* __restrict__ ws,
                                                    int64_t strideWs,
                                                    double* __restrict__ tri_out,
                                                    const int* __restrict__ run_flag,
                                                    int* __restrict__ zero_flag) {
  if (run_flag && run_flag[blockIdx.x] == 0) return;
  if (zero_flag && threadIdx.x == 0 && blockIdx.x == 0) *zero_flag = 0;
  const int64_t bi = blockIdx.x;
  const int n = ns ? ns[bi] : n0;
  const double* __restrict__ Ain = Ain0 + bi * strideA;
  double* __restrict__ w = w0 + bi * strideW;
  double* __restrict__ Vout = Vout0 + bi * strideV;
  double* __restrict__ Qa = ws + bi * strideWs;
  double* __restrict__ Qb = Qa + (int64_t)nmax * nmax;
  double* __restrict__ Vr = Qb + (int64_t)nmax * nmax;
  __shared__ double big[EN * ELD];  // stage 1: A; stage 2: W (ld n); stage 3: reflectors (ld n)
  __shared__ double dv[EN], ev[EN], tauv[EN], pv[EN];
  __shared__ double lam[EN], Ds[EN], zs[EN], taus[EN], zh[EN], invn[EN], val[EN];
  __shared__ double rc[EN], rs[EN], dnc[EN], znc[EN], dorgv[EN];
  __shared__ int perm[EN], defl[EN], ndidx[EN], rootidx[EN], orig[EN], colpos[EN];
  __shared__ int rp[EN], rq[EN], iperm[EN];
  __shared__ int nrot[EN / 2], Kb[EN / 2], needser[EN / 2];
  __shared__ unsigned long long bmaxv[EN / 2];
  __shared__ double red[ENT / 64];
  __shared__ double s_scale;

  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;

  // ---------------------------------------------------------------- scaling
  double mx = 0.0;
  for (int idx = t; idx < n * n; idx += ENT) {
    const int i = idx % n, j = idx / n;
    mx = fmax(mx, fabs(Ain[i + (int64_t)j * lda]));
  }
  mx = wave_max(mx);
  if (lane == 0) red[wv] = mx;
  __syncthreads();
  if (t == 0) {
    double m = 0.0;
    for (int k = 0; k < ENT / 64; ++k) m = fmax(m, red[k]);
    s_scale = (m > 0.0 && isfinite(m)) ? m : 1.0;
  }
  __syncthreads();
  const double scale = s_scale, iscale = 1.0 / scale;
  double* A = big;
  for (int idx = t; idx < n * n; idx += ENT) {
    const int i = idx % n, j = idx / n;
    A[i + j * ELD] = 0.5 * (Ain[i + (int64_t)j * lda] + Ain[j + (int64_t)i * lda]) * iscale;
  }
  if (t < EN) {
    tauv[t] = 0.0;
    perm[t] = t;  // every index array starts in range
    iperm[t] = 0;
    colpos[t] = 0;
  }
  __syncthreads();

  FPLS_EIG_STAMP(0);
  // ------------------------------------------------ stage 1: tridiagonalisation
  // Only the lower triangle of the trailing block is kept current.  Wave 0
  // updates the first trailing column itself and goes straight on to the next
  // reflector, so a column costs two barriers.
  auto householder = [&](int j) {  // wave 0 only
    const int i0 = j + 1 + lane, i1 = j + 1 + lane + 64;
    const double x0 = (i0 < n && i0 > j + 1) ? A[i0 + j * ELD] : 0.0;
    const double x1 = (i1 < n) ? A[i1 + j * ELD] : 0.0;
    const double xn2 = wave_sum(x0 * x0 + x1 * x1);
    const double alpha = A[(j + 1) + j * ELD];
    double beta = alpha, tau = 0.0, sv = 0.0;
    if (xn2 > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tau = (beta - alpha) / beta;
      sv = 1.0 / (alpha - beta);
    }
    if (i0 < n && i0 > j + 1) A[i0 + j * ELD] = x0 * sv;
    if (i1 < n) A[i1 + j * ELD] = x1 * sv;
    if (lane == 0) {
      dv[j] = A[j + j * ELD];
      ev[j] = beta;
      tauv[j] = tau;
      A[(j + 1) + j * ELD] = 1.0;
    }
  };
  if (wv == 0 && n >= 3) householder(0);
  __syncthreads();
  for (int j = 0; j + 2 < n; ++j) {
    const double tau = tauv[j];
    if (tau != 0.0) {  // uniform branch; H_j = I leaves A22 unchanged
      // p = tau A22 v, 8 lanes per row i: A22(i, k) is A(i, k) for k <= i
      // (row i) and A(k, i) for k > i (column i)
      {
        const int i = j + 1 + (t >> 3), g = t & 7;
        double s = 0.0;
        if (i < n) {
          int k = j + 1 + g;
          for (; k <= i; k += 8) s += A[i + k * ELD] * A[k + j * ELD];
          for (; k < n; k += 8) s += A[k + i * ELD] * A[k + j * ELD];
        }
        s = sum8(s);
        if (i < n && g == 0) pv[i] = tau * s;
      }
      __syncthreads();
      // K = tau/2 p^T v (every wave computes the same sum)
      const int i0 = j + 1 + lane, i1 = j + 1 + lane + 64;
      double s = 0.0;
      if (i0 < n) s += pv[i0] * A[i0 + j * ELD];
      if (i1 < n) s += pv[i1] * A[i1 + j * ELD];
      const double K = 0.5 * tau * wave_sum(s);
      if (wv == 0) {
        // first trailing column (kk = j + 1), rows ii >= kk
        const int kk = j + 1;
        const double vk = 1.0, wk = pv[kk] - K;
        for (int ii = kk + lane; ii < n; ii += 64) {
          const double vi = A[ii + j * ELD], wi = pv[ii] - K * vi;
          A[ii + kk * ELD] -= vi * wk + wi * vk;
        }
      } else {
        // remaining columns kk >= j + 2: 30 groups of 32 lanes
        const int u = t - 64, gx = u & 31, gy = u >> 5;
        for (int kk = j + 2 + gy; kk < n; kk += 30) {
          const double vk = A[kk + j * ELD], wk = pv[kk] - K * vk;
          for (int ii = kk + gx; ii < n; ii += 32) {
            const double vi = A[ii + j * ELD], wi = pv[ii] - K * vi;
            A[ii + kk * ELD] -= vi * wk + wi * vk;
          }
        }
      }
    }
    if (wv == 0 && j + 3 < n) {
      // the wave's own column writes must land before it reads them back
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      householder(j + 1);
    }
    __syncthreads();
  }
  if (t == 0) {
    if (n >= 2) {
      dv[n - 2] = A[(n - 2) + (n - 2) * ELD];
      ev[n - 2] = A[(n - 1) + (n - 2) * ELD];
    }
    dv[n - 1] = A[(n - 1) + (n - 1) * ELD];
  }
  // reflectors to the workspace (stage 2 reuses the LDS)
  const int tx = t & 31, ty = t >> 5;
  for (int j = ty; j < n; j += 32) {
    for (int i = tx; i < n; i += 32) {
      const int idx = i + j * n;
      Vr[idx] = (j + 2 < n && i > j) ? A[i + j * ELD] : 0.0;
      Qa[idx] = (i == j) ? 1.0 : 0.0;
      Qb[idx] = 0.0;
    }
  }
  __syncthreads();
  if (tri_out) {
    for (int i = t; i < n; i += ENT) {
      tri_out[i] = dv[i];
      tri_out[n + i] = i + 1 < n ? ev[i] : 0.0;
      tri_out[2 * n + i] = tauv[i];
    }
    if (t == 0) tri_out[3 * n] = scale;
    return;
  }

  FPLS_EIG_STAMP(1);
  // ------------------------------------------------ stage 2: divide and conquer
  // 1x1 leaves after splitting T at every off-diagonal: T = blockdiag + sum of
  // |e_s| v_s v_s^T with v_s = e_s + sign(e_s) e_{s+1}
  if (t < n) {
    double l = dv[t];
    if (t > 0) l -= fabs(ev[t - 1]);
    if (t + 1 < n) l -= fabs(ev[t]);
    lam[t] = l;
  }
  __syncthreads();
  double* W = big;
  double* Qs = Qa;
  double* Qd = Qb;
  for (int lg = 1; (1 << (lg - 1)) < n; ++lg) {
    const int size = 1 << (lg - 1), bsz = 1 << lg;
    const int nblk = (n + bsz - 1) >> lg;
    // (a) sorted merge of the two halves, z from the boundary rows of Q
    if (t < n) {
      const int a = (t >> lg) << lg, m = a + size, b = min(a + bsz, n);
      if (m < n) {
        const double beta = ev[m - 1];
        const double sg = beta >= 0.0 ? 1.0 : -1.0;
        const double v = lam[t];
        int rank;
        double z;
        if (t < m) {
          rank = (t - a) + (bound(lam, m, b, v, false) - m);
          z = Qs[(m - 1) + (int64_t)t * n];
        } else {
          rank = (t - m) + (bound(lam, a, m, v, true) - a);
          z = sg * Qs[m + (int64_t)t * n];
        }
        Ds[a + rank] = v;
        zs[a + rank] = z * 0.70710678118654752440;
        perm[a + rank] = t;
        iperm[t] = rank;
      }
    }
    __syncthreads();
    FPLS_EIG_STAMP(2);
    // (b) deflation.  A parallel pre-pass finds the block max |d| (for the
    //     tolerance), the tiny-z flags, and tests every surviving z against its
    //     surviving predecessor with the unrotated values.  When no pair of a
    //     merge passes that test the serial scan would rotate nothing, so the
    //     survivors are packed in parallel; otherwise one thread runs the
    //     serial scan (dlaed2's order).  K survivors go to dnc / znc.
    if (t < nblk) {
      bmaxv[t] = 0ull;
      needser[t] = 0;
    }
    __syncthreads();
    if (t < n && ((t >> lg) << lg) + size < n)
      atomicMax(&bmaxv[t >> lg], (unsigned long long)__double_as_longlong(fabs(Ds[t])));
    __syncthreads();
    auto tol_of = [&](int bi) {
      const int m = (bi << lg) + size;
      const double rho2 = 2.0 * fabs(ev[m - 1]);
      return 8.0 * EEPS * fmax(__longlong_as_double((long long)bmaxv[bi]), rho2);
    };
    if (t < n) {
      const int bi = t >> lg, m = (bi << lg) + size;
      if (m < n) defl[t] = (2.0 * fabs(ev[m - 1]) * fabs(zs[t]) <= tol_of(bi)) ? 1 : 0;
    }
    __syncthreads();
    if (t < n) {
      const int bi = t >> lg, a = bi << lg, m = a + size;
      if (m < n && !defl[t]) {
        int pv = t - 1;
        while (pv >= a && defl[pv]) --pv;
        if (pv >= a) {
          const double zp = zs[pv], zq = zs[t];
          const double it = 1.0 / sqrt(zp * zp + zq * zq);
          if (fabs((Ds[t] - Ds[pv]) * (zq * it) * (zp * it)) <= tol_of(bi)) needser[bi] = 1;
        }
      }
    }
    __syncthreads();
    if (t < nblk) {
      const int a = t << lg, m = a + size, b = min(a + bsz, n), mm = b - a;
      int K = 0, nr = 0;
      if (m < n && needser[t]) {
        const double tol = tol_of(t);
        const double rho2 = 2.0 * fabs(ev[m - 1]);
        // the state of the current candidate (prev) stays in registers; entries
        // q are read before anything writes them, so the loads pipeline
        int prev = -1;
        double dprev = 0.0, zprev = 0.0;
        for (int q = 0; q < mm; ++q) {
          const double dq = Ds[a + q], zq = zs[a + q];
          defl[a + q] = 0;
          if (rho2 * fabs(zq) <= tol) {
            defl[a + q] = 1;
            continue;
          }
          if (prev >= 0) {
            const double tt = sqrt(zprev * zprev + zq * zq), it = 1.0 / tt;
            const double c = zq * it, sn = -zprev * it;
            if (fabs((dq - dprev) * c * sn) <= tol) {
              // rotate (prev, q): z_prev -> 0, z_q -> tt; prev deflates
              const double dp2 = c * c * dprev + sn * sn * dq;
              const double dq2 = sn * sn * dprev + c * c * dq;
              Ds[a + prev] = dp2;
              zs[a + prev] = 0.0;
              defl[a + prev] = 1;
              rp[a + nr] = prev;
              rq[a + nr] = q;
              rc[a + nr] = c;
              rs[a + nr] = sn;
              ++nr;
              Ds[a + q] = dq2;
              zs[a + q] = tt;
              prev = q;
              dprev = dq2;
              zprev = tt;
              continue;
            }
            // prev survives
            ndidx[a + K] = prev;
            rootidx[a + prev] = K;
            dnc[a + K] = dprev;
            znc[a + K] = zprev;
            ++K;
          }
          prev = q;
          dprev = dq;
          zprev = zq;
        }
        if (prev >= 0) {
          ndidx[a + K] = prev;
          rootidx[a + prev] = K;
          dnc[a + K] = dprev;
          znc[a + K] = zprev;
          ++K;
        }
      } else if (m < n) {
        for (int q = 0; q < mm; ++q) K += defl[a + q] ? 0 : 1;
      }
      Kb[t] = K;
      nrot[t] = nr;
    }
    if (t < n) {
      const int bi = t >> lg, a = bi << lg, m = a + size;
      if (m < n && !needser[bi] && !defl[t]) {
        int r = 0;
        for (int q = a; q < t; ++q) r += defl[q] ? 0 : 1;
        ndidx[a + r] = t - a;
        rootidx[t] = r;
        dnc[a + r] = Ds[t];
        znc[a + r] = zs[t];
      }
    }
    __syncthreads();
    const int slot = t >> 3, g = t & 7;
    const int sbi = slot >> lg, sa = sbi << lg, sm = sa + size;
    const int sj = slot - sa;
    const bool spair = slot < n && sm < n;
    const int sK = spair ? Kb[sbi] : 0;
    const double* Dn = dnc + sa;
    const double* Zn = znc + sa;
    FPLS_EIG_STAMP(3);
    // (c) secular roots, 8 lanes per root
    if (spair && sj < sK) {
      const int K = sK, j = sj;
      const double rho2 = 2.0 * fabs(ev[sm - 1]);
      int o;
      double lo, hi;
      const double dj = Dn[j];
      if (j < K - 1) {
        const double delta = 0.5 * (Dn[j + 1] - dj);
        double s = 0.0;
        for (int k = g; k < K; k += 8) s += Zn[k] * Zn[k] * rcp_nr((Dn[k] - dj) - delta);
        const double f = 1.0 + rho2 * sum8(s);
        if (f >= 0.0) {
          o = j;
          lo = 0.0;
          hi = delta;
        } else {
          o = j + 1;
          lo = -delta;
          hi = 0.0;
        }
      } else {
        double s = 0.0;
        for (int k = g; k < K; k += 8) s += Zn[k] * Zn[k];
        o = j;
        lo = 0.0;
        hi = rho2 * sum8(s);
      }
      const double dorg = Dn[o];
      double x = 0.5 * (lo + hi);
      for (int it = 0; it < 200; ++it) {
        double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0, eab = 0.0;
        for (int k = g; k < K; k += 8) {
          const double zk = Zn[k];
          const double r = zk * rcp_nr((Dn[k] - dorg) - x);
          const double zr = zk * r, rr = r * r;
          if (k <= j) {
            psi += zr;
            dpsi += rr;
          } else {
            phi += zr;
            dphi += rr;
          }
          eab += fabs(zr);
        }
        psi = rho2 * sum8(psi);
        dpsi = rho2 * sum8(dpsi);
        phi = rho2 * sum8(phi);
        dphi = rho2 * sum8(dphi);
        eab = rho2 * sum8(eab);
        const double f = 1.0 + psi + phi;
        if (fabs(f) <= 2.0 * EEPS * (1.0 + eab)) break;
        if (f < 0.0) lo = x;
        else hi = x;
        if (hi - lo <= 2.0 * EEPS * fmax(fabs(lo), fabs(hi))) break;
        // rational model: psi ~ a1 + b1/(dl - h), phi ~ a2 + b2/(dr - h)
        const double dl = (dj - dorg) - x;
        const double b1 = dpsi * dl * dl, a1 = psi - b1 / dl;
        double h;
        if (j < K - 1) {
          const double dr = (Dn[j + 1] - dorg) - x;
          const double b2 = dphi * dr * dr, a2 = phi - b2 / dr;
          const double c = 1.0 + a1 + a2;
          const double qb = -(c * (dl + dr) + b1 + b2);
          const double qc = c * dl * dr + b1 * dr + b2 * dl;
          if (c == 0.0) {
            h = qb != 0.0 ? -qc / qb : NAN;
          } else {
            const double disc = fmax(qb * qb - 4.0 * c * qc, 0.0);
            const double q = -0.5 * (qb + copysign(sqrt(disc), qb));
            const double h1 = q / c, h2 = q != 0.0 ? qc / q : NAN;
            h = (h1 > dl && h1 < dr) ? h1 : h2;
          }
        } else {
          const double c = 1.0 + a1 + phi;
          h = c != 0.0 ? dl + b1 / c : NAN;
        }
        double xn = x + h;
        if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
        x = xn;
      }
      if (g == 0) {
        orig[sa + j] = o;
        taus[sa + j] = x;
        dorgv[sa + j] = dorg;
      }
    }
    __syncthreads();
    FPLS_EIG_STAMP(4);
    // (d) Gu-Eisenstat z-hat, 8 lanes per surviving pole i:
    //     zhat_i^2 = prod_j (lam_j - d_i) / (rho prod_{k != i} (d_k - d_i))
    if (spair && sj < sK) {
      const int K = sK, i = sj;
      const double rho2 = 2.0 * fabs(ev[sm - 1]);
      const double di = Dn[i];
      // lam_j - d_i = (D_o(j) - d_i) + tau_j
      double pr = 1.0;
      for (int k = g; k < K; k += 8) {
        if (k < i) pr *= ((dorgv[sa + k] - di) + taus[sa + k]) * rcp_nr(Dn[k] - di);
        else if (k > i) pr *= ((dorgv[sa + k - 1] - di) + taus[sa + k - 1]) * rcp_nr(Dn[k] - di);
      }
      pr = prod8(pr) * (((dorgv[sa + K - 1] - di) + taus[sa + K - 1]) / rho2);
      if (g == 0) zh[sa + i] = copysign(sqrt(fmax(pr, 0.0)), Zn[i]);
    }
    __syncthreads();
    FPLS_EIG_STAMP(5);
    // (e) eigenvector norms and output eigenvalues (sorted-input index space)
    if (spair) {
      if (sj < sK) {
        const double dorg = dorgv[sa + sj], tj = taus[sa + sj];
        double s = 0.0;
        for (int i = g; i < sK; i += 8) {
          const double u = zh[sa + i] * rcp_nr((Dn[i] - dorg) - tj);
          s += u * u;
        }
        s = sum8(s);
        if (g == 0) {
          invn[sa + sj] = 1.0 / sqrt(s);
          val[sa + ndidx[sa + sj]] = dorg + tj;
        }
      }
      if (g == 0 && sa + sj < n && defl[sa + sj]) val[sa + sj] = Ds[sa + sj];
    }
    __syncthreads();
    // (f) output order (ascending, ties by index) and the small matrix W
    if (spair) {
      const int mm = min(sa + bsz, n) - sa;
      const double vc = val[slot];
      int r = 0;
      for (int k = g; k < mm; k += 8) {
        const double vk = val[sa + k];
        r += (vk < vc || (vk == vc && k < sj)) ? 1 : 0;
      }
      r = isum8(r);
      if (g == 0) colpos[slot] = r;
    }
    for (int idx = t; idx < (n << lg); idx += ENT) {
      const int c = idx >> lg, ql = idx & (bsz - 1);
      const int a = (c >> lg) << lg, q = a + ql;
      if (q >= n || a + size >= n) continue;
      const int cl = c - a;
      double wqc;
      if (defl[c]) {
        wqc = (ql == cl) ? 1.0 : 0.0;
      } else if (defl[q]) {
        wqc = 0.0;
      } else {
        const int jr = rootidx[c], i = rootidx[q];
        wqc = zh[a + i] * rcp_nr((Ds[q] - dorgv[a + jr]) - taus[a + jr]) * invn[a + jr];
      }
      W[q + c * n] = wqc;
    }
    __syncthreads();
    FPLS_EIG_STAMP(6);
    // (g) deflation rotations folded into W: W <- G_1 (G_2 (... W)), rows p, q
    {
      int maxr = 0;
      for (int k = 0; k < nblk; ++k) maxr = max(maxr, nrot[k]);
      for (int r = maxr - 1; r >= 0; --r) {
        if (t < n) {
          const int bi = t >> lg, a = bi << lg;
          if (a + size < n && r < nrot[bi]) {
            const int c = t;  // column of W
            const int p = a + rp[a + r], q = a + rq[a + r];
            const double cs = rc[a + r], sn = rs[a + r];
            const double wp = W[p + c * n], wq = W[q + c * n];
            W[p + c * n] = cs * wp - sn * wq;
            W[q + c * n] = sn * wp + cs * wq;
          }
        }
        __syncthreads();
      }
    }
    FPLS_EIG_STAMP(7);
    // (h) Q_new(:, a + colpos(c)) = Q(:, perm) W(:, c) over the rows of the block;
    //     unpaired trailing blocks are copied.  Two inner steps per batch of
    //     global loads keep several L2 requests in flight per thread.
    {
      const int tpd = (bsz + 3) >> 2;  // 4x4 tiles per dimension of a full block
      const int ntile = nblk * tpd * tpd;
      for (int tile = t; tile < ntile; tile += ENT) {
        const int bi = tile / (tpd * tpd), tt = tile - bi * (tpd * tpd);
        const int a = bi << lg, m = a + size, b = min(a + bsz, n), mm = b - a;
        const int r0 = (tt % tpd) * 4, c0 = (tt / tpd) * 4;
        if (r0 >= mm || c0 >= mm) continue;
        const int rlim = min(4, mm - r0), clim = min(4, mm - c0);
        if (m >= n) {
          for (int cc = c0; cc < c0 + clim; ++cc)
            for (int rr = r0; rr < r0 + rlim; ++rr)
              Qd[(a + rr) + (int64_t)(a + cc) * n] = Qs[(a + rr) + (int64_t)(a + cc) * n];
          continue;
        }
        double acc[4][4];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int y = 0; y < 4; ++y) acc[x][y] = 0.0;
        // Q is block diagonal over the two halves: rows of one half only meet
        // the columns of that half.  Column k of Q pairs with row iperm(k) of W.
        const int ra = a + r0, rz = a + r0 + rlim;
        const int kb = ra >= m ? m : a, ke = rz <= m ? m : b;
        for (int k0 = kb; k0 < ke; k0 += 2) {
          double qv[2][4];
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int k = min(k0 + u, ke - 1);
            const double* col = Qs + (int64_t)k * n + ra;
#pragma unroll
            for (int x = 0; x < 4; ++x) qv[u][x] = (x < rlim && k0 + u < ke) ? col[x] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int q = a + iperm[min(k0 + u, ke - 1)];
            double wvv[4];
#pragma unroll
            for (int y = 0; y < 4; ++y) wvv[y] = y < clim ? W[q + (a + c0 + y) * n] : 0.0;
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
              for (int y = 0; y < 4; ++y) acc[x][y] += qv[u][x] * wvv[y];
          }
        }
        for (int y = 0; y < clim; ++y) {
          const int oc = a + colpos[a + c0 + y];
          for (int x = 0; x < rlim; ++x) Qd[(a + r0 + x) + (int64_t)oc * n] = acc[x][y];
        }
      }
    }
    __syncthreads();
    // eigenvalues of the merged blocks in ascending order
    if (t < n) {
      const int a = (t >> lg) << lg;
      if (a + size < n) lam[a + colpos[t]] = val[t];
    }
    {
      double* tmp = Qs;
      Qs = Qd;
      Qd = tmp;
    }
    __syncthreads();
    FPLS_EIG_STAMP(8);
  }

  FPLS_EIG_STAMP(9);
  // ------------------------------------------------ stage 3: back-transformation
  double* R = big;
  for (int idx = t; idx < n * n; idx += ENT) R[idx] = Vr[idx];
  __syncthreads();
  {
    const int c = t >> 3, g = t & 7;
    constexpr int RPL = EN / 8;  // rows per lane
    double z[RPL];
#pragma unroll
    for (int ii = 0; ii < RPL; ++ii) z[ii] = 0.0;
    if (c < n) {
#pragma unroll
      for (int ii = 0; ii < RPL; ++ii) {
        const int r = g + 8 * ii;
        z[ii] = r < n ? Qs[r + (int64_t)c * n] : 0.0;
      }
    }
    // reflector rows <= j are stored as zeros, so only rows >= n are masked
    bool ok[RPL];
#pragma unroll
    for (int ii = 0; ii < RPL; ++ii) ok[ii] = g + 8 * ii < n;
    for (int j = n - 3; j >= 0; --j) {
      const double tj = tauv[j];
      if (tj == 0.0) continue;
      const double* Rj = R + j * n + g;
      double rv[RPL];
#pragma unroll
      for (int ii = 0; ii < RPL; ++ii) rv[ii] = ok[ii] ? Rj[8 * ii] : 0.0;
      double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ii = 0; ii < RPL; ++ii) s4[ii & 3] = fma(rv[ii], z[ii], s4[ii & 3]);
      const double s = tj * sum8((s4[0] + s4[1]) + (s4[2] + s4[3]));
#pragma unroll
      for (int ii = 0; ii < RPL; ++ii) z[ii] = fma(-rv[ii], s, z[ii]);
    }
    if (c < n) {
      const int oc = n - 1 - c;
#pragma unroll
      for (int ii = 0; ii < RPL; ++ii) {
        const int r = g + 8 * ii;
        if (r < n) Vout[r + (int64_t)oc * n] = z[ii];
      }
      if (g == 0) w[oc] = lam[c] * scale;
    }
  }
  FPLS_EIG_STAMP(10);
}

void sym_eig(Ctx* ctx, const double* A, int64_t lda, int n, double* w, double* V, int kwant) {
  sym_eig_batched(ctx, A, lda, 0, nullptr, n, 1, w, 0, V, 0, n, kwant);
}

void sym_eig_batched(Ctx* ctx, const double* A, int64_t lda, int64_t strideA, const int* ns_dev,
                     int nmax, int count, double* w, int64_t strideW, double* V,
                     int64_t strideV, int n0, int kwant, bool dc_only) {
  FPLS_REQUIRE(nmax >= 1 && nmax <= kJacobiMax && count >= 1, FPLS_ERR_ARG,
               "sym_eig_batched: orders must be in 1..128");
  if (!ns_dev && n0 == 0) n0 = nmax;
  if (!ns_dev) FPLS_REQUIRE(n0 >= 1 && n0 <= nmax, FPLS_ERR_ARG, "sym_eig: order out of range");
  // tridiagonalisation, then every eigenpair from bisection and the twisted
  // factorisation on its own workgroup (k_eig_ii.hip); a problem whose
  // spectrum the inverse-free vectors cannot resolve (flag set) is solved
  // again by divide and conquer
  const int nfix = ns_dev ? nmax : n0;
  DBuf<double> ws(ctx, (size_t)3 * nmax * nmax * count);
  if (dc_only) {
    launch(ctx, "sym_eig_dc", k_sym_eig_dc, dim3((unsigned)count), dim3(ENT), 0, A, lda,
           strideA, ns_dev, nfix, nmax, w, strideW, V, strideV, ws.get(),
           (int64_t)3 * nmax * nmax, (double*)nullptr, (const int*)nullptr, (int*)nullptr);
    return;
  }
  DBuf<int> flags(ctx, (size_t)count);
  sym_eig_tri(ctx, A, lda, strideA, ns_dev, nfix, nmax, count, w, strideW, V, strideV, kwant,
              flags.get());
  launch(ctx, "sym_eig_dc", k_sym_eig_dc, dim3((unsigned)count), dim3(ENT), 0, A, lda, strideA,
         ns_dev, nfix, nmax, w, strideW, V, strideV, ws.get(),
         (int64_t)3 * nmax * nmax, (double*)nullptr, (const int*)flags.get(), (int*)nullptr);
}

}  // namespace fpls
