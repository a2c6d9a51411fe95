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

namespace fpls {

namespace {

constexpr int EN = 128;           // largest order
constexpr int ELD = EN + 1;       // LDS leading dimension of A in stage 1
constexpr int ENT = 1024;         // threads
constexpr double EEPS = DBL_EPSILON;

__device__ __forceinline__ double sum8(double v) {
  // butterfly over the 8-lane group: every step adds the same two operands on
  // both partners, so all 8 lanes end with bitwise identical sums
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  return v;
}
__device__ __forceinline__ double prod8(double v) {
  v *= __shfl_xor(v, 1, 64);
  v *= __shfl_xor(v, 2, 64);
  v *= __shfl_xor(v, 4, 64);
  return v;
}
__device__ __forceinline__ int isum8(int v) {
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

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
__global__ __launch_bounds__(ENT) void k_sym_eig_dc(const double* __restrict__ Ain, int64_t lda,
                                                    int n, double* __restrict__ w,
                                                    double* __restrict__ Vout,
                                                    double* __restrict__ Qa,
                                                    double* __restrict__ Qb,
                                                    double* __restrict__ Vr) {
  __shared__ double big[EN * ELD];  // stage 1: A; stage 2: W (ld n); stage 3: reflectors (ld n)
  __shared__ double dv[EN], ev[EN], tauv[EN], pv[EN];
  __shared__ double lam[EN], Ds[EN], zs[EN], taus[EN], zh[EN], invn[EN], val[EN];
  __shared__ double rc[EN], rs[EN];
  __shared__ int perm[EN], defl[EN], ndidx[EN], rootidx[EN], orig[EN], colpos[EN];
  __shared__ int rp[EN], rq[EN];
  __shared__ int nrot[EN / 2], Kb[EN / 2];
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
    colpos[t] = 0;
  }
  __syncthreads();

  // ------------------------------------------------ stage 1: tridiagonalisation
  for (int j = 0; j + 2 < n; ++j) {
    const int m = n - 1 - j;
    if (wv == 0) {
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
    }
    __syncthreads();
    const double tau = tauv[j];
    if (tau == 0.0) continue;  // H_j = I: A22 unchanged (uniform branch)
    // p = tau A22 v, 8 lanes per row, rows read as columns (A22 symmetric)
    {
      const int i = j + 1 + (t >> 3), g = t & 7;
      double s = 0.0;
      if (i < n) {
        for (int k = j + 1 + g; k < n; k += 8) s += A[k + i * ELD] * A[k + j * ELD];
      }
      s = sum8(s);
      if (i < n && g == 0) pv[i] = tau * s;
    }
    __syncthreads();
    {
      // K = tau/2 p^T v (every wave computes the same butterfly sum)
      const int i0 = j + 1 + lane, i1 = j + 1 + lane + 64;
      double s = 0.0;
      if (i0 < n) s += pv[i0] * A[i0 + j * ELD];
      if (i1 < n) s += pv[i1] * A[i1 + j * ELD];
      const double K = 0.5 * tau * wave_sum(s);
      for (int idx = t; idx < m * m; idx += ENT) {
        const int i = j + 1 + idx % m, k = j + 1 + idx / m;
        const double vi = A[i + j * ELD], vk = A[k + j * ELD];
        const double wi = pv[i] - K * vi, wk = pv[k] - K * vk;
        A[i + k * ELD] -= vi * wk + wi * vk;
      }
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
  for (int idx = t; idx < n * n; idx += ENT) {
    const int i = idx % n, j = idx / n;
    Vr[idx] = (j + 2 < n && i > j) ? A[i + j * ELD] : 0.0;
    Qa[idx] = (i == j) ? 1.0 : 0.0;
    Qb[idx] = 0.0;
  }
  __syncthreads();

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
  for (int size = 1; size < n; size *= 2) {
    const int bsz = 2 * size;
    const int nblk = (n + bsz - 1) / bsz;
    // (a) sorted merge of the two halves, z from the boundary rows of Q
    if (t < n) {
      const int bi = t / bsz, a = bi * bsz, m = a + size, b = min(a + bsz, n);
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
      }
    }
    __syncthreads();
    // (b) deflation, one thread per merge (serial scan in sorted order)
    if (t < nblk) {
      const int a = t * bsz, m = a + size, b = min(a + bsz, n), mm = b - a;
      int K = 0, nr = 0;
      if (m < n) {
        const double rho2 = 2.0 * fabs(ev[m - 1]);
        double dmax = 0.0;
        for (int q = 0; q < mm; ++q) dmax = fmax(dmax, fabs(Ds[a + q]));
        const double tol = 8.0 * EEPS * fmax(dmax, rho2);
        int prev = -1;
        for (int q = 0; q < mm; ++q) {
          defl[a + q] = 0;
          if (rho2 * fabs(zs[a + q]) <= tol) {
            defl[a + q] = 1;
            continue;
          }
          if (prev < 0) {
            prev = q;
            continue;
          }
          double s = zs[a + prev], c = zs[a + q];
          const double tt = hypot(s, c), dd = Ds[a + q] - Ds[a + prev];
          c /= tt;
          s = -s / tt;
          if (fabs(dd * c * s) <= tol) {
            zs[a + q] = tt;
            zs[a + prev] = 0.0;
            const double dp = Ds[a + prev], dq = Ds[a + q];
            Ds[a + prev] = c * c * dp + s * s * dq;
            Ds[a + q] = s * s * dp + c * c * dq;
            defl[a + prev] = 1;
            rp[a + nr] = prev;
            rq[a + nr] = q;
            rc[a + nr] = c;
            rs[a + nr] = s;
            ++nr;
          }
          prev = q;
        }
        for (int q = 0; q < mm; ++q) {
          if (!defl[a + q]) {
            ndidx[a + K] = q;
            rootidx[a + q] = K;
            ++K;
          }
        }
      }
      Kb[t] = K;
      nrot[t] = nr;
    }
    __syncthreads();
    // (c) secular roots, 8 lanes per root
    {
      const int slot = t >> 3, g = t & 7;
      const int bi = slot / bsz, a = bi * bsz, m = a + size;
      const int j = slot - a;
      const bool act = slot < n && m < n && j < Kb[bi];
      if (act) {
        const int K = Kb[bi];
        const double rho2 = 2.0 * fabs(ev[m - 1]);
        auto Dn = [&](int k) { return Ds[a + ndidx[a + k]]; };
        auto Zn = [&](int k) { return zs[a + ndidx[a + k]]; };
        int o;
        double lo, hi;
        const double dj = Dn(j);
        if (j < K - 1) {
          const double delta = 0.5 * (Dn(j + 1) - dj);
          double s = 0.0;
          for (int k = g; k < K; k += 8) {
            const double zk = Zn(k);
            s += zk * zk / ((Dn(k) - dj) - delta);
          }
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
          for (int k = g; k < K; k += 8) {
            const double zk = Zn(k);
            s += zk * zk;
          }
          o = j;
          lo = 0.0;
          hi = rho2 * sum8(s);
        }
        const double dorg = Dn(o);
        double x = 0.5 * (lo + hi);
        for (int it = 0; it < 200; ++it) {
          double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0, eab = 0.0;
          for (int k = g; k < K; k += 8) {
            const double zk = Zn(k);
            const double den = (Dn(k) - dorg) - x;
            const double r = zk / den;
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
            const double dr = (Dn(j + 1) - dorg) - x;
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
          orig[a + j] = o;
          taus[a + j] = x;
        }
      }
    }
    __syncthreads();
    // (d) Gu-Eisenstat z-hat, 8 lanes per non-deflated index
    {
      const int slot = t >> 3, g = t & 7;
      const int bi = slot / bsz, a = bi * bsz, m = a + size;
      const int i = slot - a;
      const bool act = slot < n && m < n && i < Kb[bi];
      if (act) {
        const int K = Kb[bi];
        const double rho2 = 2.0 * fabs(ev[m - 1]);
        const double di = Ds[a + ndidx[a + i]];
        auto Dn = [&](int k) { return Ds[a + ndidx[a + k]]; };
        // -delta(i, j) = lam_j - d_i = (D_o(j) - d_i) + tau_j
        auto lmd = [&](int jj) { return (Dn(orig[a + jj]) - di) + taus[a + jj]; };
        double pr = 1.0;
        for (int k = g; k < K; k += 8) {
          if (k < i) pr *= lmd(k) / (Dn(k) - di);
          else if (k > i) pr *= lmd(k - 1) / (Dn(k) - di);
        }
        pr = prod8(pr) * (lmd(K - 1) / rho2);
        if (g == 0) zh[a + i] = copysign(sqrt(fmax(pr, 0.0)), zs[a + ndidx[a + i]]);
      }
    }
    __syncthreads();
    // (e) eigenvector norms and output eigenvalues (sorted-input index space)
    {
      const int slot = t >> 3, g = t & 7;
      const int bi = slot / bsz, a = bi * bsz, m = a + size, b = min(a + bsz, n);
      const int jr = slot - a;
      if (slot < n && m < n) {
        const int K = Kb[bi];
        if (jr < K) {
          const double dorg = Ds[a + ndidx[a + orig[a + jr]]];
          const double tj = taus[a + jr];
          double s = 0.0;
          for (int i = g; i < K; i += 8) {
            const double del = (Ds[a + ndidx[a + i]] - dorg) - tj;
            const double u = zh[a + i] / del;
            s += u * u;
          }
          s = sum8(s);
          if (g == 0) {
            invn[a + jr] = 1.0 / sqrt(s);
            val[a + ndidx[a + jr]] = dorg + tj;
          }
        }
        if (g == 0 && jr < b - a && defl[a + jr]) val[a + jr] = Ds[a + jr];
      }
    }
    __syncthreads();
    // (f) output order (ascending, ties by index) and the small matrix W
    {
      const int slot = t >> 3, g = t & 7;
      const int bi = slot / bsz, a = bi * bsz, m = a + size, b = min(a + bsz, n);
      const int c = slot - a, mm = b - a;
      if (slot < n && m < n) {
        const double vc = val[a + c];
        int r = 0;
        for (int k = g; k < mm; k += 8) {
          const double vk = val[a + k];
          r += (vk < vc || (vk == vc && k < c)) ? 1 : 0;
        }
        r = isum8(r);
        if (g == 0) colpos[a + c] = r;
      }
    }
    for (int idx = t; idx < n * n; idx += ENT) {
      const int q = idx % n, c = idx / n;  // W(q, c), block-local indices stored at (a+q, a+c)
      const int bi = q / bsz;
      if (c / bsz != bi) continue;
      const int a = bi * bsz, m = a + size;
      if (m >= n) continue;
      const int ql = q - a, cl = c - a;
      double wqc;
      if (defl[a + cl]) {
        wqc = (ql == cl) ? 1.0 : 0.0;
      } else if (defl[a + ql]) {
        wqc = 0.0;
      } else {
        const int jr = rootidx[a + cl], i = rootidx[a + ql];
        const double dorg = Ds[a + ndidx[a + orig[a + jr]]];
        const double del = (Ds[a + ql] - dorg) - taus[a + jr];
        wqc = zh[a + i] / del * invn[a + jr];
      }
      W[q + c * n] = wqc;
    }
    __syncthreads();
    // (g) deflation rotations folded into W: W <- G_1 (G_2 (... W)), rows p, q
    {
      int maxr = 0;
      for (int k = 0; k < nblk; ++k) maxr = max(maxr, nrot[k]);
      for (int r = maxr - 1; r >= 0; --r) {
        for (int idx = t; idx < n; idx += ENT) {
          const int bi = idx / bsz, a = bi * bsz, m = a + size;
          if (m >= n || r >= nrot[bi]) continue;
          const int c = idx;  // column of W
          const int p = a + rp[a + r], q = a + rq[a + r];
          const double cs = rc[a + r], sn = rs[a + r];
          const double wp = W[p + c * n], wq = W[q + c * n];
          W[p + c * n] = cs * wp - sn * wq;
          W[q + c * n] = sn * wp + cs * wq;
        }
        __syncthreads();
      }
    }
    // (h) Q_new(:, a + colpos(c)) = Q(:, perm) W(:, c) over the rows of the block;
    //     unpaired trailing blocks are copied
    {
      const int tpd = (bsz + 3) / 4;  // 4x4 tiles per dimension of a full block
      const int ntile = nblk * tpd * tpd;
      for (int tile = t; tile < ntile; tile += ENT) {
        const int bi = tile / (tpd * tpd), tt = tile % (tpd * tpd);
        const int a = bi * bsz, m = a + size, b = min(a + bsz, n), mm = b - a;
        const int r0 = (tt % tpd) * 4, c0 = (tt / tpd) * 4;
        if (r0 >= mm || c0 >= mm) continue;
        if (m >= n) {
          for (int cc = c0; cc < min(c0 + 4, mm); ++cc)
            for (int rr = r0; rr < min(r0 + 4, mm); ++rr)
              Qd[(a + rr) + (int64_t)(a + cc) * n] = Qs[(a + rr) + (int64_t)(a + cc) * n];
          continue;
        }
        double acc[4][4];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int y = 0; y < 4; ++y) acc[x][y] = 0.0;
        const int rlim = min(4, mm - r0), clim = min(4, mm - c0);
        for (int q = 0; q < mm; ++q) {
          const double* col = Qs + (int64_t)perm[a + q] * n + a + r0;
          double qv[4], wvv[4];
#pragma unroll
          for (int x = 0; x < 4; ++x) qv[x] = x < rlim ? col[x] : 0.0;
#pragma unroll
          for (int y = 0; y < 4; ++y) wvv[y] = y < clim ? W[(a + q) + (a + c0 + y) * n] : 0.0;
#pragma unroll
          for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[x][y] += qv[x] * wvv[y];
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
      const int bi = t / bsz, a = bi * bsz, m = a + size;
      if (m < n) lam[a + colpos[t]] = val[t];
    }
    {
      double* tmp = Qs;
      Qs = Qd;
      Qd = tmp;
    }
    __syncthreads();
  }

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
    for (int j = n - 3; j >= 0; --j) {
      const double tj = tauv[j];
      if (tj == 0.0) continue;
      double s = 0.0;
#pragma unroll
      for (int ii = 0; ii < RPL; ++ii) {
        const int r = g + 8 * ii;
        if (r > j && r < n) s += R[r + j * n] * z[ii];
      }
      s = tj * sum8(s);
#pragma unroll
      for (int ii = 0; ii < RPL; ++ii) {
        const int r = g + 8 * ii;
        if (r > j && r < n) z[ii] -= R[r + j * n] * s;
      }
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
}

void sym_eig(Ctx* ctx, const double* A, int64_t lda, int n, double* w, double* V) {
  FPLS_REQUIRE(n >= 1 && n <= kJacobiMax, FPLS_ERR_ARG,
               "sym_eig: order must be in 1..128 for the one-workgroup solver");
  DBuf<double> ws(ctx, (size_t)3 * n * n);
  double* Qa = ws.get();
  double* Qb = Qa + (size_t)n * n;
  double* Vr = Qb + (size_t)n * n;
  launch(ctx, "sym_eig_dc", k_sym_eig_dc, dim3(1), dim3(ENT), 0, A, lda, n, w, V, Qa, Qb, Vr);
}

}  // namespace fpls
