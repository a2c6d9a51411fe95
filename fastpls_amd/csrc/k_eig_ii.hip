// Eigenpairs of a symmetric tridiagonal matrix T (order n <= 128, from the
// one-workgroup Householder reduction of k_eig.hip) with one wave per
// eigenvalue, spread over the chip instead of one workgroup:
//
//  1. eigenvalues by multisection: each wave brackets eigenvalue k between
//     points where the Sturm count (number of eigenvalues below x, from the
//     LDL^T recurrence q_i = d_i - x - e_{i-1}^2 / q_{i-1}) is <= k and > k;
//     its 64 lanes evaluate 64 interior points per step, so the bracket
//     shrinks 65-fold per step (Barth, Martin & Wilkinson 1967, bisection on
//     the Sturm sequence; the pivmin guard and the Gershgorin start of
//     LAPACK's dstebz);
//  2. eigenvectors by inverse iteration (LAPACK's dstein: the partially
//     pivoted factorisation of T - lambda I and the perturbed solve of
//     dlagtf / dlagts, three solves per vector, eigenvalues closer than
//     1e-3 |T| form a cluster whose vectors are orthogonalised against each
//     other by one wave in turn), normalised with the largest entry positive;
//  3. the back-transformation by the Householder reflectors, per vector.
//
// Orthogonality between clusters is then within eps |T| / gap <= 2e-13; a
// cluster of more than CL_MAX eigenvalues sets *fallback for the divide and
// conquer solver (k_eig.hip), which handles any clustering.
#include <cfloat>
#include <cmath>

#include "fpls_internal.h"
#include "fpls_wave.h"

namespace fpls {

namespace {

constexpr int II_N = 128;   // largest order
constexpr int II_W = 4;     // waves per workgroup
constexpr int CL_MAX = 8;   // largest cluster inverse iteration takes
constexpr double II_EPS = DBL_EPSILON * 0.5;  // unit roundoff (dlamch('E'))

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

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Number of eigenvalues of T below x (LDL^T signs, with the pivmin guard).
__device__ __forceinline__ int sturm_count(const double* d, const double* ee, int n, double x,
                                           double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0.0 ? 1 : 0;
  for (int i = 1; i < n; ++i) {
    q = (d[i] - x) - ee[i - 1] * rcp2(q);
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0 ? 1 : 0;
  }
  return c;
}

// Gershgorin interval, |T| bound and pivmin shared by both kernels (every
// lane of a wave gets the same values).
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

// ---------------------------------------------------------------- eigenvalues
// One workgroup per eigenvalue k: its 4 waves evaluate 256 interior points of
// the bracket per step (257-fold shrink), combined through LDS ballots.
constexpr int BI_T = 256, BI_PTS = BI_T;
__global__ __launch_bounds__(BI_T) void k_trd_eigval(int n, const double* __restrict__ tri,
                                                     double* __restrict__ lam) {
  __shared__ double d[II_N], e[II_N], ee[II_N];
  __shared__ unsigned long long above_w[BI_T / 64];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    d[i] = tri[i];
    e[i] = tri[n + i];
    ee[i] = tri[n + i] * tri[n + i];
  }
  __syncthreads();
  const int k = blockIdx.x, t = threadIdx.x, wv = t >> 6;
  const TriBounds b = tri_bounds(d, e, ee, n);
  // count(lo) <= k < count(hi)
  double lo = b.gl, hi = b.gu;
  for (int it = 0; it < 12; ++it) {
    const double width = hi - lo;
    // dstebz's tolerance: absolute eps |T|, relative 2 eps, at least pivmin
    const double tol = fmax(II_EPS * b.tnorm,
                            fmax(b.pivmin, 2.0 * II_EPS * fmax(fabs(lo), fabs(hi))));
    if (width <= tol) break;  // uniform over the workgroup
    const double step = width / (double)(BI_PTS + 1);
    const double x = lo + step * (double)(t + 1);
    const int c = sturm_count(d, ee, n, x, b.pivmin);
    const unsigned long long above = __ballot(c > k);
    if ((t & 63) == 0) above_w[wv] = above;
    __syncthreads();
    // first point with more than k eigenvalues below it; the one before it
    int f = BI_PTS;
    for (int q = 0; q < BI_T / 64; ++q) {
      const unsigned long long m = above_w[q];
      if (m) {
        f = 64 * q + __ffsll((long long)m) - 1;
        break;
      }
    }
    __syncthreads();
    const double xf = f < BI_PTS ? lo + step * (double)(f + 1) : hi;
    const double xb = f > 0 ? lo + step * (double)f : lo;
    lo = xb;
    hi = xf;
  }
  if (t == 0) lam[k] = 0.5 * (lo + hi);
}

// ---------------------------------------------------------------- eigenvectors
// Cheap deterministic start vector in (-1, 1) for eigenvalue index k.
__device__ __forceinline__ double start_entry(int k, int i) {
  unsigned h = (unsigned)(k * 7919 + i * 104729 + 12345);
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return (double)(h >> 8) * (2.0 / 16777216.0) - 1.0;
}

__global__ __launch_bounds__(64 * II_W) void k_trd_eigvec(int n, const double* __restrict__ tri,
                                                          const double* __restrict__ lam,
                                                          const double* __restrict__ refl,
                                                          double* __restrict__ w,
                                                          double* __restrict__ V,
                                                          int* __restrict__ fallback) {
  __shared__ double d[II_N], e[II_N], ee[II_N], tau[II_N];
  __shared__ double fa[II_W][II_N], fb[II_W][II_N], fc[II_W][II_N], fd[II_W][II_N];
  __shared__ double xv[II_W][II_N];
  __shared__ int fin[II_W][II_N];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    d[i] = tri[i];
    e[i] = tri[n + i];
    ee[i] = tri[n + i] * tri[n + i];
    tau[i] = tri[2 * n + i];
  }
  __syncthreads();
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int k0 = blockIdx.x * II_W + wv;
  if (k0 >= n) return;
  const TriBounds b = tri_bounds(d, e, ee, n);
  const double ortol = 1e-3 * b.onenrm, tight = 1e-9 * b.onenrm;
  // clusters (gaps <= 1e-3 |T|) are orthogonalised afterwards by their first
  // member (k_trd_clorth); a tight or wide cluster goes to divide and conquer
  {
    int lo = k0, hi = k0;
    while (lo > 0 && lam[lo] - lam[lo - 1] <= ortol) --lo;
    while (hi + 1 < n && lam[hi + 1] - lam[hi] <= ortol) ++hi;
    bool bad = hi - lo + 1 > CL_MAX;
    for (int q = lo; q < hi && !bad; ++q) bad = lam[q + 1] - lam[q] <= tight;
    if (bad) {
      if (lane == 0) *fallback = 1;
      return;
    }
  }
  const double scale = tri[3 * n];
  double* a = fa[wv];
  double* bb = fb[wv];
  double* c = fc[wv];
  double* d2 = fd[wv];
  double* x = xv[wv];
  int* in = fin[wv];
  const double dtpcrt = sqrt(0.1 / (double)n);
  {
    const int k = k0;
    const double xj = lam[k];
    // factorise T - xj I = P L U (dlagtf) on lane 0: the running diagonal
    // and super-diagonal entries carried in registers, factors to LDS; the
    // pivot test |c| / scale2 <= |a| / scale1 compared without divisions
    if (lane == 0) {
      double aq = d[0] - xj;
      double bq = n > 1 ? e[0] : 0.0;
      double scale1 = fabs(aq) + fabs(bq);
      for (int q = 0; q + 1 < n; ++q) {
        const double cq = e[q];
        double a1 = d[q + 1] - xj;
        const double bq1 = q + 2 < n ? e[q + 1] : 0.0;
        const double scale2 = fabs(cq) + fabs(a1) + fabs(bq1);
        double aout = aq, bout = bq, cout = 0.0, d2q = 0.0, bnext = bq1;
        int inq = 0;
        if (cq == 0.0) {
          scale1 = scale2;
        } else if (fabs(cq) * scale1 <= fabs(aq) * scale2) {
          scale1 = scale2;
          cout = cq * rcp2(aq);
          a1 = fma(-cout, bq, a1);
        } else {
          inq = 1;
          const double mult = aq * rcp2(cq);
          aout = cq;
          const double temp = a1;
          a1 = fma(-mult, temp, bq);
          d2q = bq1;
          bnext = -mult * bq1;
          bout = temp;
          cout = mult;
        }
        a[q] = aout;
        bb[q] = bout;
        c[q] = cout;
        d2[q] = d2q;
        in[q] = inq;
        aq = a1;
        bq = bnext;
      }
      a[n - 1] = aq;
      bb[n - 1] = 0.0;
      c[n - 1] = 0.0;
      d2[n - 1] = 0.0;
      in[n - 1] = 0;
    }
    wave_sync();
    // perturbation size of dlagts (job -1)
    double tl = 0.0;
    for (int i = lane; i < n; i += 64) {
      tl = fmax(tl, fabs(a[i]));
      if (i >= 1) tl = fmax(tl, fabs(bb[i - 1]));
      if (i >= 2) tl = fmax(tl, fabs(d2[i - 2]));
    }
    tl = wave_max_d(tl) * II_EPS;
    if (tl == 0.0) tl = II_EPS;
    const double an = wave_max_d(lane == 0 ? fabs(a[n - 1]) : 0.0);
    for (int i = lane; i < n; i += 64) x[i] = start_entry(k, i);
    wave_sync();
    int nrmchk = 0;
    for (int its = 0; its < 5; ++its) {
      // x *= n |T|_1 max(eps, |u_nn|) / |x|_1
      double s1 = 0.0;
      for (int i = lane; i < n; i += 64) s1 += fabs(x[i]);
      s1 = wave_sum_d(s1);
      const double scl = (double)n * b.onenrm * fmax(II_EPS, an) / s1;
      for (int i = lane; i < n; i += 64) x[i] *= scl;
      wave_sync();
      if (lane == 0) {
        // solve (T - xj I) y = x (dlagts, job -1: tiny pivots perturbed);
        // the last solved entries are carried in registers
        double cur = x[0];
        for (int q = 1; q < n; ++q) {
          const double xq = x[q], cm = c[q - 1];
          if (in[q - 1] == 0) {
            x[q - 1] = cur;
            cur = fma(-cm, cur, xq);
          } else {
            x[q - 1] = xq;
            cur = fma(-cm, xq, cur);
          }
        }
        x[n - 1] = cur;
        double x1 = 0.0, x2 = 0.0;
        for (int q = n - 1; q >= 0; --q) {
          double temp = fma(-d2[q], x2, fma(-bb[q], x1, x[q]));
          double ak = a[q];
          double pert = copysign(tl, ak);
          for (int guard = 0; guard < 200; ++guard) {
            const double absak = fabs(ak);
            if (absak < 1.0) {
              if (absak < DBL_MIN) {
                if (absak == 0.0 || fabs(temp) * DBL_MIN > absak) {
                  ak += pert;
                  pert *= 2.0;
                  continue;
                }
                temp *= 1.0 / DBL_MIN;
                ak *= 1.0 / DBL_MIN;
              } else if (fabs(temp) > absak * (1.0 / DBL_MIN)) {
                ak += pert;
                pert *= 2.0;
                continue;
              }
            }
            break;
          }
          const double y = temp * rcp2(ak);
          x[q] = y;
          x2 = x1;
          x1 = y;
        }
      }
      wave_sync();
      double xm = 0.0;
      for (int i = lane; i < n; i += 64) xm = fmax(xm, fabs(x[i]));
      xm = wave_max_d(xm);
      // one more solve after dstein's growth criterion is met (its EXTRA)
      if (xm < dtpcrt) continue;
      if (++nrmchk >= 2) break;
    }
    // unit length, largest entry positive
    double amax = -1.0, av = 0.0, s2 = 0.0;
    int ai = n;
    for (int i = lane; i < n; i += 64) {
      const double v = x[i], u = fabs(v);
      s2 = fma(v, v, s2);
      if (u > amax) {
        amax = u;
        av = v;
        ai = i;
      }
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
    s2 = wave_sum_d(s2);
    const double inv = (av < 0.0 ? -1.0 : 1.0) / sqrt(s2);
    for (int i = lane; i < n; i += 64) x[i] *= inv;
    wave_sync();
    // tridiagonal eigenvector k (ascending) is output column n - 1 - k
    const int oc = n - 1 - k;
    for (int i = lane; i < n; i += 64) V[i + (int64_t)oc * n] = x[i];
    if (lane == 0) w[oc] = lam[k] * scale;
  }
}

// Modified Gram-Schmidt inside each cluster of eigenvalues closer than
// 1e-3 |T| (one wave per cluster, led by its first member), in ascending
// eigenvalue order: the members were computed independently from their own
// eigenvalues, each accurate to eps |T| / gap, and are made orthogonal here.
__global__ __launch_bounds__(64 * II_W) void k_trd_clorth(int n, const double* __restrict__ tri,
                                                          const double* __restrict__ lam,
                                                          double* __restrict__ V,
                                                          const int* __restrict__ fallback) {
  __shared__ double d[II_N], e[II_N], ee[II_N];
  if (*fallback) return;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    d[i] = tri[i];
    e[i] = tri[n + i];
    ee[i] = tri[n + i] * tri[n + i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int k0 = blockIdx.x * II_W + (threadIdx.x >> 6);
  if (k0 >= n) return;
  const TriBounds b = tri_bounds(d, e, ee, n);
  const double ortol = 1e-3 * b.onenrm;
  if (k0 > 0 && lam[k0] - lam[k0 - 1] <= ortol) return;
  int cl = 1;
  while (k0 + cl < n && lam[k0 + cl] - lam[k0 + cl - 1] <= ortol) ++cl;
  if (cl == 1) return;
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

// Back-transformation V <- H_0 ... H_{n-3} V for every eigenvector at once:
// the reflectors (column j of refl: support j + 1 .. n - 1, v(j + 1) = 1)
// staged in LDS, one column of V per 8 lanes (16 rows each) in registers,
// reflectors applied from the last to the first.
constexpr int BT_T = 128, BT_COLS = BT_T / 8, BT_RPL = II_N / 8;
__global__ __launch_bounds__(BT_T) void k_trd_backtf(int n, const double* __restrict__ tri,
                                                     const double* __restrict__ refl,
                                                     double* __restrict__ V) {
  __shared__ double R[II_N * II_N];
  __shared__ double tau[II_N];
  {
    // eight loads in flight per thread
    const int nn = n * n;
    for (int base = threadIdx.x; base < nn; base += 8 * BT_T) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int idx = base + u * BT_T;
        v[u] = idx < nn ? refl[idx] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int idx = base + u * BT_T;
        if (idx < nn) R[idx] = v[u];
      }
    }
  }
  for (int i = threadIdx.x; i < n; i += BT_T) tau[i] = tri[2 * n + i];
  __syncthreads();
  const int col = blockIdx.x * BT_COLS + (threadIdx.x >> 3), g = threadIdx.x & 7;
  const bool live = col < n;
  double z[BT_RPL];
#pragma unroll
  for (int ii = 0; ii < BT_RPL; ++ii) {
    const int r = g + 8 * ii;
    z[ii] = live && r < n ? V[r + (int64_t)col * n] : 0.0;
  }
  // reflector j + 1's entries are loaded while reflector j is applied; a
  // zero tau applies as a zero update
  double rv[BT_RPL];
  int j = n - 3;
  if (j >= 0) {
#pragma unroll
    for (int ii = 0; ii < BT_RPL; ++ii) rv[ii] = g + 8 * ii < n ? R[j * n + g + 8 * ii] : 0.0;
  }
  for (; j >= 0; --j) {
    const double tj = tau[j];
    double rn[BT_RPL];
    const int jn = j > 0 ? j - 1 : 0;
#pragma unroll
    for (int ii = 0; ii < BT_RPL; ++ii) rn[ii] = g + 8 * ii < n ? R[jn * n + g + 8 * ii] : 0.0;
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ii = 0; ii < BT_RPL; ++ii) s4[ii & 3] = fma(rv[ii], z[ii], s4[ii & 3]);
    const double sj = tj * grp8_sum((s4[0] + s4[1]) + (s4[2] + s4[3]));
#pragma unroll
    for (int ii = 0; ii < BT_RPL; ++ii) {
      z[ii] = fma(-rv[ii], sj, z[ii]);
      rv[ii] = rn[ii];
    }
  }
  if (live) {
#pragma unroll
    for (int ii = 0; ii < BT_RPL; ++ii) {
      const int r = g + 8 * ii;
      if (r < n) V[r + (int64_t)col * n] = z[ii];
    }
  }
}

}  // namespace

void tridiag_eig_ii(Ctx* ctx, int n, const double* tri, const double* refl, double* w, double* V,
                    int* fallback) {
  FPLS_REQUIRE(n >= 1 && n <= II_N, FPLS_ERR_ARG, "tridiag_eig_ii: order must be 1..128");
  DBuf<double> lam(ctx, n);
  const unsigned blocks = (unsigned)((n + II_W - 1) / II_W);
  launch(ctx, "sym_eig_bisect", k_trd_eigval, dim3((unsigned)n), dim3(BI_T), 0, n, tri,
         lam.get());
  launch(ctx, "sym_eig_invit", k_trd_eigvec, dim3(blocks), dim3(64 * II_W), 0, n, tri,
         (const double*)lam.get(), refl, w, V, fallback);
  launch(ctx, "sym_eig_clorth", k_trd_clorth, dim3(blocks), dim3(64 * II_W), 0, n, tri,
         (const double*)lam.get(), V, (const int*)fallback);
  if (n >= 3)
    launch(ctx, "sym_eig_backtf", k_trd_backtf, dim3((unsigned)((n + BT_COLS - 1) / BT_COLS)),
           dim3(BT_T), 0, n, tri, refl, V);
}

}  // namespace fpls
