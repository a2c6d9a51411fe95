// Cross-validated prediction: pls_cv_predict_compiled (src/fastPLS.cpp:6615-7151)
// with the reference's group -> fold map (:6719-6770): leave one group out, or
// k folds over a permutation of the groups (stratified by the class of each
// group's first row for classification) drawn from R's uniform generator as
// set.seed(seed) leaves it (R/main.R:3883).  X stays
// resident; each fold gathers its train / test rows on the device, fits with
// seed + f, and predicts every ncomp slice from the one fit: compact argmax
// classes (:4115-4168), the LDA head (:6975-7004) or the candidate-kNN head
// (:7005-7050), or regression predictions.  Status codes per fold: 1 ok,
// 2 empty test fold, 3 no training rows, 4 degenerate classification fold.
#include <algorithm>
#include <map>
#include <set>
#include <vector>

#include "model.h"

namespace fpls {

template <typename T>
__global__ void k_gather_rows(const T* __restrict__ X, int64_t ld, int64_t cols,
                              const int* __restrict__ idx, int64_t nr, T* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nr * cols;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t % nr, j = t / nr;
    out[r + j * nr] = X[(int64_t)idx[r] + j * ld];
  }
}

static unsigned g1(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

// ---------------------------------------------------------------- folds
// R's default uniform generator after set.seed(seed) (R's RNG.c: RNG_Init,
// MT_genrand, fixup): the seed is scrambled by 50 steps of s <- 69069 s + 1,
// the next 625 steps fill the position slot and the 624 Mersenne-Twister
// words, and each draw is a tempered MT19937 word times 2^-32, kept inside
// (0, 1).
struct RUnif {
  uint32_t mt[624];
  int mti = 624;
  explicit RUnif(int32_t seed) {
    uint32_t s = (uint32_t)seed;
    for (int j = 0; j < 50; ++j) s = 69069u * s + 1u;
    s = 69069u * s + 1u;  // the position slot, reset to 624
    for (int j = 0; j < 624; ++j) {
      s = 69069u * s + 1u;
      mt[j] = s;
    }
  }
  double next() {
    if (mti >= 624) {
      for (int k = 0; k < 624; ++k) {
        const uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
        mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
      }
      mti = 0;
    }
    uint32_t y = mt[mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    const double x = (double)y * 2.3283064365386963e-10;
    const double half = 0.5 * 2.328306437080797e-10;
    if (x <= 0.0) return half;
    if (1.0 - x <= 0.0) return 1.0 - half;
    return x;
  }
};

// samplewithoutreplace (src/fastPLS.cpp:4498-4509): size draws, each taking
// the element at floor(u * rest) out of the remaining ones.
static std::vector<int> sample_without_replace(std::vector<int> yy, int size, RUnif& rng) {
  std::vector<int> xx(size);
  int rest = (int)yy.size();
  for (int ii = 0; ii < size; ++ii) {
    const int it = (int)(rng.next() * rest);
    xx[ii] = yy[it];
    yy.erase(yy.begin() + it);
    --rest;
  }
  return xx;
}

// Row -> fold (0-based) of pls_cv_predict_compiled (src/fastPLS.cpp:6666-6775);
// labels (1-based, may be null) stratify the permutation by the class of each
// group's first row.  Returns the fold count.
static int assign_folds(const int32_t* constrain, int64_t n, int kfold, const int32_t* labels,
                        int n_classes, int32_t seed, std::vector<int>& fold) {
  const bool loo_requested = kfold < 0;
  if (!loo_requested && kfold < 2) kfold = 2;
  std::set<int32_t> uniq(constrain, constrain + n);
  std::map<int32_t, int> gidx;
  for (int32_t g : uniq) gidx.emplace(g, (int)gidx.size());
  const int ngroups = (int)uniq.size();
  std::vector<int> group_of(n);
  for (int64_t i = 0; i < n; ++i) group_of[i] = gidx[constrain[i]];
  std::vector<int> group_fold(ngroups, 0);
  int nfold;
  if (loo_requested || kfold >= ngroups) {
    nfold = std::max(ngroups, 1);
    for (int j = 0; j < ngroups; ++j) group_fold[j] = j;
  } else {
    nfold = kfold;
    RUnif rng(seed);
    if (labels && n_classes > 1) {
      std::vector<int64_t> first(ngroups, -1);
      for (int64_t i = 0; i < n; ++i)
        if (first[group_of[i]] < 0) first[group_of[i]] = i;
      std::vector<std::vector<int>> by_class(n_classes);
      for (int j = 0; j < ngroups; ++j) {
        int cls = labels[first[j]];
        cls = std::min(std::max(cls, 1), n_classes);
        by_class[cls - 1].push_back(j);
      }
      for (int cls = 0; cls < n_classes; ++cls) {
        const int ng = (int)by_class[cls].size();
        if (ng < 1) continue;
        std::vector<int> frame(ng);
        for (int q = 0; q < ng; ++q) frame[q] = q + 1;
        const std::vector<int> perm = sample_without_replace(frame, ng, rng);
        for (int pos = 0; pos < ng; ++pos) group_fold[by_class[cls][perm[pos] - 1]] = pos % kfold;
      }
    } else {
      std::vector<int> frame(ngroups);
      for (int q = 0; q < ngroups; ++q) frame[q] = q + 1;
      const std::vector<int> perm = sample_without_replace(frame, ngroups, rng);
      for (int j = 0; j < ngroups; ++j) group_fold[j] = (perm[j] - 1) % kfold;
    }
  }
  fold.resize(n);
  for (int64_t i = 0; i < n; ++i) fold[i] = group_fold[group_of[i]];
  return nfold;
}

// Gathered copy of the rows idx of A (same dtype, packed ld = nr).
static DMat gather(Ctx* c, const DMat& A, const std::vector<int>& idx, DBuf<int>& didx,
                   DBuf<unsigned char>& hold) {
  const int64_t nr = (int64_t)idx.size();
  didx = DBuf<int>(c, std::max<int64_t>(nr, 1));
  h2d(c, didx.get(), idx.data(), nr * sizeof(int));
  const size_t es = A.dtype == FPLS_F32 ? 4 : 8;
  hold = DBuf<unsigned char>(c, std::max<size_t>(1, nr * A.cols * es));
  if (A.dtype == FPLS_F32)
    launch(c, "gather_rows", k_gather_rows<float>, dim3(g1(nr * A.cols)), dim3(256), 0, A.f(),
           A.ld, A.cols, (const int*)didx.get(), nr, (float*)hold.get());
  else
    launch(c, "gather_rows", k_gather_rows<double>, dim3(g1(nr * A.cols)), dim3(256), 0, A.d(),
           A.ld, A.cols, (const int*)didx.get(), nr, (double*)hold.get());
  DMat d;
  d.p = hold.get();
  d.dtype = A.dtype;
  d.rows = nr;
  d.cols = A.cols;
  d.ld = std::max<int64_t>(nr, 1);
  return d;
}

// Ypred slices (w x m x L) of a fitted model: T = Xs R (kmax) once, then
// T_mc Q_mc^T (SIMPLS) or T_mc W_a (PLSSVD), plus mY.
static void predict_cube(Ctx* c, const Model* M, const DMat& X, double* out) {
  const int64_t w = X.rows, m = M->m;
  const int kmax = M->kmax;
  DBuf<double> T(c, (size_t)w * kmax);
  if (X.dtype == FPLS_F32)
    gemm_nn_f32(c, M->precision, w, kmax, X.cols, X.f(), X.ld, M->mX.get(), M->R_predict_f32.get(),
                M->p, nullptr, nullptr, T.get(), w);
  else
    project_scores(c, M, X, kmax, nullptr, T.get(), w);
  void add_row(Ctx*, double*, int64_t, int64_t, int64_t, const double*);
  for (size_t a = 0; a < M->ncomp.size(); ++a) {
    const int mc = std::min(M->ncomp[a], kmax);
    double* Ya = out + a * w * m;
    if (M->method == FPLS_METHOD_SIMPLS)
      gemm_f64(c, false, true, w, m, mc, 1.0, T.get(), w, nullptr, M->Q.get(), m, nullptr, 0.0, Ya,
               w);
    else
      gemm_f64(c, false, false, w, m, mc, 1.0, T.get(), w, nullptr,
               M->W_latent.get() + a * (size_t)kmax * m, kmax, nullptr, 0.0, Ya, w);
    add_row(c, Ya, w, w, m, M->mY.get());
  }
}

}  // namespace fpls

using namespace fpls;

extern "C" int fpls_cv_predict(fpls_ctx* ctx, const fpls_mat* Xm, const fpls_mat* Ym,
                               const fpls_ivec* yv, int32_t n_classes, const int32_t* constrain,
                               const int32_t* ncomp, int32_t n_ncomp, int32_t kfold,
                               int32_t method, const fpls_opts* opts, int32_t classifier,
                               double lda_ridge, int32_t knn_k, double tau, double alpha,
                               int32_t top_m, int32_t* fold_out, int32_t* status_out,
                               int32_t status_cap, int32_t* n_folds_out, int32_t* class_pred,
                               double* ypred) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    fpls_opts o;
    fpls_opts_default(&o);
    if (opts) o = *opts;
    FPLS_REQUIRE(method == FPLS_METHOD_PLSSVD || method == FPLS_METHOD_SIMPLS ||
                     method == FPLS_METHOD_OPLS || method == FPLS_METHOD_KERNELPLS,
                 FPLS_ERR_ARG, "method must be 1=plssvd, 3=simpls_fast, 4=opls, or 5=kernelpls");
    // kernel PLS folds fit the direct SIMPLS core (src/fastPLS.cpp:6887-6891)
    const bool opls = method == FPLS_METHOD_OPLS;
    if (method == FPLS_METHOD_KERNELPLS) method = FPLS_METHOD_SIMPLS;
    DBuf<unsigned char> hx, hy;
    DMat X = stage(c, Xm, hx, "cv: X");
    const int64_t n = X.rows;
    FPLS_REQUIRE(n >= 2, FPLS_ERR_ARG, "Xdata must contain at least two samples");
    FPLS_REQUIRE(constrain != nullptr, FPLS_ERR_ARG, "constrain must have one value per sample");
    FPLS_REQUIRE(ncomp && n_ncomp >= 1, FPLS_ERR_ARG, "ncomp must contain at least one value");
    const bool classification = (yv && yv->ptr);
    std::vector<int32_t> yh;
    DMat Y;
    int64_t m = 0;
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
      FPLS_REQUIRE(classifier == 0, FPLS_ERR_ARG, "LDA / cKNN classifiers need class labels");
      Y = stage(c, Ym, hy, "cv: Y");
      FPLS_REQUIRE(Y.rows == n, FPLS_ERR_ARG, "Ydata must have the same number of rows as Xdata");
      m = Y.cols;
    }
    std::vector<int> nc(ncomp, ncomp + n_ncomp);
    const int L = n_ncomp;
    if (method == FPLS_METHOD_PLSSVD) {
      const int max_rank = (int)std::min<int64_t>(n, std::min<int64_t>(X.cols, m));
      for (int& v : nc) v = std::min(std::max(v, 1), max_rank);
    } else {
      for (int& v : nc) v = std::max(v, 1);
    }
    // groups -> folds (arma::unique sorts ascending)
    std::vector<int> fold;
    const int nfold = assign_folds(constrain, n, kfold, classification ? yh.data() : nullptr,
                                   (int)n_classes, (int32_t)o.seed, fold);
    FPLS_REQUIRE(status_out && status_cap >= nfold, FPLS_ERR_ARG,
                 "status buffer smaller than the number of folds");
    if (fold_out)
      for (int64_t i = 0; i < n; ++i) fold_out[i] = fold[i] + 1;
    if (n_folds_out) *n_folds_out = nfold;
    std::vector<int32_t> cp(classification ? (size_t)n * L : 0, 0);
    const bool store_scores = !classification || ypred != nullptr;
    std::vector<double> yp(store_scores ? (size_t)n * m * L : 0, 0.0);
    std::vector<double> yhost;
    if (!classification) {
      yhost.resize((size_t)n * m);
      for (int64_t j = 0; j < m; ++j) {
        if (Y.dtype == FPLS_F64)
          FPLS_HIP(hipMemcpyAsync(yhost.data() + j * n, Y.d() + j * Y.ld, n * sizeof(double),
                                  hipMemcpyDeviceToHost, c->stream));
        else {
          std::vector<float> tmp(n);
          FPLS_HIP(hipMemcpyAsync(tmp.data(), Y.f() + j * Y.ld, n * sizeof(float),
                                  hipMemcpyDeviceToHost, c->stream));
          sync(c);
          for (int64_t i = 0; i < n; ++i) yhost[i + j * n] = tmp[i];
        }
      }
      sync(c);
    }
    for (int f = 0; f < nfold; ++f) {
      std::vector<int> te, tr;
      for (int64_t i = 0; i < n; ++i) (fold[i] == f ? te : tr).push_back((int)i);
      if (te.empty()) {
        status_out[f] = 2;
        continue;
      }
      auto write_fallback_rows = [&](int cls, const std::vector<double>& row) {
        for (int s = 0; s < L; ++s)
          for (int i : te) {
            if (classification) cp[i + (size_t)s * n] = cls > 0 ? cls : yh[i];
            if (store_scores)
              for (int64_t j = 0; j < m; ++j)
                yp[i + j * n + (size_t)s * n * m] =
                    row.empty() ? (classification ? (yh[i] == j + 1 ? 1.0 : 0.0)
                                                  : yhost[i + j * n])
                                : row[j];
          }
      };
      if (tr.empty()) {  // status 3: predictions are the observed responses
        write_fallback_rows(0, {});
        status_out[f] = 3;
        continue;
      }
      if (classification) {
        std::vector<int> cnt(m, 0);
        for (int i : tr) cnt[yh[i] - 1]++;
        int active = 0, last = -1;
        for (int j = 0; j < m; ++j)
          if (cnt[j] > 0) {
            ++active;
            last = j;
          }
        if (active <= 1) {  // status 4: the single training class (one-hot row)
          std::vector<double> row(m, 0.0);
          if (last >= 0) row[last] = 1.0;
          write_fallback_rows(last + 1, row);
          status_out[f] = 4;
          continue;
        }
      }
      DBuf<int> dtr, dte, dyt;
      DBuf<unsigned char> hxtr, hxte, hytr;
      DMat Xtr = gather(c, X, tr, dtr, hxtr);
      DMat Xte = gather(c, X, te, dte, hxte);
      DMat Ytr;
      const int32_t* ytr = nullptr;
      std::vector<int32_t> ytr_h;
      if (classification) {
        ytr_h.resize(tr.size());
        for (size_t r = 0; r < tr.size(); ++r) ytr_h[r] = yh[tr[r]];
        dyt = DBuf<int>(c, tr.size());
        h2d(c, dyt.get(), ytr_h.data(), tr.size() * sizeof(int32_t));
        ytr = dyt.get();
      } else {
        Ytr = gather(c, Y, tr, dtr, hytr);
        // dtr now describes Y's rows too: both gathers use the same indices
      }
      Model M;
      fpls_opts of = o;
      of.seed = o.seed + (uint32_t)f;
      of.fit = 0;
      of.store_b = 0;
      // OPLS: filter the training fold, apply it to the held-out rows, fit
      // SIMPLS with scaling "none" (src/fastPLS.cpp:6876-6886)
      DBuf<double> fX, fXte, fmX, fvX, fWo, fPo;
      if (opls) {
        int used = 0;
        fX = DBuf<double>(c, (size_t)std::max<int64_t>(Xtr.rows * Xtr.cols, 1));
        opls_filter_dev(c, Xtr, classification ? nullptr : &Ytr, ytr, m, std::max(o.opls_north, 0),
                        o.scaling, fX.get(), std::max<int64_t>(Xtr.rows, 1), fmX, fvX, fWo, fPo,
                        used);
        fXte = DBuf<double>(c, (size_t)std::max<int64_t>(Xte.rows * Xte.cols, 1));
        opls_apply_dev(c, Xte, fmX.get(), fvX.get(), fWo.get(), fPo.get(), used, fXte.get(),
                       std::max<int64_t>(Xte.rows, 1));
        Xtr.p = fX.get();
        Xtr.dtype = FPLS_F64;
        Xtr.ld = std::max<int64_t>(Xtr.rows, 1);
        Xte.p = fXte.get();
        Xte.dtype = FPLS_F64;
        Xte.ld = std::max<int64_t>(Xte.rows, 1);
        of.scaling = 3;
      }
      if (method == FPLS_METHOD_PLSSVD)
        plssvd_fit_dev(c, &M, Xtr, Ytr, ytr, m, nc.data(), L, of);
      else
        simpls_fit_dev(c, &M, Xtr, Ytr, ytr, m, nc.data(), L, of);
      const int64_t w = (int64_t)te.size();
      std::vector<int32_t> lab((size_t)w * L, 0);
      if (classification && classifier == 1) {
        // LDA on the fold's latent scores (src/fastPLS.cpp:6975-7004)
        int kmax = *std::max_element(nc.begin(), nc.end());
        kmax = std::min(kmax, M.kmax);
        DBuf<double> Ttr(c, (size_t)Xtr.rows * kmax), Tte(c, (size_t)w * kmax);
        DBuf<float> Ttrf, Ttef;
        DMat Tt, Tq;
        if (Xtr.dtype == FPLS_F32) {
          Ttrf = DBuf<float>(c, (size_t)Xtr.rows * kmax);
          Ttef = DBuf<float>(c, (size_t)w * kmax);
          project_scores(c, &M, Xtr, kmax, Ttrf.get(), nullptr, Xtr.rows);
          project_scores(c, &M, Xte, kmax, Ttef.get(), nullptr, w);
          Tt.p = Ttrf.get();
          Tq.p = Ttef.get();
          Tt.dtype = Tq.dtype = FPLS_F32;
        } else {
          project_scores(c, &M, Xtr, kmax, nullptr, Ttr.get(), Xtr.rows);
          project_scores(c, &M, Xte, kmax, nullptr, Tte.get(), w);
          Tt.p = Ttr.get();
          Tq.p = Tte.get();
          Tt.dtype = Tq.dtype = FPLS_F64;
        }
        Tt.rows = Xtr.rows;
        Tt.cols = kmax;
        Tt.ld = Xtr.rows;
        Tq.rows = w;
        Tq.cols = kmax;
        Tq.ld = w;
        std::vector<int> ks;
        for (int v : nc) ks.push_back(std::min(v, kmax));
        DBuf<double> counts;
        std::vector<LdaSlice> sl;
        lda_train_core(c, Tt, nullptr, 0, ytr, (int32_t)m, ks, lda_ridge, M.precision, counts, sl);
        DBuf<double> sc(c, (size_t)w * m);
        DBuf<int32_t> dl(c, (size_t)w * L);
        for (int s = 0; s < L; ++s) {
          const int kk = ks[s];
          if (Tq.dtype == FPLS_F32) {
            DBuf<double> linT(c, (size_t)kk * m);
            transpose_f64(c, sl[s].linear.get(), m, m, kk, linT.get(), kk);
            DBuf<float> linTf(c, (size_t)kk * m);
            f64_to_f32(c, linTf.get(), linT.get(), (int64_t)kk * m);
            gemm_nn_f32(c, M.precision, w, m, kk, Tq.f(), w, nullptr, linTf.get(), kk, nullptr,
                        nullptr, sc.get(), w);
          } else {
            gemm_f64(c, false, true, w, m, kk, 1.0, Tq.d(), w, nullptr, sl[s].linear.get(), m,
                     nullptr, 0.0, sc.get(), w);
          }
          argmax_rows(c, sc.get(), w, w, m, sl[s].constants.get(), dl.get() + (size_t)s * w);
        }
        to_host_i32(c, lab.data(), dl.get(), (size_t)w * L);
      } else if (classification && classifier == 2) {
        // candidate kNN on the fold's normalised score space (:7005-7050)
        int kmax = *std::max_element(nc.begin(), nc.end());
        kmax = std::min(kmax, M.kmax);
        DBuf<float> Ttrf(c, (size_t)Xtr.rows * kmax), Ttef(c, (size_t)w * kmax);
        DBuf<double> td;
        if (Xtr.dtype == FPLS_F32) {
          project_scores(c, &M, Xtr, kmax, Ttrf.get(), nullptr, Xtr.rows);
          project_scores(c, &M, Xte, kmax, Ttef.get(), nullptr, w);
        } else {
          DBuf<double> a(c, (size_t)Xtr.rows * kmax), b(c, (size_t)w * kmax);
          project_scores(c, &M, Xtr, kmax, nullptr, a.get(), Xtr.rows);
          project_scores(c, &M, Xte, kmax, nullptr, b.get(), w);
          f64_to_f32(c, Ttrf.get(), a.get(), Xtr.rows * (int64_t)kmax);
          f64_to_f32(c, Ttef.get(), b.get(), w * (int64_t)kmax);
        }
        DMat Tt, Tq;
        Tt.p = Ttrf.get();
        Tt.dtype = FPLS_F32;
        Tt.rows = Xtr.rows;
        Tt.cols = kmax;
        Tt.ld = Xtr.rows;
        Tq = Tt;
        Tq.p = Ttef.get();
        Tq.rows = w;
        Tq.ld = w;
        CkClassIndex ci;
        stable_class_sort(c, ytr, Xtr.rows, (int)m, ci);
        DBuf<int32_t> dl(c, (size_t)w * L);
        for (int s = 0; s < L; ++s) {
          const int kk = std::min(nc[s], kmax);
          DBuf<float> S(c, (size_t)Xtr.rows * kk), Tn(c, (size_t)Xtr.rows * kk);
          score_space(c, &M, Tt, kk, S.get());
          rownorm_rows(c, S.get(), Xtr.rows, kk, ci.d_rows.get(), Tn.get());
          DBuf<double> cen;
          DBuf<float> cenf;
          class_centroids(c, Tn.get(), kk, ci, (int)m, cen, cenf);
          DBuf<float> Sq(c, (size_t)w * kk), Tqn(c, (size_t)w * kk);
          score_space(c, &M, Tq, kk, Sq.get());
          rownorm_rows(c, Sq.get(), w, kk, nullptr, Tqn.get());
          cknn_predict_core(c, M.precision, Tqn.get(), w, kk, Tn.get(), ci.off, cenf.get(), (int)m,
                            nullptr, 1, top_m, knn_k, tau, alpha, dl.get() + (size_t)s * w,
                            nullptr, w);
        }
        to_host_i32(c, lab.data(), dl.get(), (size_t)w * L);
      } else {
        DBuf<double> cube(c, (size_t)w * m * L);
        predict_cube(c, &M, Xte, cube.get());
        std::vector<double> hc((size_t)w * m * L);
        to_host(c, hc.data(), cube.get(), hc.size());
        for (int s = 0; s < L; ++s)
          for (int64_t r = 0; r < w; ++r) {
            const int i = te[r];
            if (classification) {  // first maximum (index_max)
              int best = 0;
              double bv = hc[r + (size_t)s * w * m];
              for (int64_t j = 1; j < m; ++j) {
                const double v = hc[r + j * w + (size_t)s * w * m];
                if (v > bv) {
                  bv = v;
                  best = (int)j;
                }
              }
              lab[r + (size_t)s * w] = best + 1;
            }
            if (store_scores)
              for (int64_t j = 0; j < m; ++j)
                yp[i + j * n + (size_t)s * n * m] = hc[r + j * w + (size_t)s * w * m];
          }
      }
      if (classification)
        for (int s = 0; s < L; ++s)
          for (int64_t r = 0; r < w; ++r) cp[te[r] + (size_t)s * n] = lab[r + (size_t)s * w];
      status_out[f] = 1;
    }
    if (classification && class_pred) std::copy(cp.begin(), cp.end(), class_pred);
    if (store_scores && ypred) std::copy(yp.begin(), yp.end(), ypred);
  });
}

extern "C" int fpls_cv_folds(const int32_t* constrain, int64_t n, int32_t kfold,
                             const int32_t* labels, int32_t n_classes, uint32_t seed,
                             int32_t* fold_out, int32_t* n_folds_out) {
  if (!constrain || !fold_out || n < 1) return FPLS_ERR_ARG;
  if (labels)
    for (int64_t i = 0; i < n; ++i)
      if (labels[i] < 1 || labels[i] > n_classes) return FPLS_ERR_LABEL_RANGE;
  try {
    std::vector<int> fold;
    const int nfold = assign_folds(constrain, n, kfold, labels, (int)n_classes, (int32_t)seed,
                                   fold);
    for (int64_t i = 0; i < n; ++i) fold_out[i] = fold[i] + 1;
    if (n_folds_out) *n_folds_out = nfold;
  } catch (const std::exception&) {
    return FPLS_ERR_ARG;
  }
  return FPLS_OK;
}
