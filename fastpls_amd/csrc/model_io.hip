// A device model from the reference's model-list fields, so that a model fitted
// elsewhere (the reference's CPU path, a saved R object) predicts through this
// library: R, mX, vX, mY, ncomp and the compact factors that
// compact_prediction_weights (src/fastPLS.cpp:3737-3836) reads -- Q for SIMPLS,
// W_latent or C_latent + Q for PLSSVD -- plus the optional dense B
// (the B_stored policy, src/fastPLS.cpp:63-90).  The projection cache
// (R_predict = R / vX, R_offset = mX R_predict, R/main.R:1019-1041) is built on
// the device as after a fit.
#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "model.h"

using namespace fpls;

extern "C" int fpls_model_import(fpls_ctx* ctx, const fpls_model_fields* f, fpls_model** out) {
  if (!f || !out) return FPLS_ERR_ARG;
  *out = nullptr;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    FPLS_REQUIRE(f->method == FPLS_METHOD_SIMPLS || f->method == FPLS_METHOD_PLSSVD, FPLS_ERR_ARG,
                 "model_import: unknown method");
    FPLS_REQUIRE(f->p >= 1 && f->m >= 1 && f->kmax >= 1 && f->n_ncomp >= 1 && f->ncomp,
                 FPLS_ERR_ARG, "model_import: empty model");
    FPLS_REQUIRE(f->R && f->mX && f->vX && f->mY, FPLS_ERR_ARG,
                 "model_import: R, mX, vX and mY are required");
    const int64_t p = f->p, m = f->m;
    const int kmax = f->kmax, L = f->n_ncomp;
    std::vector<int> nc(f->ncomp, f->ncomp + L);
    for (int v : nc)
      FPLS_REQUIRE(v >= 1 && v <= kmax, FPLS_ERR_ARG,
                   "ncomp exceeds latent rank for compact class prediction");
    auto mdl = std::make_unique<fpls_model>();
    Model* M = &mdl->m;
    M->ctx = c;
    M->method = f->method;
    M->n = 0;
    M->p = p;
    M->m = m;
    M->kmax = kmax;
    M->n_components = kmax;
    M->ncomp = nc;
    auto up = [&](DBuf<double>& dst, const double* src, size_t count) {
      dst = DBuf<double>(c, count);
      h2d(c, dst.get(), src, sizeof(double) * count);
    };
    up(M->R, f->R, (size_t)p * kmax);
    up(M->mX, f->mX, (size_t)p);
    up(M->vX, f->vX, (size_t)p);
    up(M->mY, f->mY, (size_t)m);
    if (f->Q) up(M->Q, f->Q, (size_t)m * kmax);
    if (f->method == FPLS_METHOD_SIMPLS) {
      FPLS_REQUIRE(f->Q != nullptr, FPLS_ERR_ARG,
                   "Model `Q` is not compatible with compact class prediction");
    } else if (f->W_latent) {
      up(M->W_latent, f->W_latent, (size_t)kmax * m * L);
      if (f->C_latent) up(M->C_latent, f->C_latent, (size_t)kmax * kmax * L);
    } else {
      FPLS_REQUIRE(f->C_latent && f->Q, FPLS_ERR_ARG,
                   "Compact class prediction requires compact low-rank factors");
      up(M->C_latent, f->C_latent, (size_t)kmax * kmax * L);
      // W_a = C_a[:mc, :mc] Q[:, :mc]^T in the first mc rows, zeros below
      M->W_latent = DBuf<double>(c, (size_t)kmax * m * L);
      fill_f64(c, M->W_latent.get(), 0.0, (int64_t)kmax * m * L);
      for (int a = 0; a < L; ++a)
        gemm_f64(c, false, true, nc[a], m, nc[a], 1.0, M->C_latent.get() + (size_t)a * kmax * kmax,
                 kmax, nullptr, M->Q.get(), m, nullptr, 0.0,
                 M->W_latent.get() + (size_t)a * kmax * m, kmax);
    }
    if (f->B) {
      up(M->B, f->B, (size_t)p * m * L);
      M->b_stored = true;
    }
    build_projection_cache(c, M);
    sync(c);
    *out = mdl.release();
  });
}
