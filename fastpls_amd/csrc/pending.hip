// Entry points declared in the ABI whose device paths are not built yet.
// They fail loudly with FPLS_ERR_STATE (never a silent CPU fallback).
#include "model.h"

using namespace fpls;

static int not_built(fpls_ctx* ctx, const char* what) {
  if (ctx) ctx->c.err = std::string(what) + ": not available in this build";
  return FPLS_ERR_STATE;
}

extern "C" {
int fpls_rsvd_xprod(fpls_ctx* ctx, const fpls_mat*, const fpls_mat*, int32_t, const fpls_opts*,
                    double*, double*, double*, int32_t*) {
  return not_built(ctx, "fpls_rsvd_xprod");
}
int fpls_class_sums(fpls_ctx* ctx, const fpls_mat*, const fpls_ivec*, int32_t, double*, double*,
                    int64_t*) {
  return not_built(ctx, "fpls_class_sums");
}
}
