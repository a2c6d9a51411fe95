// Label-aware streaming primitive: class sums and column sums of X from int32
// labels in one HBM pass each (the rowsum() accumulation of the label-aware
// stream model, R/main.R:563-588), plus exact integer class counts.
#include <vector>

#include "model.h"

using namespace fpls;

extern "C" int fpls_class_sums(fpls_ctx* ctx, const fpls_mat* Xm, const fpls_ivec* yv,
                               int32_t n_classes, double* class_sums, double* col_sums_out,
                               int64_t* counts) {
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    DBuf<unsigned char> hx;
    DMat X = stage(c, Xm, hx, "class_sums: X");
    FPLS_REQUIRE(n_classes >= 1, FPLS_ERR_ARG, "class_sums: n_classes must be >= 1");
    DBuf<int32_t> hl;
    const int32_t* y = stage_labels(c, yv, X.rows, hl, "class_sums: y");
    const int64_t n = X.rows, p = X.cols;
    DBuf<int64_t> cnt(c, n_classes);
    DBuf<int32_t> bad(c, 1);
    label_counts(c, y, n, n_classes, cnt.get(), bad.get());
    int32_t hbad = 0;
    to_host_i32(c, &hbad, bad.get(), 1);
    FPLS_REQUIRE(hbad == 0, FPLS_ERR_LABEL_RANGE, "labels must be encoded as 1..n_classes");
    DBuf<double> S(c, (size_t)p * n_classes), cs(c, p);
    if (X.dtype == FPLS_F32)
      gemm_tn_f32(c, FPLS_PREC_FP32, p, n_classes, n, X.f(), X.ld, nullptr, nullptr, 0, y, nullptr,
                  S.get(), p);
    else
      gemm_f64(c, true, false, p, n_classes, n, 1.0, X.d(), X.ld, nullptr, nullptr, n, nullptr,
               0.0, S.get(), p, y);
    col_sums(c, X, cs.get());
    if (class_sums) to_host(c, class_sums, S.get(), (size_t)p * n_classes);
    if (col_sums_out) to_host(c, col_sums_out, cs.get(), p);
    if (counts) {
      FPLS_HIP(hipMemcpyAsync(counts, cnt.get(), n_classes * sizeof(int64_t),
                              hipMemcpyDeviceToHost, c->stream));
      sync(c);
    }
  });
}
