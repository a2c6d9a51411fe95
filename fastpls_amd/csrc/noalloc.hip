// Allocation-free LDA prediction for callers that cannot call hipMalloc
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// allocation).  Every temporary of the prediction (projected test scores,
// discriminant score block, padded operands, split-K partials) is carved from
// one static device array per GPU, `scratch`, instead of the context pool.
#include <mutex>

#include "model.h"

namespace fpls {

// 64 MiB per device, held until the process ends.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
hipError_t g_rc[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};
}  // namespace

void scratch_setup(Ctx* ctx) {
  const int d = ctx->device;
  ctx->scratch_ptr = nullptr;
  ctx->scratch_bytes = 0;
  if (d < 0 || d >= kMaxDevices) {
    ctx->scratch_err = "device index beyond the scratch table";
    return;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_done[d]) {
    // the symbol's address on the current device (each GPU holds its own copy)
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != d) e = hipSetDevice(d);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) (void)hipGetLastError();
    g_ptr[d] = e == hipSuccess ? p : nullptr;
    g_rc[d] = e;
    g_done[d] = true;
  }
  if (g_ptr[d]) {
    ctx->scratch_ptr = g_ptr[d];
    ctx->scratch_bytes = kScratchBytes;
  } else {
    ctx->scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_rc[d]);
  }
}

// Rows of an allocation-free LDA prediction that fit the scratch array:
// projected scores (w x k, FP32 or FP64) and the score block (w x C FP64) per
// row, plus the row-independent operands (padded R_predict and linear terms,
// FP32 copies) with 1 MiB of slack.
static int64_t noalloc_max_rows(const Model* M, int dtype) {
  const int64_t k = M->lda.back().k, C = M->n_classes, p = M->p;
  // per row: the projected scores, the score block and one split-K slab of it
  const int64_t per_row = k * (dtype == FPLS_F32 ? 4 : 8) + 12 * C;
  // padded operands and FP32 copies of the projection, of the latent form
  // (linear terms) and of the direct form (R_predict lin^T), plus slack
  const int64_t fixed = 4 * (p + 64) * (k + 128) + 4 * p + 12 * k * C +
                        4 * (k + 64) * (C + 128) + 12 * p * C + 4 * (p + 64) * (C + 128) +
                        4 * p + (int64_t(1) << 20);
  const int64_t room = (int64_t)kScratchBytes - fixed;
  return room > 0 ? room / per_row : 0;
}

// Arena scope: the context allocates from the scratch array until it ends.
struct ArenaScope {
  Ctx* c;
  explicit ArenaScope(Ctx* ctx) : c(ctx) {
    c->arena_base = static_cast<unsigned char*>(c->scratch_ptr);
    c->arena_cap = c->scratch_bytes;
    c->arena_used = 0;
    c->arena_blocks.clear();
  }
  ~ArenaScope() {
    c->arena_base = nullptr;
    c->arena_cap = 0;
    c->arena_used = 0;
    c->arena_blocks.clear();
  }
};

}  // namespace fpls

using namespace fpls;

extern "C" int fpls_lda_predict_noalloc(fpls_ctx* ctx, const fpls_model* model,
                                        const fpls_mat* Xm, int32_t* labels, double* scores) {
  if (!model) return FPLS_ERR_ARG;
  return fpls_guard(ctx, [&] {
    Ctx* c = &ctx->c;
    const Model* M = &model->m;
    FPLS_REQUIRE(c->scratch_ptr != nullptr, FPLS_ERR_HIP,
                 "lda_predict_noalloc: no scratch array on this device (" + c->scratch_err + ")");
    FPLS_REQUIRE(!M->lda.empty(), FPLS_ERR_STATE,
                 "this model does not contain fitted LDA classifier parameters");
    FPLS_REQUIRE(Xm != nullptr && Xm->loc == FPLS_DEVICE, FPLS_ERR_ARG,
                 "lda_predict_noalloc: Xtest must be a device matrix");
    FPLS_REQUIRE(labels != nullptr, FPLS_ERR_ARG, "lda_predict_noalloc: labels are NULL");
    DBuf<unsigned char> unused;
    DMat X = stage(c, Xm, unused, "lda_predict_noalloc: X");
    FPLS_REQUIRE(X.cols == M->p, FPLS_ERR_ARG,
                 "lda_predict_noalloc: X has the wrong number of columns");
    const int64_t w = X.rows, C = M->n_classes;
    const int64_t wmax = noalloc_max_rows(M, X.dtype);
    FPLS_REQUIRE(w <= wmax, FPLS_ERR_ARG,
                 "lda_predict_noalloc: " + std::to_string(w) + " test rows exceed the " +
                     std::to_string(kScratchBytes >> 20) + " MiB scratch limit of " +
                     std::to_string(wmax) + " rows for this model");
    if (w == 0) return;
    ArenaScope arena(c);
    std::function<void(int, const double*)> cb;
    if (scores)
      cb = [&](int a, const double* sc) {
        FPLS_HIP(hipMemcpyAsync(scores + (size_t)a * w * C, sc, (size_t)w * C * sizeof(double),
                                hipMemcpyDeviceToDevice, c->stream));
      };
    lda_predict_core(c, M, X, labels, cb);
    // the next call reuses the scratch array: finish before returning
    sync(c);
  });
}
