// Internal runtime of the gfx950 fastPLS library: context, HBM pool,
// launch/timing wrapper, error plumbing and the kernel entry points shared
// between translation units.  Nothing here crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <unordered_map>
#include <vector>

#include "../../include/fastpls_hip.h"

namespace fpls {

// ---------------------------------------------------------------- errors
struct Error : public std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

#define FPLS_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      ::fpls::fail(FPLS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define FPLS_REQUIRE(cond, code, msg)        \
  do {                                       \
    if (!(cond)) ::fpls::fail((code), (msg)); \
  } while (0)

// ---------------------------------------------------------------- pool
// HBM pool: blocks are cached by size so that repeated fits allocate nothing
// after the first call.
// Caching HBM pool: blocks are kept after their first hipMalloc and handed
// out again (smallest free block that fits without wasting more than half of
// it); lookups are O(log blocks).
struct Pool {
  struct Block {
    void* ptr;
    size_t bytes;
    bool used;
  };
  std::vector<Block> blocks;
  std::unordered_map<void*, size_t> index;  // ptr -> blocks[i]
  std::multimap<size_t, size_t> free_by_size;
  void* get(size_t bytes);
  void put(void* p);
  void release_all();
  size_t total_bytes() const;
};

struct Ctx;

// Device buffer borrowed from the context pool (RAII).
template <typename T>
struct DBuf {
  Ctx* ctx = nullptr;
  T* ptr = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(Ctx* c, size_t count);
  ~DBuf();
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : ctx(o.ctx), ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { reset(); ctx = o.ctx; ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
    return *this;
  }
  void reset();
  T* get() const { return ptr; }
};

// ---------------------------------------------------------------- context
struct KStat {
  int64_t launches = 0;
  double total_ms = 0.0;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  Pool pool;
  bool timing = false;
  std::map<std::string, KStat> kstats;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> spare_events;
  int num_cus = 256;
  // Deferred checks: kernels record failure counts in device flag slots, and
  // one read-back at the end of an entry point raises the first failing check,
  // so a fit makes no host round trip half way through.
  struct Check {
    int slot, count, code, fail_value;  // fail_value 0: fails when != 0
    std::string msg;
  };
  int32_t* flags_dev = nullptr;
  // Ticket counter of the last-block reductions (zero between launches: the
  // last block of each launch resets it).
  unsigned* ticket_dev = nullptr;
  unsigned* ticket();
  int32_t* flags_host = nullptr;
  int nflags = 0;
  std::vector<Check> checks;
  // Captured launch plan of the batched cross-validation (cvb.hip), kept while
  // the shapes and X stay the same.
  std::shared_ptr<void> cv_plan;
  // Autotuner of the skinny projection tiles (k_gemm2.hip): per shape key the
  // chosen (tile height mb, split-K count) and its time in microseconds.
  // tune_mode 1: the candidates are timed on a shape's first call (not while
  // the stream is captured or in the fixed-arena mode); 0: the built-in rule.
  // force_mb / force_splits (> 0) pin a configuration.
  struct TuneChoice {
    int mb = 2;
    int splits = 1;
    double us = 0.0;
    bool tuned = false;
  };
  int tune_mode = 1;
  bool tune_log = false;
  int force_mb = 0, force_splits = 0;
  std::map<std::string, TuneChoice> tune_table;
  int32_t* defer_check(int count, int code, const std::string& msg, int fail_value = 0);
  void flush_checks();
  // Device -> host copies completed by the next flush (the entry point's one
  // synchronisation): staged in pinned memory, copied to dst after the sync.
  struct HostCopy {
    void* dst;
    size_t off, bytes;
  };
  unsigned char* stage_host = nullptr;
  size_t stage_cap = 0, stage_used = 0;
  std::vector<HostCopy> host_copies;
  void to_host_later(void* dst, const void* src, size_t bytes);
  void complete_host_copies(bool synced);
  // Persistent zero buffer of integer class counts (zero between launches:
  // the kernel that uses it drains it).
  unsigned long long* zcounts_dev = nullptr;
  static constexpr int kZeroCounts = 16384;
  unsigned long long* zero_counts();
  void drop_checks();
  // Fixed-arena mode of the allocation-free entry points: while arena_base is
  // set, alloc() carves 256-byte aligned blocks from the device range
  // [arena_base, arena_base + arena_cap) as a stack instead of the pool (a
  // freed block returns to the arena once every block above it is freed too),
  // and running out raises FPLS_ERR_ARG.
  struct ArenaBlock {
    size_t off, bytes;
    bool live;
  };
  unsigned char* arena_base = nullptr;
  size_t arena_cap = 0, arena_used = 0, arena_peak = 0;
  std::vector<ArenaBlock> arena_blocks;
  // Per-device address of the library's static scratch array (set up when
  // the context is created; null if that failed, with the reason below).
  void* scratch_ptr = nullptr;
  size_t scratch_bytes = 0;
  std::string scratch_err;
  void* alloc(size_t bytes);
  void free(void* p);
  // Byte budget for the split-K FP32 partials of one product (smaller inside
  // the fixed arena, so that the slab count, not the call, gives way).
  size_t partial_budget(size_t normal = size_t(256) << 20) const {
    if (!arena_base) return normal;
    return (arena_cap - arena_used) / 2;
  }
  hipEvent_t event();
  void collect_timing();
};

template <typename T>
DBuf<T>::DBuf(Ctx* c, size_t count) : ctx(c), n(count) {
  ptr = count ? static_cast<T*>(c->alloc(count * sizeof(T))) : nullptr;
}
template <typename T>
DBuf<T>::~DBuf() { reset(); }
template <typename T>
void DBuf<T>::reset() {
  if (ptr && ctx) ctx->free(ptr);
  ptr = nullptr;
  n = 0;
}

// Launch wrapper: brackets the launch with HIP events when timing is on.
template <typename Kernel, typename... Args>
void launch(Ctx* ctx, const char* name, Kernel k, dim3 grid, dim3 block, size_t shmem,
            Args... args) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ctx->timing) {
    e0 = ctx->event();
    e1 = ctx->event();
    FPLS_HIP(hipEventRecord(e0, ctx->stream));
  }
  hipLaunchKernelGGL(k, grid, block, shmem, ctx->stream, args...);
  FPLS_HIP(hipGetLastError());
  if (ctx->timing) {
    FPLS_HIP(hipEventRecord(e1, ctx->stream));
    ctx->pending.push_back({name, {e0, e1}});
  }
}

// ---------------------------------------------------------------- views
// A device-resident column-major matrix (after staging).
struct DMat {
  const void* p = nullptr;
  int dtype = FPLS_F64;
  int64_t rows = 0, cols = 0, ld = 0;
  const float* f() const { return static_cast<const float*>(p); }
  const double* d() const { return static_cast<const double*>(p); }
};
// A column-major FP64 device array as a DMat.
inline DMat dview(const double* p, int64_t rows, int64_t cols, int64_t ld) {
  DMat d;
  d.p = p;
  d.dtype = FPLS_F64;
  d.rows = rows;
  d.cols = cols;
  d.ld = ld;
  return d;
}

// Stage a caller matrix on the device (host buffers are copied into `hold`).
DMat stage(Ctx* ctx, const fpls_mat* m, DBuf<unsigned char>& hold, const char* what);
// Stage labels; validates 1..C when check is set.
const int32_t* stage_labels(Ctx* ctx, const fpls_ivec* y, int64_t n, DBuf<int32_t>& hold,
                            const char* what);

// Host -> device copy that returns only when the source may be reused (the
// small bookkeeping arrays copied this way live on the host stack).
inline void h2d(Ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  FPLS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  FPLS_HIP(hipStreamSynchronize(ctx->stream));
}
void to_host(Ctx* ctx, double* dst, const double* src, size_t n);
void to_host_i32(Ctx* ctx, int32_t* dst, const int32_t* src, size_t n);
void to_device(Ctx* ctx, double* dst, const double* src, size_t n);
void sync(Ctx* ctx);

// ---------------------------------------------------------------- kernels (k_blas.hip)
void fill_f64(Ctx* ctx, double* x, double v, int64_t n);
void fill_f32(Ctx* ctx, float* x, float v, int64_t n);
void fill_i32(Ctx* ctx, int32_t* x, int32_t v, int64_t n);
void copy_i32(Ctx* ctx, int32_t* d, const int32_t* s, int64_t n);
void copy_f64(Ctx* ctx, double* dst, const double* src, int64_t n);
// Copy a rows x cols block between column-major buffers with different ld.
void copy2d_f64(Ctx* ctx, double* dst, int64_t ldd, const double* src, int64_t lds,
                int64_t rows, int64_t cols);
// B (cols x rows, ldb) = A^T (A: rows x cols, lda)
void transpose_f64(Ctx* ctx, const double* A, int64_t lda, int64_t rows, int64_t cols,
                   double* B, int64_t ldb);
void f64_to_f32(Ctx* ctx, float* dst, const double* src, int64_t n);
void f32_to_f64(Ctx* ctx, double* dst, const float* src, int64_t n);

// General FP64 GEMM: C = alpha op(A) op(B) + beta C, with optional per-stored-
// column shifts subtracted from A and B on load (centering without writing
// the centred matrix).  shiftA indexes the stored columns of A.
void gemm_f64(Ctx* ctx, bool transA, bool transB, int64_t M, int64_t N, int64_t K, double alpha,
              const double* A, int64_t lda, const double* shiftA, const double* B, int64_t ldb,
              const double* shiftB, double beta, double* C, int64_t ldc,
              const int32_t* labelsB = nullptr, const char* name = nullptr);

// FP32-input contractions on MFMA.
//  nn: C(M x N) = (A - 1 shiftA^T) B - 1 offset^T     A: M x K (ld lda), B: K x N FP32
//      out FP32 (Cf) or FP64 (Cd), whichever is non-null.
//  tn: C(M x N) = (A - 1 shiftA^T)^T (B - 1 shiftB^T)  A: K x M, B: K x N (K long)
//      FP32 MFMA partials per K-slab, reduced in FP64 into C (FP64).
void gemm_nn_f32(Ctx* ctx, int precision, int64_t M, int64_t N, int64_t K, const float* A,
                 int64_t lda, const double* shiftA, const float* B, int64_t ldb,
                 const double* offset, float* Cf, double* Cd, int64_t ldc,
                 bool c_rowmajor = false);
// The same with its operands prepared once by the caller: B zero-padded to
// nn_pad_k(K) x nn_pad_n(N) (ld nn_pad_k(K), from nn_pad_operand) and, with
// shiftA, its FP32 copy shift32 (nn_shift_operand).  A model keeps these for
// the projections it repeats.
void gemm_nn_f32_padded(Ctx* ctx, int precision, int64_t M, int64_t N, int64_t K,
                        const float* A, int64_t lda, const double* shiftA,
                        const float* shift32, const float* B, const double* offset, float* Cf,
                        double* Cd, int64_t ldc, bool c_rowmajor = false);
int64_t nn_pad_k(int64_t K);
int64_t nn_pad_n(int64_t N);
DBuf<float> nn_pad_operand(Ctx* ctx, const double* B, int64_t ldb, bool trans, int64_t K,
                           int64_t N);
DBuf<float> nn_shift_operand(Ctx* ctx, const double* shift, int64_t K);
// C (M x N col-major FP32, ld M) = A B^T, A: M x K and B: N x K both row-major.
void gemm_rr_f32(Ctx* ctx, int precision, int64_t M, int64_t N, int64_t K, const float* A,
                 int64_t lda, const float* B, int64_t ldb, float* C);
//      labelsB (1-based codes) replaces B by the one-hot matrix onehot(y) (K x N)
//      generated on the fly: the dense response is never formed.
void gemm_tn_f32(Ctx* ctx, int precision, int64_t M, int64_t N, int64_t K, const float* A,
                 int64_t lda, const double* shiftA, const float* B, int64_t ldb,
                 const int32_t* labelsB, const double* shiftB, double* C, int64_t ldc);

// The dense-B TN product with an explicit split-K count (gemm_tn_f32 picks
// tn_default_splits), and the split-K skinny NN slabs of the batched fits.
int64_t tn_default_splits(const Ctx* ctx, int64_t M, int64_t N, int64_t K);
void gemm_tn_dense(Ctx* ctx, int precision, int64_t M, int64_t N, int64_t K, const float* A,
                   int64_t lda, const double* shiftA, const float* B, int64_t ldb,
                   const int32_t* labelsB, const double* shiftB, double* C, int64_t ldc,
                   int64_t splits);
void proj_splitk(Ctx* ctx, int precision, int64_t M, int N, int64_t K, const float* A,
                 int64_t lda, const float* shA32, const float* Bpad, int64_t ldb, int64_t kchunk,
                 int splits, float* part);
// Skinny TN (N <= 128) with a dense zero-padded B: FP32 (Bf) or split into
// bf16 hi / lo planes (Bh, Bl; [n][k], ld ldb >= K rounded up to 32).
void tn_skinny(Ctx* ctx, int precision, int64_t M, int N, int64_t K, const float* A, int64_t lda,
               const double* shA, const float* Bf, const __bf16* Bh, const __bf16* Bl,
               int64_t ldb, int64_t splits, double* C, int64_t ldc);
// Device Gaussian streams (k_rng.hip): column s of out (len values, ld) is the
// std::mt19937(seeds[s]) + std::normal_distribution<double>(0, 1) stream.
void gaussian_dev(Ctx* ctx, const uint32_t* seeds_dev, int n_streams, int64_t len, double* out,
                  int64_t ld);
// Matrix-vector products on a (possibly centred) data matrix.
//  y(M) = (A - 1 shift^T) x      (A: M x K col-major; FP32 or FP64)
//  y(K) = (A - 1 shift^T)^T x
void gemv_n(Ctx* ctx, const DMat& A, const double* shift, const double* x, double* y);
// y = (A - 1 shift^T) x, then y /= |y| and also (n2 entries) /= |y| in the same
// launch (a zero or non-finite norm sets flags[0]; nothing is normalised once
// it is set).  False when the shape does not fit the one-launch form
// (rows <= 2048, 16-byte vector rows) and nothing was launched.
bool gemv_n_normalize(Ctx* ctx, const DMat& A, const double* shift, const double* x, double* y,
                      double* also, int64_t n2, int32_t* flags);
void gemv_t(Ctx* ctx, const DMat& A, const double* shift, const double* x, double* y);

// ---------------------------------------------------------------- kernels (k_stats.hip)
// Column sums of X (FP64 accumulation, fixed order).
void col_sums(Ctx* ctx, const DMat& X, double* sums);
// Column sums of squared deviations from `center`.
void col_ssd(Ctx* ctx, const DMat& X, const double* center, double* ssd);
// One pass of either (center null: sums; else squared deviations) finished in
// the same kernel: out = sum / div, then post 1: sqrt, post 2: sqrt with zero
// or non-finite mapped to 1; ones (optional) is set to 1 in every column.
void col_stat(Ctx* ctx, const DMat& X, const double* center, double* out, double div, int post,
              double* ones = nullptr);
// Class sums of an FP32 X (n x p, ld) from labels y in 1..C in one pass (any
// C up to 1663): S (p x C, ld lds) = X^T onehot(y), or with shiftB the centred
// X^T (onehot(y) - 1 shiftB^T) (independent of any centring of X); mean (p,
// optional) gets the column means of X and ones (p, optional) the value 1.
// Returns false (nothing launched) when C is too large.
bool class_sums_f32(Ctx* ctx, int64_t n, int64_t p, const float* X, int64_t ld, const int32_t* y,
                    int C, const double* shiftB, double* S, int64_t lds, double* mean,
                    double* ones);
// The centred class cross-product S (p x C, ld lds) = (X - 1 m^T)^T
// (onehot(y) - 1 mY^T) of an FP32 X for C <= 127 on the one-hot MFMA tiles in
// one pass over X, which also yields the column means (mean, optional) and
// ones (p, optional, set to 1).  Returns false when C is too large.
bool class_cross_cov_f32(Ctx* ctx, int precision, int64_t p, int64_t C, int64_t n,
                         const float* X, int64_t ld, const int32_t* y, const double* mY,
                         double* S, int64_t lds, double* mean, double* ones);
// Class counts from 1-based labels; *bad counts labels outside 1..C (cleared
// first unless bad_zeroed: the slots of deferred checks are zero already).
void label_counts(Ctx* ctx, const int32_t* y, int64_t n, int32_t C, int64_t* counts,
                  int32_t* bad, bool bad_zeroed = false);
// FP64 class proportions prop[c] = count_c / n and counts (optional) in one
// launch; *bad (zero on entry) counts labels outside 1..C.  Returns false
// (nothing launched) when the per-block partials would be too large.
bool label_counts_prop(Ctx* ctx, const int32_t* y, int64_t n, int32_t C, double* prop,
                       double* counts, int32_t* bad);

// ---------------------------------------------------------------- kernels (k_dense.hip)
// Symmetric eigen-decomposition (FP64), n <= kJacobiMax: tridiagonal path
// with the divide-and-conquer solver as fallback.  Eigenvalues descending in
// w, eigenvectors in V (ld n).  kwant (> 0): only the kwant largest pairs
// are used by the caller, so only their accuracy decides the fallback (the
// others are returned too, possibly less orthogonal); <= 0: all of them.
constexpr int kJacobiMax = 128;
void sym_eig(Ctx* ctx, const double* A, int64_t lda, int n, double* w, double* V,
             int kwant = 0);
// count independent problems (order ns_dev[b] <= nmax, or n0 for every
// problem when ns_dev is null; A + b strideA with ld lda); eigenvalues
// descending at w + b strideW, eigenvectors at V + b strideV (ld = the
// problem's order).
// dc_only: straight to the divide-and-conquer solver (spectra known to be
// clustered, such as the pair blocks of the block Jacobi).
void sym_eig_batched(Ctx* ctx, const double* A, int64_t lda, int64_t strideA, const int* ns_dev,
                     int nmax, int count, double* w, int64_t strideW, double* V,
                     int64_t strideV, int n0 = 0, int kwant = 0, bool dc_only = false);
// The tridiagonal path of sym_eig_batched (k_eig_ii.hip): Householder
// reduction with the matrix in registers, bisection and twisted-factorisation
// eigenvectors one workgroup per eigenvalue, Gram-Schmidt inside clusters.
// flags[b] (device) is set for a problem whose clusters or residuals need the
// divide-and-conquer solver.
void sym_eig_tri(Ctx* ctx, const double* A, int64_t lda, int64_t strideA, const int* ns_dev,
                 int n0, int nmax, int count, double* w, int64_t strideW, double* V,
                 int64_t strideV, int kwant, int* flags);
// Any order: block Jacobi over pairs of <= 64-row blocks, each pair
// diagonalised by the one-workgroup solver (k_jacobi.hip).  Same outputs as
// sym_eig (eigenvalues descending in w, eigenvectors in V, ld n).
void sym_eig_large(Ctx* ctx, const double* A, int64_t lda, int n, double* w, double* V,
                   int kwant = 0);
// Orders 129 .. 2048 on the whole chip (k_eig_large.hip): blocked Householder
// tridiagonalisation, the K = kwant (or n) largest eigenpairs by bisection and
// twisted factorisation, symmetric re-orthogonalisation and a blocked
// back-transformation.  Writes w[0 .. K) descending and V[:, 0 .. K) (ld n);
// returns false (outputs unspecified) when a check fails and the caller must
// use another method.
bool sym_eig_chip(Ctx* ctx, const double* A, int64_t lda, int n, double* w, double* V,
                  int kwant);
// spd_solve_slices for orders above 128, through the eigendecomposition.
void spd_solve_large(Ctx* ctx, const double* A, int64_t lda, const std::vector<int>& ks,
                     int n_slices, int mode, const double* dvals, double ridge, double* X,
                     int64_t ldx, int64_t slice_stride, double* lambda_out, int32_t* status);
// Cholesky-based solves of leading blocks: for each slice s, solve
// (A[:k_s,:k_s] + diag_shift_s I) X = RHS, with LU fallback.  Used for the
// PLSSVD latent coefficients and the LDA inverse covariance.
//   mode 0: A_k X = diag(d)            mode 1: (A_k + lambda I) X = I (LDA ridge rule)
//   mode 2: (A_k + 1e-10 mean(diag A_k) I) X = diag(d)   (label-aware PLSSVD)
// d_k: device array of slice orders; status[s]: 0 Cholesky, 2 LU fallback, 1 singular.
void spd_solve_slices(Ctx* ctx, const double* A, int64_t lda, int kmax, const std::vector<int>& ks,
                      int n_slices, int mode, const double* dvals, double ridge, double* X,
                      int64_t ldx, int64_t slice_stride, double* lambda_out, int32_t* status);
// One solve of spd_solve_slices' kind, on its own matrix: X (k x k, ld ldx) =
// (A_k + shift I)^{-1} RHS by mode; lambda_out / status single entries.
struct SpdJob {
  const double* A;
  int64_t lda;
  int k, mode;
  const double* d;
  double ridge;
  double* X;
  int64_t ldx;
  double* lambda_out;
  int32_t* status;
};
constexpr int kSpdJobMax = 16;
struct SpdJobs {
  SpdJob j[kSpdJobMax];
};
// Independent solves (orders <= 128) in as few launches as possible, one
// workgroup each.
void spd_solve_jobs(Ctx* ctx, const std::vector<SpdJob>& jobs);
// The jobs spd_solve_slices would run (orders <= 128), appended to jobs.
void spd_slice_jobs(const double* A, int64_t lda, int kmax, const std::vector<int>& ks,
                    int n_slices, int mode, const double* dvals, double ridge, double* X,
                    int64_t ldx, int64_t slice_stride, double* lambda_out, int32_t* status,
                    std::vector<SpdJob>& jobs);
void scale_cols(Ctx* ctx, double* X, int64_t ld, int64_t rows, int64_t cols, const double* f,
                bool divide);
// Linv = chol(G + shift I)^{-1} (lower), with shift escalation on breakdown.
void chol_inv_lower(Ctx* ctx, const double* G, int64_t ldg, int l, int64_t rows, double shift,
                    double* Linv, int32_t* status);
void scale_rows_div(Ctx* ctx, double* X, int64_t ld, int64_t rows, int64_t cols,
                    const double* f);

// ---------------------------------------------------------------- scratch (noalloc.hip)
// Looks up (once per device, with ctx->device current) the address of the
// library's static scratch array and records it in ctx->scratch_ptr.
void scratch_setup(Ctx* ctx);

// ---------------------------------------------------------------- RNG (host)
// std::mt19937(seed) + std::normal_distribution<double>(0,1), column-major.
void gaussian_host(double* out, int64_t n, uint32_t seed);

}  // namespace fpls

// The opaque ABI structs.
struct fpls_ctx {
  fpls::Ctx c;
};
