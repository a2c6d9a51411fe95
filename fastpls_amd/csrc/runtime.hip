// Context, HBM pool, staging, timing and the small host utilities of the
// gfx950 fastPLS library.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "fpls_internal.h"

namespace fpls {

// ---------------------------------------------------------------- pool
void* Pool::get(size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  if (bytes == 0) bytes = 256;
  // smallest free block that fits without wasting more than half of it
  auto it = free_by_size.lower_bound(bytes);
  if (it != free_by_size.end() && it->first <= 2 * bytes + (1 << 20)) {
    const size_t i = it->second;
    free_by_size.erase(it);
    blocks[i].used = true;
    return blocks[i].ptr;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    // Drop cached blocks and retry once.
    (void)hipGetLastError();
    std::vector<Block> keep;
    for (const Block& b : blocks) {
      if (b.used) keep.push_back(b);
      else (void)hipFree(b.ptr);
    }
    blocks.swap(keep);
    index.clear();
    free_by_size.clear();
    for (size_t i = 0; i < blocks.size(); ++i) index[blocks[i].ptr] = i;
    FPLS_HIP(hipMalloc(&p, bytes));
  }
  index[p] = blocks.size();
  blocks.push_back({p, bytes, true});
  return p;
}

void Pool::put(void* p) {
  auto it = index.find(p);
  if (it == index.end()) return;
  Block& b = blocks[it->second];
  if (!b.used) return;
  b.used = false;
  free_by_size.emplace(b.bytes, it->second);
}

void Pool::release_all() {
  for (auto& b : blocks) (void)hipFree(b.ptr);
  blocks.clear();
  index.clear();
  free_by_size.clear();
}

size_t Pool::total_bytes() const {
  size_t t = 0;
  for (const auto& b : blocks) t += b.bytes;
  return t;
}

// ---------------------------------------------------------------- context allocation
void* Ctx::alloc(size_t bytes) {
  if (!arena_base) return pool.get(bytes);
  const size_t need = std::max<size_t>(256, (bytes + 255) & ~size_t(255));
  FPLS_REQUIRE(need <= arena_cap - arena_used, FPLS_ERR_ARG,
               "allocation-free call: its temporaries need more than the " +
                   std::to_string(arena_cap >> 20) + " MiB scratch array (" +
                   std::to_string((arena_used + need) >> 20) + " MiB asked so far)");
  void* p = arena_base + arena_used;
  arena_blocks.push_back({arena_used, need, true});
  arena_used += need;
  arena_peak = std::max(arena_peak, arena_used);
  return p;
}

void Ctx::free(void* p) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  if (arena_base && b >= arena_base && b < arena_base + arena_cap) {
    const size_t off = static_cast<size_t>(b - arena_base);
    for (size_t i = arena_blocks.size(); i-- > 0;) {
      if (arena_blocks[i].off == off && arena_blocks[i].live) {
        arena_blocks[i].live = false;
        break;
      }
    }
    while (!arena_blocks.empty() && !arena_blocks.back().live) {
      arena_used = arena_blocks.back().off;
      arena_blocks.pop_back();
    }
    return;
  }
  pool.put(p);
}

// ---------------------------------------------------------------- deferred checks
constexpr int kFlagSlots = 1024;

__global__ void k_fill_flags(int32_t* f, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = 0;
}

// The flag slots are zero between entry points: they are cleared once when
// allocated and again (one fill over the slots used) whenever a batch of
// checks is read back or dropped, so handing out a slot launches nothing.
int32_t* Ctx::defer_check(int count, int code, const std::string& msg, int fail_value) {
  FPLS_REQUIRE(count >= 1 && count <= kFlagSlots, FPLS_ERR_ARG, "too many deferred checks");
  if (!flags_dev) {
    FPLS_REQUIRE(!arena_base, FPLS_ERR_STATE,
                 "allocation-free call reached a deferred check before any other call");
    FPLS_HIP(hipMalloc(&flags_dev, kFlagSlots * sizeof(int32_t)));
    FPLS_HIP(hipHostMalloc(&flags_host, kFlagSlots * sizeof(int32_t), hipHostMallocDefault));
    fill_i32(this, flags_dev, 0, kFlagSlots);
  }
  if (nflags + count > kFlagSlots) flush_checks();
  int32_t* p = flags_dev + nflags;
  checks.push_back(Check{nflags, count, code, fail_value, msg});
  nflags += count;
  return p;
}

unsigned* Ctx::ticket() {
  if (!ticket_dev) {
    FPLS_REQUIRE(!arena_base, FPLS_ERR_STATE,
                 "allocation-free call reached a ticket counter before any other call");
    FPLS_HIP(hipMalloc(&ticket_dev, 64 * sizeof(unsigned)));
    fill_i32(this, reinterpret_cast<int32_t*>(ticket_dev), 0, 64);
  }
  return ticket_dev;
}

void Ctx::drop_checks() {
  // an entry point that failed half way: its kernels may have set slots
  if (flags_dev && nflags > 0) {
    hipLaunchKernelGGL(k_fill_flags, dim3((nflags + 255) / 256), dim3(256), 0, stream, flags_dev,
                       nflags);
    (void)hipGetLastError();
  }
  checks.clear();
  nflags = 0;
  host_copies.clear();
  stage_used = 0;
}

void Ctx::to_host_later(void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  const size_t need = (bytes + 255) & ~size_t(255);
  if (stage_used + need > stage_cap) {
    complete_host_copies(false);
    if (need > stage_cap) {
      if (stage_host) FPLS_HIP(hipHostFree(stage_host));
      stage_host = nullptr;
      stage_cap = std::max<size_t>(need, std::max<size_t>(2 * stage_cap, 1 << 20));
      FPLS_HIP(hipHostMalloc(reinterpret_cast<void**>(&stage_host), stage_cap,
                             hipHostMallocDefault));
    }
  }
  FPLS_HIP(hipMemcpyAsync(stage_host + stage_used, src, bytes, hipMemcpyDeviceToHost, stream));
  host_copies.push_back(HostCopy{dst, stage_used, bytes});
  stage_used += need;
}

void Ctx::complete_host_copies(bool synced) {
  if (host_copies.empty()) return;
  if (!synced) FPLS_HIP(hipStreamSynchronize(stream));
  for (const HostCopy& h : host_copies) std::memcpy(h.dst, stage_host + h.off, h.bytes);
  host_copies.clear();
  stage_used = 0;
}

unsigned long long* Ctx::zero_counts() {
  if (!zcounts_dev) {
    FPLS_REQUIRE(!arena_base, FPLS_ERR_STATE,
                 "allocation-free call reached the count buffer before any other call");
    FPLS_HIP(hipMalloc(&zcounts_dev, kZeroCounts * sizeof(unsigned long long)));
    fill_f64(this, reinterpret_cast<double*>(zcounts_dev), 0.0, kZeroCounts);
  }
  return zcounts_dev;
}

void Ctx::flush_checks() {
  if (checks.empty()) {
    complete_host_copies(false);
    nflags = 0;
    return;
  }
  FPLS_HIP(hipMemcpyAsync(flags_host, flags_dev, nflags * sizeof(int32_t), hipMemcpyDeviceToHost,
                          stream));
  FPLS_HIP(hipStreamSynchronize(stream));
  complete_host_copies(true);
  fill_i32(this, flags_dev, 0, nflags);
  std::vector<Check> cs;
  cs.swap(checks);
  nflags = 0;
  for (const Check& c : cs) {
    for (int i = 0; i < c.count; ++i) {
      const int32_t v = flags_host[c.slot + i];
      if (c.fail_value == 0 ? v != 0 : v == c.fail_value) fail(c.code, c.msg);
    }
  }
}

// ---------------------------------------------------------------- timing
hipEvent_t Ctx::event() {
  if (!spare_events.empty()) {
    hipEvent_t e = spare_events.back();
    spare_events.pop_back();
    return e;
  }
  hipEvent_t e;
  FPLS_HIP(hipEventCreate(&e));
  return e;
}

void Ctx::collect_timing() {
  if (pending.empty()) return;
  FPLS_HIP(hipStreamSynchronize(stream));
  for (auto& pe : pending) {
    float ms = 0.f;
    FPLS_HIP(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
    KStat& s = kstats[pe.first];
    s.launches += 1;
    s.total_ms += ms;
    spare_events.push_back(pe.second.first);
    spare_events.push_back(pe.second.second);
  }
  pending.clear();
}

// ---------------------------------------------------------------- staging
DMat stage(Ctx* ctx, const fpls_mat* m, DBuf<unsigned char>& hold, const char* what) {
  FPLS_REQUIRE(m != nullptr, FPLS_ERR_ARG, std::string(what) + ": matrix is NULL");
  FPLS_REQUIRE(m->dtype == FPLS_F32 || m->dtype == FPLS_F64, FPLS_ERR_ARG,
               std::string(what) + ": dtype must be FPLS_F32 or FPLS_F64");
  FPLS_REQUIRE(m->rows >= 0 && m->cols >= 0 && m->ld >= std::max<int64_t>(m->rows, 1),
               FPLS_ERR_ARG, std::string(what) + ": bad shape or leading dimension");
  FPLS_REQUIRE(m->rows == 0 || m->cols == 0 || m->ptr != nullptr, FPLS_ERR_ARG,
               std::string(what) + ": NULL data pointer");
  DMat d;
  d.dtype = m->dtype;
  d.rows = m->rows;
  d.cols = m->cols;
  const size_t es = m->dtype == FPLS_F32 ? 4 : 8;
  if (m->loc == FPLS_DEVICE) {
    d.p = m->ptr;
    d.ld = m->ld;
    return d;
  }
  FPLS_REQUIRE(m->loc == FPLS_HOST, FPLS_ERR_ARG, std::string(what) + ": bad location");
  hold = DBuf<unsigned char>(ctx, std::max<size_t>(1, size_t(m->rows) * m->cols * es));
  if (m->rows > 0 && m->cols > 0) {
    FPLS_HIP(hipMemcpy2DAsync(hold.get(), m->rows * es, m->ptr, m->ld * es, m->rows * es,
                              m->cols, hipMemcpyHostToDevice, ctx->stream));
  }
  d.p = hold.get();
  d.ld = std::max<int64_t>(m->rows, 1);
  return d;
}

const int32_t* stage_labels(Ctx* ctx, const fpls_ivec* y, int64_t n, DBuf<int32_t>& hold,
                            const char* what) {
  FPLS_REQUIRE(y != nullptr && y->ptr != nullptr, FPLS_ERR_ARG,
               std::string(what) + ": labels are NULL");
  FPLS_REQUIRE(y->n == n, FPLS_ERR_ARG,
               std::string(what) + ": one class label per training row is required");
  if (y->loc == FPLS_DEVICE) return y->ptr;
  hold = DBuf<int32_t>(ctx, std::max<int64_t>(n, 1));
  FPLS_HIP(hipMemcpyAsync(hold.get(), y->ptr, n * sizeof(int32_t), hipMemcpyHostToDevice,
                          ctx->stream));
  return hold.get();
}

void to_host(Ctx* ctx, double* dst, const double* src, size_t n) {
  if (!n) return;
  FPLS_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FPLS_HIP(hipStreamSynchronize(ctx->stream));
}

void to_host_i32(Ctx* ctx, int32_t* dst, const int32_t* src, size_t n) {
  if (!n) return;
  FPLS_HIP(hipMemcpyAsync(dst, src, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  FPLS_HIP(hipStreamSynchronize(ctx->stream));
}

void to_device(Ctx* ctx, double* dst, const double* src, size_t n) {
  if (!n) return;
  h2d(ctx, dst, src, n * sizeof(double));
}

void sync(Ctx* ctx) { FPLS_HIP(hipStreamSynchronize(ctx->stream)); }

// ---------------------------------------------------------------- RNG
// Same generator and distribution as the reference's sketch
// (src/svd_cpu_rsvd.cpp:10-22): the libstdc++ templates give identical draws.
void gaussian_host(double* out, int64_t n, uint32_t seed) {
  std::mt19937 rng(seed);
  std::normal_distribution<double> norm(0.0, 1.0);
  for (int64_t i = 0; i < n; ++i) out[i] = norm(rng);
}

}  // namespace fpls

using namespace fpls;

// ---------------------------------------------------------------- ABI
extern "C" {

void fpls_opts_default(fpls_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->scaling = 1;
  o->svd_method = FPLS_SVD_RSVD;
  o->oversample = 10;
  o->power = 1;
  o->seed = 1;
  o->inc_iters = 2;
  o->precision = FPLS_PREC_FP32;
  o->fit = 0;
  o->store_b = -1;
  o->opls_north = 1;
}

int fpls_available(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
    (void)hipGetLastError();
    return 0;
  }
  for (int d = 0; d < count; ++d) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) == hipSuccess &&
        std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
      return 1;
  }
  return 0;
}

int fpls_ctx_create(int device, fpls_ctx** out) {
  if (!out) return FPLS_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    (void)hipGetLastError();
    return FPLS_ERR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return FPLS_ERR_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return FPLS_ERR_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return FPLS_ERR_HIP;
  fpls_ctx* c = new fpls_ctx();
  c->c.device = device;
  c->c.num_cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return FPLS_ERR_HIP;
  }
  scratch_setup(&c->c);
  if (const char* e = std::getenv("FPLS_AUTOTUNE")) c->c.tune_mode = std::atoi(e) != 0 ? 1 : 0;
  if (const char* e = std::getenv("FPLS_AUTOTUNE_LOG")) c->c.tune_log = std::atoi(e) != 0;
  *out = c;
  return FPLS_OK;
}

int fpls_autotune_set(fpls_ctx* ctx, int32_t mode, int32_t force_mb, int32_t force_splits) {
  if (!ctx) return FPLS_ERR_ARG;
  if (force_mb != 0 && force_mb != 1 && force_mb != 2 && force_mb != 4) return FPLS_ERR_ARG;
  if (force_splits < 0) return FPLS_ERR_ARG;
  ctx->c.tune_mode = mode != 0 ? 1 : 0;
  ctx->c.force_mb = force_mb;
  ctx->c.force_splits = force_splits;
  ctx->c.tune_table.clear();
  return FPLS_OK;
}

int fpls_autotune_report(fpls_ctx* ctx, char* buf, int64_t cap) {
  if (!ctx || (!buf && cap > 0)) return FPLS_ERR_ARG;
  std::string js = "{";
  bool first = true;
  for (const auto& kv : ctx->c.tune_table) {
    char one[256];
    std::snprintf(one, sizeof(one), "%s\"%s\": {\"mb\": %d, \"rows\": %d, \"splits\": %d, "
                  "\"us\": %.2f, \"tuned\": %s}", first ? "" : ", ", kv.first.c_str(),
                  kv.second.mb, 64 * kv.second.mb, kv.second.splits, kv.second.us,
                  kv.second.tuned ? "true" : "false");
    js += one;
    first = false;
  }
  js += "}";
  if ((int64_t)js.size() + 1 > cap) return (int)std::min<size_t>(js.size() + 1, 1u << 30);
  std::memcpy(buf, js.c_str(), js.size() + 1);
  return FPLS_OK;
}

void fpls_ctx_destroy(fpls_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  (void)hipStreamSynchronize(ctx->c.stream);
  for (auto& pe : ctx->c.pending) {
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
  }
  for (auto e : ctx->c.spare_events) (void)hipEventDestroy(e);
  if (ctx->c.flags_dev) (void)hipFree(ctx->c.flags_dev);
  if (ctx->c.ticket_dev) (void)hipFree(ctx->c.ticket_dev);
  if (ctx->c.zcounts_dev) (void)hipFree(ctx->c.zcounts_dev);
  if (ctx->c.stage_host) (void)hipHostFree(ctx->c.stage_host);
  if (ctx->c.flags_host) (void)hipHostFree(ctx->c.flags_host);
  ctx->c.cv_plan.reset();
  ctx->c.pool.release_all();
  (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

const char* fpls_last_error(const fpls_ctx* ctx) { return ctx ? ctx->c.err.c_str() : ""; }

int fpls_ctx_synchronize(fpls_ctx* ctx) {
  if (!ctx) return FPLS_ERR_ARG;
  return hipStreamSynchronize(ctx->c.stream) == hipSuccess ? FPLS_OK : FPLS_ERR_HIP;
}

int fpls_timing_enable(fpls_ctx* ctx, int enable) {
  if (!ctx) return FPLS_ERR_ARG;
  try {
    ctx->c.collect_timing();
  } catch (const Error& e) {
    ctx->c.err = e.what();
    return e.code;
  }
  ctx->c.timing = enable != 0;
  ctx->c.kstats.clear();
  return FPLS_OK;
}

int fpls_kernel_time(fpls_ctx* ctx, const char* kernel, int64_t* launches, double* total_ms) {
  if (!ctx || !kernel) return FPLS_ERR_ARG;
  try {
    ctx->c.collect_timing();
  } catch (const Error& e) {
    ctx->c.err = e.what();
    return e.code;
  }
  auto it = ctx->c.kstats.find(kernel);
  if (launches) *launches = it == ctx->c.kstats.end() ? 0 : it->second.launches;
  if (total_ms) *total_ms = it == ctx->c.kstats.end() ? 0.0 : it->second.total_ms;
  return FPLS_OK;
}

}  // extern "C"
