// runtime.hip -- context lifetime, error reporting, per-kernel timing,
// scratch workspace and the device-math probe of libtrinity_amd.so.
#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "trn_api.hpp"
#include "trn_numeric.hpp"

namespace trn {

static thread_local std::string g_last_error;

[[noreturn]] void fail(int code, const std::string& msg) { throw Error{code, msg}; }
void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

int guard_impl(const char* where, const std::function<void()>& body) {
  try {
    body();
    return TRN_OK;
  } catch (const Error& e) {
    set_last_error(std::string(where) + ": " + e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error(std::string(where) + ": host allocation failed");
    return TRN_ERR_STATE;
  } catch (const std::exception& e) {
    set_last_error(std::string(where) + ": " + e.what());
    return TRN_ERR_STATE;
  } catch (...) {
    set_last_error(std::string(where) + ": unknown failure");
    return TRN_ERR_STATE;
  }
}

DeviceScope::DeviceScope(trn_ctx* ctx) {
  TRN_REQUIRE(ctx != nullptr, TRN_ERR_ARG, "null context");
  TRN_HIP(hipGetDevice(&prev_));
  if (prev_ != ctx->device) TRN_HIP(hipSetDevice(ctx->device));
  dev_ = ctx->device;
}
DeviceScope::~DeviceScope() {
  if (prev_ != dev_) (void)hipSetDevice(prev_);
}

int stat_index(trn_ctx* ctx, const char* name) {
  for (size_t i = 0; i < ctx->stats.size(); ++i)
    if (ctx->stats[i].name == name) return (int)i;
  ctx->stats.push_back(KernelStat{name, 0.0, 0, 0.0});
  return (int)ctx->stats.size() - 1;
}

static hipEvent_t pool_event(trn_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  TRN_HIP(hipEventCreate(&e));
  return e;
}

void timing_begin(trn_ctx* ctx, int, hipEvent_t* start, hipEvent_t* stop) {
  *start = pool_event(ctx);
  *stop = pool_event(ctx);
  TRN_HIP(hipEventRecord(*start, ctx->stream));
}

void timing_end(trn_ctx* ctx, int stat, hipEvent_t start, hipEvent_t stop,
                double bytes) {
  TRN_HIP(hipEventRecord(stop, ctx->stream));
  ctx->pending.push_back(PendingTiming{stat, start, stop});
  ctx->stats[stat].bytes += bytes;
  ctx->stats[stat].launches += 1;
  if (ctx->pending.size() > 4096) flush_timing(ctx);
}

unsigned long long* timing_units(trn_ctx* ctx, const char* name, double bytes_per_unit) {
  if (!timed(ctx, name)) return nullptr;
  const int st = stat_index(ctx, name);
  if (st >= kTimingUnitSlots) return nullptr;
  if (!ctx->unit_dev) {
    TRN_HIP(hipMalloc(&ctx->unit_dev, sizeof(unsigned long long) * kTimingUnitSlots));
    TRN_HIP(hipMemset(ctx->unit_dev, 0, sizeof(unsigned long long) * kTimingUnitSlots));
  }
  ctx->stats[st].bytes_per_unit = bytes_per_unit;
  return ctx->unit_dev + st;
}

void flush_timing(trn_ctx* ctx) {
  if (ctx->pending.empty()) return;
  TRN_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& p : ctx->pending) {
    float ms = 0.f;
    TRN_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
    ctx->stats[p.stat].total_ms += ms;
    ctx->event_pool.push_back(p.start);
    ctx->event_pool.push_back(p.stop);
  }
  ctx->pending.clear();
  if (ctx->unit_dev) {
    unsigned long long units[kTimingUnitSlots];
    TRN_HIP(hipMemcpy(units, ctx->unit_dev, sizeof(units), hipMemcpyDeviceToHost));
    TRN_HIP(hipMemset(ctx->unit_dev, 0, sizeof(units)));
    for (size_t st = 0; st < ctx->stats.size() && st < (size_t)kTimingUnitSlots; ++st)
      ctx->stats[st].bytes += (double)units[st] * ctx->stats[st].bytes_per_unit;
  }
}

Scratch::Scratch(trn_ctx* ctx, size_t total) : ctx_(ctx) {
  TRN_REQUIRE(ctx->ws.live == 0, TRN_ERR_STATE,
              "internal: a second workspace user while one is live");
  if (total > ctx->ws.size) {
    // the previous workspace may still be read by queued work
    TRN_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->ws.base) TRN_HIP(hipFree(ctx->ws.base));
    ctx->ws.base = nullptr;
    ctx->ws.size = 0;
    size_t sz = total + total / 4 + (1u << 20);
    TRN_HIP(hipMalloc(&ctx->ws.base, sz));
    ctx->ws.size = sz;
  }
  ctx->ws.live = 1;
}

Scratch::~Scratch() { ctx_->ws.live = 0; }

void* aux_workspace(trn_ctx* ctx, size_t bytes) {
  if (bytes > ctx->aux.size) {
    TRN_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->aux.base) TRN_HIP(hipFree(ctx->aux.base));
    ctx->aux.base = nullptr;
    ctx->aux.size = 0;
    const size_t sz = bytes + bytes / 4;
    TRN_HIP(hipMalloc(&ctx->aux.base, sz));
    ctx->aux.size = sz;
  }
  return ctx->aux.base;
}

void sync(trn_ctx* ctx) { TRN_HIP(hipStreamSynchronize(ctx->stream)); }

// Byte fill as a kernel rather than hipMemsetAsync, so that every sequence
// of launches a caller may capture into a hipGraph holds kernel nodes only
// (memset nodes have given wrong results on graph replays).  16-byte
// stores over the aligned middle, single bytes at the ends.
__global__ __launch_bounds__(256) void k_fill_bytes(uint8_t* __restrict__ p, uint32_t v,
                                                    uint64_t head, uint64_t nvec,
                                                    uint64_t bytes) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t w = v * 0x01010101u;
  uint4* q = (uint4*)(p + head);
  for (uint64_t k = tid; k < nvec; k += stride) q[k] = make_uint4(w, w, w, w);
  const uint64_t tail0 = head + 16 * nvec;
  for (uint64_t k = tid; k < head; k += stride) p[k] = (uint8_t)v;
  for (uint64_t k = tail0 + tid; k < bytes; k += stride) p[k] = (uint8_t)v;
}

void memset_async(trn_ctx* ctx, void* p, int v, size_t bytes) {
  if (!bytes) return;
  uint64_t head = (16 - ((uintptr_t)p & 15)) & 15;
  if (head > bytes) head = bytes;
  const uint64_t nvec = (bytes - head) / 16;
  uint64_t blocks = (nvec + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_bytes, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                     (uint8_t*)p, (uint32_t)(v & 0xff), head, nvec, (uint64_t)bytes);
  TRN_HIP(hipGetLastError());
}

void copy_async(trn_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes) TRN_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
}

// ---- device math probe ---------------------------------------------------
__global__ void k_debug_math(int op, int64_t n, const double* __restrict__ x,
                             const double* __restrict__ y, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r;
  switch (op) {
    case 0: r = trn::libm::log(x[i]); break;
    case 1: r = trn::libm::exp(x[i]); break;
    case 2: r = trn::libm::pow(x[i], y[i]); break;
    case 3: r = trn::cube_pow(x[i]); break;
    default: r = __builtin_sqrt(x[i]); break;
  }
  out[i] = r;
}

}  // namespace trn

using namespace trn;

extern "C" {

const char* trn_version(void) { return "trinity_amd 0.1.0 (gfx950, FP64)"; }
const char* trn_last_error(void) { return trn::last_error(); }

int trn_ctx_create(int device, trn_ctx** out) {
  return guard_impl("trn_ctx_create", [&] {
    TRN_REQUIRE(out != nullptr, TRN_ERR_ARG, "null output");
    *out = nullptr;
    int count = 0;
    TRN_HIP(hipGetDeviceCount(&count));
    TRN_REQUIRE(device >= 0 && device < count, TRN_ERR_ARG, "no such device");
    int prev = 0;
    TRN_HIP(hipGetDevice(&prev));
    TRN_HIP(hipSetDevice(device));
    trn_ctx* c = new trn_ctx();
    c->device = device;
    // a blocking stream: ordered with the null stream that host frameworks
    // use by default, so results are visible to them without extra syncs
    hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamDefault);
    if (e != hipSuccess) {
      delete c;
      (void)hipSetDevice(prev);
      fail(TRN_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    c->stream = c->own_stream;
    e = hipHostMalloc((void**)&c->pinned_i32, 4096 * sizeof(int32_t), hipHostMallocDefault);
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&c->pinned_f64, 512 * sizeof(double), hipHostMallocDefault);
    (void)hipSetDevice(prev);
    if (e != hipSuccess) {
      trn_ctx_destroy(c);
      fail(TRN_ERR_HIP, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    *out = c;
  });
}

int trn_ctx_destroy(trn_ctx* ctx) {
  if (!ctx) return TRN_OK;
  return guard_impl("trn_ctx_destroy", [&] {
    DeviceScope ds(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->pending) {
      (void)hipEventDestroy(p.start);
      (void)hipEventDestroy(p.stop);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->ws.base) (void)hipFree(ctx->ws.base);
    if (ctx->aux.base) (void)hipFree(ctx->aux.base);
    if (ctx->pinned_i32) (void)hipHostFree(ctx->pinned_i32);
    if (ctx->pinned_f64) (void)hipHostFree(ctx->pinned_f64);
    if (ctx->unit_dev) (void)hipFree(ctx->unit_dev);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
  });
}

int trn_ctx_set_stream(trn_ctx* ctx, void* hip_stream) {
  return guard_impl("trn_ctx_set_stream", [&] {
    TRN_REQUIRE(ctx != nullptr, TRN_ERR_ARG, "null context");
    flush_timing(ctx);
    ctx->stream = (hipStream_t)hip_stream;  // NULL: the null stream
  });
}

int trn_ctx_use_own_stream(trn_ctx* ctx) {
  return guard_impl("trn_ctx_use_own_stream", [&] {
    TRN_REQUIRE(ctx != nullptr, TRN_ERR_ARG, "null context");
    flush_timing(ctx);
    ctx->stream = ctx->own_stream;
  });
}

int trn_ctx_synchronize(trn_ctx* ctx) {
  return guard_impl("trn_ctx_synchronize", [&] {
    DeviceScope ds(ctx);
    sync(ctx);
  });
}

int trn_ctx_set_timing(trn_ctx* ctx, int enable) {
  return guard_impl("trn_ctx_set_timing", [&] {
    TRN_REQUIRE(ctx != nullptr, TRN_ERR_ARG, "null context");
    ctx->timing = enable != 0;
  });
}

int trn_ctx_set_timing_only(trn_ctx* ctx, const char* name) {
  return guard_impl("trn_ctx_set_timing_only", [&] {
    TRN_REQUIRE(ctx != nullptr, TRN_ERR_ARG, "null context");
    ctx->timing_only = name ? name : "";
  });
}

int trn_ctx_reset_timing(trn_ctx* ctx) {
  return guard_impl("trn_ctx_reset_timing", [&] {
    DeviceScope ds(ctx);
    flush_timing(ctx);
    ctx->stats.clear();
  });
}

int trn_ctx_timing_count(trn_ctx* ctx, int32_t* count) {
  return guard_impl("trn_ctx_timing_count", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(count != nullptr, TRN_ERR_ARG, "null output");
    flush_timing(ctx);
    *count = (int32_t)ctx->stats.size();
  });
}

int trn_ctx_timing_get(trn_ctx* ctx, int32_t index, const char** name,
                       double* total_ms, int64_t* launches, double* bytes) {
  return guard_impl("trn_ctx_timing_get", [&] {
    DeviceScope ds(ctx);
    flush_timing(ctx);
    TRN_REQUIRE(index >= 0 && (size_t)index < ctx->stats.size(), TRN_ERR_ARG,
                "timing index out of range");
    const KernelStat& s = ctx->stats[index];
    if (name) *name = s.name.c_str();
    if (total_ms) *total_ms = s.total_ms;
    if (launches) *launches = s.launches;
    if (bytes) *bytes = s.bytes;
  });
}

int trn_debug_math(trn_ctx* ctx, int32_t op, int64_t n, const double* x_d,
                   const double* y_d, double* out_d) {
  return guard_impl("trn_debug_math", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(n >= 0 && x_d && out_d && (op != 2 || y_d), TRN_ERR_ARG, "bad arguments");
    TRN_REQUIRE(op >= 0 && op <= 4, TRN_ERR_ARG, "unknown op");
    const int B = 256;
    launch(ctx, "debug_math", 0.0, k_debug_math, dim3((unsigned)((n + B - 1) / B)), dim3(B), 0,
           (int)op, n, x_d, y_d, out_d);
  });
}

}  // extern "C"
