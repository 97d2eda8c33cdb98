// trn_api.hpp -- internal declarations shared by the translation units that
// implement the C ABI.
#pragma once

#include <functional>

#include "trn_runtime.hpp"

namespace trn {

const char* last_error();
// Runs `body`, converting internal errors into a status code and a
// thread-local message prefixed with `where`.
int guard_impl(const char* where, const std::function<void()>& body);

// Makes the context's device current for the duration of an entry point.
class DeviceScope {
 public:
  explicit DeviceScope(trn_ctx* ctx);
  ~DeviceScope();

 private:
  int prev_ = 0, dev_ = 0;
};

void flush_timing(trn_ctx* ctx);
void copy_async(trn_ctx* ctx, void* dst, const void* src, size_t bytes);

// Exclusive scan of n int32 values into out (out[n] receives the total when
// out has n+1 entries and with_total is set).  In place is allowed.
// tmp: exclusive_scan_scratch(n) bytes of device scratch.
void exclusive_scan_i32(trn_ctx* ctx, const int32_t* in, int32_t* out, int64_t n,
                        bool with_total, int32_t* tmp);
size_t exclusive_scan_scratch(int64_t n);

// Z-curve order of the vertices from their coordinates (order.hip).
void spatial_order(trn_ctx* ctx, int32_t nv, const double* xy, int32_t* perm);

// Per-vertex sums of per-triangle areas over the stored 2-ring patches, in
// patch order (the patch areas of Metrics::computeGradient,
// src/hessian.cpp:53), from metric.hip's gather.
void patch_area_sums(trn_ctx* ctx, const trn_mesh* m, const double* area_t, double* tarea);

// Reductions used by several passes.
void reduce_max_i32(trn_ctx* ctx, const int32_t* in, int64_t n, int32_t* out_d);

// Per-thread LDS sorted set: element k of thread `tid` lives at
// s[k * stride]; returns false when the set is full.
__device__ __forceinline__ bool lds_sorted_insert(int* s, int stride, int& m,
                                                  int cap, int x) {
  int p = m;
  while (p > 0) {
    const int y = s[(p - 1) * stride];
    if (y <= x) {
      if (y == x) return true;
      break;
    }
    --p;
  }
  if (m >= cap) return false;
  for (int k = m; k > p; --k) s[k * stride] = s[(k - 1) * stride];
  s[p * stride] = x;
  ++m;
  return true;
}

constexpr int kMaxPatchCap = 256;  // 2-ring elements per vertex the gathers hold
constexpr int kMaxDegree = 64;     // incident triangles per vertex in smoothing

}  // namespace trn
