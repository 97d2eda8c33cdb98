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
// The same for a disjoint union of meshes: grouped by mesh (seg_of[nv] its
// mesh, seg_off[nseg + 1] the meshes' first vertices, device arrays), each
// mesh along the Z-curve of its own bounding box.
void spatial_order_segments(trn_ctx* ctx, int32_t nv, const double* xy, int32_t nseg,
                            const int32_t* seg_of, const int32_t* seg_off, int32_t* perm);
// trn_mesh_create with the vertex order of the patch slabs given (device,
// a permutation of 0..nv-1) instead of derived from coordinates.
trn_mesh* mesh_create_ordered(trn_ctx* ctx, int32_t nv, int32_t nt, const int32_t* tri,
                              const uint8_t* tags, const int32_t* order_d);

// The metric pass as trn_metric_run runs it, for the adjoint's tape: the
// gradient (grad), the raw Hessian (hess), the locally normalised metric
// (mloc) and the patch areas (tarea) of the pass are kept beside the final
// metric (tensor); returns the complexity (synchronises).
double metric_tape(trn_ctx* ctx, const trn_mesh* m, const double* xy, const double* f,
                   double target_factor, int lp_norm, double h_min, double h_max, double* tensor,
                   double* grad, double* hess, double* mloc, double* tarea);

// Internal launch sequences shared by the entry points and the batched pass.
// Each takes its device scratch from the caller (sized by the *_scratch
// function) and only queues work on ctx->stream: no allocation, no host
// synchronisation, no memset node, so a caller may capture it in a hipGraph.
// metric.hip: gradient + Hessian recovery + local normalisation into tensor
// (grad_out / hess_out optional stage outputs).
size_t recover_scratch(const trn_mesh* m);
void recover_launch(trn_ctx* ctx, const trn_mesh* m, const double* xy, const double* f,
                    double* grad_out, double* hess_out, double* tensor, int lp_norm,
                    void* scratch, size_t bytes);
// smoothing.hip: Smooth::run's preparation (records, quality cache) and its
// sweeps; logm is [nv] device scratch of the preparation.  The sweeps return
// the device array of per-sweep (commits, tasks) when want_totals is set.
void smooth_prepare_launch(trn_ctx* ctx, const trn_mesh* m, const double* xy,
                           const double* tensor, double* qualit, bool recompute_qualit,
                           double4* logm);
size_t smooth_sweep_scratch(const trn_mesh* m, int ncolors, int sweeps, int max_class);
int32_t* smooth_sweeps_launch(trn_ctx* ctx, const trn_mesh* m, double* xy, double* tensor,
                              double* qualit, const int32_t* order_d, const int32_t* class_off,
                              int ncolors, int sweeps, int max_class, void* scratch,
                              size_t scratch_size, bool want_totals);

// Shapes the autotuner (tune.hip) measures.  metric.hip: one gradient and
// one Hessian gather at the given slabs per block, bracketed by ev[0..1] and
// ev[2..3]; scratch of tune_gather_scratch bytes.  smoothing.hip: one
// first-sweep phase over a class at the given register shape, bracketed by
// ev[0..1]; the mesh's records must be prepared.
size_t tune_gather_scratch(const trn_mesh* m);
void tune_gathers(trn_ctx* ctx, const trn_mesh* m, const double* xy, const double* f,
                  int lp_norm, void* scratch, size_t bytes, int grad_waves, int hess_waves,
                  hipEvent_t ev[4]);
void tune_phase(trn_ctx* ctx, const trn_mesh* m, int32_t count, const int32_t* list,
                double* qualit, uint8_t* stale, int32_t* counters, int minw, hipEvent_t ev[2]);

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
