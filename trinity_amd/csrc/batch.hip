// batch.hip -- many independent meshes of uneven sizes per launch.
//
// The reference processes one mesh per Metrics / Smooth object
// (src/metric.cpp:59-70, src/smoothing.cpp:274-302).  A batch is the
// disjoint union of its meshes: one device topology whose vertex and
// triangle indices are offset per mesh, built once by trn_mesh_create.
// Every stage of the path is local to a mesh's vertices and triangles
// (2-ring patches, 1-rings and colour classes never cross a mesh boundary
// of a disjoint union), so the plain kernels run over the union and each
// launch covers every mesh: a 2,000-vertex mesh and a 200,000-vertex one
// share each launch's workgroups and the chip stays full without per-mesh
// launches.  Two stages are per mesh and get kernels of their own:
//
//   k_batch_cplx_part / k_batch_cplx_final   the complexity
//       (src/metric.cpp:153-185) of every mesh, summed in exactly the tree
//       the plain path uses for a mesh of that size (a fixed grid of
//       min(ceil(nt_m / 256), 1024) blocks striding over the mesh's
//       triangles, then one block over the partials), so each mesh's
//       complexity, hence its metric, is bit-identical to a plain pass;
//       the target floor(nv_m * target_factor) (:39) and the scale
//       target / complexity (:197-198) are per mesh;
//   k_batch_global   the global normalisation (:188-238) with the scale of
//       the vertex's mesh.
//
// The colouring of the union restricted to a mesh is a valid colouring of
// that mesh, and a colour class's vertices move independently, so the
// sweeps over the union give every mesh exactly what a plain pass fed that
// restricted colouring gives.  A pass is kernel launches only (no host
// synchronisation, allocation or memset node), so trn_batch_pass_graph
// captures it once in a hipGraph and replays it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "trn_api.hpp"
#include "trn_recovery.hpp"

namespace trn {

constexpr int kBatchCplxB = 256;      // the plain complexity kernel's block
constexpr int kBatchCplxMax = 1024;   // and its largest grid

// Partial complexity of block b of mesh m's plain-path grid (G_m blocks):
// thread g adds triangles toff[m] + b*256 + g + k*G_m*256 in order, then the
// block's fixed tree.  Tile k of the batch is (m, b).
__global__ __launch_bounds__(kBatchCplxB) void k_batch_cplx_part(
    int32_t ntiles, const int32_t* __restrict__ tile_mesh, const int32_t* __restrict__ tile0,
    const int32_t* __restrict__ toff, const int32_t* __restrict__ tri,
    const double2* __restrict__ xy, const double* __restrict__ tensor,
    double* __restrict__ partial) {
  __shared__ double lds[kBatchCplxB / 64];
  const int32_t k = (int32_t)blockIdx.x;
  if (k >= ntiles) return;
  const int32_t m = tile_mesh[k];
  const int32_t b = k - tile0[m];
  const int64_t g = tile0[m + 1] - tile0[m];
  const int64_t t0 = toff[m], ntm = (int64_t)toff[m + 1] - t0;
  double phi = 0.;
#pragma unroll 4
  for (int64_t t = (int64_t)b * kBatchCplxB + threadIdx.x; t < ntm; t += g * kBatchCplxB)
    phi += cplx_one(t0 + t, tri, xy, tensor);
  const double s = block_sum(phi, lds);
  if (threadIdx.x == 0) partial[k] = s;
}

// One block per mesh: the partials in the plain final kernel's order, then
// the mesh's target and scale.
__global__ __launch_bounds__(kBatchCplxB) void k_batch_cplx_final(
    int32_t nmesh, const int32_t* __restrict__ tile0, const int32_t* __restrict__ voff,
    const double* __restrict__ partial, double target_factor, double* __restrict__ cplx,
    double* __restrict__ fact) {
  __shared__ double lds[kBatchCplxB / 64];
  const int32_t m = (int32_t)blockIdx.x;
  if (m >= nmesh) return;
  const int32_t p0 = tile0[m], n = tile0[m + 1] - p0;
  double v = 0.;
  for (int k = threadIdx.x; k < n; k += blockDim.x) v += partial[p0 + k];
  const double s = block_sum(v, lds);
  if (threadIdx.x == 0) {
    const double c = 0. + s;                                        // metric.cpp:182-183
    const int32_t nvm = voff[m + 1] - voff[m];
    const double target = (double)(int)floor(nvm * target_factor);  // metric.cpp:39
    cplx[m] = c;
    fact[m] = target / c;                                           // metric.cpp:197-198
  }
}

__global__ __launch_bounds__(256) void k_batch_global(int32_t nv,
                                                      const int32_t* __restrict__ mesh_of,
                                                      const double* __restrict__ fact,
                                                      double* __restrict__ tensor, double emin,
                                                      double emax) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  global_one(tensor + 3 * i, fact[mesh_of[i]], emin, emax);
}

__global__ __launch_bounds__(256) void k_copy_f64(int64_t n, const double* __restrict__ src,
                                                  double* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_copy_cplx(int32_t n, const double* __restrict__ src,
                                                   double* __restrict__ dst) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) dst[i] = src[i];
}

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace trn

using namespace trn;

struct trn_batch {
  trn_ctx* ctx = nullptr;
  int32_t nmesh = 0, nv = 0, nt = 0;
  trn_mesh* mesh = nullptr;  // the union topology
  std::vector<int32_t> voff, toff;
  int32_t* voff_d = nullptr;       // [nmesh + 1]
  int32_t* toff_d = nullptr;       // [nmesh + 1]
  int32_t* mesh_of = nullptr;      // [nv] mesh of each vertex
  int32_t ntiles = 0;
  int32_t* tile_mesh = nullptr;    // [ntiles]
  int32_t* tile0 = nullptr;        // [nmesh + 1] first tile of each mesh
  double* partial = nullptr;       // [ntiles]
  double* cplx = nullptr;          // [nmesh]
  double* fact = nullptr;          // [nmesh]
  double4* logm = nullptr;         // [nv] scratch of the smoothing preparation
  void* work = nullptr;            // recovery scratch, then the sweeps' scratch
  size_t work_size = 0;
  // colouring of the union
  int32_t* color = nullptr;
  int32_t* order = nullptr;
  int32_t ncolors = 0;
  int32_t max_class = 1;
  std::vector<int32_t> class_off;
  // captured pass
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> graph_key;
};

namespace trn {

template <typename T>
static void dalloc(T** p, size_t n) {
  TRN_HIP(hipMalloc((void**)p, sizeof(T) * (n > 0 ? n : 1)));
}

template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

static void drop_graph(trn_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_key.clear();
}

// Grows the shared scratch to what recovery and `sweeps` sweeps need.
static void ensure_work(trn_batch* b, int sweeps) {
  const size_t need = std::max(recover_scratch(b->mesh),
                               smooth_sweep_scratch(b->mesh, b->ncolors, sweeps, b->max_class));
  if (need <= b->work_size) return;
  TRN_HIP(hipStreamSynchronize(b->ctx->stream));
  dfree(b->work);
  drop_graph(b);
  TRN_HIP(hipMalloc(&b->work, need));
  b->work_size = need;
}

static void check_pass_args(const trn_batch* b, const double* xy, const double* solut,
                            const double* tensor, const double* qualit, double h_min,
                            double h_max, int sweeps) {
  TRN_REQUIRE(b, TRN_ERR_ARG, "null batch");
  TRN_REQUIRE(xy && solut && tensor && qualit, TRN_ERR_ARG, "null argument");
  TRN_REQUIRE(h_min > 0 && h_max > 0, TRN_ERR_ARG, "h_min and h_max must be positive");
  TRN_REQUIRE(sweeps >= 0 && sweeps <= 1024, TRN_ERR_ARG, "sweeps out of range");
  TRN_REQUIRE(sweeps == 0 || b->ncolors > 0, TRN_ERR_STATE,
              "batch not coloured (trn_batch_color)");
}

// Queues one pass on ctx->stream.  Kernel launches only.
static int32_t* enqueue_pass(trn_ctx* ctx, trn_batch* b, const double* xy_in, double* xy,
                             const double* solut, double tf, int p, double h_min, double h_max,
                             double* tensor, double* qualit, int sweeps, double* complexity_d,
                             bool want_totals) {
  const trn_mesh* m = b->mesh;
  if (xy_in && xy_in != xy)
    launch(ctx, "batch_copy_xy", 32.0 * b->nv, k_copy_f64, dim3(nblk(2 * (int64_t)b->nv, 256)),
           dim3(256), 0, 2 * (int64_t)b->nv, xy_in, xy);
  recover_launch(ctx, m, xy, solut, nullptr, nullptr, tensor, p, b->work, b->work_size);
  launch(ctx, "batch_cplx_part", 64.0 * b->nv, k_batch_cplx_part, dim3(b->ntiles),
         dim3(kBatchCplxB), 0, b->ntiles, (const int32_t*)b->tile_mesh, (const int32_t*)b->tile0,
         (const int32_t*)b->toff_d, (const int32_t*)m->tri, (const double2*)xy,
         (const double*)tensor, b->partial);
  launch(ctx, "batch_cplx_final", 8.0 * b->ntiles, k_batch_cplx_final, dim3(b->nmesh),
         dim3(kBatchCplxB), 0, b->nmesh, (const int32_t*)b->tile0, (const int32_t*)b->voff_d,
         (const double*)b->partial, tf, b->cplx, b->fact);
  launch(ctx, "batch_global", 52.0 * b->nv, k_batch_global, dim3(nblk(b->nv, 256)), dim3(256),
         0, b->nv, (const int32_t*)b->mesh_of, (const double*)b->fact, tensor,
         1 / (h_min * h_min), 1 / (h_max * h_max));
  if (complexity_d)
    launch(ctx, "batch_copy_cplx", 0.0, k_copy_cplx, dim3(nblk(b->nmesh, 256)), dim3(256), 0,
           b->nmesh, (const double*)b->cplx, complexity_d);
  if (sweeps == 0) return nullptr;
  smooth_prepare_launch(ctx, m, xy, tensor, qualit, true, b->logm);
  return smooth_sweeps_launch(ctx, m, xy, tensor, qualit, b->order, b->class_off.data(),
                              b->ncolors, sweeps, b->max_class, b->work, b->work_size,
                              want_totals);
}

}  // namespace trn

extern "C" {

static int batch_create(const char* name, trn_ctx* ctx, int32_t nmesh, const int32_t* nv,
                        const int32_t* nt, const int32_t* tri, const uint8_t* tags,
                        const double* xy, bool want_xy, trn_batch** out) {
  return guard_impl(name, [&] {
    TRN_REQUIRE(!want_xy || xy != nullptr, TRN_ERR_ARG, "null coordinates");
    DeviceScope ds(ctx);
    TRN_REQUIRE(out && nv && nt && tri && tags && nmesh > 0, TRN_ERR_ARG, "bad arguments");
    *out = nullptr;
    trn_batch* b = new trn_batch();
    b->ctx = ctx;
    b->nmesh = nmesh;
    try {
      b->voff.assign(nmesh + 1, 0);
      b->toff.assign(nmesh + 1, 0);
      for (int m = 0; m < nmesh; ++m) {
        TRN_REQUIRE(nv[m] > 0 && nt[m] > 0, TRN_ERR_ARG, "empty mesh in batch");
        TRN_REQUIRE((int64_t)b->voff[m] + nv[m] < INT32_MAX, TRN_ERR_ARG, "batch too large");
        b->voff[m + 1] = b->voff[m] + nv[m];
        TRN_REQUIRE((int64_t)b->toff[m] + nt[m] < INT32_MAX / 3, TRN_ERR_ARG, "batch too large");
        b->toff[m + 1] = b->toff[m] + nt[m];
      }
      b->nv = b->voff[nmesh];
      b->nt = b->toff[nmesh];
      // union triangle list with global indices; local indices are checked
      // against their own mesh (an index into another mesh would join them)
      std::vector<int32_t> gtri(3 * (size_t)b->nt);
      for (int m = 0; m < nmesh; ++m) {
        for (int64_t t = b->toff[m]; t < b->toff[m + 1]; ++t) {
          const int32_t* s = tri + 3 * t;
          if (s[0] < 0) {
            gtri[3 * t] = gtri[3 * t + 1] = gtri[3 * t + 2] = -1;
            continue;
          }
          for (int j = 0; j < 3; ++j) {
            TRN_REQUIRE(s[j] >= 0 && s[j] < nv[m], TRN_ERR_ARG,
                        "triangle index outside its mesh (mesh " + std::to_string(m) + ")");
            gtri[3 * t + j] = s[j] + b->voff[m];
          }
        }
      }
      // complexity tiles: mesh m gets the plain path's grid for its size
      std::vector<int32_t> tile0(nmesh + 1, 0), tile_mesh;
      std::vector<int32_t> mesh_of(b->nv);
      for (int m = 0; m < nmesh; ++m) {
        const int64_t ntm = nt[m];
        int64_t g = (ntm + kBatchCplxB - 1) / kBatchCplxB;
        if (g > kBatchCplxMax) g = kBatchCplxMax;
        tile0[m + 1] = tile0[m] + (int32_t)g;
        tile_mesh.insert(tile_mesh.end(), (size_t)g, m);
        std::fill(mesh_of.begin() + b->voff[m], mesh_of.begin() + b->voff[m + 1], m);
      }
      b->ntiles = tile0[nmesh];
      dalloc(&b->voff_d, nmesh + 1);
      dalloc(&b->toff_d, nmesh + 1);
      dalloc(&b->mesh_of, b->nv);
      copy_async(ctx, b->voff_d, b->voff.data(), sizeof(int32_t) * (nmesh + 1));
      copy_async(ctx, b->mesh_of, mesh_of.data(), sizeof(int32_t) * b->nv);
      if (xy) {
        // patch slabs along each mesh's own Z-curve
        double* xy_d = nullptr;
        int32_t* perm = nullptr;
        dalloc(&xy_d, 2 * (int64_t)b->nv);
        dalloc(&perm, b->nv);
        try {
          copy_async(ctx, xy_d, xy, sizeof(double) * 2 * (size_t)b->nv);
          spatial_order_segments(ctx, b->nv, xy_d, nmesh, (const int32_t*)b->mesh_of,
                                 (const int32_t*)b->voff_d, perm);
          b->mesh = mesh_create_ordered(ctx, b->nv, b->nt, gtri.data(), tags, perm);
        } catch (...) {
          (void)hipFree(xy_d);
          (void)hipFree(perm);
          throw;
        }
        sync(ctx);
        TRN_HIP(hipFree(xy_d));
        TRN_HIP(hipFree(perm));
      } else {
        const int rc = trn_mesh_create(ctx, b->nv, b->nt, gtri.data(), tags, nullptr, &b->mesh);
        if (rc != TRN_OK) fail(rc, trn_last_error());
      }
      dalloc(&b->tile_mesh, b->ntiles);
      dalloc(&b->tile0, nmesh + 1);
      dalloc(&b->partial, b->ntiles);
      dalloc(&b->cplx, nmesh);
      dalloc(&b->fact, nmesh);
      dalloc(&b->logm, b->nv);
      dalloc(&b->color, b->nv);
      dalloc(&b->order, b->nv);
      copy_async(ctx, b->toff_d, b->toff.data(), sizeof(int32_t) * (nmesh + 1));
      copy_async(ctx, b->tile_mesh, tile_mesh.data(), sizeof(int32_t) * b->ntiles);
      copy_async(ctx, b->tile0, tile0.data(), sizeof(int32_t) * (nmesh + 1));
      sync(ctx);  // the host vectors go out of scope
      ensure_work(b, 0);
    } catch (...) {
      trn_batch_destroy(b);
      throw;
    }
    *out = b;
  });
}

int trn_batch_create(trn_ctx* ctx, int32_t nmesh, const int32_t* nv, const int32_t* nt,
                     const int32_t* tri, const uint8_t* tags, trn_batch** out) {
  return batch_create("trn_batch_create", ctx, nmesh, nv, nt, tri, tags, nullptr, false, out);
}

int trn_batch_create_xy(trn_ctx* ctx, int32_t nmesh, const int32_t* nv, const int32_t* nt,
                        const int32_t* tri, const uint8_t* tags, const double* xy,
                        trn_batch** out) {
  return batch_create("trn_batch_create_xy", ctx, nmesh, nv, nt, tri, tags, xy, true, out);
}

int trn_batch_destroy(trn_batch* b) {
  if (!b) return TRN_OK;
  return guard_impl("trn_batch_destroy", [&] {
    DeviceScope ds(b->ctx);
    (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    if (b->mesh) trn_mesh_destroy(b->mesh);
    dfree(b->voff_d); dfree(b->toff_d); dfree(b->mesh_of); dfree(b->tile_mesh);
    dfree(b->tile0); dfree(b->partial); dfree(b->cplx); dfree(b->fact); dfree(b->logm);
    dfree(b->color); dfree(b->order);
    if (b->work) (void)hipFree(b->work);
    delete b;
  });
}

int trn_batch_info(const trn_batch* b, int32_t* nv_total, int32_t* nt_total,
                   const trn_mesh** mesh) {
  return guard_impl("trn_batch_info", [&] {
    TRN_REQUIRE(b, TRN_ERR_ARG, "null batch");
    if (nv_total) *nv_total = b->nv;
    if (nt_total) *nt_total = b->nt;
    if (mesh) *mesh = b->mesh;
  });
}

int trn_batch_color(trn_ctx* ctx, trn_batch* b, uint64_t seed, int32_t max_colors,
                    int32_t* ncolors) {
  return guard_impl("trn_batch_color", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(b && ncolors, TRN_ERR_ARG, "null argument");
    TRN_REQUIRE(max_colors >= 1 && max_colors <= 64, TRN_ERR_ARG, "max_colors out of range");
    std::vector<int32_t> off(max_colors + 2, 0);
    int32_t nc = 0;
    const int rc = trn_color(ctx, b->mesh, seed, max_colors, b->color, b->order, off.data(), &nc);
    if (rc != TRN_OK) fail(rc, trn_last_error());
    b->ncolors = nc;
    b->class_off.assign(off.begin(), off.begin() + nc + 1);
    b->max_class = 1;
    for (int c = 0; c < nc; ++c) b->max_class = std::max(b->max_class, off[c + 1] - off[c]);
    drop_graph(b);
    *ncolors = nc;
  });
}

int trn_batch_coloring(const trn_batch* b, const int32_t** color_d, const int32_t** order_d,
                       int32_t* class_off, int32_t* ncolors) {
  return guard_impl("trn_batch_coloring", [&] {
    TRN_REQUIRE(b, TRN_ERR_ARG, "null batch");
    if (color_d) *color_d = b->color;
    if (order_d) *order_d = b->order;
    if (class_off)
      std::copy(b->class_off.begin(), b->class_off.end(), class_off);
    if (ncolors) *ncolors = b->ncolors;
  });
}

int trn_batch_pass(trn_ctx* ctx, trn_batch* b, const double* xy_in_d, double* xy_d,
                   const double* solut_d, double target_factor, int32_t lp_norm, double h_min,
                   double h_max, double* tensor_d, double* qualit_d, int32_t sweeps,
                   double* complexity_d, int32_t* commits, int32_t* tasks) {
  return guard_impl("trn_batch_pass", [&] {
    DeviceScope ds(ctx);
    check_pass_args(b, xy_d, solut_d, tensor_d, qualit_d, h_min, h_max, sweeps);
    ensure_work(b, sweeps);
    const bool want = commits || tasks;
    const int32_t* totals = enqueue_pass(ctx, b, xy_in_d, xy_d, solut_d, target_factor, lp_norm,
                                         h_min, h_max, tensor_d, qualit_d, sweeps, complexity_d,
                                         want);
    if (totals) {
      copy_async(ctx, ctx->pinned_i32, totals, sizeof(int32_t) * 2 * (size_t)sweeps);
      sync(ctx);
      for (int s = 0; s < sweeps; ++s) {
        if (commits) commits[s] = ctx->pinned_i32[2 * s];
        if (tasks) tasks[s] = ctx->pinned_i32[2 * s + 1];
      }
    }
  });
}

int trn_batch_pass_graph(trn_ctx* ctx, trn_batch* b, const double* xy_in_d, double* xy_d,
                         const double* solut_d, double target_factor, int32_t lp_norm,
                         double h_min, double h_max, double* tensor_d, double* qualit_d,
                         int32_t sweeps) {
  return guard_impl("trn_batch_pass_graph", [&] {
    DeviceScope ds(ctx);
    check_pass_args(b, xy_d, solut_d, tensor_d, qualit_d, h_min, h_max, sweeps);
    TRN_REQUIRE(ctx->stream != nullptr, TRN_ERR_STATE,
                "a graph is captured on a created stream, not the null stream");
    TRN_REQUIRE(!ctx->timing, TRN_ERR_STATE, "per-kernel timing cannot be captured");
    ensure_work(b, sweeps);
    uint64_t dbits[3];
    std::memcpy(&dbits[0], &target_factor, 8);
    std::memcpy(&dbits[1], &h_min, 8);
    std::memcpy(&dbits[2], &h_max, 8);
    const std::vector<uint64_t> key = {
        (uint64_t)(uintptr_t)xy_in_d, (uint64_t)(uintptr_t)xy_d,  (uint64_t)(uintptr_t)solut_d,
        (uint64_t)(uintptr_t)tensor_d, (uint64_t)(uintptr_t)qualit_d,
        (uint64_t)(uintptr_t)ctx->stream, dbits[0], dbits[1], dbits[2], (uint64_t)lp_norm,
        (uint64_t)sweeps, (uint64_t)(uintptr_t)b->work};
    if (!b->exec || b->graph_key != key) {
      drop_graph(b);
      TRN_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      try {
        enqueue_pass(ctx, b, xy_in_d, xy_d, solut_d, target_factor, lp_norm, h_min, h_max,
                     tensor_d, qualit_d, sweeps, nullptr, false);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      TRN_HIP(hipStreamEndCapture(ctx->stream, &b->graph));
      TRN_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_key = key;
    }
    TRN_HIP(hipGraphLaunch(b->exec, ctx->stream));
  });
}

}  // extern "C"
