// metric.hip -- L2-projection gradient/Hessian recovery and metric
// construction on gfx950, replacing Metrics::run (src/metric.cpp:59-70).
//
// Layout and kernels (V vertices, T triangles, all FP64):
//   k_tri_grad    per triangle: area A_t, A_t*g_t (g_t = element gradient)
//                 (src/hessian.cpp:42-67)
//   k_vtx_grad    per vertex: sum A_t and A_t*g_t over the ascending 2-ring
//                 patch, divide (src/hessian.cpp:75-82; patch =
//                 Mesh::getVicinity(i,2)); the patches are stored with the
//                 mesh in sliced ELL form, one wave per slab, so the patch
//                 ids are read with coalesced loads
//   k_tri_hess    per triangle: A_t*delta_t (src/hessian.cpp:108-143)
//   k_vtx_hess    per vertex: patch-ordered sums / total area, symmetrise
//                 (src/hessian.cpp:141-152), optionally fused with the local
//                 L^p normalisation (src/metric.cpp:100-150)
//   k_tri_cplx    per triangle: A_t * mean sqrt(det M), deterministic
//                 two-level tree sum (src/metric.cpp:153-185)
//   k_vtx_global  per vertex: eigen, scale by target/complexity, clamp,
//                 rebuild (src/metric.cpp:188-238)
// The per-triangle terms do not depend on the vertex whose patch contains
// the triangle, so computing them once per triangle and summing the stored
// products in patch order reproduces the reference's roundings exactly.
// The element bodies live in trn_recovery.hpp (shared with batch.hip).
#include <algorithm>
#include <cmath>

#include "trn_recovery.hpp"

namespace trn {

constexpr int kGatherB = 64;  // one wave per block for the LDS-set gathers
constexpr int kCplxBlocks = 1024;
constexpr int kCplxB = 256;

__global__ __launch_bounds__(256) void k_tri_grad(int32_t nt, const int32_t* __restrict__ tri,
                                                  const double2* __restrict__ xy,
                                                  const double* __restrict__ f,
                                                  double* __restrict__ area_t,
                                                  double2* __restrict__ ag_t) {
  const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  tri_grad_one(t, tri, xy, f, area_t, ag_t);
}

static_assert(kGatherB == kPatchSlab, "a gather wave reads one slab of stored patches");

// The gathers run one wave per slab, 1, 2 or 4 slabs per block (the mesh's
// tuned shape, as many as fit kGatherLdsMax bytes of LDS); each wave stages
// its slab's union in its own LDS region and waits only at the block's one
// barrier.  One slab per block is the default (4 per block: 1.19 -> 1.21 ms
// for the Hessian gather of a 4096^2 grid, 0.86 -> 0.95 ms for the
// gradient's).
constexpr int kGatherWaves = 4;  // the most slabs per block
constexpr size_t kGatherLdsMax = 64 * 1024;

// lds[k] = src[U[k]] for the nu union entries of a slab, by its 64 lanes in
// batches of kStageBatch independent loads per lane.
constexpr int kStageBatch = 8;

template <class T>
__device__ __forceinline__ void stage_union(const int lane, const int32_t* __restrict__ U,
                                            const int nu, const T* __restrict__ src, T* lds) {
  for (int k0 = lane; k0 < nu; k0 += kStageBatch * kGatherB) {
    int32_t t[kStageBatch];
#pragma unroll
    for (int c = 0; c < kStageBatch; ++c) {
      const int k = k0 + c * kGatherB;
      t[c] = k < nu ? U[k] : 0;
    }
    T v[kStageBatch];
#pragma unroll
    for (int c = 0; c < kStageBatch; ++c) v[c] = src[t[c]];  // t = 0 past the end: in bounds
#pragma unroll
    for (int c = 0; c < kStageBatch; ++c)
      if (k0 + c * kGatherB < nu) lds[k0 + c * kGatherB] = v[c];
  }
}

// Fused staging: the union's per-triangle terms computed in place from the
// corners of its triangles (C: three vertex ids per union entry, stored with
// the mesh) with the same arithmetic as k_tri_grad / k_tri_hess, so the
// sums are bit-identical; kFusedBatch* entries per lane in flight.  Lanes
// past the end read vertex 0 and store nothing.  (4096^2 jittered grid on
// its Z-curve slabs, ms per pass with both batches at 3: 9.54, at 2: 9.37,
// at 4: 9.49; gradient / Hessian batches 1 / 2: 9.32, 2 / 1: 9.37, 2 / 3:
// 9.42; gradient gather alone 0.77 ms at 3, 0.69 at 2, 0.64 at 1.)
constexpr int kFusedBatchGrad = 1;
constexpr int kFusedBatchHess = 2;

__device__ __forceinline__ void stage_grad_terms(const int lane, const int32_t* __restrict__ C,
                                                 const int nu, const double2* __restrict__ xy,
                                                 const double* __restrict__ f, double* lds_a,
                                                 double2* lds_g) {
  for (int k0 = lane; k0 < nu; k0 += kFusedBatchGrad * kGatherB) {
    int32_t n[kFusedBatchGrad][3];
#pragma unroll
    for (int c = 0; c < kFusedBatchGrad; ++c) {
      const int k = k0 + c * kGatherB;
      const bool in = k < nu;
#pragma unroll
      for (int j = 0; j < 3; ++j) n[c][j] = in ? C[3 * k + j] : 0;
    }
#pragma unroll
    for (int c = 0; c < kFusedBatchGrad; ++c) {
      double a;
      double2 ag;
      grad_terms(xy[n[c][0]], xy[n[c][1]], xy[n[c][2]], f[n[c][0]], f[n[c][1]], f[n[c][2]], a,
                 ag);
      const int k = k0 + c * kGatherB;
      if (k < nu) {
        lds_a[k] = a;
        lds_g[k] = ag;
      }
    }
  }
}

__device__ __forceinline__ void stage_hess_terms(const int lane, const int32_t* __restrict__ C,
                                                 const int nu, const double2* __restrict__ xy,
                                                 const double2* __restrict__ grad,
                                                 double4* lds_h) {
  for (int k0 = lane; k0 < nu; k0 += kFusedBatchHess * kGatherB) {
    int32_t n[kFusedBatchHess][3];
#pragma unroll
    for (int c = 0; c < kFusedBatchHess; ++c) {
      const int k = k0 + c * kGatherB;
      const bool in = k < nu;
#pragma unroll
      for (int j = 0; j < 3; ++j) n[c][j] = in ? C[3 * k + j] : 0;
    }
#pragma unroll
    for (int c = 0; c < kFusedBatchHess; ++c) {
      const double4 r = hess_terms(xy[n[c][0]], xy[n[c][1]], xy[n[c][2]], grad[n[c][0]],
                                   grad[n[c][1]], grad[n[c][2]]);
      const int k = k0 + c * kGatherB;
      if (k < nu) lds_h[k] = r;
    }
  }
}

// The lane's column of its slab's local patch indices, rows < `rows` (the
// slab's row count, wave-uniform, <= kRegRows); entries past the lane's
// own patch are padding and never used.
__device__ __forceinline__ void load_column(const uint16_t* __restrict__ col, const int rows,
                                            int (&loc)[kRegRows]) {
#pragma unroll
  for (int e = 0; e < kRegRows; ++e) loc[e] = e < rows ? (int)col[e * kPatchSlab] : 0;
}

// Slab, wave and lane of the calling thread.
struct GatherSlot {
  uint32_t b;
  int w, lane;
};

__device__ __forceinline__ GatherSlot gather_slot() {
  GatherSlot g;
  g.w = (int)(threadIdx.x >> 6);
  g.lane = (int)(threadIdx.x & 63);
  g.b = xcd_block() * (blockDim.x >> 6) + (uint32_t)g.w;
  return g;
}

// Gathers over the stored patches.  A slab with a union (union_n[b] >= 0)
// first stages the per-triangle terms of its union in LDS with coalesced
// loads (the union is ascending, so its triangles come in runs) while its
// lanes load their local patch indices, then each vertex sums its patch
// from LDS in patch order; a slab without one gathers from global memory.
// wave_bytes: LDS bytes per wave, ag terms (16 B) then areas (8 B) of up to
// max_union entries.
// FUSED: union slabs compute their terms from tri, xy and f instead of
// reading area_t / ag_t (which only slabs without a union then need).
template <bool WITH_GRAD, bool FUSED>
__global__ __launch_bounds__(kGatherB * kGatherWaves) void k_vtx_grad(
    int32_t nv, int32_t nslab, const int32_t* __restrict__ slab_v,
    const int32_t* __restrict__ patch_n, const int32_t* __restrict__ patch_row,
    const int32_t* __restrict__ patch_ell, const int32_t* __restrict__ union_n,
    const int32_t* __restrict__ union_off, const int32_t* __restrict__ union_ids,
    const uint16_t* __restrict__ patch_local, int max_union, int wave_bytes,
    const int32_t* __restrict__ union_cu, const int32_t* __restrict__ union_tri,
    const double2* __restrict__ xy, const double* __restrict__ f,
    const double* __restrict__ area_t,
    const double2* __restrict__ ag_t, double2* __restrict__ grad, double* __restrict__ tarea) {
  extern __shared__ double2 lds_all[];
  const GatherSlot g = gather_slot();
  const bool live = g.b < (uint32_t)nslab;
  double2* lds_g = (double2*)((char*)lds_all + (size_t)g.w * wave_bytes);
  double* lds_a = (double*)(lds_g + max_union);
  const int pos = (int)(g.b * kGatherB + g.lane);
  const int nu = live ? union_n[g.b] : -1;
  int32_t r0 = 0, rows = 0;
  int loc[kRegRows];
  if (nu >= 0) {
    const int32_t* U = union_ids + union_off[g.b];
    r0 = patch_row[g.b];
    rows = patch_row[g.b + 1] - r0;
    if (rows <= kRegRows) load_column(patch_local + (int64_t)r0 * kPatchSlab + g.lane, rows, loc);
    if (FUSED) {
      stage_grad_terms(g.lane, union_tri + 3 * (int64_t)union_cu[g.b], nu, xy, f, lds_a, lds_g);
    } else {
      stage_union(g.lane, U, nu, area_t, lds_a);
      if (WITH_GRAD) stage_union(g.lane, U, nu, ag_t, lds_g);
    }
  }
  __syncthreads();
  if (!live || pos >= nv) return;
  const int i = slab_v ? slab_v[pos] : pos;
  if (nu >= 0) {
    if (rows <= kRegRows)
      vtx_grad_reg<WITH_GRAD>(i, loc, patch_n[i], (const double*)lds_a, (const double2*)lds_g,
                              grad, tarea);
    else
      vtx_grad_one<WITH_GRAD>(i, patch_local + (int64_t)r0 * kPatchSlab + g.lane, kPatchSlab,
                              patch_n[i], (const double*)lds_a, (const double2*)lds_g, grad,
                              tarea);
    return;
  }
  vtx_grad_one<WITH_GRAD>(i, patch_ell + (int64_t)patch_row[g.b] * kPatchSlab + g.lane,
                          kPatchSlab, patch_n[i], area_t, ag_t, grad, tarea);
}

__global__ __launch_bounds__(256) void k_tri_hess(int32_t nt, const int32_t* __restrict__ tri,
                                                  const double2* __restrict__ xy,
                                                  const double2* __restrict__ grad,
                                                  double4* __restrict__ ah_t) {
  const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  tri_hess_one(t, tri, xy, grad, ah_t);
}

// wave_bytes: LDS bytes per wave, the ah terms (32 B) of up to max_union
// entries.  Union slabs compute their terms from tri, xy and grad; ah_t is
// read by slabs without a union only.
template <bool NORMALIZE>
__global__ __launch_bounds__(kGatherB * kGatherWaves) void k_vtx_hess(
    int32_t nv, int32_t nslab, const int32_t* __restrict__ slab_v,
    const int32_t* __restrict__ patch_n, const int32_t* __restrict__ patch_row,
    const int32_t* __restrict__ patch_ell, const int32_t* __restrict__ union_n,
    const int32_t* __restrict__ union_off, const int32_t* __restrict__ union_ids,
    const uint16_t* __restrict__ patch_local, int wave_bytes,
    const int32_t* __restrict__ union_cu, const int32_t* __restrict__ union_tri,
    const double2* __restrict__ xy, const double2* __restrict__ grad,
    const double4* __restrict__ ah_t, const double* __restrict__ tarea,
    double* __restrict__ hess, double* __restrict__ out, double* __restrict__ sdet, double ex) {
  extern __shared__ double4 lds_all4[];
  const GatherSlot g = gather_slot();
  const bool live = g.b < (uint32_t)nslab;
  double4* lds_h = (double4*)((char*)lds_all4 + (size_t)g.w * wave_bytes);
  const int pos = (int)(g.b * kGatherB + g.lane);
  const int nu = live ? union_n[g.b] : -1;
  int32_t r0 = 0, rows = 0;
  int loc[kRegRows];
  if (nu >= 0) {
    r0 = patch_row[g.b];
    rows = patch_row[g.b + 1] - r0;
    if (rows <= kRegRows) load_column(patch_local + (int64_t)r0 * kPatchSlab + g.lane, rows, loc);
    stage_hess_terms(g.lane, union_tri + 3 * (int64_t)union_cu[g.b], nu, xy, grad, lds_h);
  }
  __syncthreads();
  if (!live || pos >= nv) return;
  const int i = slab_v ? slab_v[pos] : pos;
  if (nu >= 0) {
    if (rows <= kRegRows)
      vtx_hess_reg<NORMALIZE>(i, loc, patch_n[i], (const double4*)lds_h, tarea, hess, out, sdet,
                              ex);
    else
      vtx_hess_one<NORMALIZE>(i, patch_local + (int64_t)r0 * kPatchSlab + g.lane, kPatchSlab,
                              patch_n[i], (const double4*)lds_h, tarea, hess, out, sdet, ex);
    return;
  }
  vtx_hess_one<NORMALIZE>(i, patch_ell + (int64_t)patch_row[g.b] * kPatchSlab + g.lane,
                          kPatchSlab, patch_n[i], ah_t, tarea, hess, out, sdet, ex);
}

__global__ __launch_bounds__(256) void k_normalize_local(int32_t nv, double* __restrict__ tensor,
                                                         double ex) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  double t0 = tensor[3 * i], t1 = tensor[3 * i + 1], t2 = tensor[3 * i + 2];
  normalize_local(t0, t1, t2, ex);
  tensor[3 * i] = t0;
  tensor[3 * i + 1] = t1;
  tensor[3 * i + 2] = t2;
}

// Fixed-shape deterministic sum: thread g of the fixed grid adds triangles
// g, g+G, ... in order; waves reduce with a fixed butterfly; blocks write
// partials that one block adds in a fixed tree.  sdet: the vertices' terms
// from the normalising Hessian gather (8 bytes per corner read instead of
// the 24 of its tensor; the same values), else the tensor.
__global__ __launch_bounds__(kCplxB) void k_tri_cplx(int32_t nt, const int32_t* __restrict__ tri,
                                                     const double2* __restrict__ xy,
                                                     const double* __restrict__ tensor,
                                                     const double* __restrict__ sdet,
                                                     double* __restrict__ partial) {
  __shared__ double lds[kCplxB / 64];
  double phi = 0.;
  if (sdet) {
#pragma unroll 4
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt;
         t += (int64_t)gridDim.x * blockDim.x)
      phi += cplx_one_sdet(t, tri, xy, sdet);
  } else {
#pragma unroll 4
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt;
         t += (int64_t)gridDim.x * blockDim.x)
      phi += cplx_one(t, tri, xy, tensor);
  }
  const double b = block_sum(phi, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

__global__ __launch_bounds__(kCplxB) void k_cplx_final(const double* __restrict__ partial, int n,
                                                       double* __restrict__ out) {
  __shared__ double lds[kCplxB / 64];
  double v = 0.;
  for (int k = threadIdx.x; k < n; k += blockDim.x) v += partial[k];
  const double b = block_sum(v, lds);
  if (threadIdx.x == 0) out[0] = 0. + b;  // metric.cpp:182-183
}

// Metrics::normalizeGlobally for every vertex (src/metric.cpp:188-238);
// fact = target / complexity is formed on the device (:197-198).
__global__ __launch_bounds__(256) void k_vtx_global(int32_t nv, double* __restrict__ tensor,
                                                    const double* __restrict__ cplx,
                                                    double target, double fact_given,
                                                    double emin, double emax) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const double fact = cplx ? target / cplx[0] : fact_given;
  global_one(tensor + 3 * i, fact, emin, emax);
}

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

static void check_mesh(const trn_mesh* mesh) {
  TRN_REQUIRE(mesh != nullptr, TRN_ERR_ARG, "null mesh");
}

// Algorithmic HBM bytes per unit (each array read once / written once per
// stage; int32 indices, FP64 values; CSR lists 4 B offset + 4 B per entry).
constexpr double kBytesGradPerV = 120.0;   // xy16 f8 tri24 v2t28 v2v28 | g16
constexpr double kBytesHessPerV = 136.0;   // xy16 g16 tri24 v2t28 v2v28 | M24
constexpr double kBytesCplxPerV = 64.0;    // M24 xy16 tri24
constexpr double kBytesGlobPerV = 48.0;    // M24 | M24

struct GradBuffers {
  double* area_t = nullptr;
  double2* ag_t = nullptr;
  double4* ah_t = nullptr;
  double* tarea = nullptr;
  double* partial = nullptr;
  double* cplx = nullptr;
  double* sdet = nullptr;   // per-vertex complexity terms, when carved
  bool sdet_ready = false;  // written by the last normalising Hessian gather
};

static size_t grad_scratch_bytes(const trn_mesh* m) {
  return scratch_bytes(sizeof(double) * m->nt) + scratch_bytes(sizeof(double2) * m->nt) +
         scratch_bytes(sizeof(double4) * m->nt) + scratch_bytes(sizeof(double) * m->nv) +
         scratch_bytes(sizeof(double) * kCplxBlocks) + scratch_bytes(sizeof(double) * 4) +
         scratch_bytes(sizeof(double) * m->nv);
}

static GradBuffers carve_grad(Scratch& sc, const trn_mesh* m) {
  GradBuffers b;
  b.area_t = sc.carve<double>(m->nt);
  b.ag_t = sc.carve<double2>(m->nt);
  b.ah_t = sc.carve<double4>(m->nt);
  b.tarea = sc.carve<double>(m->nv);
  b.partial = sc.carve<double>(kCplxBlocks);
  b.cplx = sc.carve<double>(4);
  b.sdet = sc.carve<double>(m->nv);
  return b;
}


// Launch shape of a gather: waves (slabs) per block, LDS bytes per wave and
// per block.  per_entry: LDS bytes per union entry of one wave.
struct GatherShape {
  int nslab, waves, wave_bytes;
  size_t shmem;
  dim3 grid, block;
};

static GatherShape gather_shape(const trn_mesh* m, size_t per_entry, int waves) {
  GatherShape g;
  g.nslab = (int)nblk(m->nv, kPatchSlab);
  g.wave_bytes = (int)(((per_entry * (size_t)m->max_union) + 15) & ~(size_t)15);
  g.waves = std::min(std::max(waves, 1), kGatherWaves);
  while (g.waves > 1 && (size_t)g.waves * g.wave_bytes > kGatherLdsMax) g.waves >>= 1;
  g.shmem = (size_t)g.waves * g.wave_bytes;
  g.grid = dim3(nblk(g.nslab, g.waves));
  g.block = dim3(kGatherB * g.waves);
  return g;
}

// The gradient gathers' LDS holds the union's ag (16 B) and area (8 B)
// terms.  fused: union slabs compute them from tri, xy and f.
static void launch_vtx_grad(trn_ctx* ctx, const char* name, double bytes, bool with_grad,
                            bool fused, const trn_mesh* m, const double* xy, const double* f,
                            const double* area_t, const double2* ag_t, double2* grad,
                            double* tarea, int waves) {
  const GatherShape g = gather_shape(m, sizeof(double2) + sizeof(double), waves);
  auto kern = fused ? k_vtx_grad<true, true>
                    : (with_grad ? k_vtx_grad<true, false> : k_vtx_grad<false, false>);
  launch(ctx, name, bytes, kern, g.grid, g.block, g.shmem, m->nv, g.nslab, m->slab_v,
         (const int32_t*)m->patch_n, (const int32_t*)m->patch_row,
         (const int32_t*)m->patch_ell, (const int32_t*)m->union_n,
         (const int32_t*)m->union_off, (const int32_t*)m->union_ids,
         (const uint16_t*)m->patch_local, m->max_union, g.wave_bytes,
         (const int32_t*)m->union_cu, (const int32_t*)m->union_tri, (const double2*)xy, f,
         area_t, ag_t, grad, tarea);
}

// The per-triangle kernels run only for meshes with slabs too large for a
// union (the union slabs compute their terms while staging).
static void run_gradient(trn_ctx* ctx, const trn_mesh* m, const double* xy, const double* f,
                         double* grad, GradBuffers& b) {
  if (!m->union_all)
    launch(ctx, "tri_grad", 0.0, k_tri_grad, dim3(nblk(m->nt, 256)), dim3(256), 0, m->nt,
           (const int32_t*)m->tri, (const double2*)xy, f, b.area_t, b.ag_t);
  launch_vtx_grad(ctx, "vtx_grad", kBytesGradPerV * m->nv, true, true, m, xy, f, b.area_t,
                  b.ag_t, (double2*)grad, b.tarea, m->tune.grad_waves);
}

static void launch_vtx_hess(trn_ctx* ctx, const trn_mesh* m, const double* xy,
                            const double* grad, const double4* ah_t, const double* tarea,
                            double* hess, double* out, double* sdet, bool normalize, double ex,
                            int waves) {
  const GatherShape g = gather_shape(m, sizeof(double4), waves);
  auto kern = normalize ? k_vtx_hess<true> : k_vtx_hess<false>;
  launch(ctx, normalize ? "vtx_hess_norm" : "vtx_hess", kBytesHessPerV * m->nv, kern, g.grid,
         g.block, g.shmem, m->nv, g.nslab, m->slab_v, (const int32_t*)m->patch_n,
         (const int32_t*)m->patch_row, (const int32_t*)m->patch_ell,
         (const int32_t*)m->union_n, (const int32_t*)m->union_off,
         (const int32_t*)m->union_ids, (const uint16_t*)m->patch_local, g.wave_bytes,
         (const int32_t*)m->union_cu, (const int32_t*)m->union_tri, (const double2*)xy,
         (const double2*)grad, ah_t, tarea, hess, out, normalize ? sdet : nullptr, ex);
}

static void run_hessian(trn_ctx* ctx, const trn_mesh* m, const double* xy, const double* grad,
                        double* hess, double* out, bool normalize, double ex, GradBuffers& b) {
  if (!m->union_all)
    launch(ctx, "tri_hess", 0.0, k_tri_hess, dim3(nblk(m->nt, 256)), dim3(256), 0, m->nt,
           (const int32_t*)m->tri, (const double2*)xy, (const double2*)grad, b.ah_t);
  launch_vtx_hess(ctx, m, xy, grad, b.ah_t, b.tarea, hess, out, b.sdet, normalize, ex,
                  m->tune.hess_waves);
  b.sdet_ready = normalize && b.sdet;
}

// ---- shapes the autotuner measures ----------------------------------------

size_t tune_gather_scratch(const trn_mesh* m) {
  return scratch_bytes(sizeof(double) * m->nt) + scratch_bytes(sizeof(double2) * m->nt) +
         scratch_bytes(sizeof(double4) * m->nt) + scratch_bytes(sizeof(double) * m->nv) +
         scratch_bytes(sizeof(double2) * m->nv) + scratch_bytes(sizeof(double) * 3 * m->nv);
}

// One gradient and one Hessian gather (with the local normalisation) at the
// given shapes, bracketed by ev[0..1] and ev[2..3]; the per-triangle
// kernels that meshes with union-less slabs need run outside the brackets.
void tune_gathers(trn_ctx* ctx, const trn_mesh* m, const double* xy, const double* f,
                  int lp_norm, void* scratch, size_t bytes, int grad_waves, int hess_waves,
                  hipEvent_t ev[4]) {
  Carve c(scratch, bytes);
  GradBuffers b;
  b.area_t = c.carve<double>(m->nt);
  b.ag_t = c.carve<double2>(m->nt);
  b.ah_t = c.carve<double4>(m->nt);
  b.tarea = c.carve<double>(m->nv);
  double2* grad = c.carve<double2>(m->nv);
  double* tensor = c.carve<double>(3 * (size_t)m->nv);
  if (!m->union_all)
    launch(ctx, "tri_grad", 0.0, k_tri_grad, dim3(nblk(m->nt, 256)), dim3(256), 0, m->nt,
           (const int32_t*)m->tri, (const double2*)xy, f, b.area_t, b.ag_t);
  TRN_HIP(hipEventRecord(ev[0], ctx->stream));
  launch_vtx_grad(ctx, "vtx_grad", 0.0, true, true, m, xy, f, b.area_t, b.ag_t, grad, b.tarea,
                  grad_waves);
  TRN_HIP(hipEventRecord(ev[1], ctx->stream));
  if (!m->union_all)
    launch(ctx, "tri_hess", 0.0, k_tri_hess, dim3(nblk(m->nt, 256)), dim3(256), 0, m->nt,
           (const int32_t*)m->tri, (const double2*)xy, (const double2*)grad, b.ah_t);
  TRN_HIP(hipEventRecord(ev[2], ctx->stream));
  launch_vtx_hess(ctx, m, xy, (const double*)grad, b.ah_t, b.tarea, nullptr, tensor, nullptr,
                  true, metric_exponent(lp_norm), hess_waves);
  TRN_HIP(hipEventRecord(ev[3], ctx->stream));
}

static void run_complexity(trn_ctx* ctx, const trn_mesh* m, const double* xy,
                           const double* tensor, GradBuffers& b) {
  int blocks = (int)nblk(m->nt, kCplxB);
  if (blocks > kCplxBlocks) blocks = kCplxBlocks;
  launch(ctx, "tri_cplx", kBytesCplxPerV * m->nv, k_tri_cplx, dim3(blocks), dim3(kCplxB), 0,
         m->nt, (const int32_t*)m->tri, (const double2*)xy, tensor,
         b.sdet_ready ? (const double*)b.sdet : nullptr, b.partial);
  launch(ctx, "cplx_final", 8.0 * blocks, k_cplx_final, dim3(1), dim3(kCplxB), 0,
         (const double*)b.partial, blocks, b.cplx);
}

// ---- internal launch sequence (also used by the batched pass) ------------

size_t recover_scratch(const trn_mesh* m) {
  return scratch_bytes(sizeof(double) * m->nt) + scratch_bytes(sizeof(double2) * m->nt) +
         scratch_bytes(sizeof(double4) * m->nt) + scratch_bytes(sizeof(double) * m->nv) +
         scratch_bytes(sizeof(double2) * m->nv);
}

// Gradient and Hessian recovery with the local normalisation fused
// (Metrics::computeGradient, computeHessian, normalizeLocally) into tensor.
void recover_launch(trn_ctx* ctx, const trn_mesh* m, const double* xy, const double* f,
                    double* grad_out, double* hess_out, double* tensor, int lp_norm,
                    void* scratch, size_t bytes) {
  Carve c(scratch, bytes);
  GradBuffers b;
  b.area_t = c.carve<double>(m->nt);
  b.ag_t = c.carve<double2>(m->nt);
  b.ah_t = c.carve<double4>(m->nt);
  b.tarea = c.carve<double>(m->nv);
  b.partial = nullptr;
  b.cplx = nullptr;
  double* grad = grad_out ? grad_out : (double*)c.carve<double2>(m->nv);
  run_gradient(ctx, m, xy, f, grad, b);
  run_hessian(ctx, m, xy, grad, hess_out, tensor, true, metric_exponent(lp_norm), b);
}

double metric_tape(trn_ctx* ctx, const trn_mesh* m, const double* xy, const double* f,
                   double target_factor, int lp_norm, double h_min, double h_max, double* tensor,
                   double* grad, double* hess, double* mloc, double* tarea) {
  check_mesh(m);
  Scratch sc(ctx, grad_scratch_bytes(m));
  GradBuffers b = carve_grad(sc, m);
  b.tarea = tarea;
  run_gradient(ctx, m, xy, f, grad, b);
  run_hessian(ctx, m, xy, grad, hess, mloc, true, metric_exponent(lp_norm), b);
  copy_async(ctx, tensor, mloc, sizeof(double) * 3 * (size_t)m->nv);
  run_complexity(ctx, m, xy, tensor, b);
  const double target = (double)(int)std::floor(m->nv * target_factor);  // metric.cpp:39
  launch(ctx, "vtx_global", kBytesGlobPerV * m->nv, k_vtx_global, dim3(nblk(m->nv, 256)),
         dim3(256), 0, m->nv, tensor, (const double*)b.cplx, target, 0.0, 1 / (h_min * h_min),
         1 / (h_max * h_max));
  copy_async(ctx, ctx->pinned_f64, b.cplx, sizeof(double));
  sync(ctx);
  return ctx->pinned_f64[0];
}

}  // namespace trn

using namespace trn;

extern "C" {

int trn_metric_run(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                   const double* solut_d, double target_factor, int32_t lp_norm, double h_min,
                   double h_max, double* tensor_d, double* complexity, double* gradient_d,
                   double* hessian_d) {
  return guard_impl("trn_metric_run", [&] {
    DeviceScope ds(ctx);
    check_mesh(mesh);
    TRN_REQUIRE(xy_d && solut_d && tensor_d, TRN_ERR_ARG, "null array");
    TRN_REQUIRE(h_min > 0 && h_max > 0, TRN_ERR_ARG, "h_min and h_max must be positive");
    const trn_mesh* m = mesh;
    Scratch sc(ctx, grad_scratch_bytes(m) + scratch_bytes(sizeof(double2) * m->nv));
    GradBuffers b = carve_grad(sc, m);
    double* grad = gradient_d ? gradient_d : (double*)sc.carve<double2>(m->nv);
    run_gradient(ctx, m, xy_d, solut_d, grad, b);
    run_hessian(ctx, m, xy_d, grad, hessian_d, tensor_d, true, metric_exponent(lp_norm), b);
    run_complexity(ctx, m, xy_d, tensor_d, b);
    const double target = (double)(int)std::floor(m->nv * target_factor);  // metric.cpp:39
    launch(ctx, "vtx_global", kBytesGlobPerV * m->nv, k_vtx_global, dim3(nblk(m->nv, 256)),
           dim3(256), 0, m->nv, tensor_d, (const double*)b.cplx, target, 0.0,
           1 / (h_min * h_min), 1 / (h_max * h_max));
    if (complexity) {
      copy_async(ctx, ctx->pinned_f64, b.cplx, sizeof(double));
      sync(ctx);
      *complexity = ctx->pinned_f64[0];
    }
  });
}

int trn_gradient(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d, const double* solut_d,
                 double* gradient_d) {
  return guard_impl("trn_gradient", [&] {
    DeviceScope ds(ctx);
    check_mesh(mesh);
    TRN_REQUIRE(xy_d && solut_d && gradient_d, TRN_ERR_ARG, "null array");
    Scratch sc(ctx, grad_scratch_bytes(mesh));
    GradBuffers b = carve_grad(sc, mesh);
    run_gradient(ctx, mesh, xy_d, solut_d, gradient_d, b);
  });
}

int trn_hessian(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d, const double* gradient_d,
                double* tensor_d) {
  return guard_impl("trn_hessian", [&] {
    DeviceScope ds(ctx);
    check_mesh(mesh);
    TRN_REQUIRE(xy_d && gradient_d && tensor_d, TRN_ERR_ARG, "null array");
    Scratch sc(ctx, grad_scratch_bytes(mesh));
    GradBuffers b = carve_grad(sc, mesh);
    // total stencil areas: the same ordered sums as the gradient pass
    launch(ctx, "tri_area", 0.0, k_tri_grad, dim3(nblk(mesh->nt, 256)), dim3(256), 0, mesh->nt,
           (const int32_t*)mesh->tri, (const double2*)xy_d, (const double*)nullptr, b.area_t,
           b.ag_t);
    launch_vtx_grad(ctx, "vtx_area", 0.0, false, false, mesh, xy_d, nullptr, b.area_t, b.ag_t,
                    nullptr, b.tarea, mesh->tune.grad_waves);
    run_hessian(ctx, mesh, xy_d, gradient_d, nullptr, tensor_d, false, 0.0, b);
  });
}

int trn_normalize_local(trn_ctx* ctx, int32_t nv, double* tensor_d, int32_t lp_norm) {
  return guard_impl("trn_normalize_local", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(nv >= 0 && tensor_d, TRN_ERR_ARG, "bad arguments");
    launch(ctx, "normalize_local", 48.0 * nv, k_normalize_local, dim3(nblk(nv, 256)), dim3(256),
           0, nv, tensor_d, metric_exponent(lp_norm));
  });
}

int trn_complexity(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d, const double* tensor_d,
                   double* complexity) {
  return guard_impl("trn_complexity", [&] {
    DeviceScope ds(ctx);
    check_mesh(mesh);
    TRN_REQUIRE(xy_d && tensor_d && complexity, TRN_ERR_ARG, "null argument");
    Scratch sc(ctx, grad_scratch_bytes(mesh));
    GradBuffers b = carve_grad(sc, mesh);
    run_complexity(ctx, mesh, xy_d, tensor_d, b);
    copy_async(ctx, ctx->pinned_f64, b.cplx, sizeof(double));
    sync(ctx);
    *complexity = ctx->pinned_f64[0];
  });
}

int trn_normalize_global(trn_ctx* ctx, int32_t nv, double* tensor_d, double scale_fact,
                         double h_min, double h_max) {
  return guard_impl("trn_normalize_global", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(nv >= 0 && tensor_d, TRN_ERR_ARG, "bad arguments");
    TRN_REQUIRE(h_min > 0 && h_max > 0, TRN_ERR_ARG, "h_min and h_max must be positive");
    launch(ctx, "vtx_global", kBytesGlobPerV * nv, k_vtx_global, dim3(nblk(nv, 256)), dim3(256),
           0, nv, tensor_d, (const double*)nullptr, 0.0, scale_fact, 1 / (h_min * h_min),
           1 / (h_max * h_max));
  });
}

}  // extern "C"
