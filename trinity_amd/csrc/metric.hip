// metric.hip -- L2-projection gradient/Hessian recovery and metric
// construction on gfx950, replacing Metrics::run (src/metric.cpp:59-70).
//
// Layout and kernels (V vertices, T triangles, all FP64):
//   k_tri_grad    per triangle: area A_t, A_t*g_t (g_t = element gradient)
//                 (src/hessian.cpp:42-67)
//   k_vtx_grad    per vertex: ascending 2-ring patch built in a per-thread
//                 LDS set, sum A_t and A_t*g_t in patch order, divide
//                 (src/hessian.cpp:75-82; patch = Mesh::getVicinity(i,2))
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
#include <cmath>

#include "trn_api.hpp"
#include "trn_numeric.hpp"

namespace trn {

constexpr int kGatherB = 64;  // one wave per block for the LDS-set gathers
constexpr int kCplxBlocks = 1024;
constexpr int kCplxB = 256;

// Builds the ascending 2-ring element patch of vertex i in the thread's LDS
// set (element k at s[k*kGatherB]).  Returns the size; ok=false on overflow.
__device__ __forceinline__ int build_patch(const int i, const int32_t* __restrict__ v2t_off,
                                           const int32_t* __restrict__ v2t,
                                           const int32_t* __restrict__ v2v_off,
                                           const int32_t* __restrict__ v2v, int* s,
                                           const int cap, bool& ok) {
  int m = 0;
  ok = true;
  const int32_t b0 = v2t_off[i], e0 = v2t_off[i + 1];
  for (int32_t k = b0; k < e0; ++k) ok = ok && lds_sorted_insert(s, kGatherB, m, cap, v2t[k]);
  const int32_t nb0 = v2v_off[i], ne0 = v2v_off[i + 1];
  for (int32_t a = nb0; a < ne0; ++a) {
    const int32_t j = v2v[a];
    const int32_t bj = v2t_off[j], ej = v2t_off[j + 1];
    for (int32_t k = bj; k < ej; ++k) ok = ok && lds_sorted_insert(s, kGatherB, m, cap, v2t[k]);
  }
  return m;
}

__global__ __launch_bounds__(256) void k_tri_grad(int32_t nt, const int32_t* __restrict__ tri,
                                                  const double2* __restrict__ xy,
                                                  const double* __restrict__ f,
                                                  double* __restrict__ area_t,
                                                  double2* __restrict__ ag_t) {
  const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int32_t n0 = tri[3 * t], n1 = tri[3 * t + 1], n2 = tri[3 * t + 2];
  if (n0 < 0) return;  // deleted triangles never appear in a patch
  double a, d[6];
  tri_geometry(xy[n0], xy[n1], xy[n2], a, d);
  area_t[t] = a;
  if (f == nullptr) return;  // area only
  const double f0 = f[n0], f1 = f[n1], f2 = f[n2];
  double g0 = 0., g1 = 0.;  // hessian.cpp:36, accumulated j = 0,1,2 (:65-67)
  g0 += f0 * d[0];
  g1 += f0 * d[1];
  g0 += f1 * d[2];
  g1 += f1 * d[3];
  g0 += f2 * d[4];
  g1 += f2 * d[5];
  ag_t[t] = make_double2(a * g0, a * g1);  // the products of hessian.cpp:78
}

template <bool WITH_GRAD>
__global__ __launch_bounds__(kGatherB) void k_vtx_grad(
    int32_t nv, const int32_t* __restrict__ v2t_off, const int32_t* __restrict__ v2t,
    const int32_t* __restrict__ v2v_off, const int32_t* __restrict__ v2v,
    const double* __restrict__ area_t, const double2* __restrict__ ag_t,
    double2* __restrict__ grad, double* __restrict__ tarea, int cap, int32_t* err) {
  extern __shared__ int lds[];
  const int i = (int)(xcd_block() * kGatherB + threadIdx.x);
  if (i >= nv) return;
  int* s = lds + threadIdx.x;
  bool ok;
  const int m = build_patch(i, v2t_off, v2t, v2v_off, v2v, s, cap, ok);
  if (!ok) {
    atomicOr(err, 1);
    return;
  }
  double ta = 0., gx = 0., gy = 0.;
  for (int e = 0; e < m; ++e) {
    const int t = s[e * kGatherB];
    ta += area_t[t];       // hessian.cpp:53
    if (WITH_GRAD) {
      const double2 r = ag_t[t];
      gx += r.x;           // hessian.cpp:76-78
      gy += r.y;
    }
  }
  if (WITH_GRAD) grad[i] = make_double2(gx / ta, gy / ta);  // hessian.cpp:81-82
  tarea[i] = ta;
}

__global__ __launch_bounds__(256) void k_tri_hess(int32_t nt, const int32_t* __restrict__ tri,
                                                  const double2* __restrict__ xy,
                                                  const double2* __restrict__ grad,
                                                  double4* __restrict__ ah_t) {
  const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int32_t n0 = tri[3 * t], n1 = tri[3 * t + 1], n2 = tri[3 * t + 2];
  if (n0 < 0) return;
  double a, d[6];
  tri_geometry(xy[n0], xy[n1], xy[n2], a, d);
  const double2 g0 = grad[n0], g1 = grad[n1], g2 = grad[n2];
  double l0 = 0., l1 = 0., l2 = 0., l3 = 0.;  // hessian.cpp:99, :131-136
  l0 += g0.x * d[0]; l1 += g0.x * d[1]; l2 += g0.y * d[0]; l3 += g0.y * d[1];
  l0 += g1.x * d[2]; l1 += g1.x * d[3]; l2 += g1.y * d[2]; l3 += g1.y * d[3];
  l0 += g2.x * d[4]; l1 += g2.x * d[5]; l2 += g2.y * d[4]; l3 += g2.y * d[5];
  double4 r;
  r.x = a * l0;  // the products of hessian.cpp:143
  r.y = a * l1;
  r.z = a * l2;
  r.w = a * l3;
  ah_t[t] = r;
}

// Metrics::normalizeLocally for one tensor (src/metric.cpp:113-148).
__device__ __forceinline__ void normalize_local(double& t0, double& t1, double& t2,
                                                const double ex) {
  double val[2], vec[4], h[4];
  eigen(t0, t1, t1, t2, val, vec);
  val[0] = dmax(fabs(val[0]), kEps);
  val[1] = dmax(fabs(val[1]), kEps);
  const double det = val[0] * val[1];
  const double scale_loc = pow(det, ex);
  val[0] *= scale_loc;
  val[1] *= scale_loc;
  rebuild(val, vec, h);
  t0 = h[0];
  t1 = h[1];
  t2 = h[3];
}

template <bool NORMALIZE>
__global__ __launch_bounds__(kGatherB) void k_vtx_hess(
    int32_t nv, const int32_t* __restrict__ v2t_off, const int32_t* __restrict__ v2t,
    const int32_t* __restrict__ v2v_off, const int32_t* __restrict__ v2v,
    const double4* __restrict__ ah_t, const double* __restrict__ tarea,
    double* __restrict__ hess, double* __restrict__ out, double ex, int cap, int32_t* err) {
  extern __shared__ int lds[];
  const int i = (int)(xcd_block() * kGatherB + threadIdx.x);
  if (i >= nv) return;
  int* s = lds + threadIdx.x;
  bool ok;
  const int m = build_patch(i, v2t_off, v2t, v2v_off, v2v, s, cap, ok);
  if (!ok) {
    atomicOr(err, 1);
    return;
  }
  double h0 = 0., h1 = 0., h2 = 0., h3 = 0.;  // hessian.cpp:93
  for (int e = 0; e < m; ++e) {
    const double4 r = ah_t[s[e * kGatherB]];
    h0 += r.x;
    h1 += r.y;
    h2 += r.z;
    h3 += r.w;
  }
  const double ta = tarea[i];  // same sum as the gradient's (hessian.cpp:120)
  h0 /= ta;
  h1 /= ta;
  h2 /= ta;
  h3 /= ta;
  double t0 = h0, t1 = 0.5 * (h1 + h2), t2 = h3;  // hessian.cpp:150-152
  if (hess) {
    hess[3 * (int64_t)i] = t0;
    hess[3 * (int64_t)i + 1] = t1;
    hess[3 * (int64_t)i + 2] = t2;
  }
  if (NORMALIZE) normalize_local(t0, t1, t2, ex);
  out[3 * (int64_t)i] = t0;
  out[3 * (int64_t)i + 1] = t1;
  out[3 * (int64_t)i + 2] = t2;
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
// partials that one block adds in a fixed tree.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = 0.;
  if (threadIdx.x == 0) {
    r = lds[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r += lds[w];
  }
  return r;
}

__global__ __launch_bounds__(kCplxB) void k_tri_cplx(int32_t nt, const int32_t* __restrict__ tri,
                                                     const double2* __restrict__ xy,
                                                     const double* __restrict__ tensor,
                                                     double* __restrict__ partial) {
  __shared__ double lds[kCplxB / 64];
  double phi = 0.;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t v0 = tri[3 * t], v1 = tri[3 * t + 1], v2 = tri[3 * t + 2];
    if (v0 < 0) continue;
    const double2 p0 = xy[v0], p1 = xy[v1], p2 = xy[v2];
    const double s0 = p1.x - p0.x;  // metric.cpp:163-167
    const double s1 = p2.x - p0.x;
    const double s2 = p1.y - p0.y;
    const double s3 = p2.y - p0.y;
    const double area = 0.5 * (s0 * s3 - s1 * s2);
    double rho = 0.;  // metric.cpp:171-179
    const int32_t vs[3] = {v0, v1, v2};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double* m = tensor + 3 * (int64_t)vs[j];
      const double det = m[0] * m[2] - m[1] * m[1];
      rho += __builtin_sqrt(det);
    }
    rho /= 3;
    phi += rho * area;
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
  double val[2], vec[4], h[4];
  const double t0 = tensor[3 * i], t1 = tensor[3 * i + 1], t2 = tensor[3 * i + 2];
  eigen(t0, t1, t1, t2, val, vec);
  val[0] *= fact;
  val[1] *= fact;
#pragma unroll
  for (int j = 0; j < 2; ++j) {  // metric.cpp:213-216
    val[j] = dmin(val[j], emin);
    val[j] = dmax(val[j], emax);
  }
  rebuild(val, vec, h);
  tensor[3 * i] = h[0];
  tensor[3 * i + 1] = h[1];
  tensor[3 * i + 2] = h[3];
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
  double* area_t;
  double2* ag_t;
  double4* ah_t;
  double* tarea;
  double* partial;
  double* cplx;
  int32_t* err;
};

static size_t grad_scratch_bytes(const trn_mesh* m) {
  return scratch_bytes(sizeof(double) * m->nt) + scratch_bytes(sizeof(double2) * m->nt) +
         scratch_bytes(sizeof(double4) * m->nt) + scratch_bytes(sizeof(double) * m->nv) +
         scratch_bytes(sizeof(double) * kCplxBlocks) + scratch_bytes(sizeof(double) * 4) +
         scratch_bytes(sizeof(int32_t) * 4);
}

static GradBuffers carve_grad(Scratch& sc, const trn_mesh* m) {
  GradBuffers b;
  b.area_t = sc.carve<double>(m->nt);
  b.ag_t = sc.carve<double2>(m->nt);
  b.ah_t = sc.carve<double4>(m->nt);
  b.tarea = sc.carve<double>(m->nv);
  b.partial = sc.carve<double>(kCplxBlocks);
  b.cplx = sc.carve<double>(4);
  b.err = sc.carve<int32_t>(4);
  return b;
}

static size_t gather_shmem(const trn_mesh* m) { return sizeof(int) * kGatherB * (size_t)m->patch_cap; }

static void run_gradient(trn_ctx* ctx, const trn_mesh* m, const double* xy, const double* f,
                         double* grad, GradBuffers& b) {
  launch(ctx, "tri_grad", 0.0, k_tri_grad, dim3(nblk(m->nt, 256)), dim3(256), 0, m->nt,
         (const int32_t*)m->tri, (const double2*)xy, f, b.area_t, b.ag_t);
  launch(ctx, "vtx_grad", kBytesGradPerV * m->nv, k_vtx_grad<true>, dim3(nblk(m->nv, kGatherB)),
         dim3(kGatherB), gather_shmem(m), m->nv, (const int32_t*)m->v2t_off,
         (const int32_t*)m->v2t, (const int32_t*)m->v2v_off, (const int32_t*)m->v2v,
         (const double*)b.area_t, (const double2*)b.ag_t, (double2*)grad, b.tarea, m->patch_cap,
         b.err);
}

static void run_hessian(trn_ctx* ctx, const trn_mesh* m, const double* xy, const double* grad,
                        double* hess, double* out, bool normalize, double ex, GradBuffers& b) {
  launch(ctx, "tri_hess", 0.0, k_tri_hess, dim3(nblk(m->nt, 256)), dim3(256), 0, m->nt,
         (const int32_t*)m->tri, (const double2*)xy, (const double2*)grad, b.ah_t);
  if (normalize)
    launch(ctx, "vtx_hess_norm", kBytesHessPerV * m->nv, k_vtx_hess<true>,
           dim3(nblk(m->nv, kGatherB)), dim3(kGatherB), gather_shmem(m), m->nv,
           (const int32_t*)m->v2t_off, (const int32_t*)m->v2t, (const int32_t*)m->v2v_off,
           (const int32_t*)m->v2v, (const double4*)b.ah_t, (const double*)b.tarea, hess, out, ex,
           m->patch_cap, b.err);
  else
    launch(ctx, "vtx_hess", kBytesHessPerV * m->nv, k_vtx_hess<false>,
           dim3(nblk(m->nv, kGatherB)), dim3(kGatherB), gather_shmem(m), m->nv,
           (const int32_t*)m->v2t_off, (const int32_t*)m->v2t, (const int32_t*)m->v2v_off,
           (const int32_t*)m->v2v, (const double4*)b.ah_t, (const double*)b.tarea, hess, out, ex,
           m->patch_cap, b.err);
}

static void run_complexity(trn_ctx* ctx, const trn_mesh* m, const double* xy,
                           const double* tensor, GradBuffers& b) {
  int blocks = (int)nblk(m->nt, kCplxB);
  if (blocks > kCplxBlocks) blocks = kCplxBlocks;
  launch(ctx, "tri_cplx", kBytesCplxPerV * m->nv, k_tri_cplx, dim3(blocks), dim3(kCplxB), 0,
         m->nt, (const int32_t*)m->tri, (const double2*)xy, tensor, b.partial);
  launch(ctx, "cplx_final", 8.0 * blocks, k_cplx_final, dim3(1), dim3(kCplxB), 0,
         (const double*)b.partial, blocks, b.cplx);
}

static void check_err(trn_ctx* ctx, GradBuffers& b) {
  copy_async(ctx, ctx->pinned_i32, b.err, sizeof(int32_t));
  sync(ctx);
  TRN_REQUIRE(ctx->pinned_i32[0] == 0, TRN_ERR_STATE, "2-ring patch larger than measured");
}

static double metric_exponent(int32_t p) { return p ? -1. / (2 * p + 2) : 0.5; }  // metric.cpp:44

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
    memset_async(ctx, b.err, 0, sizeof(int32_t));
    run_gradient(ctx, m, xy_d, solut_d, grad, b);
    run_hessian(ctx, m, xy_d, grad, hessian_d, tensor_d, true, metric_exponent(lp_norm), b);
    run_complexity(ctx, m, xy_d, tensor_d, b);
    const double target = (double)(int)std::floor(m->nv * target_factor);  // metric.cpp:39
    launch(ctx, "vtx_global", kBytesGlobPerV * m->nv, k_vtx_global, dim3(nblk(m->nv, 256)),
           dim3(256), 0, m->nv, tensor_d, (const double*)b.cplx, target, 0.0,
           1 / (h_min * h_min), 1 / (h_max * h_max));
    if (complexity) {
      copy_async(ctx, ctx->pinned_f64, b.cplx, sizeof(double));
      check_err(ctx, b);
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
    memset_async(ctx, b.err, 0, sizeof(int32_t));
    run_gradient(ctx, mesh, xy_d, solut_d, gradient_d, b);
    check_err(ctx, b);
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
    memset_async(ctx, b.err, 0, sizeof(int32_t));
    // total stencil areas: the same ordered sums as the gradient pass
    launch(ctx, "tri_area", 0.0, k_tri_grad, dim3(nblk(mesh->nt, 256)), dim3(256), 0, mesh->nt,
           (const int32_t*)mesh->tri, (const double2*)xy_d, (const double*)nullptr, b.area_t,
           b.ag_t);
    launch(ctx, "vtx_area", 0.0, k_vtx_grad<false>, dim3(nblk(mesh->nv, kGatherB)),
           dim3(kGatherB), gather_shmem(mesh), mesh->nv, (const int32_t*)mesh->v2t_off,
           (const int32_t*)mesh->v2t, (const int32_t*)mesh->v2v_off, (const int32_t*)mesh->v2v,
           (const double*)b.area_t, (const double2*)b.ag_t, (double2*)nullptr, b.tarea,
           mesh->patch_cap, b.err);
    run_hessian(ctx, mesh, xy_d, gradient_d, nullptr, tensor_d, false, 0.0, b);
    check_err(ctx, b);
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
