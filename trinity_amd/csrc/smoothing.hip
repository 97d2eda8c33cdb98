// smoothing.hip -- coloured metric-aware Laplacian smoothing on gfx950,
// replacing Smooth::run after its colouring (src/smoothing.cpp:274-302).
//
//   k_quality_cache  per triangle: Mesh::computeQuality (Smooth::cacheQuality,
//                    src/smoothing.cpp:145-155)
//   k_smooth_phase   one launch per colour class and sweep: every thread
//                    applies Smooth::moveSmartLaplacian (:40-126) to one
//                    vertex.  Vertices of one colour share no triangle and
//                    read only their own 1-ring, so a class is an
//                    independent set and the launch order inside it does not
//                    change the result; colours run in ascending order
//                    (Smooth::movePoints, :166-175), as in the reference.
#include <cfloat>

#include "trn_api.hpp"
#include "trn_numeric.hpp"

namespace trn {

constexpr int kSmoothB = 64;
constexpr int kCounterSlots = 64;

__global__ __launch_bounds__(256) void k_quality_cache(int32_t nt, const int32_t* __restrict__ tri,
                                                       const double2* __restrict__ xy,
                                                       const double* __restrict__ tensor,
                                                       double* __restrict__ qualit) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int32_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
  if (a < 0) return;  // Mesh::isActiveElem (src/mesh.cpp:375)
  const double2 pa = xy[a], pb = xy[b], pc = xy[c];
  const double A[2] = {pa.x, pa.y}, B[2] = {pb.x, pb.y}, C[2] = {pc.x, pc.y};
  qualit[t] = quality(A, B, C, tensor + 3 * (int64_t)a, tensor + 3 * (int64_t)b,
                      tensor + 3 * (int64_t)c);
}

// Smooth::moveSmartLaplacian (src/smoothing.cpp:40-126) for vertex i.
// q holds the thread's per-incident-triangle qualities (stride kSmoothB).
__device__ int move_vertex(const int i, double2* __restrict__ xy, double* __restrict__ tensor,
                           double* __restrict__ qualit, const int32_t* __restrict__ tri,
                           const int32_t* __restrict__ v2t_off, const int32_t* __restrict__ v2t,
                           const int32_t* __restrict__ v2v_off, const int32_t* __restrict__ v2v,
                           double* q) {
  const int32_t t0 = v2t_off[i], deg = v2t_off[i + 1] - t0;
  const int32_t n0 = v2v_off[i], nb = v2v_off[i + 1] - n0;

  double q_min = DBL_MAX;  // smoothing.cpp:48-51
  for (int k = 0; k < deg; ++k) q_min = dmin(q_min, qualit[v2t[t0 + k]]);

  const double2 pini = xy[i];
  const double P[2] = {pini.x, pini.y};
  const double T[3] = {tensor[3 * (int64_t)i], tensor[3 * (int64_t)i + 1],
                       tensor[3 * (int64_t)i + 2]};
  double opt0 = 0, opt1 = 0;
  double lm[4] = {0., 0., 0., 0.};
  for (int k = 0; k < nb; ++k) {  // smoothing.cpp:61-79 and numeric.cpp:113-172
    const int32_t j = v2v[n0 + k];
    const double2 pj = xy[j];
    const double Pb[2] = {pj.x, pj.y};
    const double* tb = tensor + 3 * (int64_t)j;
    const double Tb[3] = {tb[0], tb[1], tb[2]};
    const double len = riemann_dist(P, Pb, T, Tb);
    opt0 += P[0] + ((Pb[0] - P[0]) / len);
    opt1 += P[1] + ((Pb[1] - P[1]) / len);
    double c[4];
    tensor_log(Tb[0], Tb[1], Tb[2], c);
    lm[0] += c[0];
    lm[1] += c[1];
    lm[2] += c[2];
    lm[3] += c[3];
  }
  opt0 /= nb;
  opt1 /= nb;
  double TN[3];
  tensor_exp_mean(lm, nb, TN);  // smoothing.cpp:84

  int valid = 0;
  double w = 1;
  double px = P[0], py = P[1];
  for (int retry = 0; retry < 5; ++retry) {  // smoothing.cpp:96-114
    px = (1 - w) * P[0] + w * opt0;
    py = (1 - w) * P[1] + w * opt1;
    valid = 1;
    for (int k = 0; valid && k < deg; ++k) {
      const int32_t t = v2t[t0 + k];
      const int32_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
      double pa[2], pb[2], pc[2], ma[3], mb[3], mc[3];
      const int32_t vv[3] = {a, b, c};
      double* pp[3] = {pa, pb, pc};
      double* mm[3] = {ma, mb, mc};
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        if (vv[r] == i) {
          pp[r][0] = px; pp[r][1] = py;
          mm[r][0] = TN[0]; mm[r][1] = TN[1]; mm[r][2] = TN[2];
        } else {
          const double2 p = xy[vv[r]];
          pp[r][0] = p.x; pp[r][1] = p.y;
          const double* m = tensor + 3 * (int64_t)vv[r];
          mm[r][0] = m[0]; mm[r][1] = m[1]; mm[r][2] = m[2];
        }
      }
      valid = 0;
      if (ccw(pa, pb, pc)) {
        const double qk = quality(pa, pb, pc, ma, mb, mc);
        q[k * kSmoothB] = qk;
        valid = qk > q_min;
      }
    }
    w *= 0.5;
    if (valid) break;
  }
  if (valid) {  // smoothing.cpp:116-123
    for (int k = 0; k < deg; ++k) qualit[v2t[t0 + k]] = q[k * kSmoothB];
    xy[i] = make_double2(px, py);
    tensor[3 * (int64_t)i] = TN[0];
    tensor[3 * (int64_t)i + 1] = TN[1];
    tensor[3 * (int64_t)i + 2] = TN[2];
  }
  return valid;
}

__global__ __launch_bounds__(kSmoothB) void k_smooth_phase(
    int32_t count, const int32_t* __restrict__ order, double2* __restrict__ xy,
    double* __restrict__ tensor, double* __restrict__ qualit, const int32_t* __restrict__ tri,
    const uint8_t* __restrict__ tags, const int32_t* __restrict__ v2t_off,
    const int32_t* __restrict__ v2t, const int32_t* __restrict__ v2v_off,
    const int32_t* __restrict__ v2v, int32_t* __restrict__ counters) {
  extern __shared__ double qs[];
  const int64_t idx = (int64_t)blockIdx.x * kSmoothB + threadIdx.x;
  int commit = 0, task = 0;
  if (idx < count) {
    const int32_t i = order[idx];
    // Smooth::preProcess (:134-138): boundary vertices are never moved
    if (!(tags[i] & 4) && v2t_off[i + 1] > v2t_off[i]) {
      task = 1;
      commit = move_vertex(i, xy, tensor, qualit, tri, v2t_off, v2t, v2v_off, v2v,
                           qs + threadIdx.x);
    }
  }
  // counters are spread over kCounterSlots addresses: one hot address
  // serialises the atomics of every wave in L2
  const uint64_t mc = __ballot(commit), mt = __ballot(task);
  if ((threadIdx.x & 63) == 0) {
    int32_t* slot = counters + 2 * (blockIdx.x & (kCounterSlots - 1));
    if (mc) atomicAdd(&slot[0], (int32_t)__popcll(mc));
    if (mt) atomicAdd(&slot[1], (int32_t)__popcll(mt));
  }
}

// counters[s][slot][2] -> out[s][2]
__global__ void k_counter_reduce(int32_t sweeps, const int32_t* __restrict__ counters,
                                 int32_t* __restrict__ out) {
  const int s = threadIdx.x;
  if (s >= sweeps) return;
  int c = 0, t = 0;
  for (int k = 0; k < kCounterSlots; ++k) {
    c += counters[(s * kCounterSlots + k) * 2];
    t += counters[(s * kCounterSlots + k) * 2 + 1];
  }
  out[2 * s] = c;
  out[2 * s + 1] = t;
}

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Algorithmic bytes of one smoothing sweep per vertex: xy16 M24 v2v28 v2t28
// tri24 qualit16 colour-order4 read, xy16 M24 qualit16 written.
constexpr double kBytesSweepPerV = 196.0;

}  // namespace trn

using namespace trn;

extern "C" {

int trn_quality_cache(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                      const double* tensor_d, double* qualit_d) {
  return guard_impl("trn_quality_cache", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(mesh && xy_d && tensor_d && qualit_d, TRN_ERR_ARG, "null argument");
    launch(ctx, "quality_cache", 0.0, k_quality_cache, dim3(nblk(mesh->nt, 256)), dim3(256), 0,
           mesh->nt, (const int32_t*)mesh->tri, (const double2*)xy_d, tensor_d, qualit_d);
  });
}

int trn_smooth_run(trn_ctx* ctx, const trn_mesh* mesh, double* xy_d, double* tensor_d,
                   double* qualit_d, const int32_t* order_d, const int32_t* class_off,
                   int32_t ncolors, int32_t sweeps, int32_t recompute_qualit, int32_t* commits,
                   int32_t* tasks) {
  return guard_impl("trn_smooth_run", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(mesh && xy_d && tensor_d && qualit_d && order_d && class_off, TRN_ERR_ARG,
                "null argument");
    TRN_REQUIRE(ncolors >= 0 && ncolors <= 64 && sweeps >= 0, TRN_ERR_ARG,
                "bad ncolors or sweeps");
    for (int c = 0; c < ncolors; ++c)
      TRN_REQUIRE(class_off[c] >= 0 && class_off[c] <= class_off[c + 1] &&
                      class_off[c + 1] <= mesh->nv,
                  TRN_ERR_ARG, "class offsets out of range");
    TRN_REQUIRE(sweeps <= 1024, TRN_ERR_ARG, "at most 1024 sweeps per call");
    const trn_mesh* m = mesh;
    const size_t ncount = 2 * (size_t)kCounterSlots * (size_t)(sweeps + 1);
    Scratch sc(ctx, scratch_bytes(sizeof(int32_t) * ncount) +
                        scratch_bytes(sizeof(int32_t) * 2 * (size_t)(sweeps + 1)));
    int32_t* counters = sc.carve<int32_t>(ncount);
    int32_t* totals = sc.carve<int32_t>(2 * (size_t)(sweeps + 1));
    memset_async(ctx, counters, 0, sizeof(int32_t) * ncount);
    if (recompute_qualit)
      launch(ctx, "quality_cache", 0.0, k_quality_cache, dim3(nblk(m->nt, 256)), dim3(256), 0,
             m->nt, (const int32_t*)m->tri, (const double2*)xy_d, (const double*)tensor_d,
             qualit_d);
    const size_t shmem = sizeof(double) * kSmoothB * (size_t)(m->max_deg_t > 0 ? m->max_deg_t : 1);
    for (int s = 0; s < sweeps; ++s) {
      for (int c = 0; c < ncolors; ++c) {
        const int32_t count = class_off[c + 1] - class_off[c];
        launch(ctx, "smooth_phase", kBytesSweepPerV * count, k_smooth_phase,
               dim3(nblk(count, kSmoothB)), dim3(kSmoothB), shmem, count,
               order_d + class_off[c], (double2*)xy_d, tensor_d, qualit_d,
               (const int32_t*)m->tri, (const uint8_t*)m->tags, (const int32_t*)m->v2t_off,
               (const int32_t*)m->v2t, (const int32_t*)m->v2v_off, (const int32_t*)m->v2v,
               counters + 2 * kCounterSlots * s);
      }
    }
    if ((commits || tasks) && sweeps > 0) {
      launch(ctx, "counter_reduce", 0.0, k_counter_reduce, dim3(1), dim3(1024), 0, sweeps,
             (const int32_t*)counters, totals);
      copy_async(ctx, ctx->pinned_i32, totals, sizeof(int32_t) * 2 * (size_t)sweeps);
      sync(ctx);
      for (int s = 0; s < sweeps; ++s) {
        if (commits) commits[s] = ctx->pinned_i32[2 * s];
        if (tasks) tasks[s] = ctx->pinned_i32[2 * s + 1];
      }
    }
  });
}

}  // extern "C"
