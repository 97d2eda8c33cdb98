// trn_recovery.hpp -- per-element bodies of the recovery and metric stages
// (src/hessian.cpp, src/metric.cpp), shared by the per-stage kernels of
// metric.hip and the persistent batch kernel of batch.hip so that both run
// the same arithmetic in the same order.
#pragma once

#include "trn_api.hpp"
#include "trn_numeric.hpp"

namespace trn {

// Metrics::Metrics, src/metric.cpp:44 (note p = 0 gives +1/2).
inline double metric_exponent(int32_t p) { return p ? -1. / (2 * p + 2) : 0.5; }

// Builds the ascending 2-ring element patch of vertex i (Mesh::getVicinity
// (i, 2), src/mesh.cpp:390-434) in the thread's LDS set, element k at
// s[k * stride].  Returns the size; ok = false on overflow.
__device__ __forceinline__ int build_patch(const int i, const int32_t* __restrict__ v2t_off,
                                           const int32_t* __restrict__ v2t,
                                           const int32_t* __restrict__ v2v_off,
                                           const int32_t* __restrict__ v2v, int* s,
                                           const int stride, const int cap, bool& ok) {
  int m = 0;
  ok = true;
  const int32_t b0 = v2t_off[i], e0 = v2t_off[i + 1];
  for (int32_t k = b0; k < e0; ++k) ok = ok && lds_sorted_insert(s, stride, m, cap, v2t[k]);
  const int32_t nb0 = v2v_off[i], ne0 = v2v_off[i + 1];
  for (int32_t a = nb0; a < ne0; ++a) {
    const int32_t j = v2v[a];
    const int32_t bj = v2t_off[j], ej = v2t_off[j + 1];
    for (int32_t k = bj; k < ej; ++k) ok = ok && lds_sorted_insert(s, stride, m, cap, v2t[k]);
  }
  return m;
}

// Per-triangle gradient terms (src/hessian.cpp:42-67) from the corners'
// coordinates and field values: area and the products area * element
// gradient that the vertex sums add (:78).
__device__ __forceinline__ void grad_terms(const double2 p0, const double2 p1, const double2 p2,
                                           const double f0, const double f1, const double f2,
                                           double& a, double2& ag) {
  double d[6];
  tri_geometry(p0, p1, p2, a, d);
  double g0 = 0., g1 = 0.;  // hessian.cpp:36, accumulated j = 0,1,2 (:65-67)
  g0 += f0 * d[0];
  g1 += f0 * d[1];
  g0 += f1 * d[2];
  g1 += f1 * d[3];
  g0 += f2 * d[4];
  g1 += f2 * d[5];
  ag = make_double2(a * g0, a * g1);
}

// grad_terms of triangle t into area_t / ag_t.  f == nullptr stores the
// area only.
__device__ __forceinline__ void tri_grad_one(const int64_t t, const int32_t* __restrict__ tri,
                                             const double2* __restrict__ xy,
                                             const double* __restrict__ f,
                                             double* __restrict__ area_t,
                                             double2* __restrict__ ag_t) {
  const int32_t n0 = tri[3 * t], n1 = tri[3 * t + 1], n2 = tri[3 * t + 2];
  if (n0 < 0) return;  // deleted triangles never appear in a patch
  if (f == nullptr) {
    double a, d[6];
    tri_geometry(xy[n0], xy[n1], xy[n2], a, d);
    area_t[t] = a;
    return;
  }
  double a;
  double2 ag;
  grad_terms(xy[n0], xy[n1], xy[n2], f[n0], f[n1], f[n2], a, ag);
  area_t[t] = a;
  ag_t[t] = ag;
}

// Patch entries are visited in batches of kPatchBatch: the batch's indices
// are loaded first, then its terms, so that a lane has that many
// independent loads in flight instead of one dependent pair at a time; the
// terms are still added one by one in patch order.
constexpr int kPatchBatch = 8;

// Per-vertex gradient (src/hessian.cpp:75-82) from the patch in the LDS set:
// total area and area-weighted element gradients summed in patch order.
template <bool WITH_GRAD, class Idx>
__device__ __forceinline__ void vtx_grad_one(const int i, const Idx* s, const int stride,
                                             const int m, const double* __restrict__ area_t,
                                             const double2* __restrict__ ag_t,
                                             double2* __restrict__ grad,
                                             double* __restrict__ tarea) {
  double ta = 0., gx = 0., gy = 0.;
  for (int e0 = 0; e0 < m; e0 += kPatchBatch) {
    int t[kPatchBatch];
#pragma unroll
    for (int c = 0; c < kPatchBatch; ++c) t[c] = e0 + c < m ? (int)s[(e0 + c) * stride] : 0;
    double a[kPatchBatch];
    double2 r[kPatchBatch];
#pragma unroll
    for (int c = 0; c < kPatchBatch; ++c) {  // t = 0 past the patch end: in bounds
      a[c] = area_t[t[c]];
      if (WITH_GRAD) r[c] = ag_t[t[c]];
    }
#pragma unroll
    for (int c = 0; c < kPatchBatch; ++c)
      if (e0 + c < m) {
        ta += a[c];  // hessian.cpp:53
        if (WITH_GRAD) {
          gx += r[c].x;  // hessian.cpp:76-78
          gy += r[c].y;
        }
      }
  }
  if (WITH_GRAD) grad[i] = make_double2(gx / ta, gy / ta);  // hessian.cpp:81-82
  tarea[i] = ta;
}

// Stored patches of at most kRegRows entries keep their local indices in
// registers: a lane loads its whole column at once, while the union is being
// staged, and the sums then read LDS only.  Entries are added in patch order.
constexpr int kRegRows = 32;

template <bool WITH_GRAD>
__device__ __forceinline__ void vtx_grad_reg(const int i, const int (&loc)[kRegRows],
                                             const int m, const double* area,
                                             const double2* ag, double2* __restrict__ grad,
                                             double* __restrict__ tarea) {
  double ta = 0., gx = 0., gy = 0.;
#pragma unroll
  for (int e = 0; e < kRegRows; ++e)
    if (e < m) {
      ta += area[loc[e]];  // hessian.cpp:53
      if (WITH_GRAD) {
        const double2 r = ag[loc[e]];
        gx += r.x;  // hessian.cpp:76-78
        gy += r.y;
      }
    }
  if (WITH_GRAD) grad[i] = make_double2(gx / ta, gy / ta);  // hessian.cpp:81-82
  tarea[i] = ta;
}

// Per-triangle Hessian terms: area * delta (src/hessian.cpp:108-143) from
// the corners' coordinates and recovered gradients.
__device__ __forceinline__ double4 hess_terms(const double2 p0, const double2 p1,
                                              const double2 p2, const double2 g0,
                                              const double2 g1, const double2 g2) {
  double a, d[6];
  tri_geometry(p0, p1, p2, a, d);
  double l0 = 0., l1 = 0., l2 = 0., l3 = 0.;  // hessian.cpp:99, :131-136
  l0 += g0.x * d[0]; l1 += g0.x * d[1]; l2 += g0.y * d[0]; l3 += g0.y * d[1];
  l0 += g1.x * d[2]; l1 += g1.x * d[3]; l2 += g1.y * d[2]; l3 += g1.y * d[3];
  l0 += g2.x * d[4]; l1 += g2.x * d[5]; l2 += g2.y * d[4]; l3 += g2.y * d[5];
  double4 r;
  r.x = a * l0;  // the products of hessian.cpp:143
  r.y = a * l1;
  r.z = a * l2;
  r.w = a * l3;
  return r;
}

__device__ __forceinline__ void tri_hess_one(const int64_t t, const int32_t* __restrict__ tri,
                                             const double2* __restrict__ xy,
                                             const double2* __restrict__ grad,
                                             double4* __restrict__ ah_t) {
  const int32_t n0 = tri[3 * t], n1 = tri[3 * t + 1], n2 = tri[3 * t + 2];
  if (n0 < 0) return;
  ah_t[t] = hess_terms(xy[n0], xy[n1], xy[n2], grad[n0], grad[n1], grad[n2]);
}

// Metrics::normalizeLocally for one tensor (src/metric.cpp:113-148).
__device__ __forceinline__ void normalize_local(double& t0, double& t1, double& t2,
                                                const double ex) {
  double val[2], vec[4], h[4];
  eigen(t0, t1, t1, t2, val, vec);
  val[0] = dmax(fabs(val[0]), kEps);
  val[1] = dmax(fabs(val[1]), kEps);
  const double det = val[0] * val[1];
  const double scale_loc = libm::pow(det, ex);  // metric.cpp:129
  val[0] *= scale_loc;
  val[1] *= scale_loc;
  rebuild(val, vec, h);
  t0 = h[0];
  t1 = h[1];
  t2 = h[3];
}

// A vertex's term of Metrics::computeComplexity: sqrt(det) of its tensor.
__device__ __forceinline__ double cplx_term(const double m0, const double m1, const double m2) {
  const double det = m0 * m2 - m1 * m1;  // metric.cpp:171-179
  return __builtin_sqrt(det);
}

// Patch sums -> symmetric Hessian (src/hessian.cpp:141-152), optionally
// followed by the local normalisation, stored at out (and hess if given).
// sdet (when given): the normalised tensor's cplx_term, computed once per
// vertex instead of once per incident triangle by the complexity kernel.
template <bool NORMALIZE>
__device__ __forceinline__ void hess_finish(const int i, double h0, double h1, double h2,
                                            double h3, const double* __restrict__ tarea,
                                            double* __restrict__ hess, double* __restrict__ out,
                                            double* __restrict__ sdet, const double ex) {
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
  if (NORMALIZE && sdet) sdet[i] = cplx_term(t0, t1, t2);
}

// Per-vertex Hessian (src/hessian.cpp:141-152) from the patch in the LDS set,
// optionally followed by the local normalisation.
template <bool NORMALIZE, class Idx>
__device__ __forceinline__ void vtx_hess_one(const int i, const Idx* s, const int stride,
                                             const int m, const double4* __restrict__ ah_t,
                                             const double* __restrict__ tarea,
                                             double* __restrict__ hess, double* __restrict__ out,
                                             double* __restrict__ sdet, const double ex) {
  double h0 = 0., h1 = 0., h2 = 0., h3 = 0.;  // hessian.cpp:93
  for (int e0 = 0; e0 < m; e0 += kPatchBatch) {
    int t[kPatchBatch];
#pragma unroll
    for (int c = 0; c < kPatchBatch; ++c) t[c] = e0 + c < m ? (int)s[(e0 + c) * stride] : 0;
    double4 r[kPatchBatch];
#pragma unroll
    for (int c = 0; c < kPatchBatch; ++c) r[c] = ah_t[t[c]];  // t = 0 past the end: in bounds
#pragma unroll
    for (int c = 0; c < kPatchBatch; ++c)
      if (e0 + c < m) {
        h0 += r[c].x;
        h1 += r[c].y;
        h2 += r[c].z;
        h3 += r[c].w;
      }
  }
  hess_finish<NORMALIZE>(i, h0, h1, h2, h3, tarea, hess, out, sdet, ex);
}

template <bool NORMALIZE>
__device__ __forceinline__ void vtx_hess_reg(const int i, const int (&loc)[kRegRows],
                                             const int m, const double4* ah,
                                             const double* __restrict__ tarea,
                                             double* __restrict__ hess, double* __restrict__ out,
                                             double* __restrict__ sdet, const double ex) {
  double h0 = 0., h1 = 0., h2 = 0., h3 = 0.;  // hessian.cpp:93
#pragma unroll
  for (int e = 0; e < kRegRows; ++e)
    if (e < m) {
      const double4 r = ah[loc[e]];
      h0 += r.x;
      h1 += r.y;
      h2 += r.z;
      h3 += r.w;
    }
  hess_finish<NORMALIZE>(i, h0, h1, h2, h3, tarea, hess, out, sdet, ex);
}

// One triangle's term of Metrics::computeComplexity (src/metric.cpp:161-180).
__device__ __forceinline__ double cplx_one(const int64_t t, const int32_t* __restrict__ tri,
                                           const double2* __restrict__ xy,
                                           const double* __restrict__ tensor) {
  const int32_t v0 = tri[3 * t], v1 = tri[3 * t + 1], v2 = tri[3 * t + 2];
  if (v0 < 0) return 0.;
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
    rho += cplx_term(m[0], m[1], m[2]);
  }
  rho /= 3;
  return rho * area;
}

// cplx_one with the vertices' terms read from sdet (hess_finish): the same
// values added in the same order.
__device__ __forceinline__ double cplx_one_sdet(const int64_t t, const int32_t* __restrict__ tri,
                                                const double2* __restrict__ xy,
                                                const double* __restrict__ sdet) {
  const int32_t v0 = tri[3 * t], v1 = tri[3 * t + 1], v2 = tri[3 * t + 2];
  if (v0 < 0) return 0.;
  const double2 p0 = xy[v0], p1 = xy[v1], p2 = xy[v2];
  const double r0 = sdet[v0], r1 = sdet[v1], r2 = sdet[v2];
  const double s0 = p1.x - p0.x;  // metric.cpp:163-167
  const double s1 = p2.x - p0.x;
  const double s2 = p1.y - p0.y;
  const double s3 = p2.y - p0.y;
  const double area = 0.5 * (s0 * s3 - s1 * s2);
  double rho = 0.;
  rho += r0;
  rho += r1;
  rho += r2;
  rho /= 3;
  return rho * area;
}

// Metrics::normalizeGlobally for one tensor (src/metric.cpp:204-236).
__device__ __forceinline__ void global_one(double* __restrict__ t, const double fact,
                                           const double emin, const double emax) {
  double val[2], vec[4], h[4];
  const double t0 = t[0], t1 = t[1], t2 = t[2];
  eigen(t0, t1, t1, t2, val, vec);
  val[0] *= fact;
  val[1] *= fact;
#pragma unroll
  for (int j = 0; j < 2; ++j) {  // metric.cpp:213-216
    val[j] = dmin(val[j], emin);
    val[j] = dmax(val[j], emax);
  }
  rebuild(val, vec, h);
  t[0] = h[0];
  t[1] = h[1];
  t[2] = h[3];
}

// Deterministic sums: fixed butterflies inside a wave, fixed order across.
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
    for (int w = 1; w < (int)((blockDim.x + 63) >> 6); ++w) r += lds[w];
  }
  __syncthreads();
  return r;
}

}  // namespace trn
