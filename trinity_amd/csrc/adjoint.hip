// adjoint.hip -- the metric-space quality energy and its gradient on gfx950.
//
//   E(xy, f) = sum over active triangles t of (1 - q_M(t))^2
//
// q_M is numeric::computeQuality (src/numeric.cpp:251-306) with the metric M
// that Metrics::run (src/metric.cpp:59-70) builds from the coordinates xy
// and the nodal values f.  The reference has neither the energy nor its
// gradient; the forward reuses the drop-in metric pass unchanged (so M is
// the metric trn_metric_run returns) and the backward runs in reverse
// through quality, global normalisation, complexity, local normalisation,
// Hessian recovery, gradient recovery and triangle geometry
// (trn_adjoint.hpp).  Triangle geometry is recomputed, not stored.
//
// Determinism without FP64 atomics: triangle-major passes write one
// contribution per (triangle, corner) and vertex-major passes gather them
// over the vertex's incident triangles in stencil order.  The recovery
// transposes use that triangle t lies in the 2-ring patch of vertex i
// exactly when i is a corner of t or a neighbour of one, so a triangle sums
// the per-vertex adjoints over that set, deduplicated in ascending order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "trn_adjoint.hpp"
#include "trn_recovery.hpp"
#include "trn_smooth.hpp"

namespace trn {

constexpr int kAdjB = 64;        // threads per block of the LDS-set kernels
constexpr int kRingCap = 256;    // vertices of a triangle's ring union

// per-(triangle, corner) layouts (doubles)
constexpr int kQB = 15;  // [xy adjoint 2, M adjoint 3] x 3 corners
constexpr int kGB = 6;   // vertex-gradient adjoint 2 x 3 corners
constexpr int kAD = 7;   // area adjoint, basis-gradient adjoints 6
constexpr int kFB = 3;   // nodal-value adjoint x 3 corners
// Entries of a vertex's incident list and of a triangle's ring union whose
// loads the gathers issue as one batch before summing (in the same order);
// longer lists finish in a loop.  The ring-union batch is set per kernel
// by measurement on the 4096^2 grid.
constexpr int kAdjBatchV = 8;
constexpr int kAdjBatchHess = 8;
constexpr int kAdjBatchGrad = 16;

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---- forward -----------------------------------------------------------------

__global__ __launch_bounds__(256) void k_adj_area(int32_t nt, const int32_t* __restrict__ tri,
                                                  const double2* __restrict__ xy,
                                                  double* __restrict__ area_t) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nt) tri_grad_one(t, tri, xy, nullptr, area_t, nullptr);
}

__global__ __launch_bounds__(256) void k_adj_logm(int32_t nv, const double* __restrict__ M,
                                                  double4* __restrict__ logm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) logm[i] = tensor_log4(M[3 * i], M[3 * i + 1], M[3 * i + 2]);
}

// qualities and per-block partial energies (grid-stride, fixed order)
__global__ __launch_bounds__(256) void k_adj_energy(int32_t nt, const int32_t* __restrict__ tri,
                                                    const double2* __restrict__ xy,
                                                    const double4* __restrict__ logm,
                                                    double* __restrict__ qualit,
                                                    double* __restrict__ partial) {
  __shared__ double red[4];
  double e = 0.;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (tri[3 * t] < 0) continue;
    const double q = quality_logs(t, tri, xy, logm);
    qualit[t] = q;
    e += (1 - q) * (1 - q);
  }
  const double b = block_sum(e, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// out[slot] = sum of partial[0..n) in order, times scale
// One block: thread j adds partial[j], partial[j + B], ... in order, then a
// fixed-shape block tree (wave shuffles, then waves in order) adds the
// threads' sums, so the result does not depend on timing.
constexpr int kSumB = 1024;
__global__ __launch_bounds__(kSumB) void k_adj_sum(const double* __restrict__ partial,
                                                   int32_t n, double scale,
                                                   double* __restrict__ out) {
  __shared__ double red[kSumB / 64];
  double s = 0.;
  for (int k = threadIdx.x; k < n; k += kSumB) s += partial[k];
  s = block_sum(s, red);
  if (threadIdx.x == 0) *out = s * scale;
}

// ---- backward ----------------------------------------------------------------

// B1: quality adjoint per triangle, dq = -2 (1 - q)
__global__ __launch_bounds__(256) void k_adj_quality(int32_t nt, const int32_t* __restrict__ tri,
                                                     const double2* __restrict__ xy,
                                                     const double* __restrict__ M,
                                                     const double* __restrict__ qualit,
                                                     double* __restrict__ QB) {
  const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= nt || tri[3 * t] < 0) return;
  const int32_t n0 = tri[3 * t], n1 = tri[3 * t + 1], n2 = tri[3 * t + 2];
  double g[6] = {0., 0., 0., 0., 0., 0.}, gm[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
  const double dq = -2 * (1 - qualit[t]);
  quality_adjoint(xy[n0], xy[n1], xy[n2], M + 3 * (int64_t)n0, M + 3 * (int64_t)n1,
                  M + 3 * (int64_t)n2, dq, g, gm);
  double* o = QB + kQB * t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o[5 * k] = g[2 * k];
    o[5 * k + 1] = g[2 * k + 1];
    o[5 * k + 2] = gm[3 * k];
    o[5 * k + 3] = gm[3 * k + 1];
    o[5 * k + 4] = gm[3 * k + 2];
  }
}

// B2: gather the metric adjoint, reverse the global normalisation; partial
// sums of d E / d fact per block
__global__ __launch_bounds__(256) void k_adj_global(
    int32_t nv, const int32_t* __restrict__ v2t_off, const int32_t* __restrict__ v2t,
    const int32_t* __restrict__ star, const double* __restrict__ QB,
    const double* __restrict__ Mloc, const double* __restrict__ scal, double emin, double emax,
    double* __restrict__ Mlb, double* __restrict__ partial) {
  __shared__ double red[4];
  const uint32_t lb = xcd_block();
  const int64_t i = (int64_t)lb * blockDim.x + threadIdx.x;
  double fb = 0.;
  if (i < nv) {
    double a0 = 0., a1 = 0., a2 = 0.;
    const int32_t k0 = v2t_off[i], k1 = v2t_off[i + 1];
    const double* sp[kAdjBatchV];
#pragma unroll
    for (int q = 0; q < kAdjBatchV; ++q)
      sp[q] = k0 + q < k1 ? QB + kQB * (int64_t)v2t[k0 + q] + 5 * (star[k0 + q] & 3) + 2 : QB;
    double b[kAdjBatchV][3];
#pragma unroll
    for (int q = 0; q < kAdjBatchV; ++q) {
      b[q][0] = sp[q][0];
      b[q][1] = sp[q][1];
      b[q][2] = sp[q][2];
    }
#pragma unroll
    for (int q = 0; q < kAdjBatchV; ++q) {
      if (k0 + q < k1) {
        a0 += b[q][0];
        a1 += b[q][1];
        a2 += b[q][2];
      }
    }
    for (int32_t k = k0 + kAdjBatchV; k < k1; ++k) {
      const double* s = QB + kQB * (int64_t)v2t[k] + 5 * (star[k] & 3) + 2;
      a0 += s[0];
      a1 += s[1];
      a2 += s[2];
    }
    const double* t = Mloc + 3 * i;
    Eig2 e;
    eig2(t[0], t[1], t[1], t[2], e);
    double jac[4], gam, dfa[2], bx[4], ba[4];
    global_jac(e, scal[1], emin, emax, jac, gam, dfa);
    sym_to_mat(a0, a1, a2, bx);
    spectral_adjoint(e, jac, gam, bx, ba);
    Mlb[3 * i] = ba[0];
    Mlb[3 * i + 1] = ba[1] + ba[2];
    Mlb[3 * i + 2] = ba[3];
    double b00, b11;
    eig_diag(e, bx, b00, b11);
    fb = (e.branch == 2) ? (b00 + b11) * dfa[1] : b00 * dfa[0] + b11 * dfa[1];
  }
  const double b = block_sum(fb, red);
  if (threadIdx.x == 0) partial[lb] = b;
}

// B4: complexity (sqrt det term of the vertex) and local normalisation;
// yields the adjoints of the patch sums S_i = sum a_t delta_t (as s0, s12,
// s3: the two off-diagonal entries share one) and of the patch area TA_i
__global__ __launch_bounds__(256) void k_adj_local(
    int32_t nv, const int32_t* __restrict__ v2t_off, const int32_t* __restrict__ v2t,
    const double* __restrict__ area_t, const double* __restrict__ Mloc,
    const double* __restrict__ Hs, const double* __restrict__ TA,
    const double* __restrict__ scal, double ex, const double* __restrict__ Mlb,
    double* __restrict__ Sb, double* __restrict__ TAb) {
  const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const double cb = scal[2];
  double ring = 0.;
  const int32_t k0 = v2t_off[i], k1 = v2t_off[i + 1];
  {
    double av[kAdjBatchV];
#pragma unroll
    for (int q = 0; q < kAdjBatchV; ++q) av[q] = k0 + q < k1 ? area_t[v2t[k0 + q]] : 0.;
#pragma unroll
    for (int q = 0; q < kAdjBatchV; ++q)
      if (k0 + q < k1) ring += av[q];
  }
  for (int32_t k = k0 + kAdjBatchV; k < k1; ++k) ring += area_t[v2t[k]];
  const double* m = Mloc + 3 * i;
  const double sd = __builtin_sqrt(m[0] * m[2] - m[1] * m[1]);
  const double w = cb * ring / 3 / (2 * sd);  // metric.cpp:171-179
  const double a0 = Mlb[3 * i] + w * m[2];
  const double a1 = Mlb[3 * i + 1] - 2 * w * m[1];
  const double a2 = Mlb[3 * i + 2] + w * m[0];
  const double* h = Hs + 3 * i;
  Eig2 e;
  eig2(h[0], h[1], h[1], h[2], e);
  double jac[4], gam, bx[4], ba[4];
  local_jac(e, ex, jac, gam);
  sym_to_mat(a0, a1, a2, bx);
  spectral_adjoint(e, jac, gam, bx, ba);
  const double hb0 = ba[0], hb1 = ba[1] + ba[2], hb2 = ba[3];
  const double ta = TA[i];
  Sb[3 * i] = hb0 / ta;
  Sb[3 * i + 1] = 0.5 * hb1 / ta;  // each raw off-diagonal gets half
  Sb[3 * i + 2] = hb2 / ta;
  TAb[i] = -(hb0 * h[0] + hb1 * h[1] + hb2 * h[2]) / ta;
}

// Ring union of triangle t's corners (the vertices whose 2-ring patch holds
// t), ascending, in the thread's LDS set.
__device__ __forceinline__ int ring_union(const int32_t n[3], const int32_t* __restrict__ v2v_off,
                                          const int32_t* __restrict__ v2v, int* s, int stride,
                                          int cap, bool& ok) {
  int m = 0;
  ok = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ok = ok && lds_sorted_insert(s, stride, m, cap, n[k]);
    for (int32_t a = v2v_off[n[k]]; a < v2v_off[n[k] + 1]; ++a)
      ok = ok && lds_sorted_insert(s, stride, m, cap, v2v[a]);
  }
  return m;
}

// Stored ring unions, built with the adjoint.  SIZE: per triangle, the
// size of its ring union (err set if one exceeds cap); FILL: the union written
// column-wise into the triangle's slab, every column and row of the slab
// written (-1 padding past the union or past the last triangle).
template <bool FILL>
__global__ __launch_bounds__(kAdjB) void k_tring_build(
    int32_t nt, const int32_t* __restrict__ tri, const int32_t* __restrict__ v2v_off,
    const int32_t* __restrict__ v2v, int32_t cap, int32_t* __restrict__ tr_n,
    const int32_t* __restrict__ tr_row, int32_t* __restrict__ tr_ell, int32_t* __restrict__ err) {
  extern __shared__ int lds[];
  const int64_t t = (int64_t)blockIdx.x * kAdjB + threadIdx.x;
  int* s = lds + threadIdx.x;
  int m = 0;
  bool ok = true;
  if (t < nt && tri[3 * t] >= 0) {
    const int32_t n[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
    m = ring_union(n, v2v_off, v2v, s, kAdjB, cap, ok);
  }
  if (!FILL) {
    if (!ok) atomicOr(err, 1);
    if (t < nt) tr_n[t] = ok ? m : 0;
    return;
  }
  const int32_t r0 = tr_row[blockIdx.x], w = tr_row[blockIdx.x + 1] - r0;
  int32_t* col = tr_ell + (int64_t)r0 * kAdjB + threadIdx.x;
  for (int e = 0; e < w; ++e) col[(int64_t)e * kAdjB] = (ok && e < m) ? s[e * kAdjB] : -1;
}

// Rows of each 64-triangle slab: its largest ring union (one wave per slab).
__global__ __launch_bounds__(kAdjB) void k_tring_rows(int32_t nt, const int32_t* __restrict__ n,
                                                      int32_t* __restrict__ rows) {
  const int64_t t = (int64_t)blockIdx.x * kAdjB + threadIdx.x;
  int m = t < nt ? n[t] : 0;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const int o = __shfl_xor(m, d, 64);
    m = o > m ? o : m;
  }
  if (threadIdx.x == 0) rows[blockIdx.x] = m;
}

// B5: Hessian recovery transpose per triangle
__global__ __launch_bounds__(kAdjB) void k_adj_hess_tri(
    int32_t nt, const int32_t* __restrict__ tri, const double2* __restrict__ xy,
    const double2* __restrict__ G, const double* __restrict__ Sb,
    const int32_t* __restrict__ tr_n, const int32_t* __restrict__ tr_row,
    const int32_t* __restrict__ tr_ell, double* __restrict__ AD, double* __restrict__ GB) {
  const uint32_t lb = xcd_block();
  const int64_t t = (int64_t)lb * kAdjB + threadIdx.x;
  if (t >= nt || tri[3 * t] < 0) return;
  const int32_t n[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
  const int32_t* s = tr_ell + (int64_t)tr_row[lb] * kAdjB + threadIdx.x;
  const int m = tr_n[t];
  double s0 = 0., s12 = 0., s3 = 0.;
  {
    int32_t iv[kAdjBatchHess];
#pragma unroll
    for (int e = 0; e < kAdjBatchHess; ++e) iv[e] = e < m ? s[e * kAdjB] : 0;
    double bv[kAdjBatchHess][3];
#pragma unroll
    for (int e = 0; e < kAdjBatchHess; ++e) {
      const double* b = Sb + 3 * (int64_t)iv[e];
      bv[e][0] = e < m ? b[0] : 0.;
      bv[e][1] = e < m ? b[1] : 0.;
      bv[e][2] = e < m ? b[2] : 0.;
    }
#pragma unroll
    for (int e = 0; e < kAdjBatchHess; ++e) {
      if (e < m) {
        s0 += bv[e][0];
        s12 += bv[e][1];
        s3 += bv[e][2];
      }
    }
  }
  for (int e = kAdjBatchHess; e < m; ++e) {
    const double* b = Sb + 3 * (int64_t)s[e * kAdjB];
    s0 += b[0];
    s12 += b[1];
    s3 += b[2];
  }
  double a, d[6];
  tri_basis(xy[n[0]], xy[n[1]], xy[n[2]], a, d);
  const double2 g[3] = {G[n[0]], G[n[1]], G[n[2]]};
  double l0 = 0., l1 = 0., l2 = 0., l3 = 0.;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    l0 += g[j].x * d[2 * j];
    l1 += g[j].x * d[2 * j + 1];
    l2 += g[j].y * d[2 * j];
    l3 += g[j].y * d[2 * j + 1];
  }
  double* ad = AD + kAD * t;
  double* gb = GB + kGB * t;
  ad[0] = s0 * l0 + s12 * l1 + s12 * l2 + s3 * l3;
  const double b0 = a * s0, b12 = a * s12, b3 = a * s3;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    gb[2 * j] = b0 * d[2 * j] + b12 * d[2 * j + 1];
    gb[2 * j + 1] = b12 * d[2 * j] + b3 * d[2 * j + 1];
    ad[1 + 2 * j] = b0 * g[j].x + b12 * g[j].y;
    ad[2 + 2 * j] = b12 * g[j].x + b3 * g[j].y;
  }
}

// B6: gather the vertex-gradient adjoint; adjoints of the gradient patch
// sums and the total of the patch-area adjoint
__global__ __launch_bounds__(256) void k_adj_grad_vtx(
    int32_t nv, const int32_t* __restrict__ v2t_off, const int32_t* __restrict__ v2t,
    const int32_t* __restrict__ star, const double* __restrict__ GB,
    const double2* __restrict__ G, const double* __restrict__ TA, double* __restrict__ TAb,
    double2* __restrict__ Sgb) {
  const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  double gx = 0., gy = 0.;
  const int32_t k0 = v2t_off[i], k1 = v2t_off[i + 1];
  {
    double2 gv[kAdjBatchV];
#pragma unroll
    for (int q = 0; q < kAdjBatchV; ++q) {
      gv[q] = make_double2(0., 0.);
      if (k0 + q < k1) {
        const double* s = GB + kGB * (int64_t)v2t[k0 + q] + 2 * (star[k0 + q] & 3);
        gv[q] = make_double2(s[0], s[1]);
      }
    }
#pragma unroll
    for (int q = 0; q < kAdjBatchV; ++q) {
      if (k0 + q < k1) {
        gx += gv[q].x;
        gy += gv[q].y;
      }
    }
  }
  for (int32_t k = k0 + kAdjBatchV; k < k1; ++k) {
    const double* s = GB + kGB * (int64_t)v2t[k] + 2 * (star[k] & 3);
    gx += s[0];
    gy += s[1];
  }
  const double ta = TA[i];
  const double2 g = G[i];
  Sgb[i] = make_double2(gx / ta, gy / ta);
  TAb[i] -= (gx * g.x + gy * g.y) / ta;
}

// B7: gradient recovery transpose, complexity area term and geometry per
// triangle -> coordinate adjoints (added to the quality ones) and nodal
// value adjoints per corner
__global__ __launch_bounds__(kAdjB) void k_adj_grad_tri(
    int32_t nt, const int32_t* __restrict__ tri, const double2* __restrict__ xy,
    const double* __restrict__ f, const double* __restrict__ Mloc,
    const double2* __restrict__ Sgb, const double* __restrict__ TAb,
    const int32_t* __restrict__ tr_n, const int32_t* __restrict__ tr_row,
    const int32_t* __restrict__ tr_ell, const double* __restrict__ scal,
    const double* __restrict__ AD, double* __restrict__ QB, double* __restrict__ FB) {
  const uint32_t lb = xcd_block();
  const int64_t t = (int64_t)lb * kAdjB + threadIdx.x;
  if (t >= nt || tri[3 * t] < 0) return;
  const int32_t n[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
  const int32_t* s = tr_ell + (int64_t)tr_row[lb] * kAdjB + threadIdx.x;
  const int m = tr_n[t];
  double sx = 0., sy = 0., tb = 0.;
  {
    int32_t iv[kAdjBatchGrad];
#pragma unroll
    for (int e = 0; e < kAdjBatchGrad; ++e) iv[e] = e < m ? s[e * kAdjB] : 0;
    double2 vv[kAdjBatchGrad];
    double tv[kAdjBatchGrad];
#pragma unroll
    for (int e = 0; e < kAdjBatchGrad; ++e) {
      vv[e] = e < m ? Sgb[iv[e]] : make_double2(0., 0.);
      tv[e] = e < m ? TAb[iv[e]] : 0.;
    }
#pragma unroll
    for (int e = 0; e < kAdjBatchGrad; ++e) {
      if (e < m) {
        sx += vv[e].x;
        sy += vv[e].y;
        tb += tv[e];
      }
    }
  }
  for (int e = kAdjBatchGrad; e < m; ++e) {
    const int i = s[e * kAdjB];
    const double2 v = Sgb[i];
    sx += v.x;
    sy += v.y;
    tb += TAb[i];
  }
  const double2 p0 = xy[n[0]], p1 = xy[n[1]], p2 = xy[n[2]];
  double a, d[6];
  tri_basis(p0, p1, p2, a, d);
  const double fv[3] = {f[n[0]], f[n[1]], f[n[2]]};
  double g0 = 0., g1 = 0.;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    g0 += fv[j] * d[2 * j];
    g1 += fv[j] * d[2 * j + 1];
  }
  double rho = 0.;  // metric.cpp:171-179
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double* mm = Mloc + 3 * (int64_t)n[j];
    rho += __builtin_sqrt(mm[0] * mm[2] - mm[1] * mm[1]);
  }
  rho /= 3;
  const double* ad = AD + kAD * t;
  const double ab = ad[0] + sx * g0 + sy * g1 + tb + scal[2] * rho;
  const double gbx = a * sx, gby = a * sy;
  double db[6];
  double* fb = FB + kFB * t;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    fb[j] = gbx * d[2 * j] + gby * d[2 * j + 1];
    db[2 * j] = ad[1 + 2 * j] + fv[j] * gbx;
    db[2 * j + 1] = ad[2 + 2 * j] + fv[j] * gby;
  }
  double g[6] = {0., 0., 0., 0., 0., 0.};
  tri_basis_adjoint(p0, p1, p2, ab, db, g);
  double* q = QB + kQB * t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    q[5 * k] += g[2 * k];
    q[5 * k + 1] += g[2 * k + 1];
  }
}

// B8: gather the coordinate and nodal-value adjoints per vertex
__global__ __launch_bounds__(256) void k_adj_gather(
    int32_t nv, const int32_t* __restrict__ v2t_off, const int32_t* __restrict__ v2t,
    const int32_t* __restrict__ star, const double* __restrict__ QB,
    const double* __restrict__ FB, double2* __restrict__ dxy, double* __restrict__ df) {
  const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  double gx = 0., gy = 0., gf = 0.;
  const int32_t k0 = v2t_off[i], k1 = v2t_off[i + 1];
  {
    double vx[kAdjBatchV], vy[kAdjBatchV], vf[kAdjBatchV];
#pragma unroll
    for (int q = 0; q < kAdjBatchV; ++q) {
      vx[q] = vy[q] = vf[q] = 0.;
      if (k0 + q < k1) {
        const int64_t t = v2t[k0 + q];
        const int r = star[k0 + q] & 3;
        vx[q] = QB[kQB * t + 5 * r];
        vy[q] = QB[kQB * t + 5 * r + 1];
        vf[q] = FB[kFB * t + r];
      }
    }
#pragma unroll
    for (int q = 0; q < kAdjBatchV; ++q) {
      if (k0 + q < k1) {
        gx += vx[q];
        gy += vy[q];
        gf += vf[q];
      }
    }
  }
  for (int32_t k = k0 + kAdjBatchV; k < k1; ++k) {
    const int64_t t = v2t[k];
    const int r = star[k] & 3;
    gx += QB[kQB * t + 5 * r];
    gy += QB[kQB * t + 5 * r + 1];
    gf += FB[kFB * t + r];
  }
  if (dxy) dxy[i] = make_double2(gx, gy);
  if (df) df[i] = gf;
}

// ---- gradient-based smoothing ---------------------------------------------------

// Preconditioned gradient of the gradient-based smoothing, in place: on
// the vertices that may move, g_i becomes h_i^2 g_i with h_i the mean
// length of the edges at i (a diagonal preconditioner: the gradient of an
// element quality scales as 1/h and its curvature as 1/h^2, so a plain
// gradient step is set by the smallest elements and barely moves the
// others).
__global__ __launch_bounds__(256) void k_adj_precond(int32_t nv, const uint8_t* __restrict__ tags,
                                                     const int32_t* __restrict__ v2t_off,
                                                     const int32_t* __restrict__ v2v_off,
                                                     const int32_t* __restrict__ v2v,
                                                     const double2* __restrict__ xy,
                                                     double2* __restrict__ g) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int32_t a0 = v2v_off[i], a1 = v2v_off[i + 1];
  if ((tags[i] & 4) || v2t_off[i + 1] == v2t_off[i] || a1 == a0) return;
  const double2 p = xy[i];
  double s = 0.;
  for (int32_t a = a0; a < a1; ++a) {
    const double2 q = xy[v2v[a]];
    s += __builtin_sqrt((q.x - p.x) * (q.x - p.x) + (q.y - p.y) * (q.y - p.y));
  }
  const double h = s / (a1 - a0);
  const double2 gi = g[i];
  g[i] = make_double2((h * h) * gi.x, (h * h) * gi.y);
}

// max |dE/dx| over the vertices that may move (block partials)
__global__ __launch_bounds__(256) void k_adj_gmax(int32_t nv, const uint8_t* __restrict__ tags,
                                                  const int32_t* __restrict__ v2t_off,
                                                  const double2* __restrict__ g,
                                                  double* __restrict__ partial) {
  __shared__ double red[4];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.;
  if (i < nv && !(tags[i] & 4) && v2t_off[i + 1] > v2t_off[i]) {
    const double2 d = g[i];
    v = dmax(fabs(d.x), fabs(d.y));
  }
  v = dmax(v, __shfl_xor(v, 32, 64));
  v = dmax(v, __shfl_xor(v, 16, 64));
  v = dmax(v, __shfl_xor(v, 8, 64));
  v = dmax(v, __shfl_xor(v, 4, 64));
  v = dmax(v, __shfl_xor(v, 2, 64));
  v = dmax(v, __shfl_xor(v, 1, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = dmax(r, red[w]);
    partial[blockIdx.x] = r;
  }
}

__global__ void k_adj_max(const double* __restrict__ partial, int32_t n, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double r = 0.;
  for (int k = 0; k < n; ++k) r = dmax(r, partial[k]);
  *out = r;
}

// xy_out = xy - alpha * g on the vertices that may move (Smooth::preProcess,
// src/smoothing.cpp:134-138: boundary vertices stay), copy elsewhere
__global__ __launch_bounds__(256) void k_adj_step(int32_t nv, const uint8_t* __restrict__ tags,
                                                  const int32_t* __restrict__ v2t_off,
                                                  const double2* __restrict__ xy,
                                                  const double2* __restrict__ g, double alpha,
                                                  double2* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  double2 p = xy[i];
  if (!(tags[i] & 4) && v2t_off[i + 1] > v2t_off[i]) {
    const double2 d = g[i];
    p.x -= alpha * d.x;
    p.y -= alpha * d.y;
  }
  out[i] = p;
}

// counts triangles whose signed area is not positive
__global__ __launch_bounds__(256) void k_adj_inverted(int32_t nt, const int32_t* __restrict__ tri,
                                                      const double2* __restrict__ xy,
                                                      int32_t* __restrict__ count) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int bad = 0;
  if (t < nt && tri[3 * t] >= 0) {
    const double2 p0 = xy[tri[3 * t]], p1 = xy[tri[3 * t + 1]], p2 = xy[tri[3 * t + 2]];
    const double a = 0.5 * ((p1.x - p0.x) * (p2.y - p0.y) - (p2.x - p0.x) * (p1.y - p0.y));
    bad = !(a > 0.);
  }
  const uint64_t mask = __ballot(bad);
  if ((threadIdx.x & 63) == 0 && mask) atomicAdd(count, (int32_t)__popcll(mask));
}

}  // namespace trn

using namespace trn;

struct trn_adjoint {
  trn_ctx* ctx = nullptr;
  const trn_mesh* mesh = nullptr;
  int32_t nv = 0, nt = 0;
  // tape
  double* G = nullptr;
  double* TA = nullptr;
  double* Hs = nullptr;
  double* Mloc = nullptr;
  double* M = nullptr;
  double4* logm = nullptr;
  double* area_t = nullptr;
  double* qualit = nullptr;
  double* scal = nullptr;  // C, fact, C adjoint, E, largest gradient entry
  // backward
  double* Mlb = nullptr;
  double* Sb = nullptr;
  double* TAb = nullptr;
  double2* Sgb = nullptr;
  double* QB = nullptr;
  double* GB = nullptr;
  double* AD = nullptr;
  double* FB = nullptr;
  double* partial = nullptr;
  int32_t npartial = 0;
  int32_t* err = nullptr;
  // per triangle, the ascending union of its corners' closed 1-rings (the
  // vertices whose 2-ring patch holds it), sliced ELL over 64-triangle
  // slabs like the vertex patches: entry e of triangle t = b * kAdjB + l is
  // tr_ell[(tr_row[b] + e) * kAdjB + l], e < tr_n[t], -1 padding
  int32_t* tr_n = nullptr;
  int32_t* tr_row = nullptr;
  int32_t* tr_ell = nullptr;
  // smoothing
  double* gxy = nullptr;
  double* xy_try = nullptr;
  // last forward
  const double* xy = nullptr;
  const double* f = nullptr;
  double tf = 0.8, hmin = 1e-9, hmax = 0.4;
  int32_t p = 2;
  double C = 0., fact = 0., E = 0.;
  bool have_tape = false;
};

namespace trn {

template <typename T>
static void adj_alloc(T** p, size_t n) {
  TRN_HIP(hipMalloc((void**)p, sizeof(T) * (n > 0 ? n : 1)));
}
template <typename T>
static void adj_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

static int32_t ring_cap(const trn_mesh* m) {
  return std::min(kRingCap, 3 * (1 + std::max(m->max_deg_v, 1)));
}

// Forward for the given coordinates and values; keeps the tape.
static double adj_forward(trn_ctx* ctx, trn_adjoint* a, const double* xy, const double* f,
                          double tf, int32_t p, double hmin, double hmax) {
  const trn_mesh* m = a->mesh;
  TRN_REQUIRE(hmin > 0 && hmax > 0, TRN_ERR_ARG, "h_min and h_max must be positive");
  // the metric pass keeps the gradient, the raw Hessian, the local metric
  // and the patch areas its own kernels form
  const double C = metric_tape(ctx, m, xy, f, tf, p, hmin, hmax, a->M, a->G, a->Hs, a->Mloc,
                               a->TA);
  const double target = (double)(int)std::floor(m->nv * tf);  // metric.cpp:39
  a->C = C;
  a->fact = target / C;
  ctx->pinned_f64[0] = C;
  ctx->pinned_f64[1] = a->fact;
  copy_async(ctx, a->scal, ctx->pinned_f64, 2 * sizeof(double));
  launch(ctx, "adj_area", 0.0, k_adj_area, dim3(nblk(m->nt, 256)), dim3(256), 0, m->nt,
         (const int32_t*)m->tri, (const double2*)xy, a->area_t);
  launch(ctx, "adj_logm", 0.0, k_adj_logm, dim3(nblk(m->nv, 256)), dim3(256), 0, m->nv,
         (const double*)a->M, a->logm);
  const int eb = (int)std::min<int64_t>(nblk(m->nt, 256), a->npartial);
  launch(ctx, "adj_energy", 0.0, k_adj_energy, dim3(eb), dim3(256), 0, m->nt,
         (const int32_t*)m->tri, (const double2*)xy, (const double4*)a->logm, a->qualit,
         a->partial);
  launch(ctx, "adj_sum", 0.0, k_adj_sum, dim3(1), dim3(kSumB), 0, (const double*)a->partial, eb,
         1.0, a->scal + 3);
  copy_async(ctx, ctx->pinned_f64 + 2, a->scal + 3, sizeof(double));
  copy_async(ctx, ctx->pinned_i32, a->err, sizeof(int32_t));
  sync(ctx);
  TRN_REQUIRE(ctx->pinned_i32[0] == 0, TRN_ERR_STATE, "stencil larger than measured");
  a->E = ctx->pinned_f64[2];
  a->xy = xy;
  a->f = f;
  a->tf = tf;
  a->p = p;
  a->hmin = hmin;
  a->hmax = hmax;
  a->have_tape = true;
  return a->E;
}

// Bytes of the backward per vertex, as counted for the timing report: the
// tape and adjoint arrays each read or written once (2 + 1 + 3 + 3 + 3 per
// vertex in, 3 + 3 + 1 + 2 + 3 out) and the per-triangle buffers written
// and read once (31 doubles per triangle, about 2 triangles per vertex).
constexpr double kBytesBackwardPerV = 8.0 * (12 + 12 + 2 * 2 * 31);

static void adj_backward(trn_ctx* ctx, trn_adjoint* a, double* dxy, double* df) {
  TRN_REQUIRE(a->have_tape, TRN_ERR_STATE, "no forward to differentiate");
  const trn_mesh* m = a->mesh;
  const double2* xy = (const double2*)a->xy;
  const double emin = 1 / (a->hmin * a->hmin), emax = 1 / (a->hmax * a->hmax);
  launch(ctx, "adj_quality", 0.0, k_adj_quality, dim3(nblk(m->nt, 256)), dim3(256), 0, m->nt,
         (const int32_t*)m->tri, xy, (const double*)a->M, (const double*)a->qualit, a->QB);
  const int vb = (int)nblk(m->nv, 256);
  launch(ctx, "adj_global", 0.0, k_adj_global, dim3(vb), dim3(256), 0, m->nv,
         (const int32_t*)m->v2t_off, (const int32_t*)m->v2t, (const int32_t*)m->star,
         (const double*)a->QB, (const double*)a->Mloc, (const double*)a->scal, emin, emax,
         a->Mlb, a->partial);
  // C adjoint = (d E / d fact) * (d fact / d C) = factb * (-fact / C)
  launch(ctx, "adj_sum", 0.0, k_adj_sum, dim3(1), dim3(kSumB), 0, (const double*)a->partial, vb,
         -a->fact / a->C, a->scal + 2);
  launch(ctx, "adj_local", 0.0, k_adj_local, dim3(vb), dim3(256), 0, m->nv,
         (const int32_t*)m->v2t_off, (const int32_t*)m->v2t, (const double*)a->area_t,
         (const double*)a->Mloc, (const double*)a->Hs, (const double*)a->TA,
         (const double*)a->scal, metric_exponent(a->p), (const double*)a->Mlb, a->Sb, a->TAb);
  launch(ctx, "adj_hess_tri", 0.0, k_adj_hess_tri, dim3(nblk(m->nt, kAdjB)), dim3(kAdjB), 0,
         m->nt, (const int32_t*)m->tri, xy, (const double2*)a->G, (const double*)a->Sb,
         (const int32_t*)a->tr_n, (const int32_t*)a->tr_row, (const int32_t*)a->tr_ell, a->AD,
         a->GB);
  launch(ctx, "adj_grad_vtx", 0.0, k_adj_grad_vtx, dim3(vb), dim3(256), 0, m->nv,
         (const int32_t*)m->v2t_off, (const int32_t*)m->v2t, (const int32_t*)m->star,
         (const double*)a->GB, (const double2*)a->G, (const double*)a->TA, a->TAb, a->Sgb);
  launch(ctx, "adj_grad_tri", 0.0, k_adj_grad_tri, dim3(nblk(m->nt, kAdjB)), dim3(kAdjB), 0,
         m->nt, (const int32_t*)m->tri, xy, a->f, (const double*)a->Mloc,
         (const double2*)a->Sgb, (const double*)a->TAb, (const int32_t*)a->tr_n,
         (const int32_t*)a->tr_row, (const int32_t*)a->tr_ell, (const double*)a->scal,
         (const double*)a->AD, a->QB, a->FB);
  launch(ctx, "adj_gather", kBytesBackwardPerV * m->nv, k_adj_gather, dim3(vb), dim3(256), 0,
         m->nv, (const int32_t*)m->v2t_off, (const int32_t*)m->v2t, (const int32_t*)m->star,
         (const double*)a->QB, (const double*)a->FB, (double2*)dxy, df);
}

static void check_adj_err(trn_ctx* ctx, trn_adjoint* a) {
  copy_async(ctx, ctx->pinned_i32, a->err, sizeof(int32_t));
  sync(ctx);
  TRN_REQUIRE(ctx->pinned_i32[0] == 0, TRN_ERR_STATE, "stencil larger than measured");
}

}  // namespace trn

extern "C" {

int trn_adjoint_create(trn_ctx* ctx, const trn_mesh* mesh, trn_adjoint** out) {
  return guard_impl("trn_adjoint_create", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(ctx && mesh && out, TRN_ERR_ARG, "null argument");
    *out = nullptr;
    trn_adjoint* a = new trn_adjoint();
    a->ctx = ctx;
    a->mesh = mesh;
    a->nv = mesh->nv;
    a->nt = mesh->nt;
    try {
      TRN_REQUIRE((size_t)kAdjB * mesh->patch_cap * sizeof(int) <= 65536, TRN_ERR_CAPACITY,
                  "2-ring patch too large");
      const size_t nv = mesh->nv, nt = mesh->nt;
      adj_alloc(&a->G, 2 * nv);
      adj_alloc(&a->TA, nv);
      adj_alloc(&a->Hs, 3 * nv);
      adj_alloc(&a->Mloc, 3 * nv);
      adj_alloc(&a->M, 3 * nv);
      adj_alloc(&a->logm, nv);
      adj_alloc(&a->area_t, nt);
      adj_alloc(&a->qualit, nt);
      adj_alloc(&a->scal, 8);
      adj_alloc(&a->Mlb, 3 * nv);
      adj_alloc(&a->Sb, 3 * nv);
      adj_alloc(&a->TAb, nv);
      adj_alloc(&a->Sgb, nv);
      adj_alloc(&a->QB, kQB * nt);
      adj_alloc(&a->GB, kGB * nt);
      adj_alloc(&a->AD, kAD * nt);
      adj_alloc(&a->FB, kFB * nt);
      a->npartial = (int32_t)std::max<int64_t>(nblk(nv, 256), 1024);
      adj_alloc(&a->partial, a->npartial);
      adj_alloc(&a->err, 1);
      adj_alloc(&a->gxy, 2 * nv);
      adj_alloc(&a->xy_try, 2 * nv);
      memset_async(ctx, a->err, 0, sizeof(int32_t));
      // ring unions of the triangles, stored for the recovery transposes
      const int32_t cap = ring_cap(mesh);
      const size_t lds = sizeof(int) * kAdjB * (size_t)cap;
      const int64_t nslab = (int64_t)nblk(nt, kAdjB);
      adj_alloc(&a->tr_n, nt);
      adj_alloc(&a->tr_row, (size_t)nslab + 1);
      launch(ctx, "adj_tring_size", 0.0, k_tring_build<false>, dim3((unsigned)nslab),
             dim3(kAdjB), lds, mesh->nt, (const int32_t*)mesh->tri,
             (const int32_t*)mesh->v2v_off, (const int32_t*)mesh->v2v, cap, a->tr_n,
             (const int32_t*)nullptr, (int32_t*)nullptr, a->err);
      Scratch sc(ctx, exclusive_scan_scratch(nslab));
      int32_t* tmp = sc.carve<int32_t>(exclusive_scan_scratch(nslab) / 4);
      launch(ctx, "adj_tring_rows", 0.0, k_tring_rows, dim3((unsigned)nslab), dim3(kAdjB), 0,
             mesh->nt, (const int32_t*)a->tr_n, a->tr_row);
      exclusive_scan_i32(ctx, a->tr_row, a->tr_row, nslab, true, tmp);
      copy_async(ctx, ctx->pinned_i32, a->tr_row + nslab, sizeof(int32_t));
      copy_async(ctx, ctx->pinned_i32 + 1, a->err, sizeof(int32_t));
      sync(ctx);
      TRN_REQUIRE(ctx->pinned_i32[1] == 0, TRN_ERR_CAPACITY,
                  "a triangle's ring union exceeds " + std::to_string(cap) + " vertices");
      const int64_t rows = ctx->pinned_i32[0];
      adj_alloc(&a->tr_ell, (size_t)std::max<int64_t>(rows, 1) * kAdjB);
      launch(ctx, "adj_tring_fill", 0.0, k_tring_build<true>, dim3((unsigned)nslab),
             dim3(kAdjB), lds, mesh->nt, (const int32_t*)mesh->tri,
             (const int32_t*)mesh->v2v_off, (const int32_t*)mesh->v2v, cap, a->tr_n,
             (const int32_t*)a->tr_row, a->tr_ell, a->err);
      sync(ctx);
    } catch (...) {
      trn_adjoint_destroy(a);
      throw;
    }
    *out = a;
  });
}

int trn_adjoint_destroy(trn_adjoint* a) {
  if (!a) return TRN_OK;
  return guard_impl("trn_adjoint_destroy", [&] {
    DeviceScope ds(a->ctx);
    (void)hipStreamSynchronize(a->ctx->stream);
    adj_free(a->G); adj_free(a->TA); adj_free(a->Hs); adj_free(a->Mloc); adj_free(a->M);
    adj_free(a->logm); adj_free(a->area_t); adj_free(a->qualit); adj_free(a->scal);
    adj_free(a->Mlb); adj_free(a->Sb); adj_free(a->TAb); adj_free(a->Sgb); adj_free(a->QB);
    adj_free(a->GB); adj_free(a->AD); adj_free(a->FB); adj_free(a->partial); adj_free(a->err);
    adj_free(a->gxy); adj_free(a->xy_try);
    adj_free(a->tr_n); adj_free(a->tr_row); adj_free(a->tr_ell);
    delete a;
  });
}

int trn_adjoint_forward(trn_ctx* ctx, trn_adjoint* a, const double* xy_d, const double* f_d,
                        double target_factor, int32_t lp_norm, double h_min, double h_max,
                        double* energy, double* tensor_d, double* qualit_d) {
  return guard_impl("trn_adjoint_forward", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(a && xy_d && f_d, TRN_ERR_ARG, "null argument");
    TRN_REQUIRE(h_min > 0 && h_max > 0, TRN_ERR_ARG, "h_min and h_max must be positive");
    const double E = adj_forward(ctx, a, xy_d, f_d, target_factor, lp_norm, h_min, h_max);
    if (energy) *energy = E;
    if (tensor_d) copy_async(ctx, tensor_d, a->M, sizeof(double) * 3 * (size_t)a->nv);
    if (qualit_d) copy_async(ctx, qualit_d, a->qualit, sizeof(double) * (size_t)a->nt);
  });
}

int trn_adjoint_backward(trn_ctx* ctx, trn_adjoint* a, double* dxy_d, double* df_d) {
  return guard_impl("trn_adjoint_backward", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(a != nullptr, TRN_ERR_ARG, "null adjoint");
    adj_backward(ctx, a, dxy_d, df_d);
    check_adj_err(ctx, a);
  });
}

int trn_adjoint_smooth(trn_ctx* ctx, trn_adjoint* a, double* xy_d, const double* f_d,
                       double target_factor, int32_t lp_norm, double h_min, double h_max,
                       int32_t iters, double step, double* energies, int32_t* accepted) {
  return guard_impl("trn_adjoint_smooth", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(a && xy_d && f_d, TRN_ERR_ARG, "null argument");
    TRN_REQUIRE(iters >= 0 && step > 0, TRN_ERR_ARG, "bad iteration count or step");
    const trn_mesh* m = a->mesh;
    double E = adj_forward(ctx, a, xy_d, f_d, target_factor, lp_norm, h_min, h_max);
    if (energies) energies[0] = E;
    int32_t acc = 0;
    const int vb = (int)nblk(m->nv, 256);
    for (int it = 0; it < iters; ++it) {
      adj_backward(ctx, a, a->gxy, nullptr);
      launch(ctx, "adj_precond", 0.0, k_adj_precond, dim3(vb), dim3(256), 0, m->nv,
             (const uint8_t*)m->tags, (const int32_t*)m->v2t_off, (const int32_t*)m->v2v_off,
             (const int32_t*)m->v2v, (const double2*)xy_d, (double2*)a->gxy);
      launch(ctx, "adj_gmax", 0.0, k_adj_gmax, dim3(vb), dim3(256), 0, m->nv,
             (const uint8_t*)m->tags, (const int32_t*)m->v2t_off, (const double2*)a->gxy,
             a->partial);
      launch(ctx, "adj_max", 0.0, k_adj_max, dim3(1), dim3(64), 0, (const double*)a->partial, vb,
             a->scal + 4);
      copy_async(ctx, ctx->pinned_f64, a->scal + 4, sizeof(double));
      check_adj_err(ctx, a);
      const double gmax = ctx->pinned_f64[0];
      bool moved = false;
      if (gmax > 0) {
        // largest vertex displacement `step`, halved until E decreases
        // without inverting a triangle
        double alpha = step / gmax;
        for (int tries = 0; tries < 12 && !moved; ++tries, alpha *= 0.5) {
          launch(ctx, "adj_step", 0.0, k_adj_step, dim3(vb), dim3(256), 0, m->nv,
                 (const uint8_t*)m->tags, (const int32_t*)m->v2t_off, (const double2*)xy_d,
                 (const double2*)a->gxy, alpha, (double2*)a->xy_try);
          memset_async(ctx, a->err, 0, sizeof(int32_t));
          launch(ctx, "adj_inverted", 0.0, k_adj_inverted, dim3(nblk(m->nt, 256)), dim3(256), 0,
                 m->nt, (const int32_t*)m->tri, (const double2*)a->xy_try, a->err);
          copy_async(ctx, ctx->pinned_i32, a->err, sizeof(int32_t));
          sync(ctx);
          const int inverted = ctx->pinned_i32[0];
          memset_async(ctx, a->err, 0, sizeof(int32_t));
          if (inverted) continue;
          const double Et =
              adj_forward(ctx, a, a->xy_try, f_d, target_factor, lp_norm, h_min, h_max);
          if (Et < E) {
            copy_async(ctx, xy_d, a->xy_try, sizeof(double) * 2 * (size_t)m->nv);
            E = Et;
            moved = true;
            ++acc;
          }
        }
      }
      // the tape must describe xy_d for the next backward: after a step it
      // was computed from the same values; after a failed search it holds
      // the last rejected trial
      if (moved) a->xy = xy_d;
      else E = adj_forward(ctx, a, xy_d, f_d, target_factor, lp_norm, h_min, h_max);
      if (energies) energies[it + 1] = E;
      if (!moved) {
        for (int k = it + 2; energies && k <= iters; ++k) energies[k] = E;
        break;
      }
    }
    if (accepted) *accepted = acc;
  });
}

}  // extern "C"
