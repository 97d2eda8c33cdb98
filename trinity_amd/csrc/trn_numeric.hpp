// trn_numeric.hpp -- FP64 tensor numerics of the metric-and-smoothing path,
// as __device__ functions for gfx950.
//
// Every routine reproduces the operation order of the reference routine it
// cites, so that, with -ffp-contract=off, results are bit-identical: sqrt
// and division are IEEE-exact on both sides, exp and pow are restatements of
// the host C library's (trn_libm.hpp), and log is correctly rounded, which
// the host's is on every metric eigenvalue and all but about 2e-3 of the
// edge-length ratios near 1 that approxRiemannDist takes it of.
//
// std::min(a,b) == (b < a ? b : a) and std::max(a,b) == (a < b ? b : a) are
// restated literally so that NaN and signed-zero behaviour match.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "trn_libm.hpp"

namespace trn {

constexpr double kEps = 1.e-9;   // include/trinity/header.h:50

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

// numeric::eigenDecompose, src/numeric.cpp:25-95.  vec[0..1] is vec1 and
// vec[2..3] is vec2 (rows of the eigenvector matrix; column k belongs to
// val[k]).
__device__ __forceinline__ void eigen(const double m0, const double m1,
                                      const double m2, const double m3,
                                      double val[2], double vec[4]) {
  if (m1 * m2 < kEps) {                    // numeric.cpp:30-38 (absolute)
    val[0] = m3;
    val[1] = m0;
    vec[0] = 0; vec[2] = 1;
    vec[1] = 1; vec[3] = 0;
    return;
  }
  const double trace = m0 + m3;            // numeric.cpp:44-50
  const double det = m0 * m3 - m1 * m2;
  double term = trace * 0.5;
  const double sqrt_delta = __builtin_sqrt(term * term - det);
  val[0] = term - sqrt_delta;
  val[1] = term + sqrt_delta;
  if (!__builtin_isnormal(val[0]) || !__builtin_isnormal(val[1]))
    val[0] = val[1];                       // numeric.cpp:52-64
  const double dd = m0 - m3;               // numeric.cpp:69-71, pow(.,2)=x*x
  term = __builtin_sqrt(dmax(0.25 * (dd * dd) + m1 * m2, 0.));
  if (m0 - m3 >= 0) {                      // numeric.cpp:72-82
    vec[0] = +0.5 * (m0 - m3) - term;
    vec[1] = m1;
    vec[2] = m2;
    vec[3] = -0.5 * (m0 - m3) + term;
  } else {
    vec[0] = m2;
    vec[1] = +0.5 * (m0 - m3) + term;
    vec[2] = -0.5 * (m0 - m3) - term;
    vec[3] = m1;
  }
  const double n0 = __builtin_sqrt(vec[0] * vec[0] + vec[1] * vec[1]);
  const double n1 = __builtin_sqrt(vec[2] * vec[2] + vec[3] * vec[3]);
  vec[0] /= n0; vec[2] /= n1;              // numeric.cpp:91-94
  vec[1] /= n0; vec[3] /= n1;
}

// P.diag(val).P^T with the products of metric.cpp:134-142 (also
// metric.cpp:222-231 and numeric.cpp:146-154, :203-211).
__device__ __forceinline__ void rebuild(const double val[2], const double vec[4],
                                        double h[4]) {
  const double c0 = val[0] * vec[0];
  const double c1 = val[0] * vec[2];
  const double c2 = val[1] * vec[1];
  const double c3 = val[1] * vec[3];
  h[0] = vec[0] * c0 + vec[1] * c2;
  h[1] = vec[0] * c1 + vec[1] * c3;
  h[2] = vec[2] * c0 + vec[3] * c2;
  h[3] = vec[2] * c1 + vec[3] * c3;
}

// Matrix logarithm term of numeric::interpolateTensor for one tensor
// (numeric.cpp:115-154): eigen, log of both eigenvalues, rebuild.  The four
// entries are kept separately because cache[1] and cache[2] round
// differently and both feed the later eigen decomposition.
__device__ __forceinline__ void tensor_log(const double t0, const double t1,
                                           const double t2, double out[4]) {
  double val[2], vec[4];
  eigen(t0, t1, t1, t2, val, vec);
  val[0] = libm::log(val[0]);
  val[1] = libm::log(val[1]);
  rebuild(val, vec, out);
}

// Exponential step of numeric::interpolateTensor (numeric.cpp:192-216):
// divide the accumulated log by nb, eigen, exp, rebuild.
__device__ __forceinline__ void tensor_exp_mean(double lm[4], const int nb,
                                                double r[3]) {
  lm[0] /= nb; lm[1] /= nb; lm[2] /= nb; lm[3] /= nb;
  double val[2], vec[4], c[4];
  eigen(lm[0], lm[1], lm[2], lm[3], val, vec);
  val[0] = libm::exp(val[0]);
  val[1] = libm::exp(val[1]);
  rebuild(val, vec, c);
  r[0] = c[0];
  r[1] = c[1];
  r[2] = c[3];
}

// numeric::interpolateTensor of three tensors (numeric.cpp:98-248).
__device__ __forceinline__ void interpolate3(const double* ta, const double* tb,
                                             const double* tc, double r[3]) {
  double lm[4] = {0., 0., 0., 0.}, c[4];
  tensor_log(ta[0], ta[1], ta[2], c);
  lm[0] += c[0]; lm[1] += c[1]; lm[2] += c[2]; lm[3] += c[3];
  tensor_log(tb[0], tb[1], tb[2], c);
  lm[0] += c[0]; lm[1] += c[1]; lm[2] += c[2]; lm[3] += c[3];
  tensor_log(tc[0], tc[1], tc[2], c);
  lm[0] += c[0]; lm[1] += c[1]; lm[2] += c[2]; lm[3] += c[3];
  tensor_exp_mean(lm, 3, r);
}

// pow(x, 3) as numeric.cpp:304 evaluates it: the C library's pow (which is
// not x*x*x, nor always the correctly rounded cube).
__device__ __forceinline__ double cube_pow(const double x) { return libm::cube(x); }

// x^3 rounded once: x^2 = p + e exactly (FMA error term), p*x = q + f
// exactly, x^3 = q + f + e*x with e*x below 2^-52 of q, so the final sum
// rounds correctly except in cases within ~2^-100 of a tie.  It differs
// from cube_pow by at most one ulp (glibc's pow is within 0.52 ulp), in
// about 7 of 10^4 arguments; the smoothing tries use it to decide, and
// fall back to cube_pow where that could change a decision.
__device__ __forceinline__ double cube_rn(const double x) {
  const double p = x * x;
  const double e = __builtin_fma(x, x, -p);
  const double q = p * x;
  if (!__builtin_isfinite(q)) return q;
  const double f = __builtin_fma(p, x, -q);
  return q + (f + e * x);
}

// min(perim / 3, 3 / perim) as numeric.cpp:302 takes it, with one division:
// rounding is monotonic, so for perim < 3 the rounded perim / 3 is at most
// 1 and the rounded 3 / perim at least 1 (the reverse for perim > 3), and
// equal roundings are the same value.  0, inf and NaN give the same result
// as the two-division form.
__device__ __forceinline__ double perim_mean(const double perim) {
  return perim < 3 ? perim / 3 : 3 / perim;
}

// sqrt(det j) of an averaged tensor, the factor numeric.cpp:305 applies.
__device__ __forceinline__ double mean_sqrt_det(const double j[3]) {
  return __builtin_sqrt(j[0] * j[2] - j[1] * j[1]);
}

// Metric length of the edge vector (c0, c1) under the averaged tensor j
// (numeric.cpp:286-292).
__device__ __forceinline__ double edge_len(const double j[3], const double c0,
                                           const double c1) {
  const double v0 = j[0] * c0 + j[1] * c1;
  const double v1 = j[1] * c0 + j[2] * c1;
  return __builtin_sqrt(c0 * v0 + c1 * v1);
}

// The last steps of numeric::computeQuality (numeric.cpp:303-305) from the
// area, the normalisation factor, sqrt(det j) and the perimeter.
__device__ __forceinline__ double quality_finish(const double area, const double norm,
                                                 const double sdet, const double perim) {
  return area * norm * sdet / (perim * perim);
}

// Quality from the area, the perimeter and sqrt(det j) (numeric.cpp:296-305).
// EXACT: the reference's value (its pow).  Otherwise the cube is the
// correctly rounded one, which differs from the reference's pow(x, 3) by at
// most one ulp (in about 7 of 10^4 arguments), and the quality is within 5
// ulps of the reference's (one ulp from the cube, one more rounding in each
// of the four products and quotients after it); *near is set when that
// quality lies within 2^-48 (16 ulps) of q_min, the only case where it and
// the reference's could fall on different sides of q_min in the comparison
// of smoothing.cpp:109.
template <bool EXACT = true>
__device__ __forceinline__ double quality_tail(const double* pa, const double* pb,
                                               const double* pc, const double perim,
                                               const double sdet, const double q_min = 0,
                                               bool* near = nullptr) {
  const double a0 = pb[0] - pa[0];
  const double a1 = pc[0] - pa[0];
  const double a2 = pb[1] - pa[1];
  const double a3 = pc[1] - pa[1];
  const double area = 0.5 * fabs(a0 * a3 - a1 * a2);
  const double mean = perim_mean(perim);
  const double m2 = mean * (2. - mean);
  const double norm = (12 * 1.7320508075688772) * (EXACT ? cube_pow(m2) : cube_rn(m2));
  const double q = quality_finish(area, norm, sdet, perim);
  if (!EXACT && near) *near = fabs(q - q_min) <= 0x1p-48 * q_min;
  return q;
}

// Metric-space quality from the averaged tensor j and its sqrt(det j)
// (numeric.cpp:275-305).
template <bool EXACT = true>
__device__ __forceinline__ double quality_from_mean(const double* pa,
                                                    const double* pb,
                                                    const double* pc,
                                                    const double j[3],
                                                    const double sdet,
                                                    const double q_min = 0,
                                                    bool* near = nullptr) {
  double perim = 0;
  perim += edge_len(j, pb[0] - pa[0], pb[1] - pa[1]);
  perim += edge_len(j, pc[0] - pb[0], pc[1] - pb[1]);
  perim += edge_len(j, pa[0] - pc[0], pa[1] - pc[1]);
  return quality_tail<EXACT>(pa, pb, pc, perim, sdet, q_min, near);
}

// quality_from_mean of the triangle whose corner r is the point p and whose
// corners r+1 and r+2 are a and b, with the length lab of the edge b - a
// (the edge that does not move with p) given.  Whatever r, the reference
// forms the two other edge vectors as a - p and p - b, and adds the three
// lengths in corner order (edge k joins corners k and k+1), which this does.
template <bool EXACT = true>
__device__ __forceinline__ double quality_with_corner(const double pp[3][2], const int r,
                                                      const double px, const double py,
                                                      const double2 a, const double2 b,
                                                      const double j[3], const double sdet,
                                                      const double lab, const double q_min = 0,
                                                      bool* near = nullptr) {
  const double la = edge_len(j, a.x - px, a.y - py);  // edge r
  const double lb = edge_len(j, px - b.x, py - b.y);  // edge r+2
  double perim = 0;
  perim += r == 0 ? la : (r == 1 ? lb : lab);
  perim += r == 0 ? lab : (r == 1 ? la : lb);
  perim += r == 0 ? lb : (r == 1 ? lab : la);
  return quality_tail<EXACT>(pp[0], pp[1], pp[2], perim, sdet, q_min, near);
}

template <bool EXACT = true>
__device__ __forceinline__ double quality_of_mean(const double* pa, const double* pb,
                                                  const double* pc, const double j[3],
                                                  const double q_min = 0,
                                                  bool* near = nullptr) {
  return quality_from_mean<EXACT>(pa, pb, pc, j, mean_sqrt_det(j), q_min, near);
}

// Log-Euclidean mean of three tensors from their cached matrix logarithms
// (numeric.cpp:108-216 with the per-tensor logs precomputed): the log
// matrix starts at 0 and accumulates the three terms in vertex order.
__device__ __forceinline__ void mean_from_logs(const double4 la, const double4 lb,
                                               const double4 lc, double j[3]) {
  double lm[4] = {0., 0., 0., 0.};
  lm[0] += la.x; lm[1] += la.y; lm[2] += la.z; lm[3] += la.w;
  lm[0] += lb.x; lm[1] += lb.y; lm[2] += lb.z; lm[3] += lb.w;
  lm[0] += lc.x; lm[1] += lc.y; lm[2] += lc.z; lm[3] += lc.w;
  tensor_exp_mean(lm, 3, j);
}

__device__ __forceinline__ double4 tensor_log4(const double t0, const double t1,
                                               const double t2) {
  double c[4];
  tensor_log(t0, t1, t2, c);
  double4 r;
  r.x = c[0];
  r.y = c[1];
  r.z = c[2];
  r.w = c[3];
  return r;
}

// numeric::computeQuality, src/numeric.cpp:251-306.
__device__ __forceinline__ double quality(const double* pa, const double* pb,
                                          const double* pc, const double* ta,
                                          const double* tb, const double* tc) {
  double j[3];
  interpolate3(ta, tb, tc, j);
  return quality_of_mean(pa, pb, pc, j);
}

// numeric::approxRiemannDist (4-argument form), src/numeric.cpp:469-497.
__device__ __forceinline__ double riemann_dist(const double* pa, const double* pb,
                                               const double* ta, const double* tb) {
  const double v0 = pb[0] - pa[0];
  const double v1 = pb[1] - pa[1];
  const double m0 = ta[0] * v0 + ta[1] * v1;
  const double m1 = ta[1] * v0 + ta[2] * v1;
  const double m2 = tb[0] * v0 + tb[1] * v1;
  const double m3 = tb[1] * v0 + tb[2] * v1;
  const double l0 = __builtin_sqrt(v0 * m0 + v1 * m1);
  const double l1 = __builtin_sqrt(v0 * m2 + v1 * m3);
  if (fabs(l0 - l1) < kEps) return l0;
  const double lmin = dmin(l0, l1);
  const double lmax = dmax(l0, l1);
  const double ratio = lmax / lmin;
  return lmax * (ratio - 1) / (ratio * libm::log(ratio));
}

// Mesh::isCounterclockwise, src/mesh.cpp:700-713 (absolute threshold).
__device__ __forceinline__ bool ccw(const double* pa, const double* pb,
                                    const double* pc) {
  const double c0 = pb[0] - pa[0];
  const double c1 = pc[0] - pa[0];
  const double c2 = pb[1] - pa[1];
  const double c3 = pc[1] - pa[1];
  return 0.5 * (c0 * c3 - c1 * c2) > kEps;
}

// Per-triangle geometry of the L2 projection, src/hessian.cpp:42-61 (the
// same code appears at :108-128): area and the six basis-function
// derivatives with norm = 2/area.
__device__ __forceinline__ void tri_geometry(const double2 p0, const double2 p1,
                                             const double2 p2, double& area,
                                             double d[6]) {
  const double c0 = p1.x - p0.x;
  const double c1 = p2.x - p0.x;
  const double c2 = p1.y - p0.y;
  const double c3 = p2.y - p0.y;
  area = 0.5 * (c0 * c3 - c1 * c2);
  const double norm = 2 / area;
  d[0] = norm * (p1.y - p2.y);
  d[1] = norm * (p2.x - p1.x);
  d[2] = norm * (p2.y - p0.y);
  d[3] = norm * (p0.x - p2.x);
  d[4] = norm * (p0.y - p1.y);
  d[5] = norm * (p1.x - p0.x);
}

}  // namespace trn
