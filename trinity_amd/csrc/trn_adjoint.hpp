// trn_adjoint.hpp -- reverse-mode building blocks of the metric-space quality
// energy E = sum_t (1 - q_M(t))^2 (the reference has no energy; it is
// defined on numeric::computeQuality, src/numeric.cpp:251-306, and the
// metric of Metrics::run, src/metric.cpp:59-70).
//
// Every map is differentiated as the reference computes it, branch by
// branch: the eigen solver's diagonal shortcut when the off-diagonal
// product is below EPSILON (numeric.cpp:30-38) ignores the off-diagonal
// entry, its non-normal fallback (:52-64) copies the larger eigenvalue, and
// the clamps (metric.cpp:120-121, :213-216) have zero slope where they bind.
// Functions of a symmetric 2x2 matrix, A -> P diag(mu(lambda)) P^T, are
// differentiated with the divided differences of mu (Daleckii-Krein).
//
// Stored symmetric tensors are (m0, m1, m2) = [[m0, m1], [m1, m2]]; their
// adjoints are (dE/dm0, dE/dm1, dE/dm2), i.e. the matrix adjoint
// [[a0, a1/2], [a1/2, a2]].
#pragma once

#include "trn_numeric.hpp"

namespace trn {

struct Eig2 {
  double val[2];
  double P[4];  // rows [vec1; vec2]; column k is the eigenvector of val[k]
  int branch;   // 0 regular, 1 diagonal shortcut, 2 non-normal fallback
};

__device__ __forceinline__ void eig2(const double m0, const double m1, const double m2,
                                     const double m3, Eig2& e) {
  eigen(m0, m1, m2, m3, e.val, e.P);
  if (m1 * m2 < kEps) {
    e.branch = 1;
  } else {
    const double trace = m0 + m3;
    const double det = m0 * m3 - m1 * m2;
    const double term = trace * 0.5;
    const double sd = __builtin_sqrt(term * term - det);
    e.branch = (!__builtin_isnormal(term - sd) || !__builtin_isnormal(term + sd)) ? 2 : 0;
  }
}

// Reverse of X = P diag(mu) P^T: bx (matrix adjoint of X, 4 entries) ->
// ba (matrix adjoint of A).  jac = d mu_k / d lambda_l (row-major), gam =
// (mu_0 - mu_1) / (lambda_0 - lambda_1).
__device__ __forceinline__ void spectral_adjoint(const Eig2& e, const double jac[4],
                                                 const double gam, const double bx[4],
                                                 double ba[4]) {
  const double c00 = e.P[0], c01 = e.P[2], c10 = e.P[1], c11 = e.P[3];
  const double t00 = bx[0] * c00 + bx[1] * c01, t01 = bx[2] * c00 + bx[3] * c01;
  const double t10 = bx[0] * c10 + bx[1] * c11, t11 = bx[2] * c10 + bx[3] * c11;
  const double b00 = c00 * t00 + c01 * t01;
  const double b01 = c00 * t10 + c01 * t11;
  const double b10 = c10 * t00 + c11 * t01;
  const double b11 = c10 * t10 + c11 * t11;
  double lb0 = b00 * jac[0] + b11 * jac[2];
  double lb1 = b00 * jac[1] + b11 * jac[3];
  double off = 0.5 * (b01 + b10) * gam;
  if (e.branch == 1) off = 0.;
  if (e.branch == 2) {
    lb1 += lb0;
    lb0 = 0.;
    off = 0.;
  }
  const double c0[2] = {c00, c01}, c1[2] = {c10, c11};
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int s = 0; s < 2; ++s)
      ba[2 * r + s] = lb0 * c0[r] * c0[s] + lb1 * c1[r] * c1[s] + off * (c0[r] * c1[s] + c1[r] * c0[s]);
}

// sum_k B_kk * dmu_k/dfact for B = P^T bx P, as used by the global scaling
__device__ __forceinline__ void eig_diag(const Eig2& e, const double bx[4], double& b00,
                                         double& b11) {
  const double c00 = e.P[0], c01 = e.P[2], c10 = e.P[1], c11 = e.P[3];
  b00 = c00 * (bx[0] * c00 + bx[1] * c01) + c01 * (bx[2] * c00 + bx[3] * c01);
  b11 = c10 * (bx[0] * c10 + bx[1] * c11) + c11 * (bx[2] * c10 + bx[3] * c11);
}

__device__ __forceinline__ void sym_to_mat(const double a0, const double a1, const double a2,
                                           double m[4]) {
  m[0] = a0;
  m[1] = 0.5 * a1;
  m[2] = 0.5 * a1;
  m[3] = a2;
}

// divided differences of log and exp, stable for close arguments; the CPU
// reference adjoint evaluates the same expressions with the same functions
__device__ __forceinline__ double dd_log(const double x, const double y) {
  if (x == y) return 1. / x;
  const double r = (x - y) / y;
  if (fabs(r) < 1e-4) return libm::log1p_small(r) / (x - y);
  return (libm::log(x) - libm::log(y)) / (x - y);
}
__device__ __forceinline__ double dd_exp(const double x, const double y) {
  if (x == y) return libm::exp(x);
  return libm::exp(y) * libm::expm1(x - y) / (x - y);
}

// Metrics::normalizeLocally (metric.cpp:113-148) as a spectral map.
__device__ __forceinline__ void local_jac(const Eig2& e, const double ex, double jac[4],
                                          double& gam) {
  double c[2], dc[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double a = fabs(e.val[k]);
    c[k] = dmax(a, kEps);
    dc[k] = (a < kEps) ? 0. : (e.val[k] < 0 ? -1. : 1.);
  }
  const double s = libm::pow(c[0] * c[1], ex);
  const double ds0 = ex * s / c[0] * dc[0], ds1 = ex * s / c[1] * dc[1];
  jac[0] = dc[0] * s + c[0] * ds0;
  jac[1] = c[0] * ds1;
  jac[2] = c[1] * ds0;
  jac[3] = dc[1] * s + c[1] * ds1;
  if (dc[0] != 0. && dc[0] == dc[1]) gam = dc[0] * s;
  else if (e.val[0] == e.val[1]) gam = 0.;
  else gam = s * (c[0] - c[1]) / (e.val[0] - e.val[1]);
}

// Metrics::normalizeGlobally (metric.cpp:204-231) as a spectral map; dfact
// = d mu_k / d fact.
__device__ __forceinline__ void global_jac(const Eig2& e, const double fact, const double emin,
                                           const double emax, double jac[4], double& gam,
                                           double dfact[2]) {
  double mu[2];
  bool fr[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double v = e.val[k] * fact;
    mu[k] = dmax(dmin(v, emin), emax);
    fr[k] = (v < emin) && (v > emax);
    dfact[k] = fr[k] ? e.val[k] : 0.;
  }
  jac[0] = fr[0] ? fact : 0.;
  jac[1] = 0.;
  jac[2] = 0.;
  jac[3] = fr[1] ? fact : 0.;
  if (fr[0] && fr[1]) gam = fact;
  else if (mu[0] == mu[1] || e.val[0] == e.val[1]) gam = 0.;
  else gam = (mu[0] - mu[1]) / (e.val[0] - e.val[1]);
}

// Per-triangle geometry of src/hessian.cpp:42-61 (area and the gradients of
// the linear basis functions).
__device__ __forceinline__ void tri_basis(const double2 p0, const double2 p1, const double2 p2,
                                          double& a, double d[6]) {
  tri_geometry(p0, p1, p2, a, d);
}

// Reverse of the geometry: area adjoint ab and basis-gradient adjoints
// db[6] -> coordinate adjoints g[6] (x0 y0 x1 y1 x2 y2), added.
__device__ __forceinline__ void tri_basis_adjoint(const double2 p0, const double2 p1,
                                                  const double2 p2, double ab, const double db[6],
                                                  double g[6]) {
  const double x0 = p0.x, y0 = p0.y, x1 = p1.x, y1 = p1.y, x2 = p2.x, y2 = p2.y;
  const double c0 = x1 - x0, c1 = x2 - x0, c2 = y1 - y0, c3 = y2 - y0;
  const double a = 0.5 * (c0 * c3 - c1 * c2);
  const double nrm = 2 / a;
  const double r[6] = {y1 - y2, x2 - x1, y2 - y0, x0 - x2, y0 - y1, x1 - x0};
  double nb = 0.;
  double rb[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    nb += db[k] * r[k];
    rb[k] = nrm * db[k];
  }
  ab += nb * (-2 / (a * a));
  const double cb0 = 0.5 * c3 * ab, cb3 = 0.5 * c0 * ab;
  const double cb1 = -0.5 * c2 * ab, cb2 = -0.5 * c1 * ab;
  g[3] += rb[0]; g[5] -= rb[0];
  g[4] += rb[1]; g[2] -= rb[1];
  g[5] += rb[2]; g[1] -= rb[2];
  g[0] += rb[3]; g[4] -= rb[3];
  g[1] += rb[4]; g[3] -= rb[4];
  g[2] += rb[5]; g[0] -= rb[5];
  g[2] += cb0; g[0] -= cb0;
  g[4] += cb1; g[0] -= cb1;
  g[3] += cb2; g[1] -= cb2;
  g[5] += cb3; g[1] -= cb3;
}

// Reverse of numeric::computeQuality (numeric.cpp:251-306) for one
// triangle: dq -> coordinate adjoints g[6] and stored-tensor adjoints
// gm[9] (corner order), added.
__device__ __forceinline__ void quality_adjoint(const double2 pa, const double2 pb,
                                                const double2 pc, const double* ta,
                                                const double* tb, const double* tc,
                                                const double dq, double g[6], double gm[9]) {
  const double* tk[3] = {ta, tb, tc};
  Eig2 ek[3];
  double lsum[4] = {0., 0., 0., 0.};
#pragma unroll
  for (int k = 0; k < 3; ++k) {  // numeric.cpp:113-172
    eig2(tk[k][0], tk[k][1], tk[k][1], tk[k][2], ek[k]);
    const double v[2] = {libm::log(ek[k].val[0]), libm::log(ek[k].val[1])};
    double c[4];
    rebuild(v, ek[k].P, c);
    lsum[0] += c[0]; lsum[1] += c[1]; lsum[2] += c[2]; lsum[3] += c[3];
  }
  const double lm[4] = {lsum[0] / 3, lsum[1] / 3, lsum[2] / 3, lsum[3] / 3};
  Eig2 el;
  eig2(lm[0], lm[1], lm[2], lm[3], el);
  const double ev[2] = {libm::exp(el.val[0]), libm::exp(el.val[1])};
  double jm[4];
  rebuild(ev, el.P, jm);
  const double J0 = jm[0], J1 = jm[1], J2 = jm[3];
  const double det = J0 * J2 - J1 * J1;
  const double e[3][2] = {{pb.x - pa.x, pb.y - pa.y},
                          {pc.x - pb.x, pc.y - pb.y},
                          {pa.x - pc.x, pa.y - pc.y}};
  double len[3], perim = 0.;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double vx = J0 * e[k][0] + J1 * e[k][1];
    const double vy = J1 * e[k][0] + J2 * e[k][1];
    len[k] = __builtin_sqrt(e[k][0] * vx + e[k][1] * vy);
    perim += len[k];
  }
  const double ux = pb.x - pa.x, uy = pb.y - pa.y, wx = pc.x - pa.x, wy = pc.y - pa.y;
  const double cr = ux * wy - wx * uy;
  const double area = 0.5 * fabs(cr);
  const double mean = dmin(perim / 3, 3 / perim);
  const double base = mean * (2. - mean);
  const double norm = (12 * 1.7320508075688772) * (base * base * base);
  const double sdet = __builtin_sqrt(det);
  const double p2 = perim * perim;
  const double q = area * norm * sdet / p2;
  // reverse
  const double area_b = dq * norm * sdet / p2;
  const double norm_b = dq * area * sdet / p2;
  const double det_b = dq * area * norm / (2 * sdet * p2);
  const double dmean = (perim / 3 <= 3 / perim) ? 1. / 3 : -3 / p2;
  const double dnorm = (12 * 1.7320508075688772) * 3 * base * base * (2 - 2 * mean);
  const double perim_b = dq * (-2 * q / perim) + norm_b * dnorm * dmean;
  double jb0 = det_b * J2, jb1 = -2 * det_b * J1, jb2 = det_b * J0;
  double eb[3][2];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double vx = J0 * e[k][0] + J1 * e[k][1];
    const double vy = J1 * e[k][0] + J2 * e[k][1];
    eb[k][0] = perim_b * vx / len[k];
    eb[k][1] = perim_b * vy / len[k];
    const double w = perim_b / (2 * len[k]);
    jb0 += w * e[k][0] * e[k][0];
    jb1 += w * 2 * e[k][0] * e[k][1];
    jb2 += w * e[k][1] * e[k][1];
  }
  g[0] += eb[2][0] - eb[0][0]; g[1] += eb[2][1] - eb[0][1];
  g[2] += eb[0][0] - eb[1][0]; g[3] += eb[0][1] - eb[1][1];
  g[4] += eb[1][0] - eb[2][0]; g[5] += eb[1][1] - eb[2][1];
  {
    const double crb = (cr < 0 ? -0.5 : 0.5) * area_b;
    const double uxb = crb * wy, wyb = crb * ux, wxb = -crb * uy, uyb = -crb * wx;
    g[2] += uxb; g[0] -= uxb;
    g[3] += uyb; g[1] -= uyb;
    g[4] += wxb; g[0] -= wxb;
    g[5] += wyb; g[1] -= wyb;
  }
  double jbm[4], lmb[4];
  sym_to_mat(jb0, jb1, jb2, jbm);
  {
    const double jac[4] = {ev[0], 0., 0., ev[1]};
    spectral_adjoint(el, jac, dd_exp(el.val[0], el.val[1]), jbm, lmb);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double lb[4] = {lmb[0] / 3, lmb[1] / 3, lmb[2] / 3, lmb[3] / 3};
    const double jac[4] = {1. / ek[k].val[0], 0., 0., 1. / ek[k].val[1]};
    double mb[4];
    spectral_adjoint(ek[k], jac, dd_log(ek[k].val[0], ek[k].val[1]), lb, mb);
    gm[3 * k] += mb[0];
    gm[3 * k + 1] += mb[1] + mb[2];
    gm[3 * k + 2] += mb[3];
  }
}

}  // namespace trn
