/* trinity_adjoint.c -- TEST INFRASTRUCTURE ONLY: the FP64 CPU reference of
 * the metric-space quality energy and of its reverse-mode gradient.
 *
 *   E(xy, f) = sum over active triangles t of (1 - q_M(t))^2
 *
 * with q_M = numeric::computeQuality (src/numeric.cpp:251-306) evaluated
 * with the metric M of Metrics::run (src/metric.cpp:59-70) built from the
 * coordinates xy and the nodal values f: gradient recovery
 * (src/hessian.cpp:24-83), Hessian recovery (:86-153), local normalisation
 * (metric.cpp:100-150), complexity (:153-185) and global normalisation
 * (:188-238).  The reference has no energy or gradient; this file defines
 * them on top of its functions.  The gradient differentiates the function
 * the reference computes, branch by branch: the eigen solver's diagonal
 * shortcut when the off-diagonal product is below EPSILON
 * (numeric.cpp:30-38) makes the off-diagonal entry irrelevant, its
 * non-normal fallback (:52-64) duplicates the larger eigenvalue, and the
 * clamps (metric.cpp:120-121, :213-216) have zero slope where they bind.
 *
 * Symmetric tensors are stored as (m0, m1, m2) = [[m0, m1], [m1, m2]]; the
 * adjoint of a stored tensor is (dE/dm0, dE/dm1, dE/dm2), so its matrix
 * form is [[a0, a1/2], [a1/2, a2]].  Functions of a symmetric matrix,
 * A -> P diag(mu(lambda)) P^T, are differentiated with the divided
 * differences of mu (the Daleckii-Krein formula).
 *
 * Nothing in the product loads this file; tests and the benchmark's CPU
 * baseline call it through ctypes. */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "trinity_oracle.h"

static inline double dmin(double a, double b) { return (b < a) ? b : a; }
static inline double dmax(double a, double b) { return (a < b) ? b : a; }

/* ---- spectral decomposition as the reference computes it ---------------- */

/* Eigen-decomposition of m = [[m0, m1], [m2, m3]] by orc_eigen, plus which
 * branch it took: 0 regular, 1 diagonal shortcut, 2 non-normal fallback.
 * Columns of P (row-major [vec1; vec2]) are the eigenvectors: column k =
 * (P[k], P[2 + k]). */
typedef struct {
  double val[2];
  double P[4];
  int branch;
} Eig;

static void eig(const double m[4], Eig* e) {
  orc_eigen(m, e->val, e->P, e->P + 2);
  if (m[1] * m[2] < ORC_EPSILON) {
    e->branch = 1;
  } else {
    /* recompute the two roots to see whether the fallback replaced one */
    const double trace = m[0] + m[3];
    const double det = m[0] * m[3] - m[1] * m[2];
    const double term = trace * 0.5;
    const double sd = sqrt(term * term - det);
    const double v0 = term - sd, v1 = term + sd;
    e->branch = (!isnormal(v0) || !isnormal(v1)) ? 2 : 0;
  }
}

/* Reverse of X = P diag(mu) P^T at A: given the matrix adjoint of X (bx,
 * symmetric 2x2 as 4 entries) return the matrix adjoint of A (ba).  jac =
 * d mu_k / d lambda_l (row-major 2x2) and gam = the divided difference
 * (mu_0 - mu_1) / (lambda_0 - lambda_1). */
static void spectral_adjoint(const Eig* e, const double jac[4], double gam,
                             const double bx[4], double ba[4]) {
  const double c0[2] = {e->P[0], e->P[2]}, c1[2] = {e->P[1], e->P[3]};
  /* B = P^T bx P */
  double t0[2] = {bx[0] * c0[0] + bx[1] * c0[1], bx[2] * c0[0] + bx[3] * c0[1]};
  double t1[2] = {bx[0] * c1[0] + bx[1] * c1[1], bx[2] * c1[0] + bx[3] * c1[1]};
  const double b00 = c0[0] * t0[0] + c0[1] * t0[1];
  const double b01 = c0[0] * t1[0] + c0[1] * t1[1];
  const double b10 = c1[0] * t0[0] + c1[1] * t0[1];
  const double b11 = c1[0] * t1[0] + c1[1] * t1[1];
  double lb0 = b00 * jac[0] + b11 * jac[2];
  double lb1 = b00 * jac[1] + b11 * jac[3];
  double off = 0.5 * (b01 + b10) * gam;
  if (e->branch == 1) off = 0.;          /* output ignores the off-diagonal */
  if (e->branch == 2) {                  /* lambda_0 is a copy of lambda_1 */
    lb1 += lb0;
    lb0 = 0.;
    off = 0.;
  }
  /* ba = P C P^T with C = [[lb0, off], [off, lb1]] */
  for (int r = 0; r < 2; ++r)
    for (int s = 0; s < 2; ++s)
      ba[2 * r + s] = lb0 * c0[r] * c0[s] + lb1 * c1[r] * c1[s] +
                      off * (c0[r] * c1[s] + c1[r] * c0[s]);
}

static inline void sym_to_mat(const double a[3], double m[4]) {
  m[0] = a[0]; m[1] = 0.5 * a[1]; m[2] = 0.5 * a[1]; m[3] = a[2];
}
static inline void mat_to_sym_add(const double m[4], double a[3]) {
  a[0] += m[0]; a[1] += m[1] + m[2]; a[2] += m[3];
}

/* divided differences of log and exp, stable for close arguments.  These
 * are the adjoint's own derivatives, not reference code: they call the
 * functions the device calls (shared_libm.cpp), so that both sides round
 * them identically. */
double orc_shared_log(double x);
double orc_shared_exp(double x);
double orc_shared_log1p_small(double r);
double orc_shared_expm1(double d);
static double dd_log(double x, double y) {
  if (x == y) return 1. / x;
  const double r = (x - y) / y;
  if (fabs(r) < 1e-4) return orc_shared_log1p_small(r) / (x - y);
  return (orc_shared_log(x) - orc_shared_log(y)) / (x - y);
}
static double dd_exp(double x, double y) {
  if (x == y) return orc_shared_exp(x);
  return orc_shared_exp(y) * orc_shared_expm1(x - y) / (x - y);
}

/* ---- forward pieces as the reference computes them ---------------------- */

static void tri_geom(const double* xy, const int32_t* n, double* area,
                     double d[6]) {
  const double p0 = xy[2 * (size_t)n[0]], p1 = xy[2 * (size_t)n[0] + 1];
  const double p2 = xy[2 * (size_t)n[1]], p3 = xy[2 * (size_t)n[1] + 1];
  const double p4 = xy[2 * (size_t)n[2]], p5 = xy[2 * (size_t)n[2] + 1];
  const double c0 = p2 - p0, c1 = p4 - p0, c2 = p3 - p1, c3 = p5 - p1;
  const double a = 0.5 * (c0 * c3 - c1 * c2);       /* hessian.cpp:49 */
  const double norm = 2 / a;                          /* :52-61 */
  d[0] = norm * (p3 - p5);
  d[1] = norm * (p4 - p2);
  d[2] = norm * (p5 - p1);
  d[3] = norm * (p0 - p4);
  d[4] = norm * (p1 - p3);
  d[5] = norm * (p2 - p0);
  *area = a;
}

/* Reverse of tri_geom: adds to gxy (2 per corner, corner order) the
 * contribution of the adjoints of the area (ab) and of d (db). */
static void tri_geom_adjoint(const double* xy, const int32_t* n, double ab,
                             const double db[6], double g[6]) {
  const double x0 = xy[2 * (size_t)n[0]], y0 = xy[2 * (size_t)n[0] + 1];
  const double x1 = xy[2 * (size_t)n[1]], y1 = xy[2 * (size_t)n[1] + 1];
  const double x2 = xy[2 * (size_t)n[2]], y2 = xy[2 * (size_t)n[2] + 1];
  const double c0 = x1 - x0, c1 = x2 - x0, c2 = y1 - y0, c3 = y2 - y0;
  const double a = 0.5 * (c0 * c3 - c1 * c2);
  const double nrm = 2 / a;
  const double r[6] = {y1 - y2, x2 - x1, y2 - y0, x0 - x2, y0 - y1, x1 - x0};
  double nb = 0.;
  double rb[6];
  for (int k = 0; k < 6; ++k) {
    nb += db[k] * r[k];
    rb[k] = nrm * db[k];
  }
  ab += nb * (-2 / (a * a));
  const double cb0 = 0.5 * c3 * ab, cb3 = 0.5 * c0 * ab;
  const double cb1 = -0.5 * c2 * ab, cb2 = -0.5 * c1 * ab;
  /* g = [x0, y0, x1, y1, x2, y2] */
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

/* Local normalisation of one tensor (metric.cpp:113-148) as a spectral map
 * mu(lambda): jacobian and divided difference. */
static void local_mu(const Eig* e, double ex, double mu[2], double jac[4],
                     double* gam) {
  double lam[2] = {e->val[0], e->val[1]};
  double c[2], dc[2];
  for (int k = 0; k < 2; ++k) {
    const double a = fabs(lam[k]);
    c[k] = dmax(a, ORC_EPSILON);
    dc[k] = (a < ORC_EPSILON) ? 0. : (lam[k] < 0 ? -1. : 1.);
  }
  const double s = pow(c[0] * c[1], ex);
  mu[0] = c[0] * s;
  mu[1] = c[1] * s;
  double ds[2];
  for (int l = 0; l < 2; ++l) ds[l] = ex * s / c[l] * dc[l];
  for (int k = 0; k < 2; ++k)
    for (int l = 0; l < 2; ++l)
      jac[2 * k + l] = (k == l ? dc[k] * s : 0.) + c[k] * ds[l];
  if (dc[0] != 0. && dc[0] == dc[1]) *gam = dc[0] * s;
  else if (lam[0] == lam[1]) *gam = 0.;
  else *gam = s * (c[0] - c[1]) / (lam[0] - lam[1]);
}

/* Global normalisation (metric.cpp:204-231) as a spectral map. */
static void global_mu(const Eig* e, double fact, double emin, double emax,
                      double mu[2], double jac[4], double* gam, double dfact[2]) {
  int free_[2];
  for (int k = 0; k < 2; ++k) {
    double v = e->val[k] * fact;
    double w = dmin(v, emin);
    w = dmax(w, emax);
    mu[k] = w;
    free_[k] = (v < emin) && (v > emax);
    dfact[k] = free_[k] ? e->val[k] : 0.;
  }
  jac[0] = free_[0] ? fact : 0.;
  jac[1] = jac[2] = 0.;
  jac[3] = free_[1] ? fact : 0.;
  if (free_[0] && free_[1]) *gam = fact;
  else if (mu[0] == mu[1] || e->val[0] == e->val[1]) *gam = 0.;
  else *gam = (mu[0] - mu[1]) / (e->val[0] - e->val[1]);
}

/* Reverse of numeric::computeQuality (numeric.cpp:251-306) for one
 * triangle: given dq, adds dE/d(corner coordinates) to g[6] and dE/d(corner
 * tensors) (stored form) to gm[9].  Returns q. */
static double quality_adjoint(const double* pa, const double* pb,
                              const double* pc, const double* ta,
                              const double* tb, const double* tc, double dq,
                              double g[6], double gm[9], int with_grad) {
  const double* tk[3] = {ta, tb, tc};
  Eig ek[3];
  double lk[3][4];
  double lsum[4] = {0., 0., 0., 0.};
  for (int k = 0; k < 3; ++k) {   /* numeric.cpp:113-172 */
    const double m[4] = {tk[k][0], tk[k][1], tk[k][1], tk[k][2]};
    eig(m, &ek[k]);
    double v[2] = {log(ek[k].val[0]), log(ek[k].val[1])};
    if (ek[k].branch == 2) v[0] = v[1];
    const double* P = ek[k].P;
    const double c0 = v[0] * P[0], c1 = v[0] * P[2];
    const double c2 = v[1] * P[1], c3 = v[1] * P[3];
    lk[k][0] = P[0] * c0 + P[1] * c2;
    lk[k][1] = P[0] * c1 + P[1] * c3;
    lk[k][2] = P[2] * c0 + P[3] * c2;
    lk[k][3] = P[2] * c1 + P[3] * c3;
    for (int j = 0; j < 4; ++j) lsum[j] += lk[k][j];
  }
  double lm[4];
  for (int j = 0; j < 4; ++j) lm[j] = lsum[j] / 3;   /* :192-193 */
  Eig el;
  eig(lm, &el);                                       /* :197-216 */
  const double ev0 = exp(el.val[0]), ev1 = exp(el.val[1]);
  double jm[4];
  {
    const double* P = el.P;
    const double c0 = ev0 * P[0], c1 = ev0 * P[2];
    const double c2 = ev1 * P[1], c3 = ev1 * P[3];
    jm[0] = P[0] * c0 + P[1] * c2;
    jm[1] = P[0] * c1 + P[1] * c3;
    jm[2] = P[2] * c0 + P[3] * c2;
    jm[3] = P[2] * c1 + P[3] * c3;
  }
  const double J0 = jm[0], J1 = jm[1], J2 = jm[3];
  const double det = J0 * J2 - J1 * J1;
  const double e[3][2] = {{pb[0] - pa[0], pb[1] - pa[1]},
                          {pc[0] - pb[0], pc[1] - pb[1]},
                          {pa[0] - pc[0], pa[1] - pc[1]}};
  double len[3], perim = 0.;
  for (int k = 0; k < 3; ++k) {
    const double vx = J0 * e[k][0] + J1 * e[k][1];
    const double vy = J1 * e[k][0] + J2 * e[k][1];
    len[k] = sqrt(e[k][0] * vx + e[k][1] * vy);
    perim += len[k];
  }
  const double cr = (pb[0] - pa[0]) * (pc[1] - pa[1]) - (pc[0] - pa[0]) * (pb[1] - pa[1]);
  const double area = 0.5 * fabs(cr);
  const double mean = dmin(perim / 3, 3 / perim);
  const double base = mean * (2. - mean);
  const double norm = 12 * sqrt(3.) * base * base * base;
  const double sdet = sqrt(det);
  const double q = area * norm * sdet / (perim * perim);
  if (!with_grad) return q;

  /* reverse */
  const double p2 = perim * perim;
  const double area_b = dq * norm * sdet / p2;
  const double norm_b = dq * area * sdet / p2;
  const double det_b = dq * area * norm / (2 * sdet * p2);
  const double dmean = (perim / 3 <= 3 / perim) ? 1. / 3 : -3 / p2;
  const double dnorm = 12 * sqrt(3.) * 3 * base * base * (2 - 2 * mean);
  const double perim_b = dq * (-2 * q / perim) + norm_b * dnorm * dmean;
  double jb[3] = {det_b * J2, -2 * det_b * J1, det_b * J0};
  double eb[3][2];
  for (int k = 0; k < 3; ++k) {
    const double vx = J0 * e[k][0] + J1 * e[k][1];
    const double vy = J1 * e[k][0] + J2 * e[k][1];
    eb[k][0] = perim_b * vx / len[k];
    eb[k][1] = perim_b * vy / len[k];
    const double w = perim_b / (2 * len[k]);
    jb[0] += w * e[k][0] * e[k][0];
    jb[1] += w * 2 * e[k][0] * e[k][1];
    jb[2] += w * e[k][1] * e[k][1];
  }
  /* edges: e0 = pb - pa, e1 = pc - pb, e2 = pa - pc */
  g[0] += eb[2][0] - eb[0][0]; g[1] += eb[2][1] - eb[0][1];
  g[2] += eb[0][0] - eb[1][0]; g[3] += eb[0][1] - eb[1][1];
  g[4] += eb[1][0] - eb[2][0]; g[5] += eb[1][1] - eb[2][1];
  /* area = |cr| / 2 */
  {
    const double crb = (cr < 0 ? -0.5 : 0.5) * area_b;
    const double ux = pb[0] - pa[0], uy = pb[1] - pa[1];
    const double wx = pc[0] - pa[0], wy = pc[1] - pa[1];
    /* cr = ux*wy - wx*uy */
    const double uxb = crb * wy, wyb = crb * ux, wxb = -crb * uy, uyb = -crb * wx;
    g[2] += uxb; g[0] -= uxb;
    g[3] += uyb; g[1] -= uyb;
    g[4] += wxb; g[0] -= wxb;
    g[5] += wyb; g[1] -= wyb;
  }
  /* J = exp(lm) */
  double jbm[4], lmb[4];
  sym_to_mat(jb, jbm);
  {
    const double jac[4] = {ev0, 0., 0., ev1};
    spectral_adjoint(&el, jac, dd_exp(el.val[0], el.val[1]), jbm, lmb);
  }
  /* lm = (l_a + l_b + l_c) / 3, l_k = log(M_k) */
  for (int k = 0; k < 3; ++k) {
    double lb[4], mb[4];
    for (int j = 0; j < 4; ++j) lb[j] = lmb[j] / 3;
    const double jac[4] = {1. / ek[k].val[0], 0., 0., 1. / ek[k].val[1]};
    spectral_adjoint(&ek[k], jac, dd_log(ek[k].val[0], ek[k].val[1]), lb, mb);
    mat_to_sym_add(mb, gm + 3 * k);
  }
  return q;
}

/* ---- energy and gradient ------------------------------------------------- */

typedef struct {
  int nv, nt;
  const double* xy;
  const int32_t* tri;
  const double* f;
  const int32_t *v2t_off, *v2t, *v2v_off, *v2v;
  double ex, emin, emax, target, h_min, h_max;
  int lp_norm;
  /* tape */
  double *G, *H4, *TA, *Mloc, *M;
  double C, fact;
} Tape;

static void forward(Tape* T) {
  const int nv = T->nv;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nv; ++i) {
    int32_t patch[4096];
    const int m = orc_patch2(i, T->v2t_off, T->v2t, T->v2v_off, T->v2v, patch, 4096);
    double ta = 0., gx = 0., gy = 0.;
    for (int e = 0; e < m; ++e) {
      const int32_t* n = T->tri + 3 * (size_t)patch[e];
      double a, d[6];
      tri_geom(T->xy, n, &a, d);
      ta += a;
      double g0 = 0., g1 = 0.;
      for (int j = 0; j < 3; ++j) {
        g0 += T->f[n[j]] * d[2 * j];
        g1 += T->f[n[j]] * d[2 * j + 1];
      }
      gx += a * g0;
      gy += a * g1;
    }
    T->TA[i] = ta;
    T->G[2 * (size_t)i] = gx / ta;
    T->G[2 * (size_t)i + 1] = gy / ta;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nv; ++i) {
    int32_t patch[4096];
    const int m = orc_patch2(i, T->v2t_off, T->v2t, T->v2v_off, T->v2v, patch, 4096);
    double h[4] = {0., 0., 0., 0.};
    for (int e = 0; e < m; ++e) {
      const int32_t* n = T->tri + 3 * (size_t)patch[e];
      double a, d[6];
      tri_geom(T->xy, n, &a, d);
      double dl[4] = {0., 0., 0., 0.};
      for (int j = 0; j < 3; ++j) {
        const double gx = T->G[2 * (size_t)n[j]], gy = T->G[2 * (size_t)n[j] + 1];
        dl[0] += gx * d[2 * j];
        dl[1] += gx * d[2 * j + 1];
        dl[2] += gy * d[2 * j];
        dl[3] += gy * d[2 * j + 1];
      }
      for (int j = 0; j < 4; ++j) h[j] += a * dl[j];
    }
    for (int j = 0; j < 4; ++j) T->H4[4 * (size_t)i + j] = h[j] / T->TA[i];
    double* t = T->Mloc + 3 * (size_t)i;
    t[0] = T->H4[4 * (size_t)i];
    t[1] = 0.5 * (T->H4[4 * (size_t)i + 1] + T->H4[4 * (size_t)i + 2]);
    t[2] = T->H4[4 * (size_t)i + 3];
  }
  /* local normalisation, complexity and global normalisation: the
   * restatements of metric.cpp used by orc_metric_run */
  orc_normalize_local(nv, T->Mloc, T->lp_norm);
  T->C = orc_complexity(T->nt, T->xy, T->tri, T->Mloc);
  T->fact = T->target / T->C;
  memcpy(T->M, T->Mloc, sizeof(double) * 3 * (size_t)nv);
  orc_normalize_global(nv, T->M, T->fact, T->h_min, T->h_max);
}

static double energy_of(const Tape* T, double* qualit) {
  double E = 0.;
  for (int t = 0; t < T->nt; ++t) {
    const int32_t* n = T->tri + 3 * (size_t)t;
    if (n[0] < 0) continue;
    const double q = orc_quality(T->xy + 2 * (size_t)n[0], T->xy + 2 * (size_t)n[1],
                                 T->xy + 2 * (size_t)n[2], T->M + 3 * (size_t)n[0],
                                 T->M + 3 * (size_t)n[1], T->M + 3 * (size_t)n[2]);
    if (qualit) qualit[t] = q;
    E += (1 - q) * (1 - q);
  }
  return E;
}

static int tape_alloc(Tape* T) {
  const size_t nv = (size_t)T->nv;
  T->G = malloc(sizeof(double) * 2 * nv);
  T->H4 = malloc(sizeof(double) * 4 * nv);
  T->TA = malloc(sizeof(double) * nv);
  T->Mloc = malloc(sizeof(double) * 3 * nv);
  T->M = malloc(sizeof(double) * 3 * nv);
  return T->G && T->H4 && T->TA && T->Mloc && T->M;
}

static void tape_free(Tape* T) {
  free(T->G); free(T->H4); free(T->TA); free(T->Mloc); free(T->M);
}

static void tape_init(Tape* T, int nv, int nt, const double* xy,
                      const int32_t* tri, const double* f,
                      const int32_t* v2t_off, const int32_t* v2t,
                      const int32_t* v2v_off, const int32_t* v2v,
                      double target_factor, int lp_norm, double h_min,
                      double h_max) {
  T->nv = nv; T->nt = nt; T->xy = xy; T->tri = tri; T->f = f;
  T->v2t_off = v2t_off; T->v2t = v2t; T->v2v_off = v2v_off; T->v2v = v2v;
  T->ex = orc_metric_exponent(lp_norm);
  T->lp_norm = lp_norm;
  T->h_min = h_min;
  T->h_max = h_max;
  T->emin = 1 / (h_min * h_min);
  T->emax = 1 / (h_max * h_max);
  T->target = orc_metric_target(nv, target_factor);
}

double orc_energy(int nv, int nt, const double* xy, const int32_t* tri,
                  const double* f, const int32_t* v2t_off, const int32_t* v2t,
                  const int32_t* v2v_off, const int32_t* v2v,
                  double target_factor, int lp_norm, double h_min,
                  double h_max, double* tensor, double* qualit) {
  Tape T;
  tape_init(&T, nv, nt, xy, tri, f, v2t_off, v2t, v2v_off, v2v, target_factor,
            lp_norm, h_min, h_max);
  if (!tape_alloc(&T)) { tape_free(&T); return NAN; }
  forward(&T);
  const double E = energy_of(&T, qualit);
  if (tensor) memcpy(tensor, T.M, sizeof(double) * 3 * (size_t)nv);
  tape_free(&T);
  return E;
}

/* Reverse sweep over the tape.  The seeds are adjoints entering at the
 * final metric (seed_M), the locally normalised metric (seed_Mloc) and the
 * vertex gradients (seed_G), each optional; with_quality adds the energy's
 * own adjoint.  Returns the energy when with_quality is set. */
static double backward(const Tape* T_, const double* seed_M, const double* seed_Mloc,
                       const double* seed_G, int with_quality, double* dxy,
                       double* df) {
  const Tape T = *T_;
  const int nv = T.nv, nt = T.nt;
  const double* xy = T.xy;
  const int32_t* tri = T.tri;
  const double* f = T.f;
  const int32_t *v2t_off = T.v2t_off, *v2t = T.v2t, *v2v_off = T.v2v_off, *v2v = T.v2v;
  const size_t NV = (size_t)nv;
  double* Mb = calloc(3 * NV, sizeof(double));
  double* Mlb = calloc(3 * NV, sizeof(double));
  double* Sb = calloc(4 * NV, sizeof(double));   /* d E / d (sum a delta) */
  double* Sgb = calloc(2 * NV, sizeof(double));  /* d E / d (sum a g) */
  double* TAb = calloc(NV, sizeof(double));
  double* Gb = calloc(2 * NV, sizeof(double));
  double* ab = calloc((size_t)nt, sizeof(double));
  double* db = calloc(6 * (size_t)nt, sizeof(double));
  memset(dxy, 0, sizeof(double) * 2 * NV);
  memset(df, 0, sizeof(double) * NV);
  if (seed_M) memcpy(Mb, seed_M, sizeof(double) * 3 * NV);
  if (seed_Mloc) memcpy(Mlb, seed_Mloc, sizeof(double) * 3 * NV);
  if (seed_G) memcpy(Gb, seed_G, sizeof(double) * 2 * NV);
  double E = 0.;
  /* 1. qualities */
  for (int t = 0; with_quality && t < nt; ++t) {
    const int32_t* n = tri + 3 * (size_t)t;
    if (n[0] < 0) continue;
    double g[6] = {0, 0, 0, 0, 0, 0}, gm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double q0 = orc_quality(xy + 2 * (size_t)n[0], xy + 2 * (size_t)n[1],
                                  xy + 2 * (size_t)n[2], T.M + 3 * (size_t)n[0],
                                  T.M + 3 * (size_t)n[1], T.M + 3 * (size_t)n[2]);
    E += (1 - q0) * (1 - q0);
    quality_adjoint(xy + 2 * (size_t)n[0], xy + 2 * (size_t)n[1], xy + 2 * (size_t)n[2],
                    T.M + 3 * (size_t)n[0], T.M + 3 * (size_t)n[1], T.M + 3 * (size_t)n[2],
                    -2 * (1 - q0), g, gm, 1);
    for (int k = 0; k < 3; ++k) {
      dxy[2 * (size_t)n[k]] += g[2 * k];
      dxy[2 * (size_t)n[k] + 1] += g[2 * k + 1];
      for (int c = 0; c < 3; ++c) Mb[3 * (size_t)n[k] + c] += gm[3 * k + c];
    }
  }
  /* 2. global normalisation */
  double factb = 0.;
  for (int i = 0; i < nv; ++i) {
    const double* t = T.Mloc + 3 * (size_t)i;
    const double m[4] = {t[0], t[1], t[1], t[2]};
    Eig e;
    eig(m, &e);
    double mu[2], jac[4], gam, dfa[2], bx[4], ba[4];
    global_mu(&e, T.fact, T.emin, T.emax, mu, jac, &gam, dfa);
    sym_to_mat(Mb + 3 * (size_t)i, bx);
    spectral_adjoint(&e, jac, gam, bx, ba);
    mat_to_sym_add(ba, Mlb + 3 * (size_t)i);
    /* d M / d fact = P diag(dfa) P^T */
    const double c0[2] = {e.P[0], e.P[2]}, c1[2] = {e.P[1], e.P[3]};
    const double b00 = c0[0] * (bx[0] * c0[0] + bx[1] * c0[1]) + c0[1] * (bx[2] * c0[0] + bx[3] * c0[1]);
    const double b11 = c1[0] * (bx[0] * c1[0] + bx[1] * c1[1]) + c1[1] * (bx[2] * c1[0] + bx[3] * c1[1]);
    if (e.branch == 2) factb += (b00 + b11) * dfa[1];
    else factb += b00 * dfa[0] + b11 * dfa[1];
  }
  /* 3. complexity: fact = target / C */
  const double Cb = factb * (-T.fact / T.C);
  for (int t = 0; t < nt; ++t) {
    const int32_t* n = tri + 3 * (size_t)t;
    if (n[0] < 0) continue;
    double a, d[6];
    tri_geom(xy, n, &a, d);
    double rho = 0.;
    for (int j = 0; j < 3; ++j) {
      const double* m = T.Mloc + 3 * (size_t)n[j];
      const double sd = sqrt(m[0] * m[2] - m[1] * m[1]);
      rho += sd;
      const double w = Cb * a / 3 / (2 * sd);
      Mlb[3 * (size_t)n[j]] += w * m[2];
      Mlb[3 * (size_t)n[j] + 1] += -2 * w * m[1];
      Mlb[3 * (size_t)n[j] + 2] += w * m[0];
    }
    rho /= 3;
    ab[t] += Cb * rho;
  }
  /* 4. local normalisation -> raw Hessian adjoint */
  for (int i = 0; i < nv; ++i) {
    const double* h4 = T.H4 + 4 * (size_t)i;
    const double hs[3] = {h4[0], 0.5 * (h4[1] + h4[2]), h4[3]};
    const double m[4] = {hs[0], hs[1], hs[1], hs[2]};
    Eig e;
    eig(m, &e);
    double mu[2], jac[4], gam, bx[4], ba[4];
    local_mu(&e, T.ex, mu, jac, &gam);
    sym_to_mat(Mlb + 3 * (size_t)i, bx);
    spectral_adjoint(&e, jac, gam, bx, ba);
    double hb[3] = {0., 0., 0.};
    mat_to_sym_add(ba, hb);
    const double h4b[4] = {hb[0], 0.5 * hb[1], 0.5 * hb[1], hb[2]};
    const double ta = T.TA[i];
    double tab = 0.;
    for (int c = 0; c < 4; ++c) {
      Sb[4 * (size_t)i + c] = h4b[c] / ta;
      tab -= h4b[c] * h4[c] / ta;
    }
    TAb[i] = tab;
  }
  /* 5. Hessian recovery: S_i = sum_{t in P(i)} a_t delta_t, TA_i = sum a_t */
  int32_t* patch = malloc(sizeof(int32_t) * 4096);
  for (int i = 0; i < nv; ++i) {
    const int m = orc_patch2(i, v2t_off, v2t, v2v_off, v2v, patch, 4096);
    const double* s = Sb + 4 * (size_t)i;
    for (int e = 0; e < m; ++e) {
      const int t = patch[e];
      const int32_t* n = tri + 3 * (size_t)t;
      double a, d[6];
      tri_geom(xy, n, &a, d);
      double dl[4] = {0., 0., 0., 0.};
      for (int j = 0; j < 3; ++j) {
        const double gx = T.G[2 * (size_t)n[j]], gy = T.G[2 * (size_t)n[j] + 1];
        dl[0] += gx * d[2 * j];
        dl[1] += gx * d[2 * j + 1];
        dl[2] += gy * d[2 * j];
        dl[3] += gy * d[2 * j + 1];
      }
      ab[t] += s[0] * dl[0] + s[1] * dl[1] + s[2] * dl[2] + s[3] * dl[3];
      const double lb[4] = {a * s[0], a * s[1], a * s[2], a * s[3]};
      for (int j = 0; j < 3; ++j) {
        const double gx = T.G[2 * (size_t)n[j]], gy = T.G[2 * (size_t)n[j] + 1];
        Gb[2 * (size_t)n[j]] += lb[0] * d[2 * j] + lb[1] * d[2 * j + 1];
        Gb[2 * (size_t)n[j] + 1] += lb[2] * d[2 * j] + lb[3] * d[2 * j + 1];
        db[6 * (size_t)t + 2 * j] += lb[0] * gx + lb[2] * gy;
        db[6 * (size_t)t + 2 * j + 1] += lb[1] * gx + lb[3] * gy;
      }
    }
  }
  /* 6. gradient recovery: G_i = Sg_i / TA_i */
  for (int i = 0; i < nv; ++i) {
    const double ta = T.TA[i];
    Sgb[2 * (size_t)i] = Gb[2 * (size_t)i] / ta;
    Sgb[2 * (size_t)i + 1] = Gb[2 * (size_t)i + 1] / ta;
    TAb[i] -= (Gb[2 * (size_t)i] * T.G[2 * (size_t)i] +
               Gb[2 * (size_t)i + 1] * T.G[2 * (size_t)i + 1]) / ta;
  }
  for (int i = 0; i < nv; ++i) {
    const int m = orc_patch2(i, v2t_off, v2t, v2v_off, v2v, patch, 4096);
    const double sx = Sgb[2 * (size_t)i], sy = Sgb[2 * (size_t)i + 1];
    for (int e = 0; e < m; ++e) {
      const int t = patch[e];
      const int32_t* n = tri + 3 * (size_t)t;
      double a, d[6];
      tri_geom(xy, n, &a, d);
      double g0 = 0., g1 = 0.;
      for (int j = 0; j < 3; ++j) {
        g0 += f[n[j]] * d[2 * j];
        g1 += f[n[j]] * d[2 * j + 1];
      }
      ab[t] += sx * g0 + sy * g1 + TAb[i];
      const double gbx = a * sx, gby = a * sy;
      for (int j = 0; j < 3; ++j) {
        df[n[j]] += gbx * d[2 * j] + gby * d[2 * j + 1];
        db[6 * (size_t)t + 2 * j] += f[n[j]] * gbx;
        db[6 * (size_t)t + 2 * j + 1] += f[n[j]] * gby;
      }
    }
  }
  free(patch);
  /* 7. triangle geometry */
  for (int t = 0; t < nt; ++t) {
    const int32_t* n = tri + 3 * (size_t)t;
    if (n[0] < 0) continue;
    double g[6] = {0, 0, 0, 0, 0, 0};
    tri_geom_adjoint(xy, n, ab[t], db + 6 * (size_t)t, g);
    for (int k = 0; k < 3; ++k) {
      dxy[2 * (size_t)n[k]] += g[2 * k];
      dxy[2 * (size_t)n[k] + 1] += g[2 * k + 1];
    }
  }
  free(Mb); free(Mlb); free(Sb); free(Sgb); free(TAb); free(Gb); free(ab); free(db);
  return E;
}

double orc_energy_grad(int nv, int nt, const double* xy, const int32_t* tri,
                       const double* f, const int32_t* v2t_off,
                       const int32_t* v2t, const int32_t* v2v_off,
                       const int32_t* v2v, double target_factor, int lp_norm,
                       double h_min, double h_max, double* dxy, double* df) {
  Tape T;
  tape_init(&T, nv, nt, xy, tri, f, v2t_off, v2t, v2v_off, v2v, target_factor,
            lp_norm, h_min, h_max);
  if (!tape_alloc(&T)) { tape_free(&T); return NAN; }
  forward(&T);
  const double E = backward(&T, NULL, NULL, NULL, 1, dxy, df);
  tape_free(&T);
  return E;
}
