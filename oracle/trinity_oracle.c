/*
 * trinity_oracle.c -- CPU restatement of trinity's metric-and-smoothing path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Never linked into, loaded by or
 * called from the product path in trinity_amd/.
 *
 * Build with -O2 -ffp-contract=off: the reference's results depend on the
 * exact rounding of every product and sum (a fused multiply-add changes the
 * recovered Hessian where it cancels), so contraction must stay off here,
 * exactly as in the HIP kernels.
 *
 * Reference: hobywan/trinity v1.0.3.  Each function cites the lines it
 * restates.  std::min(a,b) is (b < a ? b : a) and std::max(a,b) is
 * (a < b ? b : a); both are restated literally so that NaN and signed-zero
 * behaviour matches.  pow(x, 2) in the reference is folded to x*x by the
 * compiler (exact for a correctly rounded pow); pow(x, 3) and pow(x, y) for
 * non-integer y are real libm calls.
 */
#include "trinity_oracle.h"

#include <math.h>
#include <float.h>
#include <limits.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline double dmin(double a, double b) { return (b < a) ? b : a; }
static inline double dmax(double a, double b) { return (a < b) ? b : a; }

static int g_threads = 1;

void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int  orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------ */
/* small sort helpers                                                        */
static void isort_int(int32_t* a, int n) {
  for (int i = 1; i < n; ++i) {
    int32_t x = a[i];
    int j = i - 1;
    while (j >= 0 && a[j] > x) { a[j + 1] = a[j]; --j; }
    a[j + 1] = x;
  }
}

static int uniq_sorted(int32_t* a, int n) {
  if (n == 0) return 0;
  int m = 1;
  for (int i = 1; i < n; ++i)
    if (a[i] != a[m - 1]) a[m++] = a[i];
  return m;
}

/* ------------------------------------------------------------------------ */
/* Mesh::rebuildTopology, src/mesh.cpp:197-210 (stencil) and :244 (sort).    */
int orc_build_v2t(int nv, int nt, const int32_t* tri,
                  int32_t* v2t_off, int32_t* v2t) {
  memset(v2t_off, 0, sizeof(int32_t) * (size_t)(nv + 1));
  for (int t = 0; t < nt; ++t) {
    const int32_t* n = tri + 3 * (size_t)t;
    if (n[0] < 0) continue;                      /* mesh.cpp:200 */
    v2t_off[n[0] + 1]++; v2t_off[n[1] + 1]++; v2t_off[n[2] + 1]++;
  }
  for (int i = 0; i < nv; ++i) v2t_off[i + 1] += v2t_off[i];
  int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
  memcpy(cur, v2t_off, sizeof(int32_t) * (size_t)nv);
  /* filled in ascending t, so each list is already sorted (mesh.cpp:244) */
  for (int t = 0; t < nt; ++t) {
    const int32_t* n = tri + 3 * (size_t)t;
    if (n[0] < 0) continue;
    for (int j = 0; j < 3; ++j) v2t[cur[n[j]]++] = t;
  }
  free(cur);
  return 0;
}

/* Mesh::rebuildTopology, src/mesh.cpp:214-243: sorted-unique 1-ring.  For
 * each incident triangle the two other vertices are collected in the
 * rotational order of :225-239; the sort + unique makes the order moot. */
int64_t orc_build_v2v(int nv, const int32_t* tri,
                      const int32_t* v2t_off, const int32_t* v2t,
                      int32_t* v2v_off, int32_t* v2v, int64_t cap) {
  int64_t pos = 0;
  int32_t buf[1024];
  v2v_off[0] = 0;
  for (int i = 0; i < nv; ++i) {
    int m = 0;
    for (int k = v2t_off[i]; k < v2t_off[i + 1]; ++k) {
      const int32_t* n = tri + 3 * (size_t)v2t[k];
      if (m + 2 > 1024) return -1;
      if (n[0] == i)      { buf[m++] = n[1]; buf[m++] = n[2]; }
      else if (n[1] == i) { buf[m++] = n[2]; buf[m++] = n[0]; }
      else if (n[2] == i) { buf[m++] = n[0]; buf[m++] = n[1]; }
    }
    isort_int(buf, m);
    m = uniq_sorted(buf, m);
    if (pos + m > cap) return -1;
    memcpy(v2v + pos, buf, sizeof(int32_t) * (size_t)m);
    pos += m;
    v2v_off[i + 1] = (int32_t)pos;
  }
  return pos;
}

/* Mesh::getVicinity(id, 2), src/mesh.cpp:390-434: the element part of the
 * patch is stenc[id] united with stenc[k] for every k in vicin[id], as an
 * ascending std::set. */
int orc_patch2(int i, const int32_t* v2t_off, const int32_t* v2t,
               const int32_t* v2v_off, const int32_t* v2v,
               int32_t* out, int cap) {
  int m = 0;
  for (int k = v2t_off[i]; k < v2t_off[i + 1]; ++k) {
    if (m >= cap) return -1;
    out[m++] = v2t[k];
  }
  for (int a = v2v_off[i]; a < v2v_off[i + 1]; ++a) {
    int j = v2v[a];
    for (int k = v2t_off[j]; k < v2t_off[j + 1]; ++k) {
      if (m >= cap) return -1;
      out[m++] = v2t[k];
    }
  }
  isort_int(out, m);
  return uniq_sorted(out, m);
}

/* ------------------------------------------------------------------------ */
/* Per-triangle geometry shared by both recovery kernels:
 * src/hessian.cpp:42-61 (gradient) and :108-128 (Hessian), identical code. */
static inline void tri_geometry(const double* xy, const int32_t* n,
                                double* area, double deriv[6]) {
  double p[6];
  p[0] = xy[2 * (size_t)n[0]]; p[1] = xy[2 * (size_t)n[0] + 1];
  p[2] = xy[2 * (size_t)n[1]]; p[3] = xy[2 * (size_t)n[1] + 1];
  p[4] = xy[2 * (size_t)n[2]]; p[5] = xy[2 * (size_t)n[2] + 1];
  double c0 = p[2] - p[0];
  double c1 = p[4] - p[0];
  double c2 = p[3] - p[1];
  double c3 = p[5] - p[1];
  double a = 0.5 * (c0 * c3 - c1 * c2);           /* hessian.cpp:49 */
  double norm = 2 / a;                              /* hessian.cpp:52 */
  deriv[0] = norm * (p[3] - p[5]);                  /* hessian.cpp:56-61 */
  deriv[1] = norm * (p[4] - p[2]);
  deriv[2] = norm * (p[5] - p[1]);
  deriv[3] = norm * (p[0] - p[4]);
  deriv[4] = norm * (p[1] - p[3]);
  deriv[5] = norm * (p[2] - p[0]);
  *area = a;
}

#define ORC_PATCH_CAP 4096

/* Metrics::computeGradient, src/hessian.cpp:24-83. */
void orc_gradient(int nv, const double* xy, const int32_t* tri,
                  const double* solut,
                  const int32_t* v2t_off, const int32_t* v2t,
                  const int32_t* v2v_off, const int32_t* v2v,
                  double* gradient) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < nv; ++i) {
    int32_t patch[ORC_PATCH_CAP];
    int size = orc_patch2(i, v2t_off, v2t, v2v_off, v2v, patch, ORC_PATCH_CAP);
    double total_area = 0.;
    double gx = 0., gy = 0.;
    /* The reference stores area[] and grad[] per patch element and sums
     * area*grad afterwards in patch order (:75-78); summing in the same
     * loop with the same order gives the same roundings. */
    for (int e = 0; e < size; ++e) {
      const int32_t* n = tri + 3 * (size_t)patch[e];
      double area, d[6];
      tri_geometry(xy, n, &area, d);
      total_area += area;                           /* hessian.cpp:53 */
      double g0 = 0., g1 = 0.;                      /* hessian.cpp:36 */
      for (int j = 0; j < 3; ++j) {                 /* hessian.cpp:65-67 */
        g0 += solut[n[j]] * d[j * 2];
        g1 += solut[n[j]] * d[j * 2 + 1];
      }
      gx += area * g0;                              /* hessian.cpp:76-78 */
      gy += area * g1;
    }
    gradient[2 * (size_t)i] = gx / total_area;      /* hessian.cpp:81-82 */
    gradient[2 * (size_t)i + 1] = gy / total_area;
  }
}

/* Metrics::computeHessian, src/hessian.cpp:86-153. */
void orc_hessian(int nv, const double* xy, const int32_t* tri,
                 const double* gradient,
                 const int32_t* v2t_off, const int32_t* v2t,
                 const int32_t* v2v_off, const int32_t* v2v,
                 double* tensor) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < nv; ++i) {
    int32_t patch[ORC_PATCH_CAP];
    int size = orc_patch2(i, v2t_off, v2t, v2v_off, v2v, patch, ORC_PATCH_CAP);
    double total_area = 0.;
    double h[4] = {0., 0., 0., 0.};                 /* hessian.cpp:93 */
    for (int e = 0; e < size; ++e) {
      const int32_t* n = tri + 3 * (size_t)patch[e];
      double area, d[6];
      tri_geometry(xy, n, &area, d);
      total_area += area;                           /* hessian.cpp:120 */
      double dl[4] = {0., 0., 0., 0.};              /* hessian.cpp:99 */
      for (int j = 0; j < 3; ++j) {                 /* hessian.cpp:131-136 */
        const double gx = gradient[2 * (size_t)n[j]];
        const double gy = gradient[2 * (size_t)n[j] + 1];
        dl[0] += gx * d[j * 2];
        dl[1] += gx * d[j * 2 + 1];
        dl[2] += gy * d[j * 2];
        dl[3] += gy * d[j * 2 + 1];
      }
      for (int j = 0; j < 4; ++j)                   /* hessian.cpp:141-143 */
        h[j] += area * dl[j];
    }
    for (int j = 0; j < 4; ++j) h[j] /= total_area; /* hessian.cpp:145-146 */
    tensor[3 * (size_t)i] = h[0];                   /* hessian.cpp:150-152 */
    tensor[3 * (size_t)i + 1] = 0.5 * (h[1] + h[2]);
    tensor[3 * (size_t)i + 2] = h[3];
  }
}

/* ------------------------------------------------------------------------ */
/* numeric::eigenDecompose, src/numeric.cpp:25-95.  vec1/vec2 are the two
 * rows of the matrix whose columns are the eigenvectors of val[0], val[1]. */
void orc_eigen(const double m[4], double val[2], double vec1[2],
               double vec2[2]) {
  if (m[1] * m[2] < ORC_EPSILON) {                  /* numeric.cpp:30-38 */
    val[0] = m[3];
    val[1] = m[0];
    vec1[0] = 0; vec2[0] = 1;
    vec1[1] = 1; vec2[1] = 0;
    return;
  }
  double trace = m[0] + m[3];                       /* numeric.cpp:44-50 */
  double det = m[0] * m[3] - m[1] * m[2];
  double term = trace * 0.5;
  double sqrt_delta = sqrt(term * term - det);
  val[0] = term - sqrt_delta;
  val[1] = term + sqrt_delta;
  if (!isnormal(val[0]) || !isnormal(val[1]))       /* numeric.cpp:52-64 */
    val[0] = val[1];
  double dd = m[0] - m[3];                          /* numeric.cpp:69-71 */
  term = sqrt(dmax(0.25 * (dd * dd) + m[1] * m[2], 0.));
  if (m[0] - m[3] >= 0) {                           /* numeric.cpp:72-82 */
    vec1[0] = +0.5 * (m[0] - m[3]) - term;
    vec1[1] = m[1];
    vec2[0] = m[2];
    vec2[1] = -0.5 * (m[0] - m[3]) + term;
  } else {
    vec1[0] = m[2];
    vec1[1] = +0.5 * (m[0] - m[3]) + term;
    vec2[0] = -0.5 * (m[0] - m[3]) - term;
    vec2[1] = m[1];
  }
  double n0 = sqrt(vec1[0] * vec1[0] + vec1[1] * vec1[1]); /* :84-94 */
  double n1 = sqrt(vec2[0] * vec2[0] + vec2[1] * vec2[1]);
  vec1[0] /= n0; vec2[0] /= n1;
  vec1[1] /= n0; vec2[1] /= n1;
}

/* P.diag(val).P^T rebuild shared by metric.cpp:134-142, :222-231 and
 * numeric.cpp:146-154, :203-211.  vec = [vec1, vec2] as above. */
static inline void rebuild(const double val[2], const double vec[4],
                           double h[4]) {
  double c0 = val[0] * vec[0];
  double c1 = val[0] * vec[2];
  double c2 = val[1] * vec[1];
  double c3 = val[1] * vec[3];
  h[0] = vec[0] * c0 + vec[1] * c2;
  h[1] = vec[0] * c1 + vec[1] * c3;
  h[2] = vec[2] * c0 + vec[3] * c2;
  h[3] = vec[2] * c1 + vec[3] * c3;
}

/* numeric::interpolateTensor, src/numeric.cpp:98-248 (log-Euclidean mean). */
void orc_interpolate_tensor(const double* m, double* result, int nb) {
  double copy[4], val[2], vec[4], cache[4];
  double log_mat[4] = {0., 0., 0., 0.};
  for (int i = 0; i < nb; ++i) {                    /* numeric.cpp:113-172 */
    int j = 3 * i;
    copy[0] = m[j];
    copy[1] = m[j + 1];
    copy[2] = m[j + 1];
    copy[3] = m[j + 2];
    orc_eigen(copy, val, vec, vec + 2);
    val[0] = log(val[0]);
    val[1] = log(val[1]);
    rebuild(val, vec, cache);
    for (int k = 0; k < 4; ++k) log_mat[k] += cache[k];
  }
  for (int k = 0; k < 4; ++k) log_mat[k] /= nb;     /* numeric.cpp:192-193 */
  orc_eigen(log_mat, val, vec, vec + 2);            /* numeric.cpp:197-199 */
  val[0] = exp(val[0]);
  val[1] = exp(val[1]);
  rebuild(val, vec, cache);                         /* numeric.cpp:203-216 */
  result[0] = cache[0];
  result[1] = cache[1];
  result[2] = cache[3];
}

/* numeric::computeQuality, src/numeric.cpp:251-306. */
double orc_quality(const double* pa, const double* pb, const double* pc,
                   const double* ta, const double* tb, const double* tc) {
  double matrix[9], jacob[3], coef[6], vec[6];
  memcpy(matrix, ta, sizeof(double) * 3);
  memcpy(matrix + 3, tb, sizeof(double) * 3);
  memcpy(matrix + 6, tc, sizeof(double) * 3);
  orc_interpolate_tensor(matrix, jacob, 3);
  double det = jacob[0] * jacob[2] - jacob[1] * jacob[1];
  coef[0] = pb[0] - pa[0];                          /* numeric.cpp:279-292 */
  coef[1] = pb[1] - pa[1];
  coef[2] = pc[0] - pb[0];
  coef[3] = pc[1] - pb[1];
  coef[4] = pa[0] - pc[0];
  coef[5] = pa[1] - pc[1];
  double perim = 0;
  for (int k = 0; k < 3; ++k) {
    int i = k * 2, j = i + 1;
    vec[i] = jacob[0] * coef[i] + jacob[1] * coef[j];
    vec[j] = jacob[1] * coef[i] + jacob[2] * coef[j];
    perim += sqrt(coef[i] * vec[i] + coef[j] * vec[j]);
  }
  coef[0] = pb[0] - pa[0];                          /* numeric.cpp:296-300 */
  coef[1] = pc[0] - pa[0];
  coef[2] = pb[1] - pa[1];
  coef[3] = pc[1] - pa[1];
  double area = 0.5 * fabs(coef[0] * coef[3] - coef[1] * coef[2]);
  double mean = dmin(perim / 3, 3 / perim);         /* numeric.cpp:303-305 */
  double norm = 12 * sqrt(3.) * pow(mean * (2. - mean), 3);
  return area * norm * sqrt(det) / (perim * perim);
}

/* numeric::approxRiemannDist (4 arguments), src/numeric.cpp:469-497. */
double orc_riemann_dist(const double* pa, const double* pb,
                        const double* ta, const double* tb) {
  double v0 = pb[0] - pa[0];
  double v1 = pb[1] - pa[1];
  double m0 = ta[0] * v0 + ta[1] * v1;
  double m1 = ta[1] * v0 + ta[2] * v1;
  double m2 = tb[0] * v0 + tb[1] * v1;
  double m3 = tb[1] * v0 + tb[2] * v1;
  double l0 = sqrt(v0 * m0 + v1 * m1);
  double l1 = sqrt(v0 * m2 + v1 * m3);
  if (fabs(l0 - l1) < ORC_EPSILON) return l0;
  double lmin = dmin(l0, l1);
  double lmax = dmax(l0, l1);
  double ratio = lmax / lmin;
  return lmax * (ratio - 1) / (ratio * log(ratio));
}

/* Mesh::isCounterclockwise, src/mesh.cpp:700-713. */
int orc_ccw(const double* pa, const double* pb, const double* pc) {
  double c0 = pb[0] - pa[0];
  double c1 = pc[0] - pa[0];
  double c2 = pb[1] - pa[1];
  double c3 = pc[1] - pa[1];
  return 0.5 * (c0 * c3 - c1 * c2) > ORC_EPSILON;
}

/* ------------------------------------------------------------------------ */
/* Metrics::Metrics, src/metric.cpp:44 (note p = 0 gives +1/2). */
double orc_metric_exponent(int lp_norm) {
  return lp_norm ? -1. / (2 * lp_norm + 2) : 0.5;
}

/* src/metric.cpp:39 */
int orc_metric_target(int nv, double target_factor) {
  return (int)floor(nv * target_factor);
}

/* Metrics::normalizeLocally, src/metric.cpp:100-150. */
void orc_normalize_local(int nv, double* tensor, int lp_norm) {
  const double ex = orc_metric_exponent(lp_norm);
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < nv; ++i) {
    double* t = tensor + 3 * (size_t)i;
    double m[4] = {t[0], t[1], t[1], t[2]};
    double val[2] = {0, 0}, vec[4] = {0, 0, 0, 0}, h[4];
    orc_eigen(m, val, vec, vec + 2);
    val[0] = dmax(fabs(val[0]), ORC_EPSILON);       /* metric.cpp:120-121 */
    val[1] = dmax(fabs(val[1]), ORC_EPSILON);
    const double det = val[0] * val[1];             /* metric.cpp:128-131 */
    const double scale_loc = pow(det, ex);
    val[0] *= scale_loc;
    val[1] *= scale_loc;
    rebuild(val, vec, h);                           /* metric.cpp:134-148 */
    t[0] = h[0];
    t[1] = h[1];
    t[2] = h[3];
  }
}

/* Metrics::computeComplexity, src/metric.cpp:153-185, in the element order
 * of a single thread (the reference adds per-thread partials in an OpenMP
 * critical section, so its order depends on the thread count). */
double orc_complexity(int nt, const double* xy, const int32_t* tri,
                      const double* tensor) {
  double phi = 0.;
  for (int i = 0; i < nt; ++i) {
    const int32_t* v = tri + 3 * (size_t)i;
    double p[6];
    for (int j = 0; j < 3; ++j) {
      p[2 * j] = xy[2 * (size_t)v[j]];
      p[2 * j + 1] = xy[2 * (size_t)v[j] + 1];
    }
    double s0 = p[2] - p[0];
    double s1 = p[4] - p[0];
    double s2 = p[3] - p[1];
    double s3 = p[5] - p[1];
    const double area = 0.5 * (s0 * s3 - s1 * s2);
    double rho = 0.;
    for (int j = 0; j < 3; ++j) {
      const double* t = tensor + 3 * (size_t)v[j];
      const double det = t[0] * t[2] - t[1] * t[1];
      rho += sqrt(det);
    }
    rho /= 3;
    phi += rho * area;
  }
  return 0. + phi;                                  /* metric.cpp:182-183 */
}

/* Metrics::normalizeGlobally, src/metric.cpp:188-238, with
 * scale_fact = target / complexity (:197-198), eigen.min = 1/h_min^2 and
 * eigen.max = 1/h_max^2 (:45-46; the names are swapped w.r.t. meaning). */
void orc_normalize_global(int nv, double* tensor, double scale_fact,
                          double h_min, double h_max) {
  const double emin = 1 / (h_min * h_min);
  const double emax = 1 / (h_max * h_max);
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < nv; ++i) {
    double* t = tensor + 3 * (size_t)i;
    double m[4] = {t[0], t[1], t[1], t[2]};
    double val[2], vec[4], h[4];
    orc_eigen(m, val, vec, vec + 2);
    val[0] *= scale_fact;
    val[1] *= scale_fact;
    for (int j = 0; j < 2; ++j) {                   /* metric.cpp:213-216 */
      val[j] = dmin(val[j], emin);
      val[j] = dmax(val[j], emax);
    }
    rebuild(val, vec, h);
    t[0] = h[0];
    t[1] = h[1];
    t[2] = h[3];
  }
}

/* Metrics::run, src/metric.cpp:59-70. */
void orc_metric_run(int nv, int nt, const double* xy, const int32_t* tri,
                    const double* solut,
                    const int32_t* v2t_off, const int32_t* v2t,
                    const int32_t* v2v_off, const int32_t* v2v,
                    double target_factor, int lp_norm,
                    double h_min, double h_max,
                    double* tensor, double* gradient, double* hessian,
                    double* complexity) {
  double* g = gradient;
  if (!g) g = (double*)malloc(sizeof(double) * 2 * (size_t)nv);
  orc_gradient(nv, xy, tri, solut, v2t_off, v2t, v2v_off, v2v, g);
  orc_hessian(nv, xy, tri, g, v2t_off, v2t, v2v_off, v2v, tensor);
  if (hessian) memcpy(hessian, tensor, sizeof(double) * 3 * (size_t)nv);
  orc_normalize_local(nv, tensor, lp_norm);
  double c = orc_complexity(nt, xy, tri, tensor);
  double fact = orc_metric_target(nv, target_factor) / c;
  orc_normalize_global(nv, tensor, fact, h_min, h_max);
  if (complexity) *complexity = c;
  if (!gradient) free(g);
}

/* ------------------------------------------------------------------------ */
/* Partit::extractColoring, src/partition.cpp:79-197, as run by one thread:
 * first-fit in index order (:94-111) never produces a conflict, so the
 * repair rounds (:126-168) do not run. */
int orc_color_firstfit(int nv, const int32_t* v2v_off, const int32_t* v2v,
                       int max_part, int32_t* color) {
  if (max_part < 6) max_part = 6;                   /* partition.cpp:27 */
  int* forbidden = (int*)malloc(sizeof(int) * (size_t)max_part);
  for (int c = 0; c < max_part; ++c) forbidden[c] = INT_MAX;
  memset(color, 0, sizeof(int32_t) * (size_t)nv);
  int nparts = 0;
  for (int i = 0; i < nv; ++i) {
    if (v2v_off[i] == v2v_off[i + 1]) continue;
    for (int a = v2v_off[i]; a < v2v_off[i + 1]; ++a) {
      int c = color[v2v[a]];
      if (c < max_part) forbidden[c] = i;
    }
    for (int c = 1; c < max_part; ++c)
      if (forbidden[c] != i) { color[i] = c; break; }
    if (color[i] > nparts) nparts = color[i];
  }
  free(forbidden);
  return nparts;
}

int orc_color_valid(int nv, const int32_t* v2v_off, const int32_t* v2v,
                    const int32_t* color, int ncolors) {
  for (int i = 0; i < nv; ++i) {
    if (v2v_off[i] == v2v_off[i + 1]) continue;
    if (color[i] < 1 || color[i] > ncolors) return 0;
    for (int a = v2v_off[i]; a < v2v_off[i + 1]; ++a)
      if (color[v2v[a]] == color[i]) return 0;
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Smooth::cacheQuality, src/smoothing.cpp:145-155 with
 * Mesh::computeQuality, src/mesh.cpp:626-644. */
void orc_quality_cache(int nt, const double* xy, const int32_t* tri,
                       const double* tensor, double* qualit) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int t = 0; t < nt; ++t) {
    const int32_t* n = tri + 3 * (size_t)t;
    if (n[0] < 0) continue;
    qualit[t] = orc_quality(xy + 2 * (size_t)n[0], xy + 2 * (size_t)n[1],
                            xy + 2 * (size_t)n[2], tensor + 3 * (size_t)n[0],
                            tensor + 3 * (size_t)n[1],
                            tensor + 3 * (size_t)n[2]);
  }
}

#define ORC_MAX_DEG 256

/* Smooth::moveSmartLaplacian, src/smoothing.cpp:40-126. */
int orc_move_vertex(int i, double* xy, double* tensor, double* qualit,
                    const int32_t* tri,
                    const int32_t* v2t_off, const int32_t* v2t,
                    const int32_t* v2v_off, const int32_t* v2v) {
  const int t0 = v2t_off[i], deg = v2t_off[i + 1] - v2t_off[i];
  const int n0 = v2v_off[i], nb = v2v_off[i + 1] - v2v_off[i];
  double q_min = DBL_MAX;                           /* smoothing.cpp:48-51 */
  for (int k = 0; k < deg; ++k) q_min = dmin(q_min, qualit[v2t[t0 + k]]);

  double metric[3 * ORC_MAX_DEG];
  double q[ORC_MAX_DEG];
  double* pa = xy + 2 * (size_t)i;
  double* ta = tensor + 3 * (size_t)i;
  double opt[2] = {0, 0};
  for (int k = 0; k < nb; ++k) {                    /* smoothing.cpp:61-79 */
    const int j = v2v[n0 + k];
    const double* pb = xy + 2 * (size_t)j;
    const double* tb = tensor + 3 * (size_t)j;
    const double len = orc_riemann_dist(pa, pb, ta, tb);
    opt[0] += pa[0] + ((pb[0] - pa[0]) / len);
    opt[1] += pa[1] + ((pb[1] - pa[1]) / len);
    memcpy(metric + 3 * k, tb, sizeof(double) * 3);
  }
  opt[0] /= nb;
  opt[1] /= nb;
  const double pini[2] = {pa[0], pa[1]};            /* smoothing.cpp:82-84 */
  const double tini[3] = {ta[0], ta[1], ta[2]};
  orc_interpolate_tensor(metric, ta, nb);

  int valid, retry = 0;                             /* smoothing.cpp:92-114 */
  double w = 1;
  do {
    valid = 1;
    pa[0] = (1 - w) * pini[0] + w * opt[0];
    pa[1] = (1 - w) * pini[1] + w * opt[1];
    for (int k = 0; valid && k < deg; ++k) {
      const int32_t* n = tri + 3 * (size_t)v2t[t0 + k];
      const double* a = xy + 2 * (size_t)n[0];
      const double* b = xy + 2 * (size_t)n[1];
      const double* c = xy + 2 * (size_t)n[2];
      valid = 0;
      if (orc_ccw(a, b, c)) {
        q[k] = orc_quality(a, b, c, tensor + 3 * (size_t)n[0],
                           tensor + 3 * (size_t)n[1],
                           tensor + 3 * (size_t)n[2]);
        valid = (q[k] > q_min);
      }
    }
    w *= 0.5;
  } while (!valid && ++retry < 5);

  if (valid) {                                      /* smoothing.cpp:116-123 */
    for (int k = 0; k < deg; ++k) qualit[v2t[t0 + k]] = q[k];
  } else {
    pa[0] = pini[0]; pa[1] = pini[1];
    ta[0] = tini[0]; ta[1] = tini[1]; ta[2] = tini[2];
  }
  return valid;
}

/* Smooth::run, src/smoothing.cpp:274-302 (preProcess :129-142,
 * movePoints :158-180), with the colouring given as input. */
void orc_smooth_run(int nv, int nt, double* xy, double* tensor,
                    double* qualit, const int32_t* tri, const uint8_t* tags,
                    const int32_t* v2t_off, const int32_t* v2t,
                    const int32_t* v2v_off, const int32_t* v2v,
                    const int32_t* color, int ncolors, int sweeps,
                    int recompute_qualit, int32_t* commits, int32_t* tasks) {
  if (recompute_qualit) orc_quality_cache(nt, xy, tri, tensor, qualit);
  /* colour classes in ascending vertex order (partition.cpp:176-194) */
  int32_t* cnt = (int32_t*)calloc((size_t)ncolors + 2, sizeof(int32_t));
  for (int i = 0; i < nv; ++i)
    if (color[i] >= 1 && color[i] <= ncolors) cnt[color[i] + 1]++;
  for (int c = 1; c <= ncolors; ++c) cnt[c + 1] += cnt[c];
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
  int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * ((size_t)ncolors + 2));
  memcpy(cur, cnt, sizeof(int32_t) * ((size_t)ncolors + 2));
  for (int i = 0; i < nv; ++i)
    if (color[i] >= 1 && color[i] <= ncolors) order[cur[color[i]]++] = i;
  for (int s = 0; s < sweeps; ++s) {
    int commit = 0, task = 0;
    for (int c = 1; c <= ncolors; ++c) {
      /* vertices of one colour share no triangle: any order gives the same
       * result, so the class may run in parallel */
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : commit, task) num_threads(g_threads)
      for (int k = cnt[c]; k < cnt[c + 1]; ++k) {
        const int i = order[k];
        if (v2t_off[i] == v2t_off[i + 1]) continue;
        if (tags[i] & 4) continue;                  /* smoothing.cpp:137 */
        commit += orc_move_vertex(i, xy, tensor, qualit, tri, v2t_off, v2t,
                                  v2v_off, v2v);
        task++;
      }
    }
    if (commits) commits[s] = commit;
    if (tasks) tasks[s] = task;
  }
  free(cnt); free(order); free(cur);
}

/* ------------------------------------------------------------------------ */
/* splitmix64 finaliser (Steele, Lea, Flood 2014), the colouring's seeded
 * vertex priority. */
uint32_t orc_color_priority(uint64_t seed, uint32_t v) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)v + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (uint32_t)(z >> 32);
}

static uint64_t mis_key(uint64_t seed, int32_t v) {
  return ((uint64_t)orc_color_priority(seed, (uint32_t)v) << 32) | (uint32_t)v;
}

static int cmp_key_desc(const void* a, const void* b) {
  const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? 1 : (x > y ? -1 : 0);
}

/* Partit::extractIndepSet (src/partition.cpp:200-306) semantics on the
 * subgraph Coarse::extractSubGraph builds (src/coarsening.cpp:307-319),
 * visited sequentially in decreasing (priority, id) order. */
int orc_indep_set(int nv, const int32_t* v2v_off, const int32_t* v2v,
                  const int32_t* nodes, int n, uint64_t seed, int32_t* in_set) {
  int8_t* state = (int8_t*)calloc((size_t)(nv > 0 ? nv : 1), 1); /* 1 sub, 2 in */
  uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
  for (int k = 0; k < n; ++k) {
    state[nodes[k]] = 1;
    keys[k] = mis_key(seed, nodes[k]);
  }
  qsort(keys, (size_t)n, sizeof(uint64_t), cmp_key_desc);
  for (int k = 0; k < n; ++k) {
    const int32_t v = (int32_t)(keys[k] & 0xFFFFFFFFu);
    int free_ = 1;
    for (int a = v2v_off[v]; a < v2v_off[v + 1]; ++a)
      if (state[v2v[a]] == 2) { free_ = 0; break; }
    if (free_) state[v] = 2;
  }
  int size = 0;
  for (int k = 0; k < n; ++k) {
    in_set[k] = state[nodes[k]] == 2;
    size += in_set[k];
  }
  free(state);
  free(keys);
  return size;
}

/* ------------------------------------------------------------------------ */
/* numeric::approxRiemannDist, 3-argument form (src/numeric.cpp:449-459). */
static double riemann_dist1(const double* pa, const double* pb, const double* t) {
  const double u0 = pb[0] - pa[0], u1 = pb[1] - pa[1];
  const double v0 = t[0] * u0 + t[1] * u1;
  const double v1 = t[1] * u0 + t[2] * u1;
  return sqrt(u0 * v0 + u1 * v1);
}

/* numeric::computeSteinerPoint, src/numeric.cpp:387-445 (coef = 1). */
void orc_steiner_point(const double* pa, const double* pb, const double* ta,
                       const double* tb, double* point, double* tensor) {
  const double l0 = riemann_dist1(pa, pb, ta);         /* :403-404 */
  const double l1 = riemann_dist1(pa, pb, tb);
  const double coef = 1;
  const double w = fabs(l0 - l1) < ORC_EPSILON            /* :408-412 */
                       ? 0.5
                       : (sqrt(coef) * sqrt(l0 * l1) - l1) / (coef * l0 - l1);
  point[0] = (1 - w) * pa[0] + (w * pb[0]);            /* :415-416 */
  point[1] = (1 - w) * pa[1] + (w * pb[1]);
  double m[6];
  memcpy(m, ta, sizeof(double) * 3);                    /* :428-430 */
  memcpy(m + 3, tb, sizeof(double) * 3);
  orc_interpolate_tensor(m, tensor, 2);
}

/* Mesh::getElemNeigh, src/mesh.cpp:437-453: the triangle other than id that
 * holds the edge j -> i, searched among i's incident triangles in order. */
static int elem_neigh(int id, int i, int j, const int32_t* tri,
                      const int32_t* v2t_off, const int32_t* v2t) {
  for (int k = v2t_off[i]; k < v2t_off[i + 1]; ++k) {
    const int t = v2t[k];
    if (t == id) continue;
    const int32_t* n = tri + 3 * (size_t)t;
    if ((n[0] == j && n[1] == i) || (n[1] == j && n[2] == i) || (n[2] == j && n[0] == i))
      return t;
  }
  return -1;
}

int orc_refine_filter(int nt, const double* xy, const int32_t* tri,
                      const double* tensor, const int32_t* v2t_off,
                      const int32_t* v2t, double l_max, int32_t* pattern,
                      int32_t* edges, int32_t* elems, int cap,
                      int32_t counts[4]) {
  int eval = 0, adds = 0, split = 0, tasks = 0;
  for (int i = 0; i < nt; ++i) {                        /* refinement.cpp:228-255 */
    pattern[i] = 0;
    const int32_t* n = tri + 3 * (size_t)i;
    if (n[0] < 0) continue;
    double len[3];
    for (int j = 0; j < 3; ++j) {
      const int k = (j + 1) % 3;
      len[j] = orc_riemann_dist(xy + 2 * (size_t)n[j], xy + 2 * (size_t)n[k],
                                tensor + 3 * (size_t)n[j], tensor + 3 * (size_t)n[k]);
    }
    for (int j = 0; j < 3; ++j) {
      const int k = (j + 1) % 3;
      if (len[j] > l_max) {
        ++pattern[i];
        const int opp = elem_neigh(i, n[j], n[k], tri, v2t_off, v2t);
        if (n[j] < n[k] || opp < 0) {
          if (split >= cap) return -1;
          edges[3 * split] = n[j];
          edges[3 * split + 1] = n[k];
          edges[3 * split + 2] = opp;
          ++split;
        }
      }
    }
    if (pattern[i]) {
      adds += pattern[i];
      elems[tasks++] = i;
    }
    ++eval;
  }
  counts[0] = eval; counts[1] = adds; counts[2] = split; counts[3] = tasks;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Host libm values (glibc) of the transcendentals on the path, for
 * comparing the device math library against what the reference links:
 * op 0 log, 1 exp, 2 pow(x, y), 3 pow(x, 3), 4 sqrt. */
void orc_libm(int op, int64_t n, const double* x, const double* y, double* out) {
  volatile double three = 3.;  /* keep pow(x, 3) a libm call */
  for (int64_t i = 0; i < n; ++i) {
    switch (op) {
      case 0: out[i] = log(x[i]); break;
      case 1: out[i] = exp(x[i]); break;
      case 2: out[i] = pow(x[i], y[i]); break;
      case 3: out[i] = pow(x[i], three); break;
      default: out[i] = sqrt(x[i]); break;
    }
  }
}
