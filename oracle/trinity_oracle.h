/*
 * trinity_oracle.h -- CPU restatement (plain C, FP64) of trinity's
 * metric-and-smoothing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in trinity_amd/.  Only the test-suite, the smoke entry
 * point and the CPU-baseline leg of the benchmark may load it.  The product
 * path never calls into it.
 *
 * Every function restates one routine of hobywan/trinity v1.0.3 and cites
 * the file:line it follows.  The operation order (accumulation order, operand
 * order of every product and sum) is kept identical to the reference so that
 * a build with -ffp-contract=off is bit-identical to the reference built with
 * the same flag (checked against oracle/_ref, see tests/test_oracle_ref.py).
 *
 * Conventions: xy[2*nv] interleaved (x,y); tri[3*nt] 0-based, counter-
 * clockwise; tensors[3*nv] = (m11, m12, m22); CSR adjacency as
 *   v2t_off[nv+1], v2t[...] : incident triangles, ascending per vertex
 *   v2v_off[nv+1], v2v[...] : 1-ring neighbours, ascending unique
 */
#ifndef TRINITY_ORACLE_H
#define TRINITY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_EPSILON 1.e-9   /* include/trinity/header.h:50 */

void   orc_set_threads(int n);
int    orc_get_threads(void);

/* ---- topology: Mesh::rebuildTopology (src/mesh.cpp:191-249) ------------ */
/* v2t: incident triangles per vertex, ascending.  Triangles whose first
 * index is negative are deleted and skipped (mesh.cpp:200).  Returns 0. */
int    orc_build_v2t(int nv, int nt, const int32_t* tri,
                     int32_t* v2t_off, int32_t* v2t);
/* v2v: sorted-unique 1-ring (mesh.cpp:214-243).  cap = capacity of v2v.
 * Returns the number of entries written, or -1 if cap is too small. */
int64_t orc_build_v2v(int nv, const int32_t* tri,
                      const int32_t* v2t_off, const int32_t* v2t,
                      int32_t* v2v_off, int32_t* v2v, int64_t cap);

/* ---- Mesh::getVicinity(i, 2) element patch (src/mesh.cpp:390-434) ------- */
/* Ascending unique union of the incident triangles of i and of its 1-ring.
 * Returns the patch size, or -1 when it exceeds cap. */
int    orc_patch2(int i, const int32_t* v2t_off, const int32_t* v2t,
                  const int32_t* v2v_off, const int32_t* v2v,
                  int32_t* out, int cap);

/* ---- recovery (src/hessian.cpp) ----------------------------------------- */
void   orc_gradient(int nv, const double* xy, const int32_t* tri,
                    const double* solut,
                    const int32_t* v2t_off, const int32_t* v2t,
                    const int32_t* v2v_off, const int32_t* v2v,
                    double* gradient);
void   orc_hessian(int nv, const double* xy, const int32_t* tri,
                   const double* gradient,
                   const int32_t* v2t_off, const int32_t* v2t,
                   const int32_t* v2v_off, const int32_t* v2v,
                   double* tensor);

/* ---- tensor numerics (src/numeric.cpp) ---------------------------------- */
void   orc_eigen(const double m[4], double val[2], double vec1[2],
                 double vec2[2]);
void   orc_interpolate_tensor(const double* m, double* result, int nb);
double orc_quality(const double* pa, const double* pb, const double* pc,
                   const double* ta, const double* tb, const double* tc);
double orc_riemann_dist(const double* pa, const double* pb,
                        const double* ta, const double* tb);
int    orc_ccw(const double* pa, const double* pb, const double* pc);

/* ---- metric (src/metric.cpp) -------------------------------------------- */
double orc_metric_exponent(int lp_norm);
void   orc_normalize_local(int nv, double* tensor, int lp_norm);
double orc_complexity(int nt, const double* xy, const int32_t* tri,
                      const double* tensor);
void   orc_normalize_global(int nv, double* tensor, double scale_fact,
                            double h_min, double h_max);
int    orc_metric_target(int nv, double target_factor);
/* Whole Metrics::run (metric.cpp:59-70).  gradient/hessian may be NULL.
 * Returns the complexity through *complexity. */
void   orc_metric_run(int nv, int nt, const double* xy, const int32_t* tri,
                      const double* solut,
                      const int32_t* v2t_off, const int32_t* v2t,
                      const int32_t* v2v_off, const int32_t* v2v,
                      double target_factor, int lp_norm,
                      double h_min, double h_max,
                      double* tensor, double* gradient, double* hessian,
                      double* complexity);

/* ---- colouring (src/partition.cpp:79-197), single-thread semantics ----- */
/* Sequential first-fit in index order with colours 1..max_part-1.
 * Vertices without neighbours keep 0.  Returns the number of colours. */
int    orc_color_firstfit(int nv, const int32_t* v2v_off, const int32_t* v2v,
                          int max_part, int32_t* color);
/* 1 when no edge joins two equal colours and every non-isolated vertex has a
 * colour in [1, ncolors]; 0 otherwise. */
int    orc_color_valid(int nv, const int32_t* v2v_off, const int32_t* v2v,
                       const int32_t* color, int ncolors);

/* ---- smoothing (src/smoothing.cpp) -------------------------------------- */
void   orc_quality_cache(int nt, const double* xy, const int32_t* tri,
                         const double* tensor, double* qualit);
int    orc_move_vertex(int i, double* xy, double* tensor, double* qualit,
                       const int32_t* tri,
                       const int32_t* v2t_off, const int32_t* v2t,
                       const int32_t* v2v_off, const int32_t* v2v);
/* Smooth::run with an externally supplied colouring (color[i] in 1..ncolors,
 * 0 = never moved).  Boundary vertices (tags bit 2) are never moved.
 * commits/tasks receive per-sweep counters (may be NULL).  If
 * recompute_qualit is non-zero the element quality cache is rebuilt first
 * (smoothing.cpp:145-155); otherwise qualit is used as given. */
void   orc_smooth_run(int nv, int nt, double* xy, double* tensor,
                      double* qualit, const int32_t* tri, const uint8_t* tags,
                      const int32_t* v2t_off, const int32_t* v2t,
                      const int32_t* v2v_off, const int32_t* v2v,
                      const int32_t* color, int ncolors, int sweeps,
                      int recompute_qualit, int32_t* commits, int32_t* tasks);

/* ---- independent set for coarsening (src/partition.cpp:200-306) ------- */
/* The colouring's seeded priority (splitmix64 finaliser, high word). */
uint32_t orc_color_priority(uint64_t seed, uint32_t v);
/* Maximal independent set of the subgraph induced by the n distinct
 * vertices `nodes` in the 1-ring graph: greedy in decreasing (priority, id)
 * order, i.e. a vertex is in the set iff no neighbour in the subgraph that
 * precedes it in that order is.  in_set[k] = 1 if nodes[k] is in the set.
 * Returns the set size. */
int    orc_indep_set(int nv, const int32_t* v2v_off, const int32_t* v2v,
                     const int32_t* nodes, int n, uint64_t seed, int32_t* in_set);

/* ---- refinement filter (src/refinement.cpp:216-296) --------------------- */
/* numeric::computeSteinerPoint (src/numeric.cpp:387-445): point of edge
 * (pa, pb) splitting it evenly in the metric, and its log-Euclidean tensor. */
void   orc_steiner_point(const double* pa, const double* pb, const double* ta,
                         const double* tb, double* point, double* tensor);
/* Refine::filterElems (src/refinement.cpp:216-265) run by one thread:
 * pattern[t] = number of edges of active triangle t longer than l_max
 * (Mesh::computeLength, 4-argument Riemann distance); edges (3 ints each:
 * v1, v2, opposite triangle or -1) lists an edge from triangle t when it is
 * long and v1 < v2 or it has no opposite triangle, in (t, edge) order;
 * elems lists the triangles with pattern > 0.  counts = {evaluated, sum of
 * patterns, edges, elems}.  Returns -1 if more than cap edges. */
int    orc_refine_filter(int nt, const double* xy, const int32_t* tri,
                         const double* tensor, const int32_t* v2t_off,
                         const int32_t* v2t, double l_max, int32_t* pattern,
                         int32_t* edges, int32_t* elems, int cap,
                         int32_t counts[4]);

/* ---- host libm probe ---------------------------------------------------- */
/* op 0 log, 1 exp, 2 pow(x, y), 3 pow(x, 3), 4 sqrt, as linked by the
 * reference (glibc). */
void   orc_libm(int op, int64_t n, const double* x, const double* y,
                double* out);

#ifdef __cplusplus
}
#endif

/* Metric-space quality energy E = sum_t (1 - q_M(t))^2 of the metric that
 * orc_metric_run builds from (xy, f), and its reverse-mode gradient
 * (trinity_adjoint.c).  tensor (3nv) and qualit (nt) are optional outputs. */
double orc_energy(int nv, int nt, const double* xy, const int32_t* tri,
                  const double* f, const int32_t* v2t_off, const int32_t* v2t,
                  const int32_t* v2v_off, const int32_t* v2v,
                  double target_factor, int lp_norm, double h_min,
                  double h_max, double* tensor, double* qualit);
double orc_energy_grad(int nv, int nt, const double* xy, const int32_t* tri,
                       const double* f, const int32_t* v2t_off,
                       const int32_t* v2t, const int32_t* v2v_off,
                       const int32_t* v2v, double target_factor, int lp_norm,
                       double h_min, double h_max, double* dxy, double* df);

#endif
