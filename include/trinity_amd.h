/*
 * trinity_amd.h -- C ABI of the MI355X (gfx950) implementation of
 * hobywan/trinity's metric-and-smoothing path.
 *
 * The reference is a C++ library whose path is driven through
 *   trinity::Metrics(Mesh*, double target, int Lp, double h_min, double h_max)
 *       .run(Stats*)                                   include/trinity/metric.h:38-44
 *   trinity::Partit(Mesh const*, int max_part).extractColoring(const Mesh*)
 *                                                      include/trinity/partition.h:46-49
 *   trinity::Smooth(Mesh*, Partit*, int level).run(Stats*)
 *                                                      include/trinity/smoothing.h:39-42
 * on the arrays of trinity::Mesh (include/trinity/mesh.h:108-149).  The
 * entry points below replace those passes one for one; the drop-in sources
 * in trinity_amd/dropin/ re-implement the three classes on top of them.
 *
 * Conventions
 *  - Every function returns TRN_OK (0) or a negative TRN_ERR_* code and
 *    never throws.  trn_last_error() gives a message for the calling thread.
 *  - Pointers suffixed _d are DEVICE pointers (caller-owned).  Arrays use
 *    the reference layout: xy[2nv] interleaved (x,y) = Mesh::geom.points,
 *    tensor[3nv] = (m11, m12, m22) = Mesh::geom.tensor, tri[3nt] 0-based
 *    = Mesh::topo.elems, tags[nv] = Mesh::sync.tags (bit 2 = boundary),
 *    qualit[nt] = Mesh::geom.qualit, solut[nv] = Mesh::geom.solut.
 *  - All work is ordered on the context's stream (trn_ctx_set_stream);
 *    functions that return host values synchronise that stream first.
 *  - Everything is FP64; results follow the reference's operation order
 *    and are built without floating-point contraction.
 */
#ifndef TRINITY_AMD_H
#define TRINITY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRN_OK            0
#define TRN_ERR_ARG      -1   /* invalid argument (null pointer, size < 0, ...) */
#define TRN_ERR_HIP      -2   /* a HIP runtime call failed */
#define TRN_ERR_CAPACITY -3   /* stencil larger than the kernels support */
#define TRN_ERR_PALETTE  -4   /* colouring needs more colours than allowed */
#define TRN_ERR_STATE    -5   /* internal inconsistency */

typedef struct trn_ctx trn_ctx;    /* device, stream, scratch, timers */
typedef struct trn_mesh trn_mesh;  /* device-resident topology (CSR) */

/* Library version string. */
const char* trn_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char* trn_last_error(void);

/* ---- context ----------------------------------------------------------- */
/* Creates a context on `device` with its own default-priority stream. */
int trn_ctx_create(int device, trn_ctx** out);
int trn_ctx_destroy(trn_ctx* ctx);
/* Routes all work to `hip_stream` (a hipStream_t; NULL = the null stream),
 * e.g. the stream a host framework allocates on. */
int trn_ctx_set_stream(trn_ctx* ctx, void* hip_stream);
/* Routes all work back to the context's own stream (the default). */
int trn_ctx_use_own_stream(trn_ctx* ctx);
int trn_ctx_synchronize(trn_ctx* ctx);
/* Per-kernel timing with hipEvents around every launch (off by default). */
int trn_ctx_set_timing(trn_ctx* ctx, int enable);
/* With timing on, time only the launches of the statistic `name` (as
 * trn_ctx_timing_get reports it, e.g. "smooth_phase"); NULL or "" times
 * every launch (the default).  The events cost a few microseconds per timed
 * launch, so a caller that wants one kernel's durations can leave the others
 * unbracketed. */
int trn_ctx_set_timing_only(trn_ctx* ctx, const char* name);
int trn_ctx_reset_timing(trn_ctx* ctx);
/* Number of timed kernels; then, per index, name, total ms, launches and
 * algorithmic bytes credited (synchronises the stream). */
int trn_ctx_timing_count(trn_ctx* ctx, int32_t* count);
int trn_ctx_timing_get(trn_ctx* ctx, int32_t index, const char** name,
                       double* total_ms, int64_t* launches, double* bytes);

/* ---- input: the parse of Mesh::load, and a binary format ------------------
 * trn_medit_info / trn_medit_read replace the parse of Mesh::load
 * (src/io.cpp:25-48 io::find, :53-114) with a multi-threaded one of the same
 * text files: sections searched in the order Vertices, Edges, Triangles,
 * Corners; "x y ref" vertices; "a b ref" edges whose ends get tag 4
 * (boundary); "a b c ref" triangles, 1-based in the file, 0-based in tri;
 * corners tagged 4|8 after the edges; the .bb field from its second line
 * (bb_path and solut may both be NULL).  Coordinates and values are the
 * reference's bit for bit (correctly rounded conversions).  Records must be
 * one per line; another layout, an index outside 1..nv or counts other than
 * nv / nt return TRN_ERR_ARG.  Host arrays: xy [2nv], tri [3nt], tags [nv],
 * solut [nv].  The topology is then built by trn_mesh_create
 * (Mesh::rebuildTopology).  No GPU is needed for these two.
 *
 * TRNM is this library's binary form of the same data: "TRNM", int32
 * version 1, int32 nv, int32 nt, f64 xy[2nv], int32 tri[3nt] (0-based),
 * u8 tags[nv], f64 solut[nv], little-endian.  trn_trnm_read fills host
 * arrays (ctx NULL) or device arrays (ctx given: read through pinned
 * buffers, each copy overlapping the next read; returns when the copies are
 * done); solut may be NULL.  trn_trnm_write writes host arrays. */
int trn_medit_info(const char* mesh_path, int32_t* nv, int32_t* nt);
int trn_medit_read(const char* mesh_path, const char* bb_path, int32_t nv, int32_t nt,
                   double* xy, int32_t* tri, uint8_t* tags, double* solut);
int trn_trnm_info(const char* path, int32_t* nv, int32_t* nt);
int trn_trnm_read(trn_ctx* ctx, const char* path, int32_t nv, int32_t nt, double* xy,
                  int32_t* tri, uint8_t* tags, double* solut);
int trn_trnm_write(const char* path, int32_t nv, int32_t nt, const double* xy,
                   const int32_t* tri, const uint8_t* tags, const double* solut);

/* ---- topology: Mesh::load + Mesh::rebuildTopology ------------------------
 * Replaces the topology part of Mesh::load (src/io.cpp:97-99, :128-129) and
 * Mesh::rebuildTopology (src/mesh.cpp:191-249): builds, on the device,
 *   v2t  incident triangles per vertex, ascending  (Mesh::topo.stenc/sync.deg)
 *   v2v  sorted-unique 1-ring                      (Mesh::topo.vicin)
 * tri (3nt, 0-based; a negative first index marks a deleted triangle as in
 * mesh.cpp:200) and tags (nv) may be host or device pointers; they are
 * copied into the mesh.  Also measures the largest 2-ring element patch
 * (Mesh::getVicinity(i,2), src/mesh.cpp:390-434) the recovery will gather.
 * xy_d (device, optional): coordinates used only for memory locality: the
 * stored 2-ring patches are grouped along a Z-curve of them (64 vertices
 * that are close in the plane share their staged terms), and
 * trn_mesh_set_class_order can visit colour classes in that order; NULL
 * keeps index order.  Results do not depend on it. */
int trn_mesh_create(trn_ctx* ctx, int32_t nv, int32_t nt, const int32_t* tri,
                    const uint8_t* tags, const double* xy_d, trn_mesh** out);
int trn_mesh_destroy(trn_mesh* mesh);
/* Order in which the colour classes (trn_color, trn_color_classes) list
 * their vertices: zcurve = 0 ascending (the default), 1 along the Z-curve of
 * the coordinates given to trn_mesh_create (TRN_ERR_ARG if none were).  The
 * classes hold the same vertices either way. */
int trn_mesh_set_class_order(trn_mesh* mesh, int32_t zcurve);
/* Sizes: nnz of v2v, max triangles per vertex, max 1-ring, max 2-ring patch. */
int trn_mesh_info(const trn_mesh* mesh, int64_t* nnz_v2v, int32_t* max_deg_t,
                  int32_t* max_deg_v, int32_t* max_patch);
/* Device pointers of the mesh-owned CSR arrays (valid until destroy). */
int trn_mesh_csr(const trn_mesh* mesh, const int32_t** v2t_off_d,
                 const int32_t** v2t_d, const int32_t** v2v_off_d,
                 const int32_t** v2v_d);

/* ---- metric: Metrics::run and its stages ----------------------------------
 * trn_metric_run replaces Metrics::run (src/metric.cpp:59-70) with
 * Metrics(mesh, target_factor, lp_norm, h_min, h_max): L2-projection
 * gradient (src/hessian.cpp:24-83) and Hessian (:86-153) over the 2-ring
 * patch, local L^p normalisation (src/metric.cpp:100-150), complexity
 * (:153-185) and global normalisation with eigenvalue clamps (:188-238).
 * Writes the metric into tensor_d [3nv].  complexity (host) may be NULL; if
 * given the call synchronises.  gradient_d [2nv] and hessian_d [3nv]
 * (symmetrised raw Hessian) are optional stage outputs (NULL = not kept). */
int trn_metric_run(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                   const double* solut_d, double target_factor,
                   int32_t lp_norm, double h_min, double h_max,
                   double* tensor_d, double* complexity,
                   double* gradient_d, double* hessian_d);
/* Metrics::computeGradient for every vertex (src/hessian.cpp:24-83). */
int trn_gradient(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                 const double* solut_d, double* gradient_d);
/* Metrics::computeHessian for every vertex (src/hessian.cpp:86-153). */
int trn_hessian(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                const double* gradient_d, double* tensor_d);
/* Metrics::normalizeLocally (src/metric.cpp:100-150), in place. */
int trn_normalize_local(trn_ctx* ctx, int32_t nv, double* tensor_d,
                        int32_t lp_norm);
/* Metrics::computeComplexity (src/metric.cpp:153-185); deterministic
 * two-level tree sum.  Synchronises. */
int trn_complexity(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                   const double* tensor_d, double* complexity);
/* Metrics::normalizeGlobally (src/metric.cpp:188-238) with
 * scale_fact = target / complexity, in place. */
int trn_normalize_global(trn_ctx* ctx, int32_t nv, double* tensor_d,
                         double scale_fact, double h_min, double h_max);

/* ---- colouring: Partit::extractColoring ------------------------------------
 * Replaces Partit::extractColoring (src/partition.cpp:79-197) with a
 * deterministic colouring of the 1-ring graph: a sequential first fit in
 * index order inside tiles of consecutive vertex ids (the reference's
 * one-thread colouring wherever an edge stays inside a tile), then the
 * conflicts on edges between tiles removed and repaired by seeded
 * speculative-greedy rounds.  Every vertex with a non-empty 1-ring gets a
 * colour in 1..max_colors such that no edge joins two equal colours (the
 * reference uses 1..max_part-1 with Partit(mesh, 8), i.e. at most 7).
 * Returns TRN_ERR_PALETTE if a vertex cannot be coloured within max_colors
 * (<= 64); see trn_color_partial for the reference's behaviour instead.
 * Also emits the colour classes (Partit::task.subset / cardin / nb.parts):
 * order_d [nv] lists the coloured vertices grouped by colour (within a
 * colour in the mesh's locality order, else ascending); class_off (host,
 * max_colors+2 entries) receives the offsets of colours 1..ncolors in
 * order_d.  The result is a pure function of (mesh, seed, tile size).
 * Synchronises. */
int trn_color(trn_ctx* ctx, const trn_mesh* mesh, uint64_t seed,
              int32_t max_colors, int32_t* color_d, int32_t* order_d,
              int32_t* class_off, int32_t* ncolors);
/* trn_color with the reference's palette behaviour (partition.cpp:104-110):
 * a vertex whose neighbours already hold every colour 1..max_colors stays
 * uncoloured (colour 0, in no class, so Smooth never moves it,
 * smoothing.cpp:170) instead of failing the call; *uncoloured (may be
 * NULL) receives the number of such vertices.  Synchronises. */
int trn_color_partial(trn_ctx* ctx, const trn_mesh* mesh, uint64_t seed,
                      int32_t max_colors, int32_t* color_d, int32_t* order_d,
                      int32_t* class_off, int32_t* ncolors, int64_t* uncoloured);
/* Tile size (vertices) of the colouring's first step on this mesh: 0
 * chooses it from the longest edge of the numbering (the default: just
 * longer, so that no edge joins two tiles of the same parity), otherwise a
 * multiple of 64 in 64..16384.  A tile shorter than the longest edge leaves
 * conflicts to the repair rounds; a longer one runs longer. */
int trn_mesh_set_color_tile(trn_mesh* mesh, int32_t tile);
/* ---- launch shapes: the autotuner ------------------------------------------
 * No counterpart in the reference (its loops have no launch shapes).
 * trn_mesh_autotune measures, on this mesh with the caller's coordinates and
 * field (lp_norm for the metric the smoothing runs on), the slabs per block
 * (1, 2, 4) of the gradient and Hessian gathers and the register shape (2,
 * 3, 4 waves per SIMD) of the smoothing phase kernel (on the largest colour
 * class, first sweep), and keeps the fastest in the mesh for every later
 * call.  shapes [3] (may be NULL) receives {gradient slabs, Hessian slabs,
 * phase waves}; times_ms [6] (may be NULL) the default and the chosen
 * shape's time of each of the three.  No shape changes a result; the
 * mesh's smoothing records are restored.  Per-kernel timing must be off.
 * Synchronises.  trn_mesh_shapes / trn_mesh_set_shapes read and set the
 * same three values (defaults 1, 1, 3). */
int trn_mesh_autotune(trn_ctx* ctx, trn_mesh* mesh, const double* xy_d, const double* solut_d,
                      int32_t lp_norm, int32_t* shapes, double* times_ms);
int trn_mesh_shapes(const trn_mesh* mesh, int32_t* shapes);
int trn_mesh_set_shapes(trn_mesh* mesh, const int32_t* shapes);

/* Colour classes of an externally supplied colouring (colour 0 = not
 * coloured), e.g. the reference's own.  class_off as above.  mesh may be
 * NULL (classes in index order) or the mesh whose locality order to use.
 * Synchronises. */
int trn_color_classes(trn_ctx* ctx, const trn_mesh* mesh, int32_t nv,
                      const int32_t* color_d, int32_t ncolors, int32_t* order_d,
                      int32_t* class_off);
/* Replaces Partit::extractIndepSet (src/partition.cpp:200-306) on the
 * subgraph Coarse::extractSubGraph builds (src/coarsening.cpp:307-319): the
 * n_nodes distinct vertices nodes_d (device) with their 1-rings, neighbours
 * outside the list never conflicting.  set_d (device, [n_nodes]) receives a
 * maximal independent set of that subgraph, in input order, *set_size its
 * size.  The set is the greedy one in decreasing (seeded priority, id)
 * order (the priority of trn_color), a pure function of (mesh, nodes,
 * seed); the reference's depends on its thread interleaving.  Synchronises. */
int trn_indep_set(trn_ctx* ctx, const trn_mesh* mesh, const int32_t* nodes_d, int32_t n_nodes,
                  uint64_t seed, int32_t* set_d, int32_t* set_size);
/* Counts edges whose end points share a colour and non-isolated vertices
 * left uncoloured or outside 1..ncolors (0 and 0 = valid).  Synchronises. */
int trn_color_check(trn_ctx* ctx, const trn_mesh* mesh, const int32_t* color_d,
                    int32_t ncolors, int64_t* bad_edges, int64_t* bad_vertices);

/* ---- refinement filter: Refine::filterElems + Steiner points -----------
 * Replaces Refine::filterElems (src/refinement.cpp:216-265) and the point
 * computation of Refine::computeSteinerPoints (:267-296;
 * Mesh::computeSteinerPoint src/mesh.cpp:684-697, numeric::computeSteinerPoint
 * src/numeric.cpp:387-445).  For every active triangle, pattern_d[t]
 * (device, [nt]) = number of its edges longer than l_max (the reference's
 * trinity::l_max = sqrt 2, header.h:67) under Mesh::computeLength; an edge
 * (v1, v2) = (n[j], n[j+1]) that is long is listed once, from the triangle
 * where v1 < v2 or from its only triangle, as (v1, v2, opposite triangle or
 * -1) in edges_d (device, [3 * edge_cap]); elems_d (device, [nt]) lists the
 * triangles with pattern > 0.  Lists are in (triangle, edge) order, the
 * order of a one-thread run of the reference.  steiner_xy_d ([2 * edge_cap])
 * and steiner_tensor_d ([3 * edge_cap]), both or neither, receive each
 * listed edge's Steiner point and tensor.  counts (host) = {evaluated
 * triangles, sum of patterns, listed edges, listed triangles} (nb.eval,
 * nb.adds, nb.split, nb.tasks); they are written before TRN_ERR_CAPACITY is
 * returned for more than edge_cap edges.  Synchronises. */
int trn_refine_filter(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                      const double* tensor_d, double l_max, int32_t* pattern_d,
                      int32_t* edges_d, int32_t* elems_d, double* steiner_xy_d,
                      double* steiner_tensor_d, int32_t edge_cap, int32_t counts[4]);

/* ---- smoothing: Smooth::run ---------------------------------------------
 * Smooth::cacheQuality (src/smoothing.cpp:145-155): qualit_d[t] =
 * Mesh::computeQuality(t) (src/mesh.cpp:626-644, numeric.cpp:251-306). */
int trn_quality_cache(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                      const double* tensor_d, double* qualit_d);
/* Replaces Mesh::computeQuality(double quality[3]) (src/mesh.cpp:647-681):
 * rebuilds qualit_d as trn_quality_cache does, then folds the active
 * triangles' qualities into the host array as the reference does:
 * quality[0] = min(quality[0], min q), quality[1] = max(quality[1], max q),
 * quality[2] = (quality[2] + sum q) / count (the reference starts its
 * partial extremes at 3 and 0).  *active (may be NULL) receives the count
 * (Mesh::nb.activ_elem).  The sum is a fixed-order tree, so the mean may
 * differ from a given OpenMP run of the reference in the last bits, as
 * runs of the reference with different thread counts do.  Synchronises. */
int trn_quality_stats(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                      const double* tensor_d, double* qualit_d, double quality[3],
                      int64_t* active);
/* Replaces Smooth::run (src/smoothing.cpp:274-302) after its colouring:
 * `sweeps` rounds (Smooth's `level`) of Smooth::movePoints (:158-180), each
 * visiting colours 1..ncolors in ascending order and applying
 * moveSmartLaplacian (:40-126) to every non-boundary vertex of the class.
 * order_d/class_off as produced by trn_color or trn_color_classes.  If
 * recompute_qualit != 0 the quality cache is rebuilt first (the reference
 * rebuilds it once per Smooth::run, :290).  commits/tasks (host, [sweeps],
 * may be NULL) receive the per-sweep accepted-move and task counts; when
 * either is given the call synchronises.  The sweeps work on a per-mesh
 * copy of the vertex data: smoothing calls on one mesh must not overlap
 * (issue them on one stream, or synchronise between them). */
int trn_smooth_run(trn_ctx* ctx, const trn_mesh* mesh, double* xy_d,
                   double* tensor_d, double* qualit_d, const int32_t* order_d,
                   const int32_t* class_off, int32_t ncolors, int32_t sweeps,
                   int32_t recompute_qualit, int32_t* commits, int32_t* tasks);
/* trn_smooth_run in two parts, equal to it when called in sequence:
 * trn_smooth_prepare loads xy_d and tensor_d into the mesh's vertex records
 * and, if recompute_qualit != 0, rebuilds the quality cache
 * (Smooth::cacheQuality, src/smoothing.cpp:145-155, :290); it needs no
 * colouring, so a caller can issue it before the colouring is done.
 * trn_smooth_run_prepared runs the sweeps on the records; every committed
 * move also writes its vertex's coordinates and metric into xy_d and
 * tensor_d, which must be the arrays given to trn_smooth_prepare (their
 * other entries already hold the records' values).  Between the two, xy_d,
 * tensor_d and qualit_d must not change and no other smoothing call may
 * use the mesh. */
int trn_smooth_prepare(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                       const double* tensor_d, double* qualit_d,
                       int32_t recompute_qualit);
int trn_smooth_run_prepared(trn_ctx* ctx, const trn_mesh* mesh, double* xy_d,
                            double* tensor_d, double* qualit_d, const int32_t* order_d,
                            const int32_t* class_off, int32_t ncolors, int32_t sweeps,
                            int32_t* commits, int32_t* tasks);

/* trn_smooth_run_prepared without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and counters as
 * trn_smooth_run_prepared (the same kernels); its scratch (per-sweep
 * counters, stale flags, live lists: about nv + 4 * largest class bytes) is
 * a fixed 64 MiB device array per GPU, held until the process ends, and a
 * mesh needing more returns TRN_ERR_CAPACITY.  Every call on one GPU shares
 * that array, so calls that overlap on one GPU (other streams, other
 * contexts, other threads) race; calls in stream order do not.  Per-kernel
 * timing must be off. */
int trn_smooth_run_prepared_noalloc(trn_ctx* ctx, const trn_mesh* mesh, double* xy_d,
                                    double* tensor_d, double* qualit_d, const int32_t* order_d,
                                    const int32_t* class_off, int32_t ncolors, int32_t sweeps,
                                    int32_t* commits, int32_t* tasks);

/* ---- batched form: many meshes of uneven sizes -----------------------------
 * No counterpart in the reference, which processes one mesh per Metrics /
 * Smooth object.  A batch is the disjoint union of nmesh meshes; one pass
 * performs, for every mesh, what Metrics::run (src/metric.cpp:59-70) and
 * Smooth::run after its colouring (src/smoothing.cpp:287-297) do, with the
 * per-mesh target floor(nv_m * target_factor) (metric.cpp:39) and per-mesh
 * complexity.  Every launch of the pass covers all meshes.  Each mesh's
 * results are bit-identical to a plain pass over that mesh (trn_metric_run,
 * then trn_smooth_run fed the batch colouring restricted to the mesh): the
 * per-mesh complexity is summed in the plain path's order. */
typedef struct trn_batch trn_batch;
/* nv[m], nt[m] (host): sizes; tri (host): the meshes' triangle arrays
 * concatenated, each with its own 0-based vertex indices; tags (host or
 * device): concatenated per-vertex tags. */
int trn_batch_create(trn_ctx* ctx, int32_t nmesh, const int32_t* nv,
                     const int32_t* nt, const int32_t* tri, const uint8_t* tags,
                     trn_batch** out);
/* trn_batch_create with the meshes' coordinates (host or device, 2 doubles
 * per vertex, concatenated like tags), used only for memory locality as in
 * trn_mesh_create: the stored 2-ring patches of each mesh are grouped along
 * the Z-curve of that mesh's own bounding box.  Results do not depend on
 * it. */
int trn_batch_create_xy(trn_ctx* ctx, int32_t nmesh, const int32_t* nv,
                        const int32_t* nt, const int32_t* tri, const uint8_t* tags,
                        const double* xy, trn_batch** out);
int trn_batch_destroy(trn_batch* batch);
/* Totals and the union topology (its CSR via trn_mesh_csr). */
int trn_batch_info(const trn_batch* batch, int32_t* nv_total, int32_t* nt_total,
                   const trn_mesh** mesh);
/* Colours all meshes at once (trn_color on the union) and keeps the classes
 * for the sweeps.  Synchronises. */
int trn_batch_color(trn_ctx* ctx, trn_batch* batch, uint64_t seed,
                    int32_t max_colors, int32_t* ncolors);
/* The kept colouring: device colour [nv] and class list [nv] (valid until
 * the next trn_batch_color or destroy), host class offsets (ncolors + 1
 * entries, may be NULL) and the colour count; any output may be NULL. */
int trn_batch_coloring(const trn_batch* batch, const int32_t** color_d,
                       const int32_t** order_d, int32_t* class_off, int32_t* ncolors);
/* One pass over every mesh.  Arrays are the concatenations over meshes in
 * batch order (device).  xy_in_d (may be NULL: xy_d holds the input) are the
 * input coordinates; the pass copies them into xy_d, which the metric stages
 * read and the sweeps move.  complexity_d (device, [nmesh], may be NULL)
 * receives the per-mesh complexities; commits/tasks (host, [sweeps], may be
 * NULL) the per-sweep counters summed over the meshes (then the call
 * synchronises).  sweeps > 0 needs trn_batch_color first. */
int trn_batch_pass(trn_ctx* ctx, trn_batch* batch, const double* xy_in_d,
                   double* xy_d, const double* solut_d, double target_factor,
                   int32_t lp_norm, double h_min, double h_max, double* tensor_d,
                   double* qualit_d, int32_t sweeps, double* complexity_d,
                   int32_t* commits, int32_t* tasks);
/* The same pass replayed from a hipGraph captured on the first call with
 * these arguments (re-captured when they, the stream or the colouring
 * change).  Needs a created stream (not the null stream) and timing off. */
int trn_batch_pass_graph(trn_ctx* ctx, trn_batch* batch, const double* xy_in_d,
                         double* xy_d, const double* solut_d,
                         double target_factor, int32_t lp_norm, double h_min,
                         double h_max, double* tensor_d, double* qualit_d,
                         int32_t sweeps);

/* ---- adjoint: gradient of a metric-space quality energy -------------------
 * No counterpart in the reference.  The energy is
 *     E(xy, f) = sum over active triangles t of (1 - q_M(t))^2
 * with q_M numeric::computeQuality (src/numeric.cpp:251-306) and M the metric
 * Metrics::run (src/metric.cpp:59-70) builds from the coordinates xy and the
 * nodal values f, i.e. exactly what trn_metric_run returns.  The backward
 * gives dE/dxy and dE/df through the global and local normalisation, the
 * complexity and the Hessian and gradient recoveries (src/hessian.cpp),
 * differentiating the reference's branches (eigen shortcuts, clamps) as
 * they are taken; triangle geometry is recomputed, not stored. */
typedef struct trn_adjoint trn_adjoint;
/* Tape and scratch for one mesh (about 200 B per vertex and 270 B per
 * triangle of device memory). */
int trn_adjoint_create(trn_ctx* ctx, const trn_mesh* mesh, trn_adjoint** out);
int trn_adjoint_destroy(trn_adjoint* adjoint);
/* Forward: builds the metric, the element qualities and E (host), keeping
 * the tape and the xy_d / f_d pointers, which must stay unchanged until the
 * backward.  tensor_d (3nv) and qualit_d (nt) are optional device copies of
 * the metric and the qualities.  Synchronises. */
int trn_adjoint_forward(trn_ctx* ctx, trn_adjoint* adjoint, const double* xy_d,
                        const double* f_d, double target_factor, int32_t lp_norm,
                        double h_min, double h_max, double* energy,
                        double* tensor_d, double* qualit_d);
/* Backward of the last forward: dE/dxy (device, 2nv) and dE/df (device, nv);
 * either may be NULL.  Synchronises. */
int trn_adjoint_backward(trn_ctx* ctx, trn_adjoint* adjoint, double* dxy_d,
                         double* df_d);
/* Gradient-based smoothing: iters steps of diagonally preconditioned
 * steepest descent on E over the vertices Smooth::run may move
 * (src/smoothing.cpp:134-138: not boundary), the direction at vertex i being
 * -h_i^2 dE/dx_i with h_i the mean length of its edges, each step with a
 * largest displacement of `step` halved until E decreases and no triangle
 * inverts.  Moves xy_d in place; energies (host, iters + 1, may be
 * NULL) receives E before and after each step; accepted (host) the number
 * of steps taken.  Stops early when no step decreases E. */
int trn_adjoint_smooth(trn_ctx* ctx, trn_adjoint* adjoint, double* xy_d,
                       const double* f_d, double target_factor, int32_t lp_norm,
                       double h_min, double h_max, int32_t iters, double step,
                       double* energies, int32_t* accepted);

/* ---- device math probe ----------------------------------------------------
 * Evaluates the device transcendental used by the path on n inputs
 * (op: 0 log, 1 exp, 2 pow(x, y), 3 pow(x, 3), 4 sqrt) -- lets tests compare
 * the device library against the host libm that the reference links. */
int trn_debug_math(trn_ctx* ctx, int32_t op, int64_t n, const double* x_d,
                   const double* y_d, double* out_d);

#ifdef __cplusplus
}
#endif
#endif /* TRINITY_AMD_H */
