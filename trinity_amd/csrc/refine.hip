// refine.hip -- the metric's consumers on the refinement side: the long-edge
// filter of Refine::filterElems (src/refinement.cpp:216-265) and the Steiner
// points of Refine::computeSteinerPoints (:267-296, Mesh::computeSteinerPoint
// src/mesh.cpp:684-697, numeric::computeSteinerPoint src/numeric.cpp:387-445).
//
// The reference fills per-thread heaps and concatenates them in thread order;
// here a count pass, two scans and a fill pass list the edges and triangles
// in (triangle, edge) order, the order a one-thread run of the reference
// gives.  Edge lengths are Mesh::computeLength (the 4-argument Riemann
// distance the smoothing uses).
#include <string>

#include "trn_api.hpp"
#include "trn_numeric.hpp"

namespace trn {

constexpr int kRefB = 256;

// Mesh::getElemNeigh (src/mesh.cpp:437-453): the triangle other than id
// holding the edge j -> i, searched among i's incident triangles in
// ascending order; -1 when there is none (a boundary edge).
__device__ __forceinline__ int32_t elem_neigh(int32_t id, int32_t i, int32_t j,
                                              const int32_t* __restrict__ tri,
                                              const int32_t* __restrict__ v2t_off,
                                              const int32_t* __restrict__ v2t) {
  for (int32_t k = v2t_off[i]; k < v2t_off[i + 1]; ++k) {
    const int32_t t = v2t[k];
    if (t == id) continue;
    const int32_t a = tri[3 * (int64_t)t], b = tri[3 * (int64_t)t + 1],
                  c = tri[3 * (int64_t)t + 2];
    if ((a == j && b == i) || (b == j && c == i) || (c == j && a == i)) return t;
  }
  return -1;
}

// numeric::approxRiemannDist, 3-argument form (src/numeric.cpp:449-459).
__device__ __forceinline__ double riemann_dist1(const double* pa, const double* pb,
                                                const double* t) {
  const double u0 = pb[0] - pa[0], u1 = pb[1] - pa[1];
  const double v0 = t[0] * u0 + t[1] * u1;
  const double v1 = t[1] * u0 + t[2] * u1;
  return __builtin_sqrt(u0 * v0 + u1 * v1);
}

// numeric::computeSteinerPoint (src/numeric.cpp:387-445) with coef = 1.
__device__ __forceinline__ void steiner_point(const double* pa, const double* pb,
                                              const double* ta, const double* tb,
                                              double point[2], double tensor[3]) {
  const double l0 = riemann_dist1(pa, pb, ta);  // :403-404
  const double l1 = riemann_dist1(pa, pb, tb);
  const double coef = 1;
  const double w = fabs(l0 - l1) < kEps  // :408-412
                       ? 0.5
                       : (__builtin_sqrt(coef) * __builtin_sqrt(l0 * l1) - l1) /
                             (coef * l0 - l1);
  point[0] = (1 - w) * pa[0] + (w * pb[0]);  // :415-416
  point[1] = (1 - w) * pa[1] + (w * pb[1]);
  double lm[4] = {0., 0., 0., 0.}, c[4];  // interpolateTensor(metric, result, 2)
  tensor_log(ta[0], ta[1], ta[2], c);
  lm[0] += c[0]; lm[1] += c[1]; lm[2] += c[2]; lm[3] += c[3];
  tensor_log(tb[0], tb[1], tb[2], c);
  lm[0] += c[0]; lm[1] += c[1]; lm[2] += c[2]; lm[3] += c[3];
  tensor_exp_mean(lm, 2, tensor);
}

// Per triangle: pattern (long edges, refinement.cpp:232-246), which of its
// edges it lists (long, and v1 < v2 or no opposite triangle), and the
// per-triangle counts the scans turn into list positions.  code[t] =
// pattern | listed-edge mask << 2 | 32 for an active triangle.
__global__ __launch_bounds__(kRefB) void k_refine_count(
    int32_t nt, const int32_t* __restrict__ tri, const double* __restrict__ xy,
    const double* __restrict__ tensor, const int32_t* __restrict__ v2t_off,
    const int32_t* __restrict__ v2t, double l_max, int32_t* __restrict__ pattern,
    int32_t* __restrict__ code, int32_t* __restrict__ nedge, int32_t* __restrict__ nbad,
    int32_t* __restrict__ sums) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int pat = 0, active = 0;
  if (t < nt) {
    const int32_t n[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
    int mask = 0;
    if (n[0] >= 0) {  // Mesh::isActiveElem
      active = 1;
      double len[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int k = (j + 1) % 3;
        len[j] = riemann_dist(xy + 2 * (int64_t)n[j], xy + 2 * (int64_t)n[k],
                              tensor + 3 * (int64_t)n[j], tensor + 3 * (int64_t)n[k]);
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int k = (j + 1) % 3;
        if (len[j] > l_max) {
          ++pat;
          if (n[j] < n[k] || elem_neigh((int32_t)t, n[j], n[k], tri, v2t_off, v2t) < 0)
            mask |= 1 << j;
        }
      }
    }
    pattern[t] = pat;
    code[t] = pat | (mask << 2) | (active << 5);
    nedge[t] = __popc(mask);
    nbad[t] = pat > 0;
  }
  // evaluated triangles and the sum of the patterns: one atomic per block
  const int s_pat = __syncthreads_count(pat >= 1) + __syncthreads_count(pat >= 2) +
                    __syncthreads_count(pat >= 3);
  const int s_act = __syncthreads_count(active);
  if (threadIdx.x == 0) {
    if (s_act) atomicAdd(&sums[0], s_act);
    if (s_pat) atomicAdd(&sums[1], s_pat);
  }
}

// Writes the listed edges (v1, v2, opposite) with their Steiner points and
// the triangles to cut, at the positions the scans gave.
__global__ __launch_bounds__(kRefB) void k_refine_fill(
    int32_t nt, const int32_t* __restrict__ tri, const double* __restrict__ xy,
    const double* __restrict__ tensor, const int32_t* __restrict__ v2t_off,
    const int32_t* __restrict__ v2t, const int32_t* __restrict__ code,
    const int32_t* __restrict__ eoff, const int32_t* __restrict__ boff,
    int32_t* __restrict__ edges, int32_t* __restrict__ elems, double* __restrict__ sxy,
    double* __restrict__ sten) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int32_t c = code[t];
  if (!(c & 32)) return;
  if (c & 3) elems[boff[t]] = (int32_t)t;
  const int mask = (c >> 2) & 7;
  if (!mask) return;
  const int32_t n[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
  int64_t e = eoff[t];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (!((mask >> j) & 1)) continue;
    const int k = (j + 1) % 3;
    edges[3 * e] = n[j];
    edges[3 * e + 1] = n[k];
    edges[3 * e + 2] = elem_neigh((int32_t)t, n[j], n[k], tri, v2t_off, v2t);
    if (sxy) {
      double p[2], m[3];
      steiner_point(xy + 2 * (int64_t)n[j], xy + 2 * (int64_t)n[k],
                    tensor + 3 * (int64_t)n[j], tensor + 3 * (int64_t)n[k], p, m);
      sxy[2 * e] = p[0];
      sxy[2 * e + 1] = p[1];
      sten[3 * e] = m[0];
      sten[3 * e + 1] = m[1];
      sten[3 * e + 2] = m[2];
    }
    ++e;
  }
}

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace trn

using namespace trn;

extern "C" {

int trn_refine_filter(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                      const double* tensor_d, double l_max, int32_t* pattern_d,
                      int32_t* edges_d, int32_t* elems_d, double* steiner_xy_d,
                      double* steiner_tensor_d, int32_t edge_cap, int32_t counts[4]) {
  return guard_impl("trn_refine_filter", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(mesh && xy_d && tensor_d && pattern_d && edges_d && elems_d && counts,
                TRN_ERR_ARG, "null argument");
    TRN_REQUIRE((steiner_xy_d == nullptr) == (steiner_tensor_d == nullptr), TRN_ERR_ARG,
                "give both Steiner outputs or neither");
    TRN_REQUIRE(edge_cap >= 0, TRN_ERR_ARG, "negative edge capacity");
    const int32_t nt = mesh->nt;
    if (nt == 0) {
      for (int k = 0; k < 4; ++k) counts[k] = 0;
      return;
    }
    Scratch sc(ctx, 5 * scratch_bytes(sizeof(int32_t) * ((size_t)nt + 1)) +
                        exclusive_scan_scratch(nt) + scratch_bytes(sizeof(int32_t) * 4));
    int32_t* code = sc.carve<int32_t>((size_t)nt + 1);
    int32_t* nedge = sc.carve<int32_t>((size_t)nt + 1);
    int32_t* nbad = sc.carve<int32_t>((size_t)nt + 1);
    int32_t* eoff = sc.carve<int32_t>((size_t)nt + 1);
    int32_t* boff = sc.carve<int32_t>((size_t)nt + 1);
    int32_t* tmp = sc.carve<int32_t>(exclusive_scan_scratch(nt) / 4);
    int32_t* sums = sc.carve<int32_t>(4);
    memset_async(ctx, sums, 0, sizeof(int32_t) * 4);
    launch(ctx, "refine_count", 0.0, k_refine_count, dim3(nblk(nt, kRefB)), dim3(kRefB), 0, nt,
           (const int32_t*)mesh->tri, xy_d, tensor_d, (const int32_t*)mesh->v2t_off,
           (const int32_t*)mesh->v2t, l_max, pattern_d, code, nedge, nbad, sums);
    exclusive_scan_i32(ctx, nedge, eoff, nt, true, tmp);
    exclusive_scan_i32(ctx, nbad, boff, nt, true, tmp);
    int32_t* h = ctx->pinned_i32;
    copy_async(ctx, h, sums, sizeof(int32_t) * 2);
    copy_async(ctx, h + 2, eoff + nt, sizeof(int32_t));
    copy_async(ctx, h + 3, boff + nt, sizeof(int32_t));
    sync(ctx);
    // refinement.cpp:256-264: evaluated, added (sum of patterns), split, tasks
    for (int k = 0; k < 4; ++k) counts[k] = h[k];
    TRN_REQUIRE(h[2] <= edge_cap, TRN_ERR_CAPACITY,
                std::to_string(h[2]) + " edges to split, capacity " + std::to_string(edge_cap));
    launch(ctx, "refine_fill", 0.0, k_refine_fill, dim3(nblk(nt, kRefB)), dim3(kRefB), 0, nt,
           (const int32_t*)mesh->tri, xy_d, tensor_d, (const int32_t*)mesh->v2t_off,
           (const int32_t*)mesh->v2t, (const int32_t*)code, (const int32_t*)eoff,
           (const int32_t*)boff, edges_d, elems_d, steiner_xy_d, steiner_tensor_d);
  });
}

}  // extern "C"
