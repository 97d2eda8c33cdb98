// topology.hip -- device construction of the mesh adjacency, replacing the
// topology part of Mesh::load / Mesh::rebuildTopology
// (src/io.cpp:97-99,128-129; src/mesh.cpp:191-249):
//   v2t : incident triangles per vertex, ascending (topo.stenc + sync.deg)
//   v2v : sorted-unique 1-ring                     (topo.vicin)
// plus the size of the largest 2-ring element patch (Mesh::getVicinity,
// src/mesh.cpp:390-434), which sizes the LDS sets of the recovery gathers.
//
// Also hosts the device-wide exclusive scan and max-reduction helpers.
#include <string>

#include "trn_api.hpp"

namespace trn {

// ---------------------------------------------------------------------------
// scan: reduce-then-scan over tiles of 256 threads x 8 items

constexpr int kScanB = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanB * kScanItems;

__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wid] = x;
  __syncthreads();
  int woff = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < kScanB / 64; ++w) {
    const int s = lds[w];
    if (w < wid) woff += s;
    total += s;
  }
  __syncthreads();
  return woff + x - v;
}

__global__ __launch_bounds__(kScanB) void k_scan_reduce(const int32_t* __restrict__ in,
                                                        int64_t n, int32_t* __restrict__ bsum) {
  __shared__ int lds[kScanB / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += in[base + k];
  int total;
  block_exclusive_scan(s, lds, total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// single block: exclusive scan of nb block sums in place
__global__ __launch_bounds__(kScanB) void k_scan_top(int32_t* bsum, int64_t nb) {
  __shared__ int lds[kScanB / 64];
  const int64_t per = (nb + kScanB - 1) / kScanB;
  const int64_t b0 = threadIdx.x * per;
  int s = 0;
  for (int64_t k = 0; k < per; ++k)
    if (b0 + k < nb) s += bsum[b0 + k];
  int total;
  int off = block_exclusive_scan(s, lds, total);
  for (int64_t k = 0; k < per; ++k) {
    if (b0 + k < nb) {
      const int v = bsum[b0 + k];
      bsum[b0 + k] = off;
      off += v;
    }
  }
  if (threadIdx.x == 0) bsum[nb] = total;
}

__global__ __launch_bounds__(kScanB) void k_scan_down(const int32_t* in, int32_t* out,
                                                      int64_t n, const int32_t* __restrict__ bsum,
                                                      int64_t nb, int with_total) {
  __shared__ int lds[kScanB / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int off = block_exclusive_scan(s, lds, total) + bsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = off;
    off += v[k];
  }
  if (with_total && blockIdx.x == 0 && threadIdx.x == 0) out[n] = bsum[nb];
}

size_t exclusive_scan_scratch(int64_t n) {
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  return scratch_bytes(sizeof(int32_t) * (size_t)(nb + 1));
}

void exclusive_scan_i32(trn_ctx* ctx, const int32_t* in, int32_t* out, int64_t n,
                        bool with_total, int32_t* tmp) {
  TRN_REQUIRE(tmp != nullptr, TRN_ERR_STATE, "scan scratch not set");
  if (n <= 0) {
    if (with_total) memset_async(ctx, out, 0, sizeof(int32_t));
    return;
  }
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  launch(ctx, "scan_reduce", 4.0 * n, k_scan_reduce, dim3((unsigned)nb), dim3(kScanB), 0,
         in, n, tmp);
  launch(ctx, "scan_top", 8.0 * nb, k_scan_top, dim3(1), dim3(kScanB), 0, tmp, nb);
  launch(ctx, "scan_down", 8.0 * n, k_scan_down, dim3((unsigned)nb), dim3(kScanB), 0,
         in, out, n, (const int32_t*)tmp, nb, with_total ? 1 : 0);
}

// ---------------------------------------------------------------------------
// reductions

// One atomic per block: many waves hitting a single address serialise in L2.
__global__ __launch_bounds__(256) void k_reduce_max(const int32_t* __restrict__ in, int64_t n,
                                                    int32_t* out) {
  __shared__ int lds[4];
  int m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, in[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(out, max(max(lds[0], lds[1]), max(lds[2], lds[3])));
}

void reduce_max_i32(trn_ctx* ctx, const int32_t* in, int64_t n, int32_t* out_d) {
  memset_async(ctx, out_d, 0, sizeof(int32_t));
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 512) blocks = 512;
  launch(ctx, "reduce_max", 4.0 * n, k_reduce_max, dim3((unsigned)blocks), dim3(256), 0, in, n,
         out_d);
}

// ---------------------------------------------------------------------------
// topology kernels

// Counts triangles with an index outside [0, nv) (deleted ones skipped).
__global__ void k_validate_tri(int32_t nv, int32_t nt, const int32_t* __restrict__ tri,
                               int32_t* bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int32_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
  if (a < 0) return;
  if (a >= nv || b < 0 || b >= nv || c < 0 || c >= nv || a == b || b == c || a == c)
    atomicAdd(bad, 1);
}

// Longest edge at each vertex, |u - v| over its 1-ring.
__global__ void k_vertex_span(int32_t nv, const int32_t* __restrict__ v2v_off,
                              const int32_t* __restrict__ v2v, int32_t* __restrict__ span) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  int32_t s = 0;
  for (int32_t a = v2v_off[v]; a < v2v_off[v + 1]; ++a) {
    const int32_t d = v2v[a] - (int32_t)v;
    const int32_t ad = d < 0 ? -d : d;
    s = ad > s ? ad : s;
  }
  span[v] = s;
}

__global__ void k_fill_v2t(int32_t nt, const int32_t* __restrict__ tri, int32_t* cursor,
                           int32_t* __restrict__ v2t) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int32_t a = tri[3 * t];
  if (a < 0) return;
  v2t[atomicAdd(&cursor[a], 1)] = (int32_t)t;
  v2t[atomicAdd(&cursor[tri[3 * t + 1]], 1)] = (int32_t)t;
  v2t[atomicAdd(&cursor[tri[3 * t + 2]], 1)] = (int32_t)t;
}

// Incident triangles without a separate counting pass: every corner claims
// the next slot of its vertex's row of kV2tRow entries (cnt[v] ends as the
// degree; slots past the row are only counted).
constexpr int kV2tRow = 16;

__global__ void k_fill_v2t_rows(int32_t nt, const int32_t* __restrict__ tri,
                                int32_t* cnt, int32_t* __restrict__ rows) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int32_t a = tri[3 * t];
  if (a < 0) return;  // deleted triangle (mesh.cpp:200)
  const int32_t c[3] = {a, tri[3 * t + 1], tri[3 * t + 2]};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int32_t slot = atomicAdd(&cnt[c[k]], 1);
    if (slot < kV2tRow) rows[(int64_t)c[k] * kV2tRow + slot] = (int32_t)t;
  }
}

// The CSR from the rows when every degree fits a row: each row sorted
// ascending in registers (odd-even transposition network over constant
// indices, padding INT32_MAX; mesh.cpp:244) and written at off[v].
__global__ void k_v2t_from_rows(int32_t nv, const int32_t* __restrict__ cnt,
                                const int32_t* __restrict__ off,
                                const int32_t* __restrict__ rows, int32_t* __restrict__ v2t) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const int32_t n = cnt[v];
  const int32_t* row = rows + v * kV2tRow;
  int32_t r[kV2tRow];
#pragma unroll
  for (int k = 0; k < kV2tRow; ++k) r[k] = k < n ? row[k] : INT32_MAX;
#pragma unroll
  for (int round = 0; round < kV2tRow; ++round)
#pragma unroll
    for (int k = round & 1; k + 1 < kV2tRow; k += 2) {
      const int32_t lo = r[k] < r[k + 1] ? r[k] : r[k + 1];
      const int32_t hi = r[k] < r[k + 1] ? r[k + 1] : r[k];
      r[k] = lo;
      r[k + 1] = hi;
    }
  int32_t* out = v2t + off[v];
#pragma unroll
  for (int k = 0; k < kV2tRow; ++k)
    if (k < n) out[k] = r[k];
}

// Ascending order of each vertex's incident list (mesh.cpp:244).  Lists
// of at most kSortRegs entries are sorted in registers (an odd-even
// transposition network over constant indices, padding INT32_MAX) and
// written back once; longer ones by insertion in memory.
constexpr int kSortRegs = 8;

__global__ void k_sort_lists(int32_t nv, const int32_t* __restrict__ off, int32_t* list) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int32_t b = off[i], e = off[i + 1];
  if (e - b <= kSortRegs) {
    int32_t r[kSortRegs];
#pragma unroll
    for (int k = 0; k < kSortRegs; ++k) r[k] = b + k < e ? list[b + k] : INT32_MAX;
#pragma unroll
    for (int round = 0; round < kSortRegs; ++round)
#pragma unroll
      for (int k = round & 1; k + 1 < kSortRegs; k += 2) {
        const int32_t lo = r[k] < r[k + 1] ? r[k] : r[k + 1];
        const int32_t hi = r[k] < r[k + 1] ? r[k + 1] : r[k];
        r[k] = lo;
        r[k + 1] = hi;
      }
#pragma unroll
    for (int k = 0; k < kSortRegs; ++k)
      if (b + k < e) list[b + k] = r[k];
    return;
  }
  for (int32_t k = b + 1; k < e; ++k) {
    const int32_t x = list[k];
    int32_t j = k - 1;
    while (j >= b && list[j] > x) {
      list[j + 1] = list[j];
      --j;
    }
    list[j + 1] = x;
  }
}

// Sorted-unique 1-ring (mesh.cpp:214-243): pass 0 counts, pass 1 writes.
template <int PASS>
__global__ __launch_bounds__(64) void k_v2v(int32_t nv, const int32_t* __restrict__ tri,
                                            const int32_t* __restrict__ v2t_off,
                                            const int32_t* __restrict__ v2t, int cap,
                                            int32_t* __restrict__ cnt,
                                            const int32_t* __restrict__ v2v_off,
                                            int32_t* __restrict__ v2v, int32_t* overflow) {
  extern __shared__ int lds[];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  int* s = lds + threadIdx.x;
  const int stride = blockDim.x;
  int m = 0;
  bool ok = true;
  for (int32_t k = v2t_off[i]; k < v2t_off[i + 1]; ++k) {
    const int32_t t = v2t[k];
    const int32_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
    int32_t u, w;
    if (a == i) { u = b; w = c; }
    else if (b == i) { u = c; w = a; }
    else { u = a; w = b; }
    ok = ok && lds_sorted_insert(s, stride, m, cap, u);
    ok = ok && lds_sorted_insert(s, stride, m, cap, w);
  }
  if (!ok) {
    atomicAdd(overflow, 1);
    if (PASS == 0) cnt[i] = 0;
    return;
  }
  if (PASS == 0) {
    cnt[i] = m;
  } else {
    int32_t* out = v2v + v2v_off[i];
    for (int k = 0; k < m; ++k) out[k] = s[k * stride];
  }
}

// Counts the entries of a below zero (slabs without a union).
__global__ void k_count_negative(int64_t n, const int32_t* __restrict__ a, int32_t* count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool neg = i < n && a[i] < 0;
  const uint64_t m = __ballot(neg);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (int32_t)__popcll(m));
}

// Neighbours and triangles per neighbour whose loads k_patch_size issues
// as batches.
constexpr int kPatchBatch = 8;

// Upper bound of the 2-ring candidates: deg(i) + sum_{j in N(i)} deg(j).
__global__ void k_patch_bound(int32_t nv, const int32_t* __restrict__ v2t_off,
                              const int32_t* __restrict__ v2v_off,
                              const int32_t* __restrict__ v2v, int32_t* __restrict__ bound) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  int s = v2t_off[i + 1] - v2t_off[i];
  for (int32_t a = v2v_off[i]; a < v2v_off[i + 1]; ++a) {
    const int32_t j = v2v[a];
    s += v2t_off[j + 1] - v2t_off[j];
  }
  bound[i] = s;
}

// Exact 2-ring element patch (getVicinity(i,2), mesh.cpp:390-434) of the
// vertex at slab position pos (vertex slab_v[pos], or pos): its size, and,
// when tmp is given, the ascending patch itself in tmp[e * nv + pos]
// (e < size <= cap), for k_patch_copy.
__global__ __launch_bounds__(64) void k_patch_size(int32_t nv, const int32_t* __restrict__ slab_v,
                                                   const int32_t* __restrict__ v2t_off,
                                                   const int32_t* __restrict__ v2t,
                                                   const int32_t* __restrict__ v2v_off,
                                                   const int32_t* __restrict__ v2v, int cap,
                                                   int32_t* __restrict__ size,
                                                   int32_t* overflow, int32_t* __restrict__ tmp) {
  extern __shared__ int lds[];
  const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= nv) return;
  const int64_t i = slab_v ? slab_v[pos] : pos;
  int* s = lds + threadIdx.x;
  const int stride = blockDim.x;
  int m = 0;
  bool ok = true;
  // the loads go out in three independent batches (own triangles and
  // 1-ring; the neighbours' CSR ranges; their first kPatchBatch triangles)
  // instead of one dependent chain per neighbour; lists longer than the
  // batch finish in loops
  const int32_t o0 = v2t_off[i], o1 = v2t_off[i + 1], n0 = v2v_off[i], n1 = v2v_off[i + 1];
  int32_t own[kPatchBatch], nbr[kPatchBatch], jb[kPatchBatch], je[kPatchBatch];
  int32_t tt[kPatchBatch][kPatchBatch];
#pragma unroll
  for (int q = 0; q < kPatchBatch; ++q) {
    own[q] = o0 + q < o1 ? v2t[o0 + q] : -1;
    nbr[q] = n0 + q < n1 ? v2v[n0 + q] : -1;
  }
#pragma unroll
  for (int q = 0; q < kPatchBatch; ++q) {
    jb[q] = nbr[q] >= 0 ? v2t_off[nbr[q]] : 0;
    je[q] = nbr[q] >= 0 ? v2t_off[nbr[q] + 1] : 0;
  }
#pragma unroll
  for (int q = 0; q < kPatchBatch; ++q)
#pragma unroll
    for (int k = 0; k < kPatchBatch; ++k) tt[q][k] = jb[q] + k < je[q] ? v2t[jb[q] + k] : -1;
#pragma unroll
  for (int q = 0; q < kPatchBatch; ++q)
    if (own[q] >= 0) ok = ok && lds_sorted_insert(s, stride, m, cap, own[q]);
  for (int32_t k = o0 + kPatchBatch; k < o1; ++k)
    ok = ok && lds_sorted_insert(s, stride, m, cap, v2t[k]);
#pragma unroll
  for (int q = 0; q < kPatchBatch; ++q) {
#pragma unroll
    for (int k = 0; k < kPatchBatch; ++k)
      if (tt[q][k] >= 0) ok = ok && lds_sorted_insert(s, stride, m, cap, tt[q][k]);
    for (int32_t k = jb[q] + kPatchBatch; k < je[q]; ++k)
      ok = ok && lds_sorted_insert(s, stride, m, cap, v2t[k]);
  }
  for (int32_t a = n0 + kPatchBatch; a < n1 && ok; ++a) {
    const int32_t j = v2v[a];
    for (int32_t k = v2t_off[j]; k < v2t_off[j + 1]; ++k)
      ok = ok && lds_sorted_insert(s, stride, m, cap, v2t[k]);
  }
  if (!ok) atomicAdd(overflow, 1);
  size[i] = m;
  if (tmp && ok)
    for (int e = 0; e < m; ++e) tmp[(int64_t)e * nv + pos] = s[e * stride];
}

// Rows of each patch slab: the largest patch among its kPatchSlab vertices
// (one wave per slab).
__global__ __launch_bounds__(kPatchSlab) void k_patch_slab_rows(int32_t nv,
                                                               const int32_t* __restrict__ slab_v,
                                                               const int32_t* __restrict__ size,
                                                               int32_t* __restrict__ rows) {
  const int64_t pos = (int64_t)blockIdx.x * kPatchSlab + threadIdx.x;
  int m = pos < nv ? size[slab_v ? slab_v[pos] : pos] : 0;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const int o = __shfl_xor(m, d, 64);
    m = o > m ? o : m;
  }
  if (threadIdx.x == 0) rows[blockIdx.x] = m;
}

// Stored patches: the ascending 2-ring element set of every vertex, as
// k_patch_size builds it, written column-wise into its slab, padded with -1;
// slab position pos holds vertex slab_v[pos] (pos without slab_v).  The
// columns past the last vertex of the last slab are all padding, so every
// entry of the array is written.
__global__ __launch_bounds__(kPatchSlab) void k_patch_fill(
    int32_t nv, const int32_t* __restrict__ slab_v, const int32_t* __restrict__ v2t_off,
    const int32_t* __restrict__ v2t, const int32_t* __restrict__ v2v_off,
    const int32_t* __restrict__ v2v, int cap, const int32_t* __restrict__ row,
    int32_t* __restrict__ ell) {
  extern __shared__ int lds[];
  const int64_t pos = (int64_t)blockIdx.x * kPatchSlab + threadIdx.x;
  const int64_t i = pos < nv ? (slab_v ? slab_v[pos] : pos) : nv;
  int* s = lds + threadIdx.x;
  int m = 0;
  bool ok = true;
  if (i < nv) {
    for (int32_t k = v2t_off[i]; k < v2t_off[i + 1]; ++k)
      ok = ok && lds_sorted_insert(s, kPatchSlab, m, cap, v2t[k]);
    for (int32_t a = v2v_off[i]; a < v2v_off[i + 1] && ok; ++a) {
      const int32_t j = v2v[a];
      for (int32_t k = v2t_off[j]; k < v2t_off[j + 1]; ++k)
        ok = ok && lds_sorted_insert(s, kPatchSlab, m, cap, v2t[k]);
    }
  }
  const int32_t r0 = row[blockIdx.x], w = row[blockIdx.x + 1] - r0;
  int32_t* col = ell + (int64_t)r0 * kPatchSlab + threadIdx.x;
  for (int e = 0; e < w; ++e) col[(int64_t)e * kPatchSlab] = e < m ? s[e * kPatchSlab] : -1;
}

// Largest temporary patch buffer mesh creation allocates (beyond it the
// patches are built a second time for the ELL fill instead of copied).
constexpr size_t kPatchTmpMax = size_t(8) << 30;

// The stored patches from the ones k_patch_size kept in tmp (column-major
// by slab position, tmp[e * nv + pos]), written column-wise into their
// slabs as k_patch_fill writes them.
__global__ __launch_bounds__(kPatchSlab) void k_patch_copy(
    int32_t nv, const int32_t* __restrict__ slab_v, const int32_t* __restrict__ size,
    const int32_t* __restrict__ tmp, const int32_t* __restrict__ row, int32_t* __restrict__ ell) {
  const int64_t pos = (int64_t)blockIdx.x * kPatchSlab + threadIdx.x;
  const int64_t i = pos < nv ? (slab_v ? slab_v[pos] : pos) : nv;
  const int m = i < nv ? size[i] : 0;
  const int32_t r0 = row[blockIdx.x], w = row[blockIdx.x + 1] - r0;
  int32_t* col = ell + (int64_t)r0 * kPatchSlab + threadIdx.x;
  for (int e = 0; e < w; ++e)
    col[(int64_t)e * kPatchSlab] = e < m ? tmp[(int64_t)e * nv + pos] : -1;
}

// Start of each slab's union: the slab's own ELL block (patch_row[b] *
// kPatchSlab), which holds every entry of the slab and so any union of them.
__global__ void k_union_offsets(int32_t nslab, const int32_t* __restrict__ row,
                                int32_t* __restrict__ out) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b <= nslab) out[b] = row[b] * kPatchSlab;
}

// Union of the patches of one slab (one block of kUnionB threads per slab).
// Writes the union size (or -1 when the slab is larger than kUnionKeys
// entries or its union larger than kUnionCap), the ascending union at
// union_off[b] = the slab's ELL offset, and, for every patch entry, its
// position in the union.  The slab's entries go into an open-addressing
// hash set in LDS (the distinct ids, whatever the numbering's spread), each
// distinct id's ascending position is counted among the distinct ids
// (about 300 on a grid slab instead of its 1,536 entries), and each slot
// records its id's position, which the entries then look up.
constexpr int kUnionB = 256;
constexpr int kUnionSlots = kUnionKeys;  // hash slots: one per entry at most
constexpr int kUnionPer = kUnionKeys / kUnionB;      // entries per thread
static_assert((kUnionSlots & (kUnionSlots - 1)) == 0, "slot count must be a power of two");
static_assert(kUnionCap < 65536, "union positions are stored in 16 bits");

__device__ __forceinline__ uint32_t union_slot(const int32_t t) {
  return ((uint32_t)t * 0x9E3779B1u) >> (32 - __builtin_ctz(kUnionSlots));  // Fibonacci hashing
}

__global__ __launch_bounds__(kUnionB) void k_patch_union(
    int32_t nv, const int32_t* __restrict__ row, const int32_t* __restrict__ ell,
    int32_t* __restrict__ union_n, const int32_t* __restrict__ union_off,
    int32_t* __restrict__ union_ids, uint16_t* __restrict__ local) {
  __shared__ int32_t tab[kUnionSlots];    // -1: empty
  __shared__ uint16_t rank[kUnionSlots];  // union position of the id in the slot
  __shared__ int32_t list[kUnionCap];     // the distinct ids
  __shared__ int32_t count;
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
  const int64_t r0 = row[b];
  const int n = (row[b + 1] - (int)r0) * kPatchSlab;
  if (n > kUnionKeys) {
    if (tid == 0) union_n[b] = -1;
    return;
  }
  const int32_t* src = ell + r0 * kPatchSlab;
  const int64_t v0 = (int64_t)b * kPatchSlab;  // first vertex of the slab
  for (int k = tid; k < kUnionSlots; k += kUnionB) tab[k] = -1;
  if (tid == 0) count = 0;
  // entry k = tid + e * kUnionB, kept in registers (-1: padding, or a
  // column past the last vertex)
  int32_t ent[kUnionPer];
#pragma unroll
  for (int e = 0; e < kUnionPer; ++e) {
    const int k = tid + e * kUnionB;
    ent[e] = (k < n && v0 + (k % kPatchSlab) < nv) ? src[k] : -1;
  }
  __syncthreads();
  // insertion; the entry that claims a slot also appends its id to the
  // list of distinct ids, in any order (one LDS atomic per wave and pass)
#pragma unroll
  for (int e = 0; e < kUnionPer; ++e) {
    const int32_t t = ent[e];
    bool fresh = false;
    if (t >= 0) {
      uint32_t h = union_slot(t);
      for (;;) {  // at most kUnionKeys distinct ids in kUnionSlots slots: ends
        const int32_t old = atomicCAS(&tab[h], -1, t);
        if (old == -1) fresh = true;
        if (old == -1 || old == t) break;
        h = (h + 1) & (kUnionSlots - 1);
      }
    }
    const uint64_t m = __ballot(fresh);
    if (m) {
      int base = 0;
      if (lane == 0) base = atomicAdd(&count, __popcll(m));
      base = __shfl(base, 0, 64);
      const int at = base + __popcll(m & ((1ull << lane) - 1ull));
      if (fresh && at < kUnionCap) list[at] = t;
    }
  }
  __syncthreads();
  const int nu = count;
  if (tid == 0) union_n[b] = nu <= kUnionCap ? nu : -1;
  if (nu > kUnionCap) return;  // nu is the block's: every thread leaves
  // ascending position of each distinct id: the number of smaller ones
  // (every thread reads the same list entry at a time, an LDS broadcast;
  // about 300 ids per slab on a grid, where a sorting network would take
  // 45 barrier-separated stages)
  int32_t* U = union_ids + union_off[b];
  for (int k = tid; k < nu; k += kUnionB) {
    const int32_t t = list[k];
    int r = 0;
    for (int j = 0; j < nu; ++j) r += list[j] < t;
    U[r] = t;
    uint32_t h = union_slot(t);
    while (tab[h] != t) h = (h + 1) & (kUnionSlots - 1);  // present: ends
    rank[h] = (uint16_t)r;
  }
  __syncthreads();
  uint16_t* dst = local + r0 * kPatchSlab;
#pragma unroll
  for (int e = 0; e < kUnionPer; ++e) {
    const int k = tid + e * kUnionB;
    if (k >= n) continue;
    const int32_t t = ent[e];
    uint16_t v = 0xFFFF;
    if (t >= 0) {
      uint32_t h = union_slot(t);
      while (tab[h] != t) h = (h + 1) & (kUnionSlots - 1);
      v = rank[h];
    }
    dst[k] = v;
  }
}

// Packed offsets of the slab unions: cnt[b] = union size (0 without one).
__global__ void k_union_sizes(int64_t nslab, const int32_t* __restrict__ union_n,
                              int32_t* __restrict__ cnt) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nslab) cnt[b] = union_n[b] > 0 ? union_n[b] : 0;
}

// The corners of every union entry's triangle, slab b's from union_cu[b]
// (one block per slab).
__global__ __launch_bounds__(256) void k_union_tri(const int32_t* __restrict__ union_n,
                                                   const int32_t* __restrict__ union_off,
                                                   const int32_t* __restrict__ union_ids,
                                                   const int32_t* __restrict__ union_cu,
                                                   const int32_t* __restrict__ tri,
                                                   int32_t* __restrict__ union_tri) {
  const int b = blockIdx.x;
  const int nu = union_n[b];
  if (nu <= 0) return;
  const int32_t* U = union_ids + union_off[b];
  int32_t* C = union_tri + 3 * (int64_t)union_cu[b];
  for (int k = threadIdx.x; k < 3 * nu; k += blockDim.x)
    C[k] = tri[3 * (int64_t)U[k / 3] + k % 3];
}

// Star codes (see trn_mesh::star).  Corners are located in the sorted 1-ring
// by binary search; every corner of an incident triangle is in the 1-ring
// by construction of v2v.
__device__ __forceinline__ int ring_index(const int32_t* __restrict__ ring, int n, int32_t x) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ring[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Star codes of the incident triangles (slot r of the vertex in the
// triangle and the 1-ring indices a, b of its two other corners,
// r | a << 2 | b << 10, one per v2t entry) and the compact topology (VTopo)
// of every vertex, in one pass.  A vertex with at most kStarBatch incident
// triangles and neighbours loads its 1-ring, triangles and their corners in
// independent batches and finds a corner's 1-ring index as the count of
// smaller ring entries (the ring is sorted and holds the corner); larger
// stars search the CSR.
constexpr int kStarBatch = 8;

__global__ __launch_bounds__(256) void k_star_vtopo(int32_t nv, const uint8_t* __restrict__ tags,
                                                    const int32_t* __restrict__ tri,
                                                    const int32_t* __restrict__ v2t_off,
                                                    const int32_t* __restrict__ v2t,
                                                    const int32_t* __restrict__ v2v_off,
                                                    const int32_t* __restrict__ v2v,
                                                    int32_t* __restrict__ star,
                                                    VTopo* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int32_t t0 = v2t_off[i], deg = v2t_off[i + 1] - t0;
  const int32_t n0 = v2v_off[i], nb = v2v_off[i + 1] - n0;
  VTopo r = {};
  r.tag = tags[i];
  r.meta = (deg <= kVtopoDeg && nb <= kVtopoDeg) ? (uint8_t)(deg | nb << 4) : kVtopoGeneric;
  if (deg <= kStarBatch && nb <= kStarBatch) {
    int32_t ring[kStarBatch], tt[kStarBatch], c0[kStarBatch], c1[kStarBatch], c2[kStarBatch];
#pragma unroll
    for (int q = 0; q < kStarBatch; ++q) {
      ring[q] = q < nb ? v2v[n0 + q] : INT32_MAX;
      tt[q] = q < deg ? v2t[t0 + q] : 0;
    }
#pragma unroll
    for (int q = 0; q < kStarBatch; ++q) {
      c0[q] = q < deg ? tri[3 * (int64_t)tt[q]] : 0;
      c1[q] = q < deg ? tri[3 * (int64_t)tt[q] + 1] : 0;
      c2[q] = q < deg ? tri[3 * (int64_t)tt[q] + 2] : 0;
    }
#pragma unroll
    for (int q = 0; q < kStarBatch; ++q) {
      if (q >= deg) continue;
      const int rr = c0[q] == i ? 0 : (c1[q] == i ? 1 : 2);
      const int32_t x = rr == 0 ? c1[q] : (rr == 1 ? c2[q] : c0[q]);  // corner r + 1
      const int32_t y = rr == 0 ? c2[q] : (rr == 1 ? c0[q] : c1[q]);  // corner r + 2
      int a = 0, b = 0;
#pragma unroll
      for (int p = 0; p < kStarBatch; ++p) {
        a += ring[p] < x;
        b += ring[p] < y;
      }
      star[t0 + q] = rr | (a << 2) | (b << 10);
      if (q < kVtopoDeg) r.star[q] = (uint8_t)(rr | ((a & 7) << 2) | ((b & 7) << 5));
    }
    if (r.meta != kVtopoGeneric) {
#pragma unroll
      for (int k = 0; k < kVtopoDeg; ++k) {
        r.ring[k] = k < nb ? ring[k] : (int32_t)i;
        r.tris[k] = k < deg ? tt[k] : 0;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kVtopoDeg; ++k) {
        r.ring[k] = (int32_t)i;
        r.star[k] = 0;
      }
    }
  } else {
    const int32_t* rg = v2v + n0;
    for (int32_t k = t0; k < t0 + deg; ++k) {
      const int32_t t = v2t[k];
      const int32_t c[3] = {tri[3 * (int64_t)t], tri[3 * (int64_t)t + 1], tri[3 * (int64_t)t + 2]};
      const int rr = c[0] == i ? 0 : (c[1] == i ? 1 : 2);
      const int a = ring_index(rg, nb, c[(rr + 1) % 3]);
      const int b = ring_index(rg, nb, c[(rr + 2) % 3]);
      star[k] = rr | (a << 2) | (b << 10);
    }
#pragma unroll
    for (int k = 0; k < kVtopoDeg; ++k) r.ring[k] = (int32_t)i;
  }
  out[i] = r;
}

static unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace trn

using namespace trn;

namespace trn {

// trn_mesh_create's work; the patch slabs follow order_d when given, else the
// Z-curve of xy_d when given, else index order.  Throws on error.
static trn_mesh* mesh_create_impl(trn_ctx* ctx, int32_t nv, int32_t nt, const int32_t* tri,
                                  const uint8_t* tags, const double* xy_d,
                                  const int32_t* order_d) {
  TRN_REQUIRE(nv > 0 && nt > 0, TRN_ERR_ARG, "nv and nt must be positive");
  TRN_REQUIRE(tri != nullptr && tags != nullptr, TRN_ERR_ARG, "null tri or tags");
  TRN_REQUIRE((int64_t)nt * 3 < (int64_t)INT32_MAX, TRN_ERR_ARG, "too many triangles");
  trn_mesh* m = new trn_mesh();
  m->ctx = ctx;
  m->nv = nv;
  m->nt = nt;
  auto cleanup = [&]() { trn_mesh_destroy(m); };
  try {
    TRN_HIP(hipMalloc(&m->tri, sizeof(int32_t) * 3 * (size_t)nt));
    TRN_HIP(hipMalloc(&m->tags, (size_t)nv));
    TRN_HIP(hipMalloc(&m->v2t_off, sizeof(int32_t) * ((size_t)nv + 1)));
    TRN_HIP(hipMalloc(&m->v2t, sizeof(int32_t) * 3 * (size_t)nt));
    TRN_HIP(hipMalloc(&m->v2v_off, sizeof(int32_t) * ((size_t)nv + 1)));
    copy_async(ctx, m->tri, tri, sizeof(int32_t) * 3 * (size_t)nt);
    copy_async(ctx, m->tags, tags, (size_t)nv);

    // 0. locality order (Z-curve of the coordinates): the vertices of a
    //    patch slab, and on request the visiting order of the colour
    //    classes.  It takes the context workspace, so it runs before the
    //    workspace below is carved.
    if (xy_d || order_d) {
      TRN_HIP(hipMalloc(&m->zorder, sizeof(int32_t) * (size_t)nv));
      if (order_d)
        copy_async(ctx, m->zorder, order_d, sizeof(int32_t) * (size_t)nv);
      else
        spatial_order(ctx, nv, xy_d, m->zorder);
      m->slab_v = m->zorder;
    }

    Scratch sc(ctx, 4 * scratch_bytes(sizeof(int32_t) * ((size_t)nv + 1)) +
                        exclusive_scan_scratch(nv) + 256 * 4);
    int32_t* cnt = sc.carve<int32_t>((size_t)nv + 1);
    int32_t* cursor = sc.carve<int32_t>((size_t)nv + 1);
    int32_t* cnt2 = sc.carve<int32_t>((size_t)nv + 1);
    int32_t* bound = sc.carve<int32_t>((size_t)nv + 1);
    int32_t* flags = sc.carve<int32_t>(64);
    int32_t* scan_tmp = sc.carve<int32_t>(exclusive_scan_scratch(nv) / 4);
    int32_t* h = ctx->pinned_i32;

    // 1. validate indices (an index out of range would make the atomics
    //    below write outside their arrays)
    memset_async(ctx, flags, 0, 64 * sizeof(int32_t));
    launch(ctx, "topo_validate", 12.0 * nt, k_validate_tri, dim3(blocks_for(nt, 256)),
           dim3(256), 0, nv, nt, (const int32_t*)m->tri, flags);
    copy_async(ctx, h, flags, sizeof(int32_t));
    sync(ctx);
    TRN_REQUIRE(h[0] == 0, TRN_ERR_ARG,
                std::to_string(h[0]) + " triangles have an index outside [0, nv) or repeat a vertex");

    // 2. v2t CSR: rows of kV2tRow slots filled and counted in one pass;
    //    when a degree exceeds the row, the counts seed the two-pass fill
    //    (positions, then sorting)
    memset_async(ctx, cnt, 0, sizeof(int32_t) * (size_t)nv);
    int32_t* rows = (int32_t*)aux_workspace(ctx, sizeof(int32_t) * kV2tRow * (size_t)nv);
    launch(ctx, "topo_fill_v2t", 36.0 * nt, k_fill_v2t_rows, dim3(blocks_for(nt, 256)),
           dim3(256), 0, nt, (const int32_t*)m->tri, cnt, rows);
    reduce_max_i32(ctx, cnt, nv, flags + 1);
    copy_async(ctx, h, flags + 1, sizeof(int32_t));
    exclusive_scan_i32(ctx, cnt, m->v2t_off, nv, true, scan_tmp);
    sync(ctx);
    m->max_deg_t = h[0];
    if (m->max_deg_t <= kV2tRow) {
      launch(ctx, "topo_sort_v2t", 8.0 * nv + 24.0 * nt, k_v2t_from_rows,
             dim3(blocks_for(nv, 256)), dim3(256), 0, nv, (const int32_t*)cnt,
             (const int32_t*)m->v2t_off, (const int32_t*)rows, m->v2t);
    } else {
      copy_async(ctx, cursor, m->v2t_off, sizeof(int32_t) * (size_t)nv);
      launch(ctx, "topo_fill_v2t", 36.0 * nt, k_fill_v2t, dim3(blocks_for(nt, 256)),
             dim3(256), 0, nt, (const int32_t*)m->tri, cursor, m->v2t);
      launch(ctx, "topo_sort_v2t", 8.0 * nv + 24.0 * nt, k_sort_lists,
             dim3(blocks_for(nv, 256)), dim3(256), 0, nv, (const int32_t*)m->v2t_off, m->v2t);
    }
    TRN_REQUIRE(m->max_deg_t <= kMaxDegree, TRN_ERR_CAPACITY,
                "a vertex has " + std::to_string(m->max_deg_t) +
                    " incident triangles; at most " + std::to_string(kMaxDegree) +
                    " are supported");

    // 3. v2v CSR (count, scan, fill)
    const int capv = 2 * (m->max_deg_t > 0 ? m->max_deg_t : 1);
    const size_t shv = sizeof(int) * 64 * (size_t)capv;
    memset_async(ctx, flags + 2, 0, sizeof(int32_t));
    launch(ctx, "topo_count_v2v", 0.0, k_v2v<0>, dim3(blocks_for(nv, 64)), dim3(64), shv, nv,
           (const int32_t*)m->tri, (const int32_t*)m->v2t_off, (const int32_t*)m->v2t, capv,
           cnt2, (const int32_t*)nullptr, (int32_t*)nullptr, flags + 2);
    exclusive_scan_i32(ctx, cnt2, m->v2v_off, nv, true, scan_tmp);
    reduce_max_i32(ctx, cnt2, nv, flags + 3);
    copy_async(ctx, h, flags + 2, 2 * sizeof(int32_t));
    copy_async(ctx, h + 2, m->v2v_off + nv, sizeof(int32_t));
    sync(ctx);
    TRN_REQUIRE(h[0] == 0, TRN_ERR_STATE, "1-ring overflow");
    m->max_deg_v = h[1];
    m->nnz_v2v = h[2];
    TRN_HIP(hipMalloc(&m->v2v, sizeof(int32_t) * (size_t)(m->nnz_v2v > 0 ? m->nnz_v2v : 1)));
    launch(ctx, "topo_fill_v2v", 0.0, k_v2v<1>, dim3(blocks_for(nv, 64)), dim3(64), shv, nv,
           (const int32_t*)m->tri, (const int32_t*)m->v2t_off, (const int32_t*)m->v2t, capv,
           (int32_t*)nullptr, (const int32_t*)m->v2v_off, m->v2v, flags + 2);

    // 3b. bandwidth of the numbering (longest edge |u - v|)
    launch(ctx, "topo_span", 8.0 * nv + 4.0 * (double)m->nnz_v2v, k_vertex_span,
           dim3(blocks_for(nv, 256)), dim3(256), 0, nv, (const int32_t*)m->v2v_off,
           (const int32_t*)m->v2v, cnt);
    reduce_max_i32(ctx, cnt, nv, flags + 7);
    copy_async(ctx, h, flags + 7, sizeof(int32_t));
    sync(ctx);
    m->bandwidth = h[0];

    // 4. largest 2-ring element patch
    launch(ctx, "topo_patch_bound", 0.0, k_patch_bound, dim3(blocks_for(nv, 256)), dim3(256),
           0, nv, (const int32_t*)m->v2t_off, (const int32_t*)m->v2v_off,
           (const int32_t*)m->v2v, bound);
    reduce_max_i32(ctx, bound, nv, flags + 4);
    copy_async(ctx, h, flags + 4, sizeof(int32_t));
    sync(ctx);
    int capp = h[0] < kMaxPatchCap ? h[0] : kMaxPatchCap;
    if (capp < 1) capp = 1;
    memset_async(ctx, flags + 5, 0, sizeof(int32_t));
    // the patches themselves are kept for the ELL fill below when the
    // bound makes the buffer affordable (else they are built again there)
    int32_t* patch_tmp = nullptr;
    if ((size_t)capp * (size_t)nv * sizeof(int32_t) <= kPatchTmpMax)
      patch_tmp = (int32_t*)aux_workspace(ctx, (size_t)capp * (size_t)nv * sizeof(int32_t));
    launch(ctx, "topo_patch_size", 0.0, k_patch_size, dim3(blocks_for(nv, 64)), dim3(64),
           sizeof(int) * 64 * (size_t)capp, nv, (const int32_t*)m->slab_v,
           (const int32_t*)m->v2t_off, (const int32_t*)m->v2t, (const int32_t*)m->v2v_off,
           (const int32_t*)m->v2v, capp, bound, flags + 5, patch_tmp);
    reduce_max_i32(ctx, bound, nv, flags + 6);
    copy_async(ctx, h, flags + 5, 2 * sizeof(int32_t));
    sync(ctx);
    TRN_REQUIRE(h[0] == 0, TRN_ERR_CAPACITY,
                "a 2-ring element patch exceeds " + std::to_string(kMaxPatchCap) + " triangles");
    m->max_patch = h[1];
    m->patch_cap = m->max_patch > 0 ? m->max_patch : 1;

    // 4b. stored patches (sliced ELL) for the recovery gathers
    const int64_t nslab = (nv + kPatchSlab - 1) / kPatchSlab;
    TRN_HIP(hipMalloc(&m->patch_n, sizeof(int32_t) * (size_t)nv));
    TRN_HIP(hipMalloc(&m->patch_row, sizeof(int32_t) * (size_t)(nslab + 1)));
    copy_async(ctx, m->patch_n, bound, sizeof(int32_t) * (size_t)nv);
    launch(ctx, "topo_patch_rows", 0.0, k_patch_slab_rows, dim3((unsigned)nslab),
           dim3(kPatchSlab), 0, nv, (const int32_t*)m->slab_v, (const int32_t*)bound,
           m->patch_row);
    exclusive_scan_i32(ctx, m->patch_row, m->patch_row, nslab, true, scan_tmp);
    copy_async(ctx, h, m->patch_row + nslab, sizeof(int32_t));
    sync(ctx);
    const int64_t ell_n = (int64_t)h[0] * kPatchSlab;
    TRN_HIP(hipMalloc(&m->patch_ell, sizeof(int32_t) * (size_t)(ell_n > 0 ? ell_n : 1)));
    if (patch_tmp) {
      launch(ctx, "topo_patch_copy", 0.0, k_patch_copy, dim3((unsigned)nslab),
             dim3(kPatchSlab), 0, nv, (const int32_t*)m->slab_v, (const int32_t*)bound,
             (const int32_t*)patch_tmp, (const int32_t*)m->patch_row, m->patch_ell);
    } else {
      launch(ctx, "topo_patch_fill", 0.0, k_patch_fill, dim3((unsigned)nslab),
             dim3(kPatchSlab), sizeof(int) * kPatchSlab * (size_t)m->patch_cap, nv,
             (const int32_t*)m->slab_v, (const int32_t*)m->v2t_off, (const int32_t*)m->v2t,
             (const int32_t*)m->v2v_off, (const int32_t*)m->v2v, m->patch_cap,
             (const int32_t*)m->patch_row, m->patch_ell);
    }


    // 4c. per-slab unions of the patches (LDS staging of the gathers),
    //     each stored in its slab's ELL block: one sort per slab
    TRN_HIP(hipMalloc(&m->union_n, sizeof(int32_t) * (size_t)nslab));
    TRN_HIP(hipMalloc(&m->union_off, sizeof(int32_t) * (size_t)(nslab + 1)));
    TRN_HIP(hipMalloc(&m->union_ids, sizeof(int32_t) * (size_t)(ell_n > 0 ? ell_n : 1)));
    TRN_HIP(hipMalloc(&m->patch_local, sizeof(uint16_t) * (size_t)(ell_n > 0 ? ell_n : 1)));
    launch(ctx, "topo_union_offsets", 0.0, k_union_offsets, dim3(blocks_for(nslab + 1, 256)),
           dim3(256), 0, (int32_t)nslab, (const int32_t*)m->patch_row, m->union_off);
    launch(ctx, "topo_patch_union", 0.0, k_patch_union, dim3((unsigned)nslab), dim3(kUnionB),
           0, nv, (const int32_t*)m->patch_row, (const int32_t*)m->patch_ell, m->union_n,
           (const int32_t*)m->union_off, m->union_ids, m->patch_local);
    reduce_max_i32(ctx, m->union_n, nslab, flags + 7);
    memset_async(ctx, flags + 8, 0, sizeof(int32_t));
    launch(ctx, "topo_union_count", 0.0, k_count_negative, dim3(blocks_for(nslab, 256)),
           dim3(256), 0, nslab, (const int32_t*)m->union_n, flags + 8);
    TRN_HIP(hipMalloc(&m->union_cu, sizeof(int32_t) * (size_t)(nslab + 1)));
    launch(ctx, "topo_union_sizes", 0.0, k_union_sizes, dim3(blocks_for(nslab, 256)),
           dim3(256), 0, nslab, (const int32_t*)m->union_n, cnt2);
    exclusive_scan_i32(ctx, cnt2, m->union_cu, nslab, true, scan_tmp);
    copy_async(ctx, h + 1, flags + 7, 2 * sizeof(int32_t));
    copy_async(ctx, h + 3, m->union_cu + nslab, sizeof(int32_t));
    sync(ctx);
    m->max_union = h[1] > 0 ? h[1] : 0;
    m->union_all = h[2] == 0;
    TRN_HIP(hipMalloc(&m->union_tri, sizeof(int32_t) * 3 * (size_t)(h[3] > 0 ? h[3] : 1)));
    launch(ctx, "topo_union_tri", 0.0, k_union_tri, dim3((unsigned)nslab), dim3(256), 0,
           (const int32_t*)m->union_n, (const int32_t*)m->union_off,
           (const int32_t*)m->union_ids, (const int32_t*)m->union_cu, (const int32_t*)m->tri,
           m->union_tri);

    // 5. star codes of the incident triangles (smoothing reads corners
    //    through the 1-ring instead of the triangle array)
    TRN_REQUIRE(m->max_deg_v < 256, TRN_ERR_CAPACITY, "1-ring larger than 255 vertices");
    TRN_HIP(hipMalloc(&m->star, sizeof(int32_t) * 3 * (size_t)nt));
    TRN_HIP(hipMalloc(&m->vdata, sizeof(VData) * (size_t)nv));
    TRN_HIP(hipMalloc(&m->vt2, sizeof(double) * (size_t)nv));
    TRN_HIP(hipMalloc(&m->vtopo, sizeof(VTopo) * (size_t)nv));
    launch(ctx, "topo_star", 0.0, k_star_vtopo, dim3(blocks_for(nv, 256)), dim3(256), 0, nv,
           (const uint8_t*)m->tags, (const int32_t*)m->tri, (const int32_t*)m->v2t_off,
           (const int32_t*)m->v2t, (const int32_t*)m->v2v_off, (const int32_t*)m->v2v,
           m->star, m->vtopo);

  } catch (...) {
    cleanup();
    throw;
  }
  return m;
}

trn_mesh* mesh_create_ordered(trn_ctx* ctx, int32_t nv, int32_t nt, const int32_t* tri,
                              const uint8_t* tags, const int32_t* order_d) {
  return mesh_create_impl(ctx, nv, nt, tri, tags, nullptr, order_d);
}

}  // namespace trn

extern "C" {

int trn_mesh_create(trn_ctx* ctx, int32_t nv, int32_t nt, const int32_t* tri,
                    const uint8_t* tags, const double* xy_d, trn_mesh** out) {
  return guard_impl("trn_mesh_create", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(out != nullptr, TRN_ERR_ARG, "null output");
    *out = nullptr;
    *out = mesh_create_impl(ctx, nv, nt, tri, tags, xy_d, nullptr);
  });
}

int trn_mesh_destroy(trn_mesh* m) {
  if (!m) return TRN_OK;
  return guard_impl("trn_mesh_destroy", [&] {
    DeviceScope ds(m->ctx);
    (void)hipStreamSynchronize(m->ctx->stream);
    if (m->tri) (void)hipFree(m->tri);
    if (m->tags) (void)hipFree(m->tags);
    if (m->v2t_off) (void)hipFree(m->v2t_off);
    if (m->v2t) (void)hipFree(m->v2t);
    if (m->v2v_off) (void)hipFree(m->v2v_off);
    if (m->v2v) (void)hipFree(m->v2v);
    if (m->star) (void)hipFree(m->star);
    if (m->zorder) (void)hipFree(m->zorder);
    if (m->vdata) (void)hipFree(m->vdata);
    if (m->vt2) (void)hipFree(m->vt2);
    if (m->vtopo) (void)hipFree(m->vtopo);
    if (m->patch_n) (void)hipFree(m->patch_n);
    if (m->patch_row) (void)hipFree(m->patch_row);
    if (m->patch_ell) (void)hipFree(m->patch_ell);
    if (m->union_n) (void)hipFree(m->union_n);
    if (m->union_off) (void)hipFree(m->union_off);
    if (m->union_ids) (void)hipFree(m->union_ids);
    if (m->union_cu) (void)hipFree(m->union_cu);
    if (m->union_tri) (void)hipFree(m->union_tri);
    if (m->patch_local) (void)hipFree(m->patch_local);
    delete m;
  });
}

int trn_mesh_set_class_order(trn_mesh* mesh, int32_t zcurve) {
  return guard_impl("trn_mesh_set_class_order", [&] {
    TRN_REQUIRE(mesh != nullptr, TRN_ERR_ARG, "null mesh");
    TRN_REQUIRE(!zcurve || mesh->zorder, TRN_ERR_ARG,
                "a Z-curve class order needs coordinates at mesh creation");
    mesh->class_v = zcurve ? mesh->zorder : nullptr;
  });
}

int trn_mesh_info(const trn_mesh* m, int64_t* nnz_v2v, int32_t* max_deg_t, int32_t* max_deg_v,
                  int32_t* max_patch) {
  return guard_impl("trn_mesh_info", [&] {
    TRN_REQUIRE(m != nullptr, TRN_ERR_ARG, "null mesh");
    if (nnz_v2v) *nnz_v2v = m->nnz_v2v;
    if (max_deg_t) *max_deg_t = m->max_deg_t;
    if (max_deg_v) *max_deg_v = m->max_deg_v;
    if (max_patch) *max_patch = m->max_patch;
  });
}

int trn_mesh_csr(const trn_mesh* m, const int32_t** v2t_off_d, const int32_t** v2t_d,
                 const int32_t** v2v_off_d, const int32_t** v2v_d) {
  return guard_impl("trn_mesh_csr", [&] {
    TRN_REQUIRE(m != nullptr, TRN_ERR_ARG, "null mesh");
    if (v2t_off_d) *v2t_off_d = m->v2t_off;
    if (v2t_d) *v2t_d = m->v2t;
    if (v2v_off_d) *v2v_off_d = m->v2v_off;
    if (v2v_d) *v2v_d = m->v2v;
  });
}

}  // extern "C"
