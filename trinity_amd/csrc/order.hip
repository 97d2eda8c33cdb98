// order.hip -- spatial (Morton / Z-curve) order of the mesh vertices.
//
// Colour classes are independent sets, so the reference's result does not
// depend on the order in which a class is visited (src/smoothing.cpp:166-175).
// Visiting a class along a Z-curve of the vertex coordinates makes the 64
// vertices of a wave a compact 2-D patch whose 1-ring halo is mostly its own,
// instead of a row segment whose 3-row band is re-read by the waves of the
// rows above and below; this is what keeps the sweeps inside L2.
#include <hipcub/hipcub.hpp>

#include "trn_api.hpp"

namespace trn {

constexpr int kBoxBlocks = 512;

__global__ __launch_bounds__(256) void k_bbox_partial(int32_t nv, const double2* __restrict__ xy,
                                                      double* __restrict__ part) {
  __shared__ double s[4][4];
  double lo_x = INFINITY, lo_y = INFINITY, hi_x = -INFINITY, hi_y = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 p = xy[i];
    lo_x = fmin(lo_x, p.x);
    lo_y = fmin(lo_y, p.y);
    hi_x = fmax(hi_x, p.x);
    hi_y = fmax(hi_y, p.y);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo_x = fmin(lo_x, __shfl_xor(lo_x, o, 64));
    lo_y = fmin(lo_y, __shfl_xor(lo_y, o, 64));
    hi_x = fmax(hi_x, __shfl_xor(hi_x, o, 64));
    hi_y = fmax(hi_y, __shfl_xor(hi_y, o, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s[w][0] = lo_x;
    s[w][1] = lo_y;
    s[w][2] = hi_x;
    s[w][3] = hi_y;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      s[0][0] = fmin(s[0][0], s[k][0]);
      s[0][1] = fmin(s[0][1], s[k][1]);
      s[0][2] = fmax(s[0][2], s[k][2]);
      s[0][3] = fmax(s[0][3], s[k][3]);
    }
    for (int c = 0; c < 4; ++c) part[4 * blockIdx.x + c] = s[0][c];
  }
}

__device__ __forceinline__ uint32_t spread_bits(uint32_t x) {
  x &= 0xFFFFu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}

// The key grid has as many cells as the mesh has vertices, in the aspect
// ratio of its bounding box: sqrt(nv * ex / ey) - 1 cells along x (and the
// same for y), each coordinate rounded to the nearest cell, so that a
// structured grid, jittered or not, puts each vertex in a cell of its own and
// 64 consecutive keys cover an 8 x 8 block of vertices.  (Keys of the full
// 16-bit box resolution cut the blocks along the jitter instead: slab unions
// of a jittered 1025^2 grid then hold 4.53 triangles per vertex, 359 at most,
// against 3.64 and 264 here.)  Vertices that share a cell keep their id order.
struct KeyGrid {
  double x0, y0, sx, sy;
};

__device__ __forceinline__ double cells_along(double e, double other, int64_t nv) {
  if (!(e > 0.)) return 0.;
  const double n = other > 0. ? sqrt((double)nv * e / other) : (double)nv;
  return fmin(fmax(rint(n) - 1., 1.), 65535.);
}

__device__ __forceinline__ KeyGrid key_grid(const double* box, int64_t nv) {
  KeyGrid g;
  g.x0 = box[0];
  g.y0 = box[1];
  const double ex = box[2] - box[0], ey = box[3] - box[1];
  g.sx = ex > 0. ? cells_along(ex, ey, nv) / ex : 0.;
  g.sy = ey > 0. ? cells_along(ey, ex, nv) / ey : 0.;
  return g;
}

__device__ __forceinline__ uint32_t quantize(double v, double lo, double scale) {
  const double q = floor((v - lo) * scale + 0.5);
  if (!(q >= 0.)) return 0u;  // also NaN
  return q >= 65535. ? 65535u : (uint32_t)q;
}

__device__ __forceinline__ uint32_t zkey(const KeyGrid& g, const double2 p) {
  return spread_bits(quantize(p.x, g.x0, g.sx)) | (spread_bits(quantize(p.y, g.y0, g.sy)) << 1);
}

// The bounding box from the per-block partials, once (one block).
__global__ __launch_bounds__(256) void k_bbox_final(const double* __restrict__ part, int nparts,
                                                    double* __restrict__ box) {
  __shared__ double s[4][4];
  const int c = threadIdx.x & 3;
  double v = c < 2 ? INFINITY : -INFINITY;
  for (int b = threadIdx.x >> 2; b < nparts; b += blockDim.x >> 2)
    v = c < 2 ? fmin(v, part[4 * b + c]) : fmax(v, part[4 * b + c]);
  // lanes c, c + 4, c + 8, ... of a wave hold the same corner
#pragma unroll
  for (int o = 32; o >= 4; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = c < 2 ? fmin(v, w) : fmax(v, w);
  }
  if ((threadIdx.x & 63) < 4) s[threadIdx.x >> 6][c] = v;
  __syncthreads();
  if (threadIdx.x < 4) {
    double r = s[0][c];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = c < 2 ? fmin(r, s[w][c]) : fmax(r, s[w][c]);
    box[c] = r;
  }
}

__global__ __launch_bounds__(256) void k_morton(int32_t nv, const double2* __restrict__ xy,
                                                const double* __restrict__ box,
                                                uint32_t* __restrict__ key,
                                                int32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  key[i] = zkey(key_grid(box, nv), xy[i]);
  idx[i] = (int32_t)i;
}

void spatial_order(trn_ctx* ctx, int32_t nv, const double* xy, int32_t* perm) {
  size_t sort_bytes = 0;
  TRN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const uint32_t*)nullptr,
                                             (uint32_t*)nullptr, (const int32_t*)nullptr,
                                             (int32_t*)nullptr, nv, 0, 32, ctx->stream));
  Scratch sc(ctx, scratch_bytes(sizeof(double) * 4 * (kBoxBlocks + 1)) +
                      2 * scratch_bytes(sizeof(uint32_t) * (size_t)nv) +
                      scratch_bytes(sizeof(int32_t) * (size_t)nv) + scratch_bytes(sort_bytes));
  double* part = sc.carve<double>(4 * kBoxBlocks);
  double* box = sc.carve<double>(4);
  uint32_t* key = sc.carve<uint32_t>(nv);
  uint32_t* key2 = sc.carve<uint32_t>(nv);
  int32_t* idx = sc.carve<int32_t>(nv);
  void* tmp = sc.carve<char>(sort_bytes);
  int blocks = (int)((nv + 255) / 256);
  if (blocks > kBoxBlocks) blocks = kBoxBlocks;
  launch(ctx, "order_bbox", 16.0 * nv, k_bbox_partial, dim3(blocks), dim3(256), 0, nv,
         (const double2*)xy, part);
  launch(ctx, "order_bbox_final", 0.0, k_bbox_final, dim3(1), dim3(256), 0, (const double*)part,
         blocks, box);
  launch(ctx, "order_morton", 24.0 * nv, k_morton, dim3((unsigned)((nv + 255) / 256)), dim3(256),
         0, nv, (const double2*)xy, (const double*)box, key, idx);
  TRN_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, (const uint32_t*)key, key2,
                                             (const int32_t*)idx, perm, nv, 0, 32, ctx->stream));
}

// The order for a disjoint union of meshes: vertices grouped by mesh (seg_of
// [nv], mesh m owning [seg_off[m], seg_off[m + 1])), each mesh along the
// Z-curve of its own bounding box (one block per mesh reduces it).
__global__ __launch_bounds__(256) void k_bbox_segments(const int32_t* __restrict__ seg_off,
                                                       const double2* __restrict__ xy,
                                                       double* __restrict__ box) {
  __shared__ double s[4][4];
  const int64_t b = seg_off[blockIdx.x], e = seg_off[blockIdx.x + 1];
  double lo_x = INFINITY, lo_y = INFINITY, hi_x = -INFINITY, hi_y = -INFINITY;
  for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
    const double2 p = xy[i];
    lo_x = fmin(lo_x, p.x);
    lo_y = fmin(lo_y, p.y);
    hi_x = fmax(hi_x, p.x);
    hi_y = fmax(hi_y, p.y);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo_x = fmin(lo_x, __shfl_xor(lo_x, o, 64));
    lo_y = fmin(lo_y, __shfl_xor(lo_y, o, 64));
    hi_x = fmax(hi_x, __shfl_xor(hi_x, o, 64));
    hi_y = fmax(hi_y, __shfl_xor(hi_y, o, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s[w][0] = lo_x;
    s[w][1] = lo_y;
    s[w][2] = hi_x;
    s[w][3] = hi_y;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
      s[0][0] = fmin(s[0][0], s[k][0]);
      s[0][1] = fmin(s[0][1], s[k][1]);
      s[0][2] = fmax(s[0][2], s[k][2]);
      s[0][3] = fmax(s[0][3], s[k][3]);
    }
    for (int c = 0; c < 4; ++c) box[4 * blockIdx.x + c] = s[0][c];
  }
}

__global__ __launch_bounds__(256) void k_morton_segments(int32_t nv, const double2* __restrict__ xy,
                                                         const int32_t* __restrict__ seg_of,
                                                         const int32_t* __restrict__ seg_off,
                                                         const double* __restrict__ box,
                                                         uint64_t* __restrict__ key,
                                                         int32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int32_t g = seg_of[i];
  const uint32_t z = zkey(key_grid(box + 4 * (int64_t)g, seg_off[g + 1] - seg_off[g]), xy[i]);
  key[i] = ((uint64_t)(uint32_t)g << 32) | z;
  idx[i] = (int32_t)i;
}

void spatial_order_segments(trn_ctx* ctx, int32_t nv, const double* xy, int32_t nseg,
                            const int32_t* seg_of, const int32_t* seg_off, int32_t* perm) {
  int bits = 32;
  while (bits < 64 && ((int64_t)1 << (bits - 32)) < (int64_t)nseg) ++bits;
  size_t sort_bytes = 0;
  TRN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const uint64_t*)nullptr,
                                             (uint64_t*)nullptr, (const int32_t*)nullptr,
                                             (int32_t*)nullptr, nv, 0, bits, ctx->stream));
  Scratch sc(ctx, scratch_bytes(sizeof(double) * 4 * (size_t)nseg) +
                      2 * scratch_bytes(sizeof(uint64_t) * (size_t)nv) +
                      scratch_bytes(sizeof(int32_t) * (size_t)nv) + scratch_bytes(sort_bytes));
  double* box = sc.carve<double>(4 * (size_t)nseg);
  uint64_t* key = sc.carve<uint64_t>(nv);
  uint64_t* key2 = sc.carve<uint64_t>(nv);
  int32_t* idx = sc.carve<int32_t>(nv);
  void* tmp = sc.carve<char>(sort_bytes);
  launch(ctx, "order_bbox", 16.0 * nv, k_bbox_segments, dim3((unsigned)nseg), dim3(256), 0,
         seg_off, (const double2*)xy, box);
  launch(ctx, "order_morton", 28.0 * nv, k_morton_segments, dim3((unsigned)((nv + 255) / 256)),
         dim3(256), 0, nv, (const double2*)xy, seg_of, seg_off, (const double*)box, key, idx);
  TRN_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, (const uint64_t*)key, key2,
                                             (const int32_t*)idx, perm, nv, 0, bits, ctx->stream));
}

}  // namespace trn
