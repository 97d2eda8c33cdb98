// coloring.hip -- vertex colouring of the 1-ring graph, replacing
// Partit::extractColoring (src/partition.cpp:79-197).
//
// The reference runs a speculative first-fit in parallel (:94-111), detects
// conflicts (:113-122) and recolours conflicting vertices until none remain
// (:126-168); with several threads its result depends on the interleaving.
// Here each round is two bulk-synchronous kernels, which makes the result a
// pure function of (mesh, seed):
//   tentative: every uncoloured vertex takes the smallest colour not held by
//              an already coloured neighbour (64-bit forbidden mask);
//   resolve:   an uncoloured vertex keeps its tentative colour unless an
//              uncoloured neighbour with a higher seeded priority chose the
//              same colour; winners commit, losers retry next round.
// Two adjacent winners of one round hold different tentative colours and
// every tentative colour avoids all committed neighbours, so the colouring
// is proper after every round; the top-priority vertex of every conflict
// group commits, so each round makes progress.
//
// The colour classes (Partit::task.subset/cardin/nb.parts, :171-194) are a
// stable counting sort: per-block colour histograms, one exclusive scan in
// colour-major order, and a scatter that ranks equal colours inside a wave
// with ballots, so each class lists its vertices in ascending order.
#include <string>
#include <vector>

#include "trn_api.hpp"

namespace trn {

constexpr int kClassB = 256;

// splitmix64 finaliser (Steele, Lea, Flood 2014) as the seeded priority.
__device__ __forceinline__ uint32_t color_priority(uint64_t seed, uint32_t v) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)v + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (uint32_t)(z >> 32);
}

constexpr int kRoundSlots = 64;   // per-round loser counters, spread over addresses
constexpr int kRoundsPerSync = 8;  // rounds queued between two host checks

// Rounds are queued in batches; a round whose predecessor left no loser
// returns at once (prev_total == 0), so over-queued rounds cost a launch.
__global__ void k_color_tent(int32_t nv, const int32_t* __restrict__ v2v_off,
                             const int32_t* __restrict__ v2v, const int32_t* __restrict__ color,
                             int32_t* __restrict__ tent, int max_colors, int32_t* fail,
                             const int32_t* __restrict__ prev_total) {
  if (prev_total && *prev_total == 0) return;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const int32_t b = v2v_off[v], e = v2v_off[v + 1];
  if (color[v] != 0 || b == e) {
    tent[v] = 0;
    return;
  }
  uint64_t used = 0;
  for (int32_t a = b; a < e; ++a) {
    const int32_t c = color[v2v[a]];
    if (c > 0) used |= 1ull << (c - 1);
  }
  const uint64_t freebits = ~used;
  const int c = freebits ? __ffsll((unsigned long long)freebits) : 65;
  if (c > max_colors) {
    atomicOr(fail, 1);
    tent[v] = 0;
    return;
  }
  tent[v] = c;
}

__global__ __launch_bounds__(256) void k_color_resolve(int32_t nv,
                                                       const int32_t* __restrict__ v2v_off,
                                                       const int32_t* __restrict__ v2v,
                                                       int32_t* __restrict__ color,
                                                       const int32_t* __restrict__ tent,
                                                       uint64_t seed, int32_t* remaining,
                                                       const int32_t* __restrict__ prev_total) {
  if (prev_total && *prev_total == 0) return;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int lose = 0;
  if (v < nv) {
    const int32_t tv = tent[v];
    if (tv != 0) {
      const uint32_t pv = color_priority(seed, (uint32_t)v);
      for (int32_t a = v2v_off[v]; a < v2v_off[v + 1]; ++a) {
        const int32_t u = v2v[a];
        if (tent[u] == tv) {
          const uint32_t pu = color_priority(seed, (uint32_t)u);
          if (pu > pv || (pu == pv && u > v)) {
            lose = 1;
            break;
          }
        }
      }
      if (!lose) color[v] = tv;
    }
  }
  // one atomic per block, spread over kRoundSlots addresses (a single hot
  // address serialises in L2)
  const int losers = __syncthreads_count(lose);
  if (threadIdx.x == 0 && losers) atomicAdd(&remaining[blockIdx.x & (kRoundSlots - 1)], losers);
}

// total[0] = sum of the round's slots, or 0 when the previous round was
// already empty (its slots were never written).
__global__ void k_color_round_total(const int32_t* __restrict__ slots, int32_t* __restrict__ total) {
  int s = slots[threadIdx.x];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (threadIdx.x == 0) total[0] = s;
}

__global__ void k_color_check(int32_t nv, const int32_t* __restrict__ v2v_off,
                              const int32_t* __restrict__ v2v, const int32_t* __restrict__ color,
                              int32_t ncolors, unsigned long long* bad) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const int32_t b = v2v_off[v], e = v2v_off[v + 1];
  if (b == e) return;
  const int32_t cv = color[v];
  if (cv < 1 || cv > ncolors) atomicAdd(&bad[1], 1ull);
  for (int32_t a = b; a < e; ++a) {
    const int32_t u = v2v[a];
    if (u > v && color[u] == cv) atomicAdd(&bad[0], 1ull);
  }
}

// Per-block histogram of colours 1..ncolors, stored colour-major.
// Position k of the visiting sequence is vertex perm[k] (or k without perm).
__global__ __launch_bounds__(kClassB) void k_class_count(int32_t nv, const int32_t* __restrict__ color,
                                                         const int32_t* __restrict__ perm,
                                                         int ncolors, int32_t* __restrict__ cnt) {
  __shared__ int hist[64];
  if (threadIdx.x < 64) hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nv) {
    const int c = color[perm ? perm[k] : k];
    if (c >= 1 && c <= ncolors) atomicAdd(&hist[c - 1], 1);
  }
  __syncthreads();
  if ((int)threadIdx.x < ncolors) cnt[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = hist[threadIdx.x];
}

__global__ __launch_bounds__(kClassB) void k_class_scatter(int32_t nv, const int32_t* __restrict__ color,
                                                           const int32_t* __restrict__ perm,
                                                           int ncolors,
                                                           const int32_t* __restrict__ off,
                                                           int32_t* __restrict__ order) {
  __shared__ int wcnt[kClassB / 64][64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int k = threadIdx.x; k < (kClassB / 64) * 64; k += blockDim.x) (&wcnt[0][0])[k] = 0;
  __syncthreads();
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t v = (k < nv && perm) ? perm[k] : k;
  int c = 0;
  if (k < nv) {
    c = color[v];
    if (c < 1 || c > ncolors) c = 0;
  }
  // rank among equal colours inside the wave (lanes in ascending v order)
  uint64_t todo = __ballot(1);
  int rank = 0;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  while (todo) {
    const int leader = __ffsll((unsigned long long)todo) - 1;
    const int lc = __shfl(c, leader, 64);
    const uint64_t m = __ballot(c == lc);
    if (c == lc) rank = __popcll(m & lt);
    if (lane == leader && lc > 0) wcnt[wid][lc - 1] = __popcll(m);
    todo &= ~m;
  }
  __syncthreads();
  if (c > 0) {
    int base = off[(int64_t)(c - 1) * gridDim.x + blockIdx.x];
    for (int w = 0; w < wid; ++w) base += wcnt[w][c - 1];
    order[base + rank] = (int32_t)v;
  }
}

__global__ void k_class_offsets(const int32_t* __restrict__ off, int ncolors, int nblocks,
                                int32_t* __restrict__ out) {
  const int k = threadIdx.x;
  if (k <= ncolors) out[k] = off[(int64_t)k * nblocks];
}

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Builds order_d and class_off (host) for colours 1..ncolors, visiting the
// vertices in perm order when given.
static void build_classes(trn_ctx* ctx, int32_t nv, const int32_t* color, const int32_t* perm,
                          int ncolors, int32_t* order, int32_t* class_off) {
  const int64_t nb = nblk(nv, kClassB);
  const int64_t n = nb * ncolors;
  Scratch sc(ctx, scratch_bytes(sizeof(int32_t) * (size_t)(n + 1)) * 2 +
                      exclusive_scan_scratch(n) + scratch_bytes(sizeof(int32_t) * 80));
  int32_t* cnt = sc.carve<int32_t>(n + 1);
  int32_t* off = sc.carve<int32_t>(n + 1);
  int32_t* tmp = sc.carve<int32_t>(exclusive_scan_scratch(n) / 4);
  int32_t* offs = sc.carve<int32_t>(80);
  launch(ctx, "class_count", 4.0 * nv, k_class_count, dim3((unsigned)nb), dim3(kClassB), 0, nv,
         color, perm, ncolors, cnt);
  exclusive_scan_i32(ctx, cnt, off, n, true, tmp);
  launch(ctx, "class_scatter", 8.0 * nv, k_class_scatter, dim3((unsigned)nb), dim3(kClassB), 0,
         nv, color, perm, ncolors, (const int32_t*)off, order);
  launch(ctx, "class_offsets", 0.0, k_class_offsets, dim3(1), dim3(128), 0, (const int32_t*)off,
         ncolors, (int)nb, offs);
  copy_async(ctx, ctx->pinned_i32, offs, sizeof(int32_t) * (size_t)(ncolors + 1));
  sync(ctx);
  for (int k = 0; k <= ncolors; ++k) class_off[k] = ctx->pinned_i32[k];
}

}  // namespace trn

using namespace trn;

extern "C" {

int trn_color(trn_ctx* ctx, const trn_mesh* mesh, uint64_t seed, int32_t max_colors,
              int32_t* color_d, int32_t* order_d, int32_t* class_off, int32_t* ncolors) {
  return guard_impl("trn_color", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(mesh != nullptr && color_d && order_d && class_off && ncolors, TRN_ERR_ARG,
                "null argument");
    TRN_REQUIRE(max_colors >= 1 && max_colors <= 64, TRN_ERR_ARG, "max_colors must be in 1..64");
    const int32_t nv = mesh->nv;
    const size_t nslots = (size_t)kRoundsPerSync * kRoundSlots;
    Scratch sc(ctx, scratch_bytes(sizeof(int32_t) * (size_t)nv) + scratch_bytes(64) +
                        scratch_bytes(sizeof(int32_t) * nslots) +
                        scratch_bytes(sizeof(int32_t) * kRoundsPerSync));
    int32_t* tent = sc.carve<int32_t>(nv);
    int32_t* flags = sc.carve<int32_t>(16);
    int32_t* slots = sc.carve<int32_t>(nslots);
    int32_t* totals = sc.carve<int32_t>(kRoundsPerSync);
    memset_async(ctx, color_d, 0, sizeof(int32_t) * (size_t)nv);
    memset_async(ctx, flags, 0, sizeof(int32_t) * 16);
    int32_t* h = ctx->pinned_i32;
    int rounds = 0;
    for (;;) {
      memset_async(ctx, slots, 0, sizeof(int32_t) * nslots);
      for (int r = 0; r < kRoundsPerSync; ++r) {
        const int32_t* prev = r ? totals + r - 1 : nullptr;
        launch(ctx, "color_tent", 4.0 * nv + 4.0 * (double)mesh->nnz_v2v, k_color_tent,
               dim3(nblk(nv, 256)), dim3(256), 0, nv, (const int32_t*)mesh->v2v_off,
               (const int32_t*)mesh->v2v, (const int32_t*)color_d, tent, (int)max_colors, flags,
               prev);
        launch(ctx, "color_resolve", 4.0 * nv + 4.0 * (double)mesh->nnz_v2v, k_color_resolve,
               dim3(nblk(nv, 256)), dim3(256), 0, nv, (const int32_t*)mesh->v2v_off,
               (const int32_t*)mesh->v2v, color_d, (const int32_t*)tent, seed,
               slots + (size_t)r * kRoundSlots, prev);
        launch(ctx, "color_round", 0.0, k_color_round_total, dim3(1), dim3(kRoundSlots), 0,
               (const int32_t*)(slots + (size_t)r * kRoundSlots), totals + r);
      }
      // the largest colour, valid once no loser is left, comes back with
      // the convergence check in one synchronisation
      reduce_max_i32(ctx, color_d, nv, flags + 4);
      copy_async(ctx, h, flags, sizeof(int32_t));
      copy_async(ctx, h + 1, totals + kRoundsPerSync - 1, sizeof(int32_t));
      copy_async(ctx, h + 2, flags + 4, sizeof(int32_t));
      sync(ctx);
      rounds += kRoundsPerSync;
      TRN_REQUIRE(h[0] == 0, TRN_ERR_PALETTE,
                  "a vertex needs more than " + std::to_string(max_colors) + " colours");
      if (h[1] == 0) break;
      TRN_REQUIRE(rounds < 100000, TRN_ERR_STATE, "colouring did not converge");
    }
    const int nc = h[2];
    build_classes(ctx, nv, color_d, mesh->perm, nc, order_d, class_off);
    *ncolors = nc;
  });
}

int trn_color_classes(trn_ctx* ctx, const trn_mesh* mesh, int32_t nv, const int32_t* color_d,
                      int32_t ncolors, int32_t* order_d, int32_t* class_off) {
  return guard_impl("trn_color_classes", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(nv >= 0 && color_d && order_d && class_off, TRN_ERR_ARG, "null argument");
    TRN_REQUIRE(ncolors >= 1 && ncolors <= 64, TRN_ERR_ARG, "ncolors must be in 1..64");
    TRN_REQUIRE(!mesh || mesh->nv == nv, TRN_ERR_ARG, "nv differs from the mesh");
    build_classes(ctx, nv, color_d, mesh ? mesh->perm : nullptr, ncolors, order_d, class_off);
  });
}

int trn_color_check(trn_ctx* ctx, const trn_mesh* mesh, const int32_t* color_d, int32_t ncolors,
                    int64_t* bad_edges, int64_t* bad_vertices) {
  return guard_impl("trn_color_check", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(mesh && color_d && bad_edges && bad_vertices, TRN_ERR_ARG, "null argument");
    Scratch sc(ctx, 256);
    unsigned long long* bad = sc.carve<unsigned long long>(2);
    memset_async(ctx, bad, 0, 2 * sizeof(unsigned long long));
    launch(ctx, "color_check", 0.0, k_color_check, dim3(nblk(mesh->nv, 256)), dim3(256), 0,
           mesh->nv, (const int32_t*)mesh->v2v_off, (const int32_t*)mesh->v2v, color_d, ncolors,
           bad);
    copy_async(ctx, ctx->pinned_f64, bad, 2 * sizeof(unsigned long long));
    sync(ctx);
    const unsigned long long* hb = (const unsigned long long*)ctx->pinned_f64;
    *bad_edges = (int64_t)hb[0];
    *bad_vertices = (int64_t)hb[1];
  });
}

}  // extern "C"
