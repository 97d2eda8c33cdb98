// coloring.hip -- vertex colouring of the 1-ring graph, replacing
// Partit::extractColoring (src/partition.cpp:79-197).
//
// The reference runs a speculative first-fit in parallel (:94-111), detects
// conflicts (:113-122) and recolours conflicting vertices until none remain
// (:126-168); with one thread it is the sequential first fit in index order
// (4 colours on a triangulated grid), with several its result depends on
// the interleaving.  Here the result is a pure function of (mesh, seed,
// tile size), in three steps:
//   tiles:    the vertices are cut into tiles of consecutive ids; one wave
//             per tile runs the sequential first fit over its tile in
//             index order, seeing only neighbours inside the tile
//             (k_color_tiles).  On a mesh numbered with locality most edges
//             stay inside a tile, so this is the reference's one-thread
//             colouring almost everywhere;
//   seams:    an edge between two tiles whose ends got the same colour is
//             a conflict; the end in the later tile is uncoloured
//             (k_color_seams, k_color_drop), which leaves a proper partial
//             colouring;
//   repair:   the uncoloured vertices are coloured in speculative rounds,
//             each two bulk-synchronous kernels:
//               tentative: every uncoloured vertex takes the smallest colour
//                          not held by a coloured neighbour (64-bit mask);
//               resolve:   it keeps it unless an uncoloured neighbour with a
//                          higher seeded priority chose the same colour;
//                          winners commit, losers retry next round.
//             Two adjacent winners of one round hold different tentative
//             colours and every tentative colour avoids all committed
//             neighbours, so the colouring stays proper; the top-priority
//             vertex of every conflict group commits, so each round makes
//             progress.  A round reads the previous round's loser count
//             itself and returns at once when it is zero, so a batch of
//             rounds is queued without host checks.
// A vertex whose neighbours hold every colour 1..max_colors is either an
// error (trn_color) or, as in the reference (partition.cpp:104-110), left
// uncoloured and out of every class (trn_color_partial).
//
// The colour classes (Partit::task.subset/cardin/nb.parts, :171-194) are a
// stable counting sort: per-block colour histograms, one exclusive scan in
// colour-major order (a single block, which also gathers the convergence
// check, the number of colours and the class offsets for the host), and a
// scatter that ranks equal colours inside a wave with ballots, so each
// class lists its vertices in ascending order.  The class build is queued
// behind the batch of rounds, so a colouring that converges within one
// batch costs one host synchronisation.
#include <string>
#include <vector>

#include "trn_api.hpp"

namespace trn {

constexpr int kClassB = 1024;  // threads per block of the class build
constexpr int kClassPer = 4;   // consecutive kClassB-vertex chunks per block
constexpr int kScanB = 1024;   // threads of the single-block scan
constexpr int kMaxRows = 64;   // colours a class build handles

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

// Losers left by the previous round: the sum of its kRoundSlots counters.
// The addresses are the same in every thread, so the loads are scalar
// (one scalar-cache line per CU) and no barrier is needed.
// prev_slots == nullptr: no previous round in this batch, taken as "losers
// left".
__device__ __forceinline__ int prev_losers(const int32_t* __restrict__ prev_slots) {
  if (!prev_slots) return 1;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kRoundSlots; ++k) s += prev_slots[k];
  return s;
}

// Rounds are queued in batches; a round whose predecessor left no loser
// returns at once, so over-queued rounds cost a launch each.
// leave != 0: a vertex that finds every colour taken gives up (colour -1,
// which no neighbour counts as a colour) instead of raising the failure flag.
__global__ __launch_bounds__(256) void k_color_tent(int32_t nv, const int32_t* __restrict__ v2v_off,
                                                    const int32_t* __restrict__ v2v,
                                                    int32_t* __restrict__ color,
                                                    int32_t* __restrict__ tent, int max_colors,
                                                    int leave, int32_t* fail,
                                                    const int32_t* __restrict__ prev_slots) {
  if (prev_losers(prev_slots) == 0) return;
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
    if (leave) color[v] = -1;
    else atomicOr(fail, 1);
    tent[v] = 0;
    return;
  }
  tent[v] = c;
}

// first != 0: the first round of the colouring.  No vertex is coloured
// yet, so every vertex with a neighbour would take colour 1 (max_colors >=
// 1): the tentative kernel is skipped, tent[] is not read, and every colour
// is written (0 for losers and isolated vertices), which also clears the
// colours of an earlier call.
__global__ __launch_bounds__(256) void k_color_resolve(int32_t nv,
                                                       const int32_t* __restrict__ v2v_off,
                                                       const int32_t* __restrict__ v2v,
                                                       int32_t* __restrict__ color,
                                                       const int32_t* __restrict__ tent,
                                                       uint64_t seed, int first,
                                                       int32_t* remaining,
                                                       const int32_t* __restrict__ prev_slots) {
  if (prev_losers(prev_slots) == 0) return;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int lose = 0;
  if (v < nv) {
    int32_t tv, b = 0, e = 0;
    if (first) {
      b = v2v_off[v];
      e = v2v_off[v + 1];
      tv = b != e ? 1 : 0;
    } else {
      tv = tent[v];
      if (tv != 0) {
        b = v2v_off[v];
        e = v2v_off[v + 1];
      }
    }
    if (tv != 0) {
      const uint32_t pv = color_priority(seed, (uint32_t)v);
      for (int32_t a = b; a < e; ++a) {
        const int32_t u = v2v[a];
        if ((first ? 1 : tent[u]) == tv) {
          const uint32_t pu = color_priority(seed, (uint32_t)u);
          if (pu > pv || (pu == pv && u > v)) {
            lose = 1;
            break;
          }
        }
      }
      if (!lose) color[v] = tv;
    }
    if (first && (tv == 0 || lose)) color[v] = 0;
  }
  // one atomic per block, spread over kRoundSlots addresses (a single hot
  // address serialises in L2)
  const int losers = __syncthreads_count(lose);
  if (threadIdx.x == 0 && losers) atomicAdd(&remaining[blockIdx.x & (kRoundSlots - 1)], losers);
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

// ---- tiles and seams -------------------------------------------------------

constexpr int kTileB = 64;                    // one wave per tile
constexpr int kTileMax = 16384;               // largest tile (local ids fit 16 bits)

__device__ __forceinline__ uint64_t lane_u64(const uint32_t lo, const uint32_t hi, const int k) {
  return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hi, k) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)lo, k);
}

// Sequential first fit in index order over tile b = 2 * blockIdx.x + parity
// ([t0, t0 + tile) of the vertex ids).  Even tiles run first and see only
// their own earlier vertices; odd tiles run second and also avoid the
// colours of their neighbours in even tiles, which are final by then.
// Per 64-vertex chunk every lane gathers, for its vertex, the mask of
// colours its earlier neighbours outside the chunk hold (read from `color`,
// where the wave stores each chunk's colours as it finishes it) and the
// mask of its earlier neighbours inside the chunk; then the chunk is
// coloured in index order, in one of two ways with the same result:
//   chains: when no vertex of the chunk has two earlier neighbours inside
//           it (a mesh numbered row by row: the only one is v - 1), the
//           colour of vertex k is a function of its in-chunk parent's
//           colour x alone, x == m1 ? m2 : m1, with m1 and m2 the two
//           smallest colours absent from its outside mask.  Such functions
//           compose into the same form, so pointer jumping over the parent
//           links resolves the chunk in at most 7 shuffle rounds;
//   walk:   otherwise the wave walks the chunk in order on the scalar unit:
//           vertex k takes the smallest colour absent from its outside
//           mask and from the colours of its in-chunk neighbours, read
//           from their lanes (readlane), and is recorded in lane k.
// Isolated vertices stay 0.  A vertex that finds colours 1..max_colors
// taken is written -1 and counted in *given_up.  Edges between tiles of
// the same parity (only where the numbering has edges longer than a tile)
// are left to the seam check.
// The kernel keeps no LDS: it runs beside the recovery gathers, whose
// occupancy is bound by their LDS, and the 1-16 KB a tile would hold per
// wave (8 waves per CU on a 4096^2 grid) took gather blocks off every CU
// (the gradient gather beside it: 1.20 ms with the tiles in LDS, 1.03 ms
// without).  In-tile
// colours of earlier chunks are read back from `color`, where the block
// barrier after each chunk makes its stores visible to the block.
constexpr uint32_t kGaveUp = 0xFF;

// The smallest colour absent from `used`, or kGaveUp past the palette.
__device__ __forceinline__ uint32_t first_free(const uint64_t used, const int max_colors) {
  const int c = __builtin_ffsll((long long)~used);
  return (c == 0 || c > max_colors) ? kGaveUp : (uint32_t)c;
}

__global__ __launch_bounds__(kTileB) void k_color_tiles(int32_t nv, int32_t tile, int parity,
                                                        const int32_t* __restrict__ v2v_off,
                                                        const int32_t* __restrict__ v2v,
                                                        int max_colors,
                                                        int32_t* __restrict__ color,
                                                        int32_t* __restrict__ given_up) {
  const int64_t tb = 2 * (int64_t)blockIdx.x + parity;
  const int64_t t0 = tb * tile;
  if (t0 >= nv) return;
  const int64_t t1 = t0 + tile < nv ? t0 + tile : nv;
  const int lane = threadIdx.x;
  int nfail = 0;
  for (int64_t c0 = t0; c0 < t1; c0 += kTileB) {
    const int64_t v = c0 + lane;
    int32_t b = 0, e = 0;
    if (v < t1) {
      b = v2v_off[v];
      e = v2v_off[v + 1];
    }
    uint64_t near = 0, far = 0;
    for (int32_t a = b; a < e; ++a) {
      const int32_t u = v2v[a];
      if (u >= c0 && u < v) {
        near |= 1ull << (u - c0);
      } else if (u >= t0 && u < c0) {  // an earlier chunk of this tile
        const int32_t cu = color[u];
        if (cu > 0) far |= 1ull << (cu - 1);
      } else if (parity && (u < t0 || u >= t1) && ((u / tile) & 1) == 0) {
        const int32_t cu = color[u];  // an even tile, finished
        if (cu > 0) far |= 1ull << (cu - 1);
      }
    }
    const bool live = v < t1 && b != e;  // coloured here (not isolated, inside the tile)
    int mine = 0;
    if (__ballot(__popcll(near) > 1) == 0) {
      // chains: g(x) = x == a ? b : c applied to the colour x of lane q
      const uint32_t m1 = first_free(far, max_colors);
      const uint32_t m2 = m1 == kGaveUp ? kGaveUp
                                        : first_free(far | (1ull << (m1 - 1)), max_colors);
      uint32_t ga = m1, gb = m2, gc = m1;
      int q = near ? __builtin_ctzll(near) : -1;
      bool done = !live || q < 0;
      uint32_t x = live ? m1 : 0u;  // the colour, once done
      while (__ballot(!done)) {
        const uint32_t word = done ? (0x80000000u | x)
                                   : (ga | (gb << 8) | (gc << 16) | ((uint32_t)q << 24));
        const uint32_t other = (uint32_t)__shfl((int)word, done ? lane : q, kTileB);
        if (!done) {
          if (other & 0x80000000u) {
            const uint32_t xp = other & 0xFFu;
            x = xp == ga ? gb : gc;
            done = true;
          } else {
            const uint32_t oa = other & 0xFFu, ob = (other >> 8) & 0xFFu,
                           oc = (other >> 16) & 0xFFu;
            const uint32_t nb = ob == ga ? gb : gc, nc = oc == ga ? gb : gc;
            ga = oa;
            gb = nb;
            gc = nc;
            q = (int)((other >> 24) & 63u);
          }
        }
      }
      mine = (int)x;
      nfail += __popcll(__ballot(live && x == kGaveUp));
    } else {
      const uint64_t iso = __ballot(v < t1 && b == e);
      const uint32_t near_lo = (uint32_t)near, near_hi = (uint32_t)(near >> 32);
      const uint32_t far_lo = (uint32_t)far, far_hi = (uint32_t)(far >> 32);
      uint32_t oh_lo = 0, oh_hi = 0;  // one-hot colour of this lane's vertex
      const int n = (int)(t1 - c0 < kTileB ? t1 - c0 : kTileB);
      for (int k = 0; k < n; ++k) {  // uniform: the scalar unit walks the chunk
        int c = 0;
        if (!((iso >> k) & 1)) {
          uint64_t used = lane_u64(far_lo, far_hi, k);
          uint64_t nb = lane_u64(near_lo, near_hi, k);
          while (nb) {
            const int j = __builtin_ctzll(nb);
            nb &= nb - 1;
            used |= lane_u64(oh_lo, oh_hi, j);
          }
          c = (int)first_free(used, max_colors);
          if (c == (int)kGaveUp) ++nfail;
        }
        const uint64_t o = (c != 0 && c != (int)kGaveUp) ? 1ull << (c - 1) : 0ull;
        if (lane == k) {  // record vertex k's colour in its lane
          oh_lo = (uint32_t)o;
          oh_hi = (uint32_t)(o >> 32);
          mine = c;
        }
      }
    }
    if (v < t1) color[v] = mine == (int)kGaveUp ? -1 : mine;
    __syncthreads();  // the chunk's colours visible to the later chunks' loads
  }
  if (lane == 0 && nfail) atomicAdd(given_up, nfail);
}

// An edge between two tiles whose ends hold the same colour: the end in the
// later tile loses (lose[v] = 1).  Reads colours only, so the losers do not
// depend on timing.
__global__ __launch_bounds__(256) void k_color_seams(int32_t nv, int32_t tile,
                                                     const int32_t* __restrict__ v2v_off,
                                                     const int32_t* __restrict__ v2v,
                                                     const int32_t* __restrict__ color,
                                                     int32_t* __restrict__ lose) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const int32_t cv = color[v];
  int l = 0;
  if (cv > 0) {
    const int32_t tv = (int32_t)(v / tile);
    for (int32_t a = v2v_off[v]; a < v2v_off[v + 1]; ++a) {
      const int32_t u = v2v[a];
      if (u / tile < tv && color[u] == cv) {
        l = 1;
        break;
      }
    }
  }
  lose[v] = l;
}

// Uncolours the losers and counts them into the round slots the first
// repair round reads.
__global__ __launch_bounds__(256) void k_color_drop(int32_t nv, const int32_t* __restrict__ lose,
                                                    int32_t* __restrict__ color,
                                                    int32_t* __restrict__ slots) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int l = v < nv ? lose[v] : 0;
  if (l) color[v] = 0;
  const int n = __syncthreads_count(l);
  if (threadIdx.x == 0 && n) atomicAdd(&slots[blockIdx.x & (kRoundSlots - 1)], n);
}

__global__ void k_copy_slots(const int32_t* __restrict__ src, int32_t* __restrict__ dst) {
  if (threadIdx.x < kRoundSlots) dst[threadIdx.x] = src[threadIdx.x];
}

// Vertices that gave up (-1) back to 0 (uncoloured, in no class), counted.
__global__ __launch_bounds__(256) void k_color_unset(int32_t nv, int32_t* __restrict__ color,
                                                     int32_t* __restrict__ count) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int g = v < nv && color[v] < 0;
  if (g) color[v] = 0;
  const int n = __syncthreads_count(g);
  if (threadIdx.x == 0 && n) atomicAdd(count, n);
}

// Per-block histogram of colours 1..ncolors, stored colour-major; a block
// covers kClassPer consecutive chunks of kClassB positions.  Position k of
// the visiting sequence is vertex perm[k] (or k without perm).
__global__ __launch_bounds__(kClassB) void k_class_count(int32_t nv, const int32_t* __restrict__ color,
                                                         const int32_t* __restrict__ perm,
                                                         int ncolors, int32_t* __restrict__ cnt) {
  __shared__ int hist[64];
  if (threadIdx.x < 64) hist[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (int s = 0; s < kClassPer; ++s) {
    const int64_t k = ((int64_t)blockIdx.x * kClassPer + s) * kClassB + threadIdx.x;
    int c = 0;
    if (k < nv) {
      c = color[perm ? perm[k] : k];
      if (c < 1 || c > ncolors) c = 0;
    }
    // one LDS atomic per colour present in the wave (lanes of one colour
    // are counted with a ballot)
    uint64_t todo = __ballot(c > 0);
    while (todo) {
      const int leader = __ffsll((unsigned long long)todo) - 1;
      const int lc = __shfl(c, leader, 64);
      const uint64_t m = __ballot(c == lc);
      if (lane == leader) atomicAdd(&hist[lc - 1], __popcll(m));
      todo &= ~m;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < ncolors) cnt[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = hist[threadIdx.x];
}

// Scatter of the block's vertices into their classes, chunk by chunk, so
// that each class keeps the visiting order.
__global__ __launch_bounds__(kClassB) void k_class_scatter(int32_t nv, const int32_t* __restrict__ color,
                                                           const int32_t* __restrict__ perm,
                                                           int ncolors,
                                                           const int32_t* __restrict__ off,
                                                           int32_t* __restrict__ order) {
  __shared__ int wcnt[kClassB / 64][64];
  __shared__ int run[64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x < 64) run[threadIdx.x] = 0;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int s = 0; s < kClassPer; ++s) {
    for (int k = threadIdx.x; k < (kClassB / 64) * 64; k += blockDim.x) (&wcnt[0][0])[k] = 0;
    __syncthreads();
    const int64_t k = ((int64_t)blockIdx.x * kClassPer + s) * kClassB + threadIdx.x;
    const int64_t v = (k < nv && perm) ? perm[k] : k;
    int c = 0;
    if (k < nv) {
      c = color[v];
      if (c < 1 || c > ncolors) c = 0;
    }
    // rank among equal colours inside the wave (lanes in visiting order)
    uint64_t todo = __ballot(1);
    int rank = 0;
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
      int base = off[(int64_t)(c - 1) * gridDim.x + blockIdx.x] + run[c - 1];
      for (int w = 0; w < wid; ++w) base += wcnt[w][c - 1];
      if ((uint32_t)(base + rank) < (uint32_t)nv) order[base + rank] = (int32_t)v;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      int t = 0;
      for (int w = 0; w < kClassB / 64; ++w) t += wcnt[w][threadIdx.x];
      run[threadIdx.x] += t;
    }
    // the counts are read above before the next chunk clears them
    __syncthreads();
  }
}

// Exclusive scan, in one block, of the colour-major per-block counts cnt
// (rows x nblocks) into off (rows * nblocks + 1 entries), and the status
// words the host reads back in one copy:
//   status[0]     palette failure flag (0 without fail),
//   status[1]     losers of the last queued round (0 without last_slots),
//   status[2]     colours in use: the largest c <= rows with a vertex,
//   status[3 + c] first position of colour c + 1 in the class order, for
//                 c = 0..rows (status[3 + rows] = number of listed vertices).
// Each thread scans a contiguous chunk; the chunk totals are scanned in LDS.
__global__ __launch_bounds__(kScanB) void k_class_scan(const int32_t* __restrict__ cnt, int rows,
                                                       int64_t nblocks, int32_t* __restrict__ off,
                                                       const int32_t* __restrict__ fail,
                                                       const int32_t* __restrict__ last_slots,
                                                       int32_t* __restrict__ status) {
  __shared__ int part[kScanB];
  __shared__ int coff[kMaxRows + 1];
  const int t = threadIdx.x;
  const int64_t n = (int64_t)rows * nblocks;
  const int64_t per = (n + kScanB - 1) / kScanB;
  const int64_t b0 = per * t < n ? per * t : n;
  const int64_t b1 = b0 + per < n ? b0 + per : n;
  if (t <= rows) coff[t] = 0;
  int s = 0;
  for (int64_t k = b0; k < b1; ++k) s += cnt[k];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < kScanB; o <<= 1) {  // inclusive scan of the chunk totals
    const int x = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  int run = t ? part[t - 1] : 0;
  // the first row start at or after b0; rows start at multiples of nblocks
  int64_t next = nblocks ? (b0 + nblocks - 1) / nblocks * nblocks : n;
  for (int64_t k = b0; k < b1; ++k) {
    if (k == next) {
      coff[k / nblocks] = run;
      next += nblocks;
    }
    const int c = cnt[k];
    off[k] = run;
    run += c;
  }
  if (t == kScanB - 1) {
    off[n] = run;
    coff[rows] = run;
  }
  __syncthreads();
  if (t <= rows) status[3 + t] = coff[t];
  if (t == 0) {
    int nc = 0;
    for (int c = 0; c < rows; ++c)
      if (coff[c + 1] > coff[c]) nc = c + 1;
    status[0] = fail ? *fail : 0;
    status[2] = nc;
  }
  if (t < kRoundSlots) {
    int l = last_slots ? last_slots[t] : 0;
#pragma unroll
    for (int o = kRoundSlots / 2; o > 0; o >>= 1) l += __shfl_xor(l, o, kRoundSlots);
    if (t == 0) status[1] = l;
  }
}

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---- independent set for coarsening ---------------------------------------
// Partit::extractIndepSet (src/partition.cpp:200-306) on the subgraph that
// Coarse::extractSubGraph builds (src/coarsening.cpp:307-319): the filtered
// vertices with their 1-rings, unfiltered neighbours never conflicting.  The
// reference keeps the colour-1 class of a speculative first fit, a maximal
// independent set whose members depend on the thread interleaving.  Here
// the set is the greedy one in decreasing (seeded priority, id) order,
// reached in rounds: an undecided vertex leaves when a neighbour has joined
// and joins when no undecided neighbour outranks it.  A vertex that reads a
// neighbour's earlier state in a round only waits a round, so the result
// does not depend on timing; the top undecided vertex always decides, so
// every round makes progress.  state: 0 outside the subgraph, 1 undecided,
// 2 in the set, 3 out.
constexpr int kMisB = 256;
constexpr int kMisRoundsPerSync = 16;  // the set needs 9 to 16 rounds on grids of 10^5 to 10^7

__global__ __launch_bounds__(kMisB) void k_mis_mark(int32_t n, const int32_t* __restrict__ nodes,
                                                    int32_t nv, int32_t* __restrict__ state,
                                                    int32_t* bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t v = nodes[i];
  if (v < 0 || v >= nv) {
    atomicOr(bad, 1);
    return;
  }
  state[v] = 1;
}

__global__ __launch_bounds__(kMisB) void k_mis_round(int32_t n, const int32_t* __restrict__ nodes,
                                                     int32_t nv,
                                                     const int32_t* __restrict__ v2v_off,
                                                     const int32_t* __restrict__ v2v,
                                                     int32_t* state, uint64_t seed,
                                                     int32_t* remaining,
                                                     const int32_t* __restrict__ prev_slots) {
  if (prev_losers(prev_slots) == 0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int undecided = 0;
  if (i < n) {
    const int32_t v = nodes[i];
    if (v >= 0 && v < nv && state[v] == 1) {  // ids outside [0, nv) are reported by the host
      const uint32_t pv = color_priority(seed, (uint32_t)v);
      int join = 1, out = 0;
      for (int32_t a = v2v_off[v]; a < v2v_off[v + 1]; ++a) {
        const int32_t u = v2v[a];
        const int32_t su = state[u];
        if (su == 2) {
          out = 1;
          break;
        }
        if (su == 1) {
          const uint32_t pu = color_priority(seed, (uint32_t)u);
          if (pu > pv || (pu == pv && u > v)) join = 0;
        }
      }
      if (out) state[v] = 3;
      else if (join) state[v] = 2;
      else undecided = 1;
    }
  }
  const int left = __syncthreads_count(undecided);
  if (threadIdx.x == 0 && left) atomicAdd(&remaining[blockIdx.x & (kRoundSlots - 1)], left);
}

__global__ __launch_bounds__(kMisB) void k_mis_flags(int32_t n, const int32_t* __restrict__ nodes,
                                                     int32_t nv, const int32_t* __restrict__ state,
                                                     int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t v = nodes[i];
  flag[i] = v >= 0 && v < nv && state[v] == 2;
}

__global__ __launch_bounds__(kMisB) void k_mis_scatter(int32_t n, const int32_t* __restrict__ nodes,
                                                       const int32_t* __restrict__ flag,
                                                       const int32_t* __restrict__ pos,
                                                       int32_t* __restrict__ set) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) set[pos[i]] = nodes[i];
}

constexpr int kStatusWords = 3 + kMaxRows + 1;

struct ClassBufs {
  int32_t* cnt;
  int32_t* off;
  int32_t* status;
};

static size_t class_bytes(int32_t nv, int rows) {
  const int64_t n = (int64_t)nblk(nv, kClassB * kClassPer) * rows;
  return 2 * scratch_bytes(sizeof(int32_t) * (size_t)(n + 1)) +
         scratch_bytes(sizeof(int32_t) * kStatusWords);
}

static ClassBufs carve_classes(Scratch& sc, int32_t nv, int rows) {
  const int64_t n = (int64_t)nblk(nv, kClassB * kClassPer) * rows;
  ClassBufs b;
  b.cnt = sc.carve<int32_t>((size_t)n + 1);
  b.off = sc.carve<int32_t>((size_t)n + 1);
  b.status = sc.carve<int32_t>(kStatusWords);
  return b;
}

// Queues the class build for colours 1..rows (per-block counts, the scan,
// which also fills the status words, and the scatter into order, visiting
// the vertices in perm order when given) and the copy of the status words
// to ctx->pinned_i32.  The caller synchronises before reading them.
static void queue_classes(trn_ctx* ctx, int32_t nv, const int32_t* color, const int32_t* perm,
                          int rows, const int32_t* fail, const int32_t* last_slots,
                          int32_t* order, const ClassBufs& b) {
  const int64_t nb = nblk(nv, kClassB * kClassPer);
  if (nb > 0)
    launch(ctx, "class_count", 4.0 * nv, k_class_count, dim3((unsigned)nb), dim3(kClassB), 0, nv,
           color, perm, rows, b.cnt);
  launch(ctx, "class_scan", 0.0, k_class_scan, dim3(1), dim3(kScanB), 0, (const int32_t*)b.cnt,
         rows, nb, b.off, fail, last_slots, b.status);
  if (nb > 0)
    launch(ctx, "class_scatter", 8.0 * nv, k_class_scatter, dim3((unsigned)nb), dim3(kClassB), 0,
           nv, color, perm, rows, (const int32_t*)b.off, order);
  copy_async(ctx, ctx->pinned_i32, b.status, sizeof(int32_t) * (size_t)(3 + rows + 1));
}

}  // namespace trn

using namespace trn;

extern "C" {

}  // extern "C"

namespace trn {

// Tile size of the colouring's first step for a mesh: large enough that a
// tile of a mesh numbered row by row holds a few rows (few seam conflicts),
// small enough for one tile per wave to fill the chip.
static int color_tile_size(const trn_mesh* m) {
  if (m->color_tile > 0) return m->color_tile;
  // no edge may join two tiles of the same parity: the tile must exceed the
  // longest edge of the numbering (rounded up to whole 64-vertex chunks);
  // past kTileMax the seams are repaired instead
  if (m->bandwidth >= kTileMax) return 4096;
  const int t = (m->bandwidth + 1 + kTileB - 1) / kTileB * kTileB;
  return t < 1024 ? 1024 : t;
}

static void color_impl(trn_ctx* ctx, const trn_mesh* mesh, uint64_t seed, int32_t max_colors,
                       bool leave, int32_t* color_d, int32_t* order_d, int32_t* class_off,
                       int32_t* ncolors, int64_t* uncoloured) {
  TRN_REQUIRE(mesh != nullptr && color_d && order_d && class_off && ncolors, TRN_ERR_ARG,
              "null argument");
  TRN_REQUIRE(max_colors >= 1 && max_colors <= kMaxRows, TRN_ERR_ARG,
              "max_colors must be in 1..64");
  const int32_t nv = mesh->nv;
  // first fit gives a vertex at most (1-ring size + 1) colours
  const int rows = max_colors < mesh->max_deg_v + 1 ? max_colors : mesh->max_deg_v + 1;
  const size_t nslots = (size_t)(kRoundsPerSync + 1) * kRoundSlots;
  Scratch sc(ctx, scratch_bytes(sizeof(int32_t) * (size_t)nv) +
                      scratch_bytes(sizeof(int32_t) * (16 + nslots)) + class_bytes(nv, rows));
  int32_t* tent = sc.carve<int32_t>(nv);
  // [0] failure (palette exhausted, trn_color), [1] vertices left
  // uncoloured (trn_color_partial), [2] tile-step give-ups, then the slots:
  // kRoundSlots for the seams, then per queued round
  int32_t* flags = sc.carve<int32_t>(16 + nslots);
  int32_t* seam_slots = flags + 16;
  int32_t* slots = seam_slots + kRoundSlots;
  const ClassBufs cb = carve_classes(sc, nv, rows);
  memset_async(ctx, flags, 0, sizeof(int32_t) * (16 + nslots));
  const int tile = color_tile_size(mesh);
  const int64_t ntiles = (nv + (int64_t)tile - 1) / tile;
  for (int parity = 0; parity < 2; ++parity) {
    const int64_t nb = (ntiles + 1 - parity) / 2;
    if (nb > 0)
      launch(ctx, "color_tiles", 4.0 * nv + 2.0 * (double)mesh->nnz_v2v, k_color_tiles,
             dim3((unsigned)nb), dim3(kTileB), 0, nv, tile, parity,
             (const int32_t*)mesh->v2v_off, (const int32_t*)mesh->v2v, (int)max_colors,
             color_d, leave ? flags + 2 : flags);
  }
  // a tile longer than every edge of the numbering leaves no edge between
  // two tiles of the same parity, hence no seam conflict to look for
  if (tile <= mesh->bandwidth) {
    launch(ctx, "color_seams", 8.0 * nv + 4.0 * (double)mesh->nnz_v2v, k_color_seams,
           dim3(nblk(nv, 256)), dim3(256), 0, nv, tile, (const int32_t*)mesh->v2v_off,
           (const int32_t*)mesh->v2v, (const int32_t*)color_d, tent);
    launch(ctx, "color_drop", 8.0 * nv, k_color_drop, dim3(nblk(nv, 256)), dim3(256), 0, nv,
           (const int32_t*)tent, color_d, seam_slots);
  }
  const int32_t* h = ctx->pinned_i32;
  // the class build right away: with no seam conflict (the usual case) one
  // synchronisation returns the loser count, the number of colours and the
  // class offsets
  const int32_t* prev = seam_slots;
  int rounds = 0;
  for (;;) {
    queue_classes(ctx, nv, color_d, mesh->class_v, rows, flags, prev, order_d, cb);
    sync(ctx);
    TRN_REQUIRE(h[0] == 0, TRN_ERR_PALETTE,
                "a vertex needs more than " + std::to_string(max_colors) + " colours");
    if (h[1] == 0) break;
    TRN_REQUIRE(rounds < 100000, TRN_ERR_STATE, "colouring did not converge");
    // repair: a batch of speculative rounds, the first reading the losers
    // left so far (the seams' or the previous batch's last round's)
    if (prev != seam_slots)
      launch(ctx, "color_carry", 0.0, k_copy_slots, dim3(1), dim3(kRoundSlots), 0, prev,
             seam_slots);
    prev = seam_slots;
    memset_async(ctx, slots, 0, sizeof(int32_t) * kRoundsPerSync * kRoundSlots);
    for (int r = 0; r < kRoundsPerSync; ++r) {
      int32_t* cur = slots + (size_t)r * kRoundSlots;
      launch(ctx, "color_tent", 4.0 * nv + 4.0 * (double)mesh->nnz_v2v, k_color_tent,
             dim3(nblk(nv, 256)), dim3(256), 0, nv, (const int32_t*)mesh->v2v_off,
             (const int32_t*)mesh->v2v, color_d, tent, (int)max_colors, (int)leave, flags, prev);
      launch(ctx, "color_resolve", 4.0 * nv + 4.0 * (double)mesh->nnz_v2v, k_color_resolve,
             dim3(nblk(nv, 256)), dim3(256), 0, nv, (const int32_t*)mesh->v2v_off,
             (const int32_t*)mesh->v2v, color_d, (const int32_t*)tent, seed, 0, cur, prev);
      prev = cur;
    }
    rounds += kRoundsPerSync;
  }
  const int nc = h[2];
  for (int k = 0; k <= nc; ++k) class_off[k] = h[3 + k];
  *ncolors = nc;
  if (leave) {  // vertices that gave up (-1, in no class) become 0, counted
    launch(ctx, "color_unset", 4.0 * nv, k_color_unset, dim3(nblk(nv, 256)), dim3(256), 0, nv,
           color_d, flags + 1);
    copy_async(ctx, ctx->pinned_i32, flags + 1, sizeof(int32_t));
    sync(ctx);
    if (uncoloured) *uncoloured = ctx->pinned_i32[0];
  }
}

}  // namespace trn

extern "C" {

int trn_color(trn_ctx* ctx, const trn_mesh* mesh, uint64_t seed, int32_t max_colors,
              int32_t* color_d, int32_t* order_d, int32_t* class_off, int32_t* ncolors) {
  return guard_impl("trn_color", [&] {
    DeviceScope ds(ctx);
    color_impl(ctx, mesh, seed, max_colors, false, color_d, order_d, class_off, ncolors,
               nullptr);
  });
}

int trn_color_partial(trn_ctx* ctx, const trn_mesh* mesh, uint64_t seed, int32_t max_colors,
                      int32_t* color_d, int32_t* order_d, int32_t* class_off, int32_t* ncolors,
                      int64_t* uncoloured) {
  return guard_impl("trn_color_partial", [&] {
    DeviceScope ds(ctx);
    color_impl(ctx, mesh, seed, max_colors, true, color_d, order_d, class_off, ncolors,
               uncoloured);
  });
}

int trn_mesh_set_color_tile(trn_mesh* mesh, int32_t tile) {
  return guard_impl("trn_mesh_set_color_tile", [&] {
    TRN_REQUIRE(mesh != nullptr, TRN_ERR_ARG, "null mesh");
    TRN_REQUIRE(tile == 0 || (tile >= 64 && tile <= kTileMax && tile % kTileB == 0),
                TRN_ERR_ARG, "tile must be 0 (automatic) or a multiple of 64 in 64..16384");
    mesh->color_tile = tile;
  });
}

int trn_color_classes(trn_ctx* ctx, const trn_mesh* mesh, int32_t nv, const int32_t* color_d,
                      int32_t ncolors, int32_t* order_d, int32_t* class_off) {
  return guard_impl("trn_color_classes", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(nv >= 0 && color_d && order_d && class_off, TRN_ERR_ARG, "null argument");
    TRN_REQUIRE(ncolors >= 1 && ncolors <= kMaxRows, TRN_ERR_ARG, "ncolors must be in 1..64");
    TRN_REQUIRE(!mesh || mesh->nv == nv, TRN_ERR_ARG, "nv differs from the mesh");
    Scratch sc(ctx, class_bytes(nv, ncolors));
    const ClassBufs cb = carve_classes(sc, nv, ncolors);
    queue_classes(ctx, nv, color_d, mesh ? mesh->class_v : nullptr, ncolors, nullptr, nullptr,
                  order_d, cb);
    sync(ctx);
    for (int k = 0; k <= ncolors; ++k) class_off[k] = ctx->pinned_i32[3 + k];
  });
}

int trn_indep_set(trn_ctx* ctx, const trn_mesh* mesh, const int32_t* nodes_d, int32_t n_nodes,
                  uint64_t seed, int32_t* set_d, int32_t* set_size) {
  return guard_impl("trn_indep_set", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(mesh && set_size && n_nodes >= 0 && (n_nodes == 0 || (nodes_d && set_d)),
                TRN_ERR_ARG, "null argument");
    *set_size = 0;
    if (n_nodes == 0) return;
    const int32_t nv = mesh->nv;
    const size_t nslots = (size_t)kMisRoundsPerSync * kRoundSlots;
    Scratch sc(ctx, scratch_bytes(sizeof(int32_t) * (size_t)nv) +
                        scratch_bytes(sizeof(int32_t) * (16 + nslots)) +
                        2 * scratch_bytes(sizeof(int32_t) * ((size_t)n_nodes + 1)) +
                        exclusive_scan_scratch(n_nodes));
    int32_t* state = sc.carve<int32_t>(nv);
    int32_t* flags = sc.carve<int32_t>(16 + nslots);  // [0] bad vertex id, then the slots
    int32_t* slots = flags + 16;
    int32_t* flag = sc.carve<int32_t>((size_t)n_nodes + 1);
    int32_t* pos = sc.carve<int32_t>((size_t)n_nodes + 1);
    int32_t* tmp = sc.carve<int32_t>(exclusive_scan_scratch(n_nodes) / 4);
    memset_async(ctx, state, 0, sizeof(int32_t) * (size_t)nv);
    memset_async(ctx, flags, 0, sizeof(int32_t) * (16 + nslots));
    launch(ctx, "mis_mark", 4.0 * n_nodes, k_mis_mark, dim3(nblk(n_nodes, kMisB)), dim3(kMisB),
           0, n_nodes, nodes_d, nv, state, flags);
    int32_t* h = ctx->pinned_i32;
    int rounds = 0;
    for (;;) {
      if (rounds) memset_async(ctx, slots, 0, sizeof(int32_t) * nslots);
      for (int r = 0; r < kMisRoundsPerSync; ++r)
        launch(ctx, "mis_round", 0.0, k_mis_round, dim3(nblk(n_nodes, kMisB)), dim3(kMisB), 0,
               n_nodes, nodes_d, nv, (const int32_t*)mesh->v2v_off, (const int32_t*)mesh->v2v,
               state, seed, slots + (size_t)r * kRoundSlots,
               r ? (const int32_t*)(slots + (size_t)(r - 1) * kRoundSlots) : nullptr);
      // the set is compacted with the batch; one synchronisation returns it
      launch(ctx, "mis_flags", 0.0, k_mis_flags, dim3(nblk(n_nodes, kMisB)), dim3(kMisB), 0,
             n_nodes, nodes_d, nv, (const int32_t*)state, flag);
      exclusive_scan_i32(ctx, flag, pos, n_nodes, true, tmp);
      launch(ctx, "mis_scatter", 0.0, k_mis_scatter, dim3(nblk(n_nodes, kMisB)), dim3(kMisB), 0,
             n_nodes, nodes_d, (const int32_t*)flag, (const int32_t*)pos, set_d);
      copy_async(ctx, h, flags, sizeof(int32_t));
      copy_async(ctx, h + 1, pos + n_nodes, sizeof(int32_t));
      copy_async(ctx, h + 2, slots + (size_t)(kMisRoundsPerSync - 1) * kRoundSlots,
                 sizeof(int32_t) * kRoundSlots);
      sync(ctx);
      TRN_REQUIRE(h[0] == 0, TRN_ERR_ARG, "a node id is outside [0, nv)");
      rounds += kMisRoundsPerSync;
      int left = 0;
      for (int k = 0; k < kRoundSlots; ++k) left += h[2 + k];
      if (left == 0) break;
      TRN_REQUIRE(rounds < 100000, TRN_ERR_STATE, "independent set did not converge");
    }
    *set_size = h[1];
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
