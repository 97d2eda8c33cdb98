// smoothing.hip -- coloured metric-aware Laplacian smoothing on gfx950,
// replacing Smooth::run after its colouring (src/smoothing.cpp:274-302).
//
//   k_tensor_log     per vertex: the matrix logarithm term of
//                    numeric::interpolateTensor (numeric.cpp:115-154), cached
//                    as 4 doubles.  The reference recomputes it for every
//                    tensor of every quality evaluation and interpolation
//                    (~60 eigen decompositions and logs per move); the
//                    cached value is the same function of the same tensor,
//                    so results are unchanged.
//   k_quality_cache  per triangle: Mesh::computeQuality (Smooth::cacheQuality,
//                    src/smoothing.cpp:145-155) from the cached logs.
//   k_smooth_phase   one launch per colour class and sweep: every thread
//                    applies Smooth::moveSmartLaplacian (:40-126) to one
//                    vertex.  Vertices of one colour share no triangle and
//                    read only their own 1-ring, so a class is an
//                    independent set and the order inside it does not change
//                    the result; colours run in ascending order
//                    (Smooth::movePoints, :166-175), as in the reference.
//                    The averaged tensor of an incident triangle depends only
//                    on tensors, so it is formed once per move (lazily, on
//                    the first try that reaches the triangle) instead of on
//                    every try; positions change per try and are re-evaluated.
//                    Triangle corners come from the mesh's star codes (the
//                    vertex's slot in the triangle and the 1-ring entries of
//                    the two other corners), so the kernel reads no
//                    triangle array.
#include <algorithm>
#include <mutex>

#include "trn_recovery.hpp"
#include "trn_smooth.hpp"

namespace trn {

constexpr int kCounterSlots = 64;
// Longest live lists a phase moves with 8 and with 2 lanes per move;
// longer ones take a lane per move (measured on the 4096^2 grid: 4 lanes
// per move in between, or the 2-lane form up to 2^20 moves, did no better).
constexpr int32_t kCoop8MaxMoves = 32768;
constexpr int32_t kCoop2MaxMoves = 262144;

__global__ __launch_bounds__(256) void k_tensor_log(int32_t nv, const double* __restrict__ tensor,
                                                    double4* __restrict__ logm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  logm[i] = tensor_log4(tensor[3 * i], tensor[3 * i + 1], tensor[3 * i + 2]);
}

__global__ __launch_bounds__(256) void k_quality_cache(int32_t nt, const int32_t* __restrict__ tri,
                                                       const double2* __restrict__ xy,
                                                       const double4* __restrict__ logm,
                                                       double* __restrict__ qualit) {
  const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  if (tri[3 * t] < 0) return;  // Mesh::isActiveElem (src/mesh.cpp:375)
  qualit[t] = quality_logs(t, tri, xy, logm);
}

// The sweeps' working copy (metric log, coordinates, metric) from the
// caller's arrays, and the logs as a dense array for the quality cache;
// replaces k_tensor_log for the sweeps.  A block stages its 256 records and
// logs in LDS and stores them with consecutive lanes on consecutive 16-byte
// chunks, so every store instruction writes whole lines (a record is 5 LDS
// chunks apart, a log 3, so the lanes writing them hit distinct banks).
constexpr int kPackB = 256;

__global__ __launch_bounds__(kPackB) void k_pack_records(int32_t nv, const double2* __restrict__ xy,
                                                         const double* __restrict__ tensor,
                                                         VData* __restrict__ vd,
                                                         double* __restrict__ t2,
                                                         double4* __restrict__ logm) {
  static_assert(sizeof(VData) == 4 * sizeof(double2), "a record is 4 chunks of 16 bytes");
  __shared__ double2 rec[kPackB * 5];
  __shared__ double2 lg[kPackB * 3];
  const int64_t v0 = (int64_t)blockIdx.x * kPackB;
  const int n = nv - v0 < kPackB ? (int)(nv - v0) : kPackB;
  const int tid = threadIdx.x;
  if (tid < n) {
    const int64_t i = v0 + tid;
    const double t0 = tensor[3 * i], t1 = tensor[3 * i + 1], t22 = tensor[3 * i + 2];
    const double4 L = tensor_log4(t0, t1, t22);
    double2* r = rec + 5 * tid;
    r[0] = make_double2(L.x, L.y);
    r[1] = make_double2(L.z, L.w);
    r[2] = xy[i];
    r[3] = make_double2(t0, t1);
    lg[3 * tid] = r[0];
    lg[3 * tid + 1] = r[1];
    t2[i] = t22;
  }
  __syncthreads();
  double2* vdc = reinterpret_cast<double2*>(vd + v0);
  for (int c = tid; c < 4 * n; c += kPackB) vdc[c] = rec[5 * (c >> 2) + (c & 3)];
  double2* lgc = reinterpret_cast<double2*>(logm + v0);
  for (int c = tid; c < 2 * n; c += kPackB) lgc[c] = lg[3 * (c >> 1) + (c & 1)];
}

// The phase over a live list of n vertices with G lanes per move, when the
// grid holds the blocks that takes (true: the block has done its part).
template <int G>
__device__ __forceinline__ bool coop_phase(
    const int32_t n, const int32_t* __restrict__ order, VData* __restrict__ vd,
    double* __restrict__ t2, const VTopo* __restrict__ vtopo, double2* __restrict__ xy_out,
    double* __restrict__ t_out, double* __restrict__ qualit, uint8_t* __restrict__ stale,
    const int32_t* __restrict__ v2t_off, const int32_t* __restrict__ v2t,
    const int32_t* __restrict__ star, const int32_t* __restrict__ v2v_off,
    const int32_t* __restrict__ v2v, int32_t* __restrict__ counters, double* lds) {
  const uint32_t vpb = blockDim.x / G;
  const uint32_t coop_blocks = (uint32_t)((n + (int32_t)vpb - 1) / (int32_t)vpb);
  if (coop_blocks > gridDim.x) return false;
  if (blockIdx.x >= coop_blocks) return true;
  const uint32_t lb = xcd_block_of(blockIdx.x, coop_blocks);
  const int sub = (int)(threadIdx.x & (G - 1));
  const int base = (int)(threadIdx.x & 63) & ~(G - 1);
  const int64_t idx = (int64_t)lb * vpb + (threadIdx.x / G);
  int commit = 0, task = 0;
  if (idx < n)
    commit = move_vertex_rec_coop<G>(order[idx], sub, base, vd, t2, vtopo, xy_out, t_out, qualit,
                                     stale, v2t_off, v2t, star, v2v_off, v2v, lds + threadIdx.x,
                                     blockDim.x, 0, &task);
  const uint64_t mc = __ballot(commit && sub == 0), mt = __ballot(task && sub == 0);
  if ((threadIdx.x & 63) == 0) {
    int32_t* slot = counters + 2 * (blockIdx.x & (kCounterSlots - 1));
    if (mc) atomicAdd(&slot[0], (int32_t)__popcll(mc));
    if (mt) atomicAdd(&slot[1], (int32_t)__popcll(mt));
  }
  return true;
}

// One colour phase.  count_d == nullptr: `order` lists `count` vertices
// (the whole class); otherwise it is a live list of *count_d <= count
// vertices written by k_class_live, the grid is sized for `count`, and the
// blocks past the live ones exit at once while the others take XCD-chunked
// ranges of the live blocks.  MINW: waves per SIMD the register allocation
// is made for (2, 3 or 4; the mesh's tuned shape), the same arithmetic.
template <int MINW>
__global__ __launch_bounds__(64, MINW) void k_smooth_phase(
    int32_t count, const int32_t* __restrict__ count_d, const int32_t* __restrict__ order,
    VData* __restrict__ vd, double* __restrict__ t2, const VTopo* __restrict__ vtopo,
    double2* __restrict__ xy_out, double* __restrict__ t_out,
    double* __restrict__ qualit, uint8_t* __restrict__ stale,
    const int32_t* __restrict__ v2t_off, const int32_t* __restrict__ v2t,
    const int32_t* __restrict__ star, const int32_t* __restrict__ v2v_off,
    const int32_t* __restrict__ v2v, int32_t* __restrict__ counters,
    unsigned long long* __restrict__ units) {
  extern __shared__ double lds[];
  int32_t n = count;
  uint32_t lb;
  if (count_d) {
    n = *count_d;
    if (units && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(units, (unsigned long long)n);
    // a short live list: several lanes per move (the phase then lasts
    // about one move's latency, which the cooperative form shortens)
    if (n <= kCoop8MaxMoves && coop_phase<8>(n, order, vd, t2, vtopo, xy_out, t_out, qualit,
                                              stale, v2t_off, v2t, star, v2v_off, v2v, counters,
                                              lds))
      return;
    if (n <= kCoop2MaxMoves && coop_phase<2>(n, order, vd, t2, vtopo, xy_out, t_out, qualit,
                                              stale, v2t_off, v2t, star, v2v_off, v2v, counters,
                                              lds))
      return;
    const uint32_t live_blocks = (uint32_t)((n + (int32_t)blockDim.x - 1) / (int32_t)blockDim.x);
    if (blockIdx.x >= live_blocks) return;
    lb = xcd_block_of(blockIdx.x, live_blocks);
  } else {
    lb = xcd_block();
  }
  const int64_t idx = (int64_t)lb * blockDim.x + threadIdx.x;
  int commit = 0, task = 0;
  if (idx < n)
    commit = move_vertex_rec(order[idx], vd, t2, vtopo, xy_out, t_out, qualit, stale, v2t_off,
                             v2t, star, v2v_off, v2v,
                             lds + threadIdx.x, blockDim.x, count_d == nullptr, &task);
  // counters are spread over kCounterSlots addresses: one hot address
  // serialises the atomics of every wave in L2
  const uint64_t mc = __ballot(commit), mt = __ballot(task);
  if ((threadIdx.x & 63) == 0) {
    int32_t* slot = counters + 2 * (blockIdx.x & (kCounterSlots - 1));
    if (mc) atomicAdd(&slot[0], (int32_t)__popcll(mc));
    if (mt) atomicAdd(&slot[1], (int32_t)__popcll(mt));
  }
}

// Live list of a colour class: the vertices not marked stale, appended one
// block of kLiveB * kLivePer class entries at a time with one atomic per
// block (a single counter takes every block's atomic in turn, so fewer,
// larger blocks finish sooner).  A block's vertices keep their class order,
// so each phase block still covers a compact region; the order of the
// blocks does not change the result, a class being an independent set.
// Stale vertices are counted as tasks.  The block's offsets are an
// exclusive scan over its (entry round, wave) groups in class order, done
// by the first kLiveGroups / 64 waves with shuffles.
constexpr int kLiveB = 1024;
constexpr int kLivePer = 8;
constexpr int kLiveWaves = kLiveB / 64;
constexpr int kLiveGroups = kLivePer * kLiveWaves;  // (entry round, wave) groups
static_assert(kLiveGroups % 64 == 0 && kLiveGroups <= kLiveB, "one scan lane per group");

__global__ __launch_bounds__(kLiveB) void k_class_live(int32_t count,
                                                       const int32_t* __restrict__ order,
                                                       const uint8_t* __restrict__ stale,
                                                       int32_t* __restrict__ live,
                                                       int32_t* __restrict__ live_count,
                                                       int32_t* __restrict__ counters) {
  __shared__ int32_t grp_live[kLiveGroups], wave_skip[kLiveWaves], part[kLiveGroups / 64], base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t first = (int64_t)xcd_block() * (kLiveB * kLivePer) + threadIdx.x;
  const uint64_t below = (1ull << lane) - 1;
  int32_t v[kLivePer];
  uint64_t ml[kLivePer];
  int nskip = 0;
#pragma unroll
  for (int j = 0; j < kLivePer; ++j) {  // entry first + j * kLiveB: class order is (j, thread)
    const int64_t idx = first + (int64_t)j * kLiveB;
    v[j] = idx < count ? order[idx] : 0;
  }
#pragma unroll
  for (int j = 0; j < kLivePer; ++j) {
    const int64_t idx = first + (int64_t)j * kLiveB;
    bool is_live = false, skipped = false;
    if (idx < count) {
      skipped = stale[v[j]] != 0;
      is_live = !skipped;
    }
    ml[j] = __ballot(is_live);
    nskip += (int)__popcll(__ballot(skipped));
    if (lane == 0) grp_live[j * kLiveWaves + wave] = (int32_t)__popcll(ml[j]);
  }
  if (lane == 0) wave_skip[wave] = nskip;
  __syncthreads();
  int32_t own = 0, incl = 0;
  if (threadIdx.x < kLiveGroups) {
    own = grp_live[threadIdx.x];
    incl = own;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    if (lane == 63) part[wave] = incl;
  }
  __syncthreads();
  if (threadIdx.x < kLiveGroups) {
    int32_t before = 0;
    for (int w = 0; w < wave; ++w) before += part[w];
    grp_live[threadIdx.x] = before + incl - own;  // exclusive prefix in class order
    if (threadIdx.x == kLiveGroups - 1) {
      const int32_t nl = before + incl;
      int32_t nsk = 0;
      for (int w = 0; w < kLiveWaves; ++w) nsk += wave_skip[w];
      base = nl ? atomicAdd(live_count, nl) : 0;
      if (nsk) atomicAdd(&counters[2 * (blockIdx.x & (kCounterSlots - 1)) + 1], nsk);
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kLivePer; ++j)
    if ((ml[j] >> lane) & 1)
      live[base + grp_live[j * kLiveWaves + wave] + (int32_t)__popcll(ml[j] & below)] = v[j];
}

// counters[s][slot][2] -> out[s][2]
__global__ void k_counter_reduce(int32_t sweeps, const int32_t* __restrict__ counters,
                                 int32_t* __restrict__ out) {
  const int s = threadIdx.x;
  if (s >= sweeps) return;
  int c = 0, t = 0;
  for (int k = 0; k < kCounterSlots; ++k) {
    c += counters[(s * kCounterSlots + k) * 2];
    t += counters[(s * kCounterSlots + k) * 2 + 1];
  }
  out[2 * s] = c;
  out[2 * s + 1] = t;
}

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Algorithmic bytes of one smoothing sweep per vertex: xy16 M24 v2v28 v2t28
// tri24 qualit16 colour-order4 read, xy16 M24 qualit16 written.
constexpr double kBytesSweepPerV = 196.0;

// Threads per block of the phase kernel and its LDS bytes per block.  Per
// thread: 5 doubles for each of at most kVtopoDeg incident triangles on the
// fast path, and 4 per incident triangle of the largest degree when the
// mesh has a vertex of more than kVtopoDeg incident triangles or neighbours
// (generic path).  At most 64 KiB per block.
static void phase_shape(const trn_mesh* m, int* block, size_t* shmem) {
  const bool generic = m->max_deg_t > kVtopoDeg || m->max_deg_v > kVtopoDeg;
  const int per_thread = std::max(5 * kVtopoDeg, generic ? 4 * m->max_deg_t : 0);
  int b = 64;
  while (b > 16 && (size_t)b * per_thread * sizeof(double) > 65536) b >>= 1;
  *block = b;
  *shmem = (size_t)b * per_thread * sizeof(double);
}

using PhaseKernel = void (*)(int32_t, const int32_t*, const int32_t*, VData*, double*,
                            const VTopo*, double2*, double*, double*, uint8_t*, const int32_t*,
                            const int32_t*, const int32_t*, const int32_t*, const int32_t*,
                            int32_t*, unsigned long long*);

static PhaseKernel phase_kernel(int minw) {
  return minw <= 2 ? k_smooth_phase<2> : (minw >= 4 ? k_smooth_phase<4> : k_smooth_phase<3>);
}

// One first-sweep phase over `count` vertices of `list` at the given shape,
// bracketed by ev[0..1] (the autotuner's measurement).
void tune_phase(trn_ctx* ctx, const trn_mesh* m, int32_t count, const int32_t* list,
                double* qualit, uint8_t* stale, int32_t* counters, int minw, hipEvent_t ev[2]) {
  int block;
  size_t shmem;
  phase_shape(m, &block, &shmem);
  TRN_HIP(hipEventRecord(ev[0], ctx->stream));
  launch(ctx, "smooth_phase", 0.0, phase_kernel(minw), dim3(nblk(count, block)), dim3(block),
         shmem, count, (const int32_t*)nullptr, list, m->vdata, m->vt2,
         (const VTopo*)m->vtopo, (double2*)nullptr, (double*)nullptr, qualit, stale,
         (const int32_t*)m->v2t_off, (const int32_t*)m->v2t, (const int32_t*)m->star,
         (const int32_t*)m->v2v_off, (const int32_t*)m->v2v, counters,
         (unsigned long long*)nullptr);
  TRN_HIP(hipEventRecord(ev[1], ctx->stream));
}

static void tensor_logs(trn_ctx* ctx, const trn_mesh* m, const double* tensor, double4* logm) {
  launch(ctx, "tensor_log", 56.0 * m->nv, k_tensor_log, dim3(nblk(m->nv, 256)), dim3(256), 0,
         m->nv, tensor, logm);
}

static void quality_cache_logs(trn_ctx* ctx, const trn_mesh* m, const double* xy,
                               const double4* logm, double* qualit) {
  launch(ctx, "quality_cache", 0.0, k_quality_cache, dim3(nblk(m->nt, 256)), dim3(256), 0, m->nt,
         (const int32_t*)m->tri, (const double2*)xy, logm, qualit);
}

// Mesh::computeQuality(double[3]) (src/mesh.cpp:647-681) reduces the
// qualities of the active triangles to their minimum, maximum, sum and
// count.  Block b takes a fixed contiguous range of triangles; sums go
// through fixed-order trees, so the result does not depend on timing.  The
// comparisons are the reference's (q_min > q, q_max < q), so a NaN quality
// changes neither extreme but makes the sum NaN, as there.
constexpr int kStatB = 256;
constexpr int kStatBlocks = 1024;

__device__ __forceinline__ void block_min_max(double& lo, double& hi, double* lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double l2 = __shfl_xor(lo, o, 64), h2 = __shfl_xor(hi, o, 64);
    lo = dmin(lo, l2);
    hi = dmax(hi, h2);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    lds[2 * wid] = lo;
    lds[2 * wid + 1] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      lo = dmin(lo, lds[2 * w]);
      hi = dmax(hi, lds[2 * w + 1]);
    }
  __syncthreads();
}

__global__ __launch_bounds__(kStatB) void k_quality_partials(int32_t nt,
                                                             const int32_t* __restrict__ tri,
                                                             const double* __restrict__ qualit,
                                                             double* __restrict__ part) {
  __shared__ double lds[2 * (kStatB / 64)];
  const int64_t chunk = ((int64_t)nt + gridDim.x - 1) / gridDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * chunk;
  const int64_t t1 = t0 + chunk < nt ? t0 + chunk : nt;
  double lo = 3., hi = 0., s = 0., c = 0.;  // mesh.cpp:649-652
  for (int64_t t = t0 + threadIdx.x; t < t1; t += kStatB) {
    if (tri[3 * t] < 0) continue;  // mesh.cpp:660-661
    const double q = qualit[t];
    if (lo > q) lo = q;
    if (hi < q) hi = q;
    s += q;
    c += 1.;
  }
  s = block_sum(s, lds);
  c = block_sum(c, lds);
  block_min_max(lo, hi, lds);
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x] = lo;
    part[4 * blockIdx.x + 1] = hi;
    part[4 * blockIdx.x + 2] = s;
    part[4 * blockIdx.x + 3] = c;
  }
}

__global__ __launch_bounds__(kStatB) void k_quality_final(int32_t nparts,
                                                          const double* __restrict__ part,
                                                          double* __restrict__ out) {
  __shared__ double lds[2 * (kStatB / 64)];
  double lo = 3., hi = 0., s = 0., c = 0.;
  for (int b = threadIdx.x; b < nparts; b += kStatB) {
    lo = dmin(lo, part[4 * b]);
    hi = dmax(hi, part[4 * b + 1]);
    s += part[4 * b + 2];
    c += part[4 * b + 3];
  }
  s = block_sum(s, lds);
  c = block_sum(c, lds);
  block_min_max(lo, hi, lds);
  if (threadIdx.x == 0) {
    out[0] = lo;
    out[1] = hi;
    out[2] = s;
    out[3] = c;
  }
}

// ---- internal launch sequences (also used by the batched pass) -----------

void smooth_prepare_launch(trn_ctx* ctx, const trn_mesh* m, const double* xy,
                           const double* tensor, double* qualit, bool recompute_qualit,
                           double4* logm) {
  launch(ctx, "pack_records", 144.0 * m->nv, k_pack_records, dim3(nblk(m->nv, kPackB)),
         dim3(kPackB), 0, m->nv, (const double2*)xy, tensor, m->vdata, m->vt2, logm);
  if (recompute_qualit) quality_cache_logs(ctx, m, xy, logm, qualit);
}

size_t smooth_sweep_scratch(const trn_mesh* m, int ncolors, int sweeps, int max_class) {
  const size_t ncount = 2 * (size_t)kCounterSlots * (size_t)(sweeps + 1);
  const size_t nlive = (size_t)(sweeps + 1) * (size_t)(ncolors + 1);
  return scratch_bytes(sizeof(int32_t) * (ncount + nlive)) +
         scratch_bytes(sizeof(int32_t) * 2 * (size_t)(sweeps + 1)) +
         scratch_bytes((size_t)m->nv) + scratch_bytes(sizeof(int32_t) * (size_t)max_class);
}

int32_t* smooth_sweeps_launch(trn_ctx* ctx, const trn_mesh* m, double* xy, double* tensor,
                              double* qualit, const int32_t* order_d, const int32_t* class_off,
                              int ncolors, int sweeps, int max_class, void* scratch,
                              size_t scratch_size, bool want_totals) {
  const size_t ncount = 2 * (size_t)kCounterSlots * (size_t)(sweeps + 1);
  const size_t nlive = (size_t)(sweeps + 1) * (size_t)(ncolors + 1);
  Carve sc(scratch, scratch_size);
  int32_t* counters = sc.carve<int32_t>(ncount + nlive);
  int32_t* live_count = counters + ncount;
  int32_t* totals = sc.carve<int32_t>(2 * (size_t)(sweeps + 1));
  uint8_t* stale = sc.carve<uint8_t>(m->nv);
  int32_t* live = sc.carve<int32_t>((size_t)(max_class > 0 ? max_class : 1));
  memset_async(ctx, counters, 0, sizeof(int32_t) * (ncount + nlive));
  int block;
  size_t shmem;
  phase_shape(m, &block, &shmem);
  for (int s = 0; s < sweeps; ++s) {
    int32_t* cs = counters + 2 * kCounterSlots * s;
    for (int c = 0; c < ncolors; ++c) {
      const int32_t count = class_off[c + 1] - class_off[c];
      if (count == 0) continue;
      const int32_t* list = order_d + class_off[c];
      const int32_t* count_d = nullptr;
      if (s > 0) {  // nothing is stale before the first sweep
        int32_t* lc = live_count + (size_t)s * (ncolors + 1) + c;
        launch(ctx, "class_live", 5.0 * count, k_class_live, dim3(nblk(count, kLiveB * kLivePer)),
               dim3(kLiveB), 0, count, list, (const uint8_t*)stale, live, lc, cs);
        list = live;
        count_d = lc;
      }
      // algorithmic bytes: per moved vertex; a live list's length is known
      // on the device only, so its launch reports it as work units
      unsigned long long* units =
          count_d ? timing_units(ctx, "smooth_phase", kBytesSweepPerV) : nullptr;
      launch(ctx, "smooth_phase", count_d ? 0.0 : kBytesSweepPerV * count,
             phase_kernel(m->tune.phase_minw), dim3(nblk(count, block)), dim3(block), shmem,
             count, count_d, list, m->vdata, m->vt2, (const VTopo*)m->vtopo, (double2*)xy,
             tensor, qualit, stale,
             (const int32_t*)m->v2t_off,
             (const int32_t*)m->v2t, (const int32_t*)m->star, (const int32_t*)m->v2v_off,
             (const int32_t*)m->v2v, cs, units);
    }
  }
  if (!want_totals || sweeps == 0) return nullptr;
  launch(ctx, "counter_reduce", 0.0, k_counter_reduce, dim3(1), dim3(1024), 0, sweeps,
         (const int32_t*)counters, totals);
  return totals;
}

// ---- allocation-free sweeps ------------------------------------------------
// Device memory of trn_smooth_run_prepared_noalloc: a file-scope array, one
// copy per GPU, held until the process ends.  Host code never uses the
// array's name as an address (it is a host-side stand-in there): the device
// address comes from hipGetSymbolAddress, once per device, with that device
// current, on the entry point's first call on it.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
constexpr int kNoAllocDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kNoAllocDevices] = {};

static void* noalloc_scratch(int device) {
  TRN_REQUIRE(device >= 0 && device < kNoAllocDevices, TRN_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    TRN_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    TRN_REQUIRE(ptr != nullptr, TRN_ERR_HIP, "hipGetSymbolAddress returned a null address");
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

}  // namespace trn

using namespace trn;

extern "C" {

int trn_quality_cache(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                      const double* tensor_d, double* qualit_d) {
  return guard_impl("trn_quality_cache", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(mesh && xy_d && tensor_d && qualit_d, TRN_ERR_ARG, "null argument");
    Scratch sc(ctx, scratch_bytes(sizeof(double4) * (size_t)mesh->nv));
    double4* logm = sc.carve<double4>(mesh->nv);
    tensor_logs(ctx, mesh, tensor_d, logm);
    quality_cache_logs(ctx, mesh, xy_d, logm, qualit_d);
  });
}

int trn_quality_stats(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                      const double* tensor_d, double* qualit_d, double quality[3],
                      int64_t* active) {
  return guard_impl("trn_quality_stats", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(mesh && xy_d && tensor_d && qualit_d && quality, TRN_ERR_ARG, "null argument");
    Scratch sc(ctx, scratch_bytes(sizeof(double4) * (size_t)mesh->nv) +
                        scratch_bytes(sizeof(double) * 4 * (size_t)kStatBlocks) +
                        scratch_bytes(sizeof(double) * 4));
    double4* logm = sc.carve<double4>(mesh->nv);
    double* part = sc.carve<double>(4 * (size_t)kStatBlocks);
    double* res = sc.carve<double>(4);
    tensor_logs(ctx, mesh, tensor_d, logm);
    quality_cache_logs(ctx, mesh, xy_d, logm, qualit_d);
    launch(ctx, "quality_partials", 8.0 * mesh->nt, k_quality_partials, dim3(kStatBlocks),
           dim3(kStatB), 0, mesh->nt, (const int32_t*)mesh->tri, (const double*)qualit_d, part);
    launch(ctx, "quality_final", 0.0, k_quality_final, dim3(1), dim3(kStatB), 0,
           (int32_t)kStatBlocks, (const double*)part, res);
    copy_async(ctx, ctx->pinned_f64, res, 4 * sizeof(double));
    sync(ctx);
    const double* h = ctx->pinned_f64;
    // mesh.cpp:671-679: fold into the caller's values, then the mean
    quality[0] = (h[0] < quality[0]) ? h[0] : quality[0];  // std::min(quality[0], q_min)
    quality[1] = (quality[1] < h[1]) ? h[1] : quality[1];  // std::max(quality[1], q_max)
    quality[2] += h[2];
    quality[2] /= h[3];
    if (active) *active = (int64_t)h[3];
  });
}

int trn_smooth_prepare(trn_ctx* ctx, const trn_mesh* mesh, const double* xy_d,
                       const double* tensor_d, double* qualit_d, int32_t recompute_qualit) {
  return guard_impl("trn_smooth_prepare", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(mesh && xy_d && tensor_d && qualit_d, TRN_ERR_ARG, "null argument");
    Scratch sc(ctx, scratch_bytes(sizeof(double4) * (size_t)mesh->nv));
    smooth_prepare_launch(ctx, mesh, xy_d, tensor_d, qualit_d, recompute_qualit != 0,
                          sc.carve<double4>(mesh->nv));
  });
}

int trn_smooth_run_prepared(trn_ctx* ctx, const trn_mesh* mesh, double* xy_d, double* tensor_d,
                            double* qualit_d, const int32_t* order_d, const int32_t* class_off,
                            int32_t ncolors, int32_t sweeps, int32_t* commits, int32_t* tasks) {
  return guard_impl("trn_smooth_run_prepared", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(mesh && xy_d && tensor_d && qualit_d && order_d && class_off, TRN_ERR_ARG,
                "null argument");
    TRN_REQUIRE(ncolors >= 0 && ncolors <= 64 && sweeps >= 0, TRN_ERR_ARG,
                "bad ncolors or sweeps");
    TRN_REQUIRE(sweeps <= 1024, TRN_ERR_ARG, "at most 1024 sweeps per call");
    for (int c = 0; c < ncolors; ++c)
      TRN_REQUIRE(class_off[c] >= 0 && class_off[c] <= class_off[c + 1] &&
                      class_off[c + 1] <= mesh->nv,
                  TRN_ERR_ARG, "class offsets out of range");
    int32_t max_class = 1;
    for (int c = 0; c < ncolors; ++c)
      max_class = std::max(max_class, class_off[c + 1] - class_off[c]);
    const size_t bytes = smooth_sweep_scratch(mesh, ncolors, sweeps, max_class);
    Scratch sc(ctx, bytes);
    void* base = sc.carve<char>(bytes);
    const bool want = commits || tasks;
    const int32_t* totals = smooth_sweeps_launch(ctx, mesh, xy_d, tensor_d, qualit_d, order_d,
                                                 class_off, ncolors, sweeps, max_class, base,
                                                 bytes, want);
    if (totals) {
      copy_async(ctx, ctx->pinned_i32, totals, sizeof(int32_t) * 2 * (size_t)sweeps);
      sync(ctx);
      for (int s = 0; s < sweeps; ++s) {
        if (commits) commits[s] = ctx->pinned_i32[2 * s];
        if (tasks) tasks[s] = ctx->pinned_i32[2 * s + 1];
      }
    }
  });
}

int trn_smooth_run_prepared_noalloc(trn_ctx* ctx, const trn_mesh* mesh, double* xy_d,
                                    double* tensor_d, double* qualit_d, const int32_t* order_d,
                                    const int32_t* class_off, int32_t ncolors, int32_t sweeps,
                                    int32_t* commits, int32_t* tasks) {
  return guard_impl("trn_smooth_run_prepared_noalloc", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(mesh && xy_d && tensor_d && qualit_d && order_d && class_off, TRN_ERR_ARG,
                "null argument");
    TRN_REQUIRE(ncolors >= 0 && ncolors <= 64 && sweeps >= 0, TRN_ERR_ARG,
                "bad ncolors or sweeps");
    TRN_REQUIRE(sweeps <= 1024, TRN_ERR_ARG, "at most 1024 sweeps per call");
    TRN_REQUIRE(!ctx->timing, TRN_ERR_STATE,
                "per-kernel timing creates events: switch it off for the allocation-free call");
    for (int c = 0; c < ncolors; ++c)
      TRN_REQUIRE(class_off[c] >= 0 && class_off[c] <= class_off[c + 1] &&
                      class_off[c + 1] <= mesh->nv,
                  TRN_ERR_ARG, "class offsets out of range");
    int32_t max_class = 1;
    for (int c = 0; c < ncolors; ++c)
      max_class = std::max(max_class, class_off[c + 1] - class_off[c]);
    const size_t bytes = smooth_sweep_scratch(mesh, ncolors, sweeps, max_class);
    TRN_REQUIRE(bytes <= kNoAllocScratchBytes, TRN_ERR_CAPACITY,
                "the allocation-free sweeps need " + std::to_string(bytes) +
                    " bytes of device scratch; its fixed array holds " +
                    std::to_string(kNoAllocScratchBytes) + " (mesh too large for this call)");
    void* base = noalloc_scratch(ctx->device);
    const bool want = commits || tasks;
    const int32_t* totals = smooth_sweeps_launch(ctx, mesh, xy_d, tensor_d, qualit_d, order_d,
                                                 class_off, ncolors, sweeps, max_class, base,
                                                 kNoAllocScratchBytes, want);
    if (totals) {
      copy_async(ctx, ctx->pinned_i32, totals, sizeof(int32_t) * 2 * (size_t)sweeps);
      sync(ctx);
      for (int s = 0; s < sweeps; ++s) {
        if (commits) commits[s] = ctx->pinned_i32[2 * s];
        if (tasks) tasks[s] = ctx->pinned_i32[2 * s + 1];
      }
    }
  });
}

int trn_smooth_run(trn_ctx* ctx, const trn_mesh* mesh, double* xy_d, double* tensor_d,
                   double* qualit_d, const int32_t* order_d, const int32_t* class_off,
                   int32_t ncolors, int32_t sweeps, int32_t recompute_qualit, int32_t* commits,
                   int32_t* tasks) {
  const int rc = trn_smooth_prepare(ctx, mesh, xy_d, tensor_d, qualit_d, recompute_qualit);
  if (rc != TRN_OK) return rc;
  return trn_smooth_run_prepared(ctx, mesh, xy_d, tensor_d, qualit_d, order_d, class_off,
                                 ncolors, sweeps, commits, tasks);
}

}  // extern "C"
