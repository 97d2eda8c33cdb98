// batch.hip -- many independent meshes of uneven sizes per launch.
//
// The batch is the disjoint union of its meshes: one device topology whose
// vertex and triangle indices are offset per mesh.  Every stage of the path
// is local to a mesh's vertices and triangles, except the complexity sum and
// the target scaling (src/metric.cpp:39, :153-238), which are per mesh.
//
// A pass (Metrics::run + the Smooth::run sweeps, for every mesh) is cut into
// work items (stage, mesh, first, count):
//   0 triangle gradient terms   1 vertex gradients      2 triangle Hessian
//   3 vertex Hessian + local normalisation             4 complexity partials
//   5 global normalisation + tensor logs               6 element qualities
//   7.. smoothing phase (sweep s, colour c) of the mesh's class segment
// and run as one launch per stage over ALL meshes: each launch is a
// persistent kernel whose workgroups take items from an atomic work queue,
// so a 2,000-vertex mesh and a 200,000-vertex one share every launch and the
// chip stays full.  Stage order comes from the launch boundaries; the
// sequence of launches (about 7 + sweeps x colours) is captured once in a
// hipGraph and replayed.
//
// A single-launch schedule that pipelines stages across meshes (an item of
// stage k waits in the kernel for its mesh's previous stage) is kept below.
// Its gain would be the launch-boundary drains, a small share of a pass.
//
// SWITCHED OFF: on MI355X both schedules stalled on an 11-mesh batch (no
// completion within the test limit) while the same sources complete, and
// match the plain pass, in a CPU emulation of the HIP runtime.  Until the
// GPU cause is shown, the pass entry points return TRN_ERR_STATE without
// launching; creation, colouring and configuration work.
//
// The element bodies are those of the per-stage kernels (trn_recovery.hpp,
// trn_smooth.hpp), so every mesh's result is the same as a plain pass over
// it, except the complexity, whose per-mesh sum runs in a different fixed
// order (both are deterministic).  No result depends on the launch shape
// (block, tiles, blocks per CU), so the autotuner changes only the time.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "trn_recovery.hpp"
#include "trn_smooth.hpp"

namespace trn {

enum Stage : int32_t {
  kTriGrad = 0,
  kVtxGrad = 1,
  kTriHess = 2,
  kVtxHess = 3,
  kCplx = 4,
  kGlobal = 5,
  kQuality = 6,
  kPhase0 = 7,
};

struct Item {
  int32_t stage, mesh, first, count;
};

constexpr int kSlots = 64;               // per-sweep counter slots
constexpr int kCplxTile = 1024;          // triangles per complexity item
constexpr int kCplxChunks = kCplxTile / 64;
constexpr double kWaitSeconds = 4.0;     // bound of one stage wait (pipelined)
constexpr bool kPipelinedEnabled = false;
// Both schedules stalled on MI355X on an 11-mesh batch (the work never
// completed) while the same sources complete, and match the plain pass, in a
// CPU emulation of the HIP runtime; the GPU cause has not been shown, so the
// pass entry points refuse instead of launching.
constexpr bool kBatchPassEnabled = false;
constexpr const char* kBatchOff =
    "the batched pass is switched off in this build: it stalled on MI355X and the cause "
    "has not been shown";

struct BatchArgs {
  // topology of the union
  const int32_t* tri;
  const uint8_t* tags;
  const int32_t* v2t_off;
  const int32_t* v2t;
  const int32_t* v2v_off;
  const int32_t* v2v;
  const int32_t* star;
  int32_t patch_cap;
  // per-mesh data
  const int32_t* voff;        // [nmesh+1]
  const int32_t* cplx_first;  // [nmesh] first complexity item of the mesh
  const int32_t* cplx_count;  // [nmesh]
  double* cplx_out;           // [nmesh] complexity (written by GLOBAL items)
  // fields
  const double2* xy_in;  // nullptr: xy already holds the input
  double2* xy;
  const double* f;
  double* tensor;
  double* qualit;
  double4* logm;
  double* area_t;
  double2* ag_t;
  double4* ah_t;
  double* tarea;
  double2* grad;
  double* partial;  // [nitems]
  // colour classes
  const int32_t* order;
  int32_t ncolors;
  // items / control
  const Item* items;
  int32_t nitems;
  int32_t nstage;
  int32_t* err;       // bit 0: patch overflow, bit 1: a wait gave up
  int32_t* queues;    // [nstage] staged queues, [nstage] the pipelined one
  int32_t* counters;  // [sweeps][kSlots][2]
  // pipelined schedule only
  const int32_t* ntiles;   // [nmesh * nstage]
  const int32_t* wait_on;  // [nmesh * nstage] done index waited on, -1 none
  int32_t* done;           // [nmesh * nstage]
  int64_t wait_ticks;      // kWaitSeconds in wall_clock64 ticks
  // parameters
  double tf, ex, emin, emax;
};

// ---- item bodies -----------------------------------------------------------
// Every thread of the workgroup calls them with the same item; bodies with a
// barrier say so.

__device__ __forceinline__ const double2* metric_xy(const BatchArgs& a) {
  // the metric stages read the input coordinates; the sweeps move a copy
  // made by the vertex-gradient stage
  return a.xy_in ? a.xy_in : (const double2*)a.xy;
}

__device__ __forceinline__ void item_tri_grad(const BatchArgs& a, const Item& I) {
  const double2* xy = metric_xy(a);
  for (int k = threadIdx.x; k < I.count; k += blockDim.x)
    tri_grad_one((int64_t)I.first + k, a.tri, xy, a.f, a.area_t, a.ag_t);
}

__device__ __forceinline__ void item_vtx_grad(const BatchArgs& a, const Item& I, int* lds) {
  int* s = lds + threadIdx.x;
  for (int k = threadIdx.x; k < I.count; k += blockDim.x) {
    const int i = I.first + k;
    bool ok;
    const int m = build_patch(i, a.v2t_off, a.v2t, a.v2v_off, a.v2v, s, blockDim.x, a.patch_cap, ok);
    if (!ok) {
      atomicOr(a.err, 1);
      continue;
    }
    vtx_grad_one<true>(i, s, blockDim.x, m, a.area_t, a.ag_t, a.grad, a.tarea);
    if (a.xy_in) a.xy[i] = a.xy_in[i];  // working copy for the sweeps
  }
}

__device__ __forceinline__ void item_tri_hess(const BatchArgs& a, const Item& I) {
  const double2* xy = metric_xy(a);
  for (int k = threadIdx.x; k < I.count; k += blockDim.x)
    tri_hess_one((int64_t)I.first + k, a.tri, xy, a.grad, a.ah_t);
}

__device__ __forceinline__ void item_vtx_hess(const BatchArgs& a, const Item& I, int* lds) {
  int* s = lds + threadIdx.x;
  for (int k = threadIdx.x; k < I.count; k += blockDim.x) {
    const int i = I.first + k;
    bool ok;
    const int m = build_patch(i, a.v2t_off, a.v2t, a.v2v_off, a.v2v, s, blockDim.x, a.patch_cap, ok);
    if (!ok) {
      atomicOr(a.err, 1);
      continue;
    }
    vtx_hess_one<true>(i, s, blockDim.x, m, a.ah_t, a.tarea, nullptr, a.tensor, a.ex);
  }
}

// Complexity partial of one item, summed in a fixed tree whatever the block
// size: 64-triangle wave sums, then the item's chunks in order.  Barrier.
__device__ __forceinline__ void item_cplx(const BatchArgs& a, const Item& I, int it,
                                          double* s_chunk) {
  const double2* xy = metric_xy(a);
  const int lane = threadIdx.x & 63;
  for (int c = threadIdx.x >> 6; c < kCplxChunks; c += blockDim.x >> 6) {
    const int k = c * 64 + lane;
    double phi = k < I.count ? cplx_one((int64_t)I.first + k, a.tri, xy, a.tensor) : 0.;
    phi = wave_sum(phi);
    if (lane == 0) s_chunk[c] = phi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.;
    for (int c = 0; c < kCplxChunks; ++c) b += s_chunk[c];
    a.partial[it] = b;
  }
}

// Global normalisation of one item's vertices with its mesh's complexity
// (the mesh's partials in item order), then the tensor logs the quality and
// smoothing stages read.  Barrier.
__device__ __forceinline__ void item_global(const BatchArgs& a, const Item& I, double* s_fact) {
  if (threadIdx.x == 0) {
    const int f0 = a.cplx_first[I.mesh], nc = a.cplx_count[I.mesh];
    double c = 0.;
    for (int q = 0; q < nc; ++q) c += a.partial[f0 + q];
    const int v0 = a.voff[I.mesh], nvm = a.voff[I.mesh + 1] - v0;
    const double target = (double)(int)floor(nvm * a.tf);  // metric.cpp:39
    *s_fact = target / c;                                    // metric.cpp:197-198
    if (a.cplx_out && I.first == v0) a.cplx_out[I.mesh] = c;
  }
  __syncthreads();
  const double fact = *s_fact;
  for (int k = threadIdx.x; k < I.count; k += blockDim.x) {
    const int64_t i = (int64_t)I.first + k;
    double* t = a.tensor + 3 * i;
    global_one(t, fact, a.emin, a.emax);
    a.logm[i] = tensor_log4(t[0], t[1], t[2]);
  }
}

__device__ __forceinline__ void item_quality(const BatchArgs& a, const Item& I) {
  for (int k = threadIdx.x; k < I.count; k += blockDim.x) {
    const int64_t t = (int64_t)I.first + k;
    if (a.tri[3 * t] >= 0) a.qualit[t] = quality_logs(t, a.tri, (const double2*)a.xy, a.logm);
  }
}

// Smoothing phase (sweep, colour) of one mesh's class segment, then one
// counter update per wave on one of kSlots addresses.
__device__ __forceinline__ void item_phase(const BatchArgs& a, const Item& I, int it, int sweep,
                                           double* lds) {
  int commit = 0, task = 0;
  for (int k = threadIdx.x; k < I.count; k += blockDim.x) {
    const int32_t i = a.order[I.first + k];
    if (!(a.tags[i] & 4) && a.v2t_off[i + 1] > a.v2t_off[i]) {  // smoothing.cpp:137
      task += 1;
      commit += move_vertex(i, a.xy, a.tensor, a.logm, a.qualit, a.v2t_off, a.v2t, a.star,
                            a.v2v_off, a.v2v, lds + threadIdx.x, blockDim.x);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    commit += __shfl_xor(commit, o, 64);
    task += __shfl_xor(task, o, 64);
  }
  if ((threadIdx.x & 63) == 0 && task) {
    int32_t* slot =
        a.counters + 2 * ((int64_t)sweep * kSlots + ((it + (threadIdx.x >> 6)) & (kSlots - 1)));
    atomicAdd(&slot[0], commit);
    atomicAdd(&slot[1], task);
  }
}

template <int ST>
__device__ __forceinline__ void run_item(const BatchArgs& a, const Item& I, int it, int sweep,
                                         double* lds, double* s_chunk, double* s_fact) {
  if constexpr (ST == kTriGrad) item_tri_grad(a, I);
  else if constexpr (ST == kVtxGrad) item_vtx_grad(a, I, (int*)lds);
  else if constexpr (ST == kTriHess) item_tri_hess(a, I);
  else if constexpr (ST == kVtxHess) item_vtx_hess(a, I, (int*)lds);
  else if constexpr (ST == kCplx) item_cplx(a, I, it, s_chunk);
  else if constexpr (ST == kGlobal) item_global(a, I, s_fact);
  else if constexpr (ST == kQuality) item_quality(a, I);
  else item_phase(a, I, it, sweep, lds);
}

// ---- staged schedule: one persistent work-queue launch per stage ----------

__global__ __launch_bounds__(256) void k_batch_reset(int32_t* __restrict__ ctrl, int32_t nctrl,
                                                     int32_t* __restrict__ counters,
                                                     int32_t ncount, int32_t* __restrict__ done,
                                                     int32_t ndone) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nctrl) ctrl[i] = 0;
  if (i < ncount) counters[i] = 0;
  if (i < ndone) done[i] = 0;
}

template <int ST>
__global__ __launch_bounds__(256) void k_batch_stage(BatchArgs a, int32_t item0, int32_t nitems,
                                                     int32_t qslot, int32_t sweep) {
  extern __shared__ double lds_raw[];
  __shared__ int s_item;
  __shared__ double s_chunk[kCplxChunks];
  __shared__ double s_fact;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(&a.queues[qslot], 1);
    __syncthreads();
    const int q = s_item;
    __syncthreads();
    if (q >= nitems) break;
    const int it = item0 + q;
    run_item<ST>(a, a.items[it], it, sweep, lds_raw, s_chunk, &s_fact);
  }
}

// ---- pipelined schedule (switched off) -------------------------------------
// One launch for the whole pass: items in stage-major order, each waiting
// for the latest earlier stage of its mesh that has items.  Items are taken
// in list order and wait only on earlier items, held by running workgroups.

__device__ __forceinline__ void wait_stage(const BatchArgs& a, int m, int k, int* s_abort) {
  const int idx = a.wait_on[m * a.nstage + k];
  if (idx < 0) return;
  const int need = a.ntiles[idx];
  const int64_t t0 = wall_clock64();
  int polls = 0;
  while (__hip_atomic_load(&a.done[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
    if (polls < 32) __builtin_amdgcn_s_sleep(8);
    else __builtin_amdgcn_s_sleep(64);
    ++polls;
    if ((polls & 15) == 0 &&
        (wall_clock64() - t0 > a.wait_ticks ||
         (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 2))) {
      atomicOr(a.err, 2);
      *s_abort = 1;
      return;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

__global__ __launch_bounds__(256) void k_batch_pipelined(BatchArgs a) {
  extern __shared__ double lds_raw[];
  __shared__ int s_item, s_abort;
  __shared__ double s_chunk[kCplxChunks];
  __shared__ double s_fact;
  const int tid = threadIdx.x;
  if (tid == 0) s_abort = 0;
  int32_t* queue = &a.queues[a.nstage];
  for (;;) {
    if (tid == 0)
      s_item = (s_abort || (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 2))
                   ? a.nitems
                   : atomicAdd(queue, 1);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= a.nitems) break;
    const Item I = a.items[it];
    if (tid == 0) wait_stage(a, I.mesh, I.stage, &s_abort);
    __syncthreads();
    if (s_abort) break;
    switch (I.stage) {
      case kTriGrad: run_item<kTriGrad>(a, I, it, 0, lds_raw, s_chunk, &s_fact); break;
      case kVtxGrad: run_item<kVtxGrad>(a, I, it, 0, lds_raw, s_chunk, &s_fact); break;
      case kTriHess: run_item<kTriHess>(a, I, it, 0, lds_raw, s_chunk, &s_fact); break;
      case kVtxHess: run_item<kVtxHess>(a, I, it, 0, lds_raw, s_chunk, &s_fact); break;
      case kCplx: run_item<kCplx>(a, I, it, 0, lds_raw, s_chunk, &s_fact); break;
      case kGlobal: run_item<kGlobal>(a, I, it, 0, lds_raw, s_chunk, &s_fact); break;
      case kQuality: run_item<kQuality>(a, I, it, 0, lds_raw, s_chunk, &s_fact); break;
      default:
        run_item<kPhase0>(a, I, it, (I.stage - kPhase0) / a.ncolors, lds_raw, s_chunk, &s_fact);
    }
    // the barrier orders every wave's writes before the release of the
    // mesh's stage count
    __syncthreads();
    if (tid == 0)
      __hip_atomic_fetch_add(&a.done[I.mesh * a.nstage + I.stage], 1, __ATOMIC_RELEASE,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
}

// start of colour c's segment of mesh m in the class list (ascending vertex
// ids, so meshes are consecutive): lower bound of voff[m]
__global__ void k_batch_segments(const int32_t* __restrict__ order,
                                 const int32_t* __restrict__ class_off_d, int ncolors,
                                 const int32_t* __restrict__ voff, int nmesh,
                                 int32_t* __restrict__ seg) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)ncolors * (nmesh + 1)) return;
  const int c = (int)(q / (nmesh + 1)), m = (int)(q % (nmesh + 1));
  int lo = class_off_d[c], hi = class_off_d[c + 1];
  const int32_t key = voff[m];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (order[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  seg[q] = lo;
}

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace trn

using namespace trn;

struct trn_batch {
  trn_ctx* ctx = nullptr;
  int32_t nmesh = 0, nv = 0, nt = 0;
  trn_mesh* mesh = nullptr;
  std::vector<int32_t> voff, toff;
  int32_t* voff_d = nullptr;
  // per-pass buffers
  double* area_t = nullptr;
  double2* ag_t = nullptr;
  double4* ah_t = nullptr;
  double* tarea = nullptr;
  double2* grad = nullptr;
  double4* logm = nullptr;
  double* cplx = nullptr;
  // colouring
  int32_t* color = nullptr;
  int32_t* order = nullptr;
  int32_t ncolors = 0;
  std::vector<int32_t> class_off;
  std::vector<int32_t> seg;  // [ncolors][nmesh+1]
  // items (rebuilt when the configuration changes)
  int32_t block = 64, tile_v = 256, tile_t = 1024, blocks_per_cu = 4;
  int32_t items_sweeps = -1, items_block = -1, items_tv = -1, items_tt = -1;
  std::vector<Item> items_h;
  std::vector<int32_t> stage_off;  // [nstage+1] first item of each stage
  std::vector<int32_t> ntiles_h, wait_h, cplx_first_h, cplx_count_h;
  Item* items = nullptr;
  double* partial = nullptr;
  int32_t* ctrl = nullptr;  // [0] err, [1 .. 2*nstage] queues
  int32_t* counters = nullptr;
  int32_t* ntiles = nullptr;
  int32_t* wait_on = nullptr;
  int32_t* done = nullptr;
  int32_t* cplx_first = nullptr;
  int32_t* cplx_count = nullptr;
  size_t cap_items = 0, cap_stage = 0, cap_counters = 0;
  int32_t nstage = 0;
  int32_t num_cus = 0;
  int64_t wait_ticks = 0;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> graph_key;
};

namespace trn {

template <typename T>
static void dalloc(T** p, size_t n) {
  TRN_HIP(hipMalloc((void**)p, sizeof(T) * (n > 0 ? n : 1)));
}

template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

static void drop_graph(trn_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_key.clear();
}

static void build_items(trn_batch* b, int sweeps) {
  if (b->items_sweeps == sweeps && b->items_block == b->block && b->items_tv == b->tile_v &&
      b->items_tt == b->tile_t)
    return;
  TRN_REQUIRE(sweeps == 0 || b->ncolors > 0, TRN_ERR_STATE, "batch not coloured");
  const int nm = b->nmesh;
  const int ns = kPhase0 + sweeps * b->ncolors;
  b->nstage = ns;
  b->items_h.clear();
  b->stage_off.assign(ns + 1, 0);
  b->ntiles_h.assign((size_t)nm * ns, 0);
  b->cplx_first_h.assign(nm, 0);
  b->cplx_count_h.assign(nm, 0);
  auto add = [&](int stage, int m, int64_t first, int64_t count, int tile) {
    for (int64_t s = 0; s < count; s += tile) {
      b->items_h.push_back(Item{stage, m, (int32_t)(first + s),
                                (int32_t)std::min<int64_t>(tile, count - s)});
      b->ntiles_h[(size_t)m * ns + stage]++;
    }
  };
  for (int st = 0; st < ns; ++st) {
    b->stage_off[st] = (int32_t)b->items_h.size();
    for (int m = 0; m < nm; ++m) {
      const int64_t v0 = b->voff[m], nvm = b->voff[m + 1] - b->voff[m];
      const int64_t t0 = b->toff[m], ntm = b->toff[m + 1] - b->toff[m];
      if (st == kCplx) {
        b->cplx_first_h[m] = (int32_t)b->items_h.size();
        add(st, m, t0, ntm, kCplxTile);  // fixed: the sums do not depend on the shape
        b->cplx_count_h[m] = (int32_t)b->items_h.size() - b->cplx_first_h[m];
      } else if (st == kTriGrad || st == kTriHess || st == kQuality) {
        add(st, m, t0, ntm, b->tile_t);
      } else if (st == kVtxGrad || st == kVtxHess || st == kGlobal) {
        add(st, m, v0, nvm, b->tile_v);
      } else {
        const int c = (st - kPhase0) % b->ncolors;
        const int32_t s0 = b->seg[(size_t)c * (nm + 1) + m];
        const int32_t s1 = b->seg[(size_t)c * (nm + 1) + m + 1];
        add(st, m, s0, s1 - s0, b->tile_v);
      }
    }
  }
  b->stage_off[ns] = (int32_t)b->items_h.size();
  // pipelined schedule: each stage waits on the latest earlier stage of its
  // mesh that has items
  b->wait_h.assign((size_t)nm * ns, -1);
  for (int m = 0; m < nm; ++m) {
    int last = -1;
    for (int st = 0; st < ns; ++st) {
      b->wait_h[(size_t)m * ns + st] = last;
      if (b->ntiles_h[(size_t)m * ns + st] > 0) last = m * ns + st;
    }
  }
  const size_t ni = b->items_h.size();
  if (ni > b->cap_items) {
    dfree(b->items);
    dfree(b->partial);
    dalloc(&b->items, ni);
    dalloc(&b->partial, ni);
    b->cap_items = ni;
  }
  const size_t nd = (size_t)nm * ns;
  if (nd + 2 * (size_t)ns + 1 > b->cap_stage) {
    dfree(b->done);
    dfree(b->ntiles);
    dfree(b->wait_on);
    dfree(b->ctrl);
    dalloc(&b->done, nd);
    dalloc(&b->ntiles, nd);
    dalloc(&b->wait_on, nd);
    dalloc(&b->ctrl, 2 * (size_t)ns + 1);
    b->cap_stage = nd + 2 * (size_t)ns + 1;
  }
  const size_t nc = 2 * (size_t)kSlots * (sweeps + 1);
  if (nc > b->cap_counters) {
    dfree(b->counters);
    dalloc(&b->counters, nc);
    b->cap_counters = nc;
  }
  if (!b->cplx_first) {
    dalloc(&b->cplx_first, nm);
    dalloc(&b->cplx_count, nm);
  }
  trn_ctx* ctx = b->ctx;
  copy_async(ctx, b->items, b->items_h.data(), sizeof(Item) * ni);
  copy_async(ctx, b->ntiles, b->ntiles_h.data(), sizeof(int32_t) * nd);
  copy_async(ctx, b->wait_on, b->wait_h.data(), sizeof(int32_t) * nd);
  copy_async(ctx, b->cplx_first, b->cplx_first_h.data(), sizeof(int32_t) * nm);
  copy_async(ctx, b->cplx_count, b->cplx_count_h.data(), sizeof(int32_t) * nm);
  sync(ctx);  // host vectors may change before the copies run otherwise
  b->items_sweeps = sweeps;
  b->items_block = b->block;
  b->items_tv = b->tile_v;
  b->items_tt = b->tile_t;
  drop_graph(b);
}

// LDS of the patch-gather stages and of the smoothing phases.
static size_t patch_shmem(const trn_mesh* m, int block) {
  return sizeof(int) * (size_t)block * m->patch_cap;
}
static size_t phase_shmem(const trn_mesh* m, int block) {
  return sizeof(double) * 4 * (size_t)block * std::max(m->max_deg_t, 1);
}

static BatchArgs make_args(trn_batch* b, const double* xy_in, double* xy, const double* f,
                           double tf, int p, double hmin, double hmax, double* tensor,
                           double* qualit) {
  BatchArgs a;
  a.tri = b->mesh->tri;
  a.tags = b->mesh->tags;
  a.v2t_off = b->mesh->v2t_off;
  a.v2t = b->mesh->v2t;
  a.v2v_off = b->mesh->v2v_off;
  a.v2v = b->mesh->v2v;
  a.star = b->mesh->star;
  a.patch_cap = b->mesh->patch_cap;
  a.voff = b->voff_d;
  a.cplx_first = b->cplx_first;
  a.cplx_count = b->cplx_count;
  a.cplx_out = b->cplx;
  a.xy_in = (const double2*)xy_in;
  a.xy = (double2*)xy;
  a.f = f;
  a.tensor = tensor;
  a.qualit = qualit;
  a.logm = b->logm;
  a.area_t = b->area_t;
  a.ag_t = b->ag_t;
  a.ah_t = b->ah_t;
  a.tarea = b->tarea;
  a.grad = b->grad;
  a.partial = b->partial;
  a.order = b->order;
  a.ncolors = b->ncolors;
  a.items = b->items;
  a.nitems = (int32_t)b->items_h.size();
  a.nstage = b->nstage;
  a.err = b->ctrl;
  a.queues = b->ctrl + 1;
  a.counters = b->counters;
  a.ntiles = b->ntiles;
  a.wait_on = b->wait_on;
  a.done = b->done;
  a.wait_ticks = b->wait_ticks;
  a.tf = tf;
  a.ex = metric_exponent(p);
  a.emin = 1 / (hmin * hmin);
  a.emax = 1 / (hmax * hmax);
  return a;
}

// Algorithmic bytes of the gather stages, counted as the plain kernels count
// them: gradient 120 and Hessian 136 per vertex, complexity 64 and global
// normalisation 48 per vertex; a smoothing phase 196 per class vertex.
static double stage_bytes(const trn_batch* b, int st) {
  switch (st) {
    case kVtxGrad: return 120.0 * b->nv;
    case kVtxHess: return 136.0 * b->nv;
    case kCplx: return 64.0 * b->nv;
    case kGlobal: return 48.0 * b->nv;
    default: return 0.0;
  }
}

// Queues one staged pass on ctx->stream: a reset, then one launch per stage
// (graph-capturable: no memset or host copy inside).
static void enqueue_pass(trn_ctx* ctx, trn_batch* b, const BatchArgs& a, int sweeps) {
  const int32_t nctrl = 1 + 2 * b->nstage;
  const int32_t ncount = 2 * kSlots * (sweeps + 1);
  launch(ctx, "batch_reset", 0.0, k_batch_reset, dim3(nblk(std::max(nctrl, ncount), 256)),
         dim3(256), 0, b->ctrl, nctrl, b->counters, ncount, (int32_t*)nullptr, 0);
  const int cap = b->num_cus * b->blocks_per_cu;
  const dim3 blk(b->block);
  const size_t lds_patch = patch_shmem(b->mesh, b->block);
  const size_t lds_phase = phase_shmem(b->mesh, b->block);
  for (int st = 0; st < b->nstage; ++st) {
    const int32_t i0 = b->stage_off[st], n = b->stage_off[st + 1] - i0;
    if (n == 0) continue;
    const dim3 g(std::min(n, cap));
    const double by = stage_bytes(b, st);
    switch (st) {
      case kTriGrad:
        launch(ctx, "batch_tri_grad", by, k_batch_stage<kTriGrad>, g, blk, 0, a, i0, n, st, 0);
        break;
      case kVtxGrad:
        launch(ctx, "batch_vtx_grad", by, k_batch_stage<kVtxGrad>, g, blk, lds_patch, a, i0, n,
               st, 0);
        break;
      case kTriHess:
        launch(ctx, "batch_tri_hess", by, k_batch_stage<kTriHess>, g, blk, 0, a, i0, n, st, 0);
        break;
      case kVtxHess:
        launch(ctx, "batch_vtx_hess", by, k_batch_stage<kVtxHess>, g, blk, lds_patch, a, i0, n,
               st, 0);
        break;
      case kCplx:
        launch(ctx, "batch_cplx", by, k_batch_stage<kCplx>, g, blk, 0, a, i0, n, st, 0);
        break;
      case kGlobal:
        launch(ctx, "batch_global", by, k_batch_stage<kGlobal>, g, blk, 0, a, i0, n, st, 0);
        break;
      case kQuality:
        launch(ctx, "batch_quality", by, k_batch_stage<kQuality>, g, blk, 0, a, i0, n, st, 0);
        break;
      default: {
        const int sweep = (st - kPhase0) / b->ncolors;
        const int c = (st - kPhase0) % b->ncolors;
        const double cnt = b->class_off[c + 1] - b->class_off[c];
        launch(ctx, "batch_phase", 196.0 * cnt, k_batch_stage<kPhase0>, g, blk, lds_phase, a, i0,
               n, st, sweep);
      }
    }
  }
}

static void check_pass(trn_ctx* ctx, trn_batch* b) {
  copy_async(ctx, ctx->pinned_i32, b->ctrl, sizeof(int32_t));
  sync(ctx);
  const int e = ctx->pinned_i32[0];
  TRN_REQUIRE(!(e & 1), TRN_ERR_STATE, "2-ring patch larger than measured");
  TRN_REQUIRE(!(e & 2), TRN_ERR_STATE, "batch pass: a stage wait exceeded its bound");
}

static void check_params(const void* xy_d, const void* solut_d, const void* tensor_d,
                         const void* qualit_d, double h_min, double h_max, int sweeps) {
  TRN_REQUIRE(xy_d && solut_d && tensor_d && qualit_d, TRN_ERR_ARG, "null argument");
  TRN_REQUIRE(h_min > 0 && h_max > 0 && sweeps >= 0 && sweeps <= 64, TRN_ERR_ARG,
              "bad parameters");
}

}  // namespace trn

extern "C" {

int trn_batch_create(trn_ctx* ctx, int32_t nmesh, const int32_t* nv, const int32_t* nt,
                     const int32_t* tri, const uint8_t* tags, trn_batch** out) {
  return guard_impl("trn_batch_create", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(out && nv && nt && tri && tags && nmesh > 0, TRN_ERR_ARG, "bad arguments");
    *out = nullptr;
    trn_batch* b = new trn_batch();
    b->ctx = ctx;
    b->nmesh = nmesh;
    try {
      b->voff.assign(nmesh + 1, 0);
      b->toff.assign(nmesh + 1, 0);
      for (int m = 0; m < nmesh; ++m) {
        TRN_REQUIRE(nv[m] > 0 && nt[m] > 0, TRN_ERR_ARG, "empty mesh in batch");
        TRN_REQUIRE((int64_t)b->voff[m] + nv[m] < INT32_MAX, TRN_ERR_ARG, "batch too large");
        b->voff[m + 1] = b->voff[m] + nv[m];
        TRN_REQUIRE((int64_t)b->toff[m] + nt[m] < INT32_MAX / 3, TRN_ERR_ARG, "batch too large");
        b->toff[m + 1] = b->toff[m] + nt[m];
      }
      b->nv = b->voff[nmesh];
      b->nt = b->toff[nmesh];
      // union triangle list with global indices; local indices are checked
      // against their own mesh (an index into another mesh would join them)
      std::vector<int32_t> gtri(3 * (size_t)b->nt);
      for (int m = 0; m < nmesh; ++m) {
        for (int64_t t = b->toff[m]; t < b->toff[m + 1]; ++t) {
          const int32_t* s = tri + 3 * t;
          if (s[0] < 0) {
            gtri[3 * t] = gtri[3 * t + 1] = gtri[3 * t + 2] = -1;
            continue;
          }
          for (int j = 0; j < 3; ++j) {
            TRN_REQUIRE(s[j] >= 0 && s[j] < nv[m], TRN_ERR_ARG,
                        "triangle index outside its mesh (mesh " + std::to_string(m) + ")");
            gtri[3 * t + j] = s[j] + b->voff[m];
          }
        }
      }
      int rc = trn_mesh_create(ctx, b->nv, b->nt, gtri.data(), tags, nullptr, &b->mesh);
      if (rc != TRN_OK) fail(rc, trn_last_error());
      dalloc(&b->voff_d, nmesh + 1);
      copy_async(ctx, b->voff_d, b->voff.data(), sizeof(int32_t) * (nmesh + 1));
      dalloc(&b->area_t, b->nt);
      dalloc(&b->ag_t, b->nt);
      dalloc(&b->ah_t, b->nt);
      dalloc(&b->tarea, b->nv);
      dalloc(&b->grad, b->nv);
      dalloc(&b->logm, b->nv);
      dalloc(&b->cplx, nmesh);
      dalloc(&b->color, b->nv);
      dalloc(&b->order, b->nv);
      hipDeviceProp_t prop;
      TRN_HIP(hipGetDeviceProperties(&prop, ctx->device));
      b->num_cus = prop.multiProcessorCount;
      int rate_khz = 0;
      TRN_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
      b->wait_ticks = (int64_t)(kWaitSeconds * 1e3 * (rate_khz > 0 ? rate_khz : 100000));
      sync(ctx);
    } catch (...) {
      trn_batch_destroy(b);
      throw;
    }
    *out = b;
  });
}

int trn_batch_destroy(trn_batch* b) {
  if (!b) return TRN_OK;
  return guard_impl("trn_batch_destroy", [&] {
    DeviceScope ds(b->ctx);
    (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    if (b->mesh) trn_mesh_destroy(b->mesh);
    dfree(b->voff_d); dfree(b->area_t); dfree(b->ag_t); dfree(b->ah_t); dfree(b->tarea);
    dfree(b->grad); dfree(b->logm); dfree(b->cplx); dfree(b->color); dfree(b->order);
    dfree(b->items); dfree(b->partial); dfree(b->ctrl); dfree(b->counters); dfree(b->ntiles);
    dfree(b->wait_on); dfree(b->done); dfree(b->cplx_first); dfree(b->cplx_count);
    delete b;
  });
}

int trn_batch_info(const trn_batch* b, int32_t* nv, int32_t* nt, const trn_mesh** mesh) {
  return guard_impl("trn_batch_info", [&] {
    TRN_REQUIRE(b != nullptr, TRN_ERR_ARG, "null batch");
    if (nv) *nv = b->nv;
    if (nt) *nt = b->nt;
    if (mesh) *mesh = b->mesh;
  });
}

int trn_batch_color(trn_ctx* ctx, trn_batch* b, uint64_t seed, int32_t max_colors,
                    int32_t* ncolors) {
  return guard_impl("trn_batch_color", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(b && ncolors, TRN_ERR_ARG, "null argument");
    std::vector<int32_t> off(66, 0);
    int32_t nc = 0;
    int rc = trn_color(ctx, b->mesh, seed, max_colors, b->color, b->order, off.data(), &nc);
    if (rc != TRN_OK) fail(rc, trn_last_error());
    b->ncolors = nc;
    b->class_off.assign(off.begin(), off.begin() + nc + 1);
    const int nm = b->nmesh;
    const size_t nseg = (size_t)nc * (nm + 1);
    Scratch sc(ctx, scratch_bytes(sizeof(int32_t) * (nc + 1)) +
                        scratch_bytes(sizeof(int32_t) * nseg));
    int32_t* off_d = sc.carve<int32_t>(nc + 1);
    int32_t* seg_d = sc.carve<int32_t>(nseg);
    copy_async(ctx, off_d, b->class_off.data(), sizeof(int32_t) * (nc + 1));
    launch(ctx, "batch_segments", 0.0, k_batch_segments, dim3(nblk((int64_t)nseg, 256)),
           dim3(256), 0, (const int32_t*)b->order, (const int32_t*)off_d, nc,
           (const int32_t*)b->voff_d, nm, seg_d);
    b->seg.assign(nseg, 0);
    copy_async(ctx, b->seg.data(), seg_d, sizeof(int32_t) * nseg);
    sync(ctx);
    b->items_sweeps = -1;  // classes changed
    *ncolors = nc;
  });
}

int trn_batch_set_config(trn_batch* b, int32_t block, int32_t tile_v, int32_t tile_t,
                         int32_t blocks_per_cu) {
  return guard_impl("trn_batch_set_config", [&] {
    TRN_REQUIRE(b != nullptr, TRN_ERR_ARG, "null batch");
    TRN_REQUIRE(block == 64 || block == 128 || block == 256, TRN_ERR_ARG,
                "block must be 64, 128 or 256");
    TRN_REQUIRE(tile_v >= 1 && tile_t >= 1 && blocks_per_cu >= 1 && blocks_per_cu <= 32,
                TRN_ERR_ARG, "bad tile or occupancy");
    TRN_REQUIRE(std::max(patch_shmem(b->mesh, block), phase_shmem(b->mesh, block)) <= 65536,
                TRN_ERR_CAPACITY, "stencil too large for this block size");
    b->block = block;
    b->tile_v = tile_v;
    b->tile_t = tile_t;
    b->blocks_per_cu = blocks_per_cu;
  });
}

int trn_batch_get_config(const trn_batch* b, int32_t* config) {
  return guard_impl("trn_batch_get_config", [&] {
    TRN_REQUIRE(b && config, TRN_ERR_ARG, "null argument");
    config[0] = b->block;
    config[1] = b->tile_v;
    config[2] = b->tile_t;
    config[3] = b->blocks_per_cu;
  });
}

int trn_batch_pass(trn_ctx* ctx, trn_batch* b, const double* xy_in_d, double* xy_d,
                   const double* solut_d, double target_factor, int32_t lp_norm, double h_min,
                   double h_max, double* tensor_d, double* qualit_d, int32_t sweeps,
                   double* complexity_d, int32_t* commits, int32_t* tasks) {
  return guard_impl("trn_batch_pass", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(b != nullptr, TRN_ERR_ARG, "null batch");
    check_params(xy_d, solut_d, tensor_d, qualit_d, h_min, h_max, sweeps);
    TRN_REQUIRE(kBatchPassEnabled, TRN_ERR_STATE, kBatchOff);
    build_items(b, sweeps);
    BatchArgs a = make_args(b, xy_in_d, xy_d, solut_d, target_factor, lp_norm, h_min, h_max,
                            tensor_d, qualit_d);
    enqueue_pass(ctx, b, a, sweeps);
    if (complexity_d) copy_async(ctx, complexity_d, b->cplx, sizeof(double) * b->nmesh);
    if (commits || tasks) {
      std::vector<int32_t> cnt(2 * (size_t)kSlots * sweeps);
      copy_async(ctx, cnt.data(), b->counters, sizeof(int32_t) * cnt.size());
      check_pass(ctx, b);
      for (int s = 0; s < sweeps; ++s) {
        int c = 0, t = 0;
        for (int k = 0; k < kSlots; ++k) {
          c += cnt[2 * ((size_t)s * kSlots + k)];
          t += cnt[2 * ((size_t)s * kSlots + k) + 1];
        }
        if (commits) commits[s] = c;
        if (tasks) tasks[s] = t;
      }
    }
  });
}

int trn_batch_pass_graph(trn_ctx* ctx, trn_batch* b, const double* xy_in_d, double* xy_d,
                         const double* solut_d, double target_factor, int32_t lp_norm,
                         double h_min, double h_max, double* tensor_d, double* qualit_d,
                         int32_t sweeps) {
  return guard_impl("trn_batch_pass_graph", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(b != nullptr && xy_in_d != nullptr, TRN_ERR_ARG, "null argument");
    check_params(xy_d, solut_d, tensor_d, qualit_d, h_min, h_max, sweeps);
    TRN_REQUIRE(kBatchPassEnabled, TRN_ERR_STATE, kBatchOff);
    TRN_REQUIRE(ctx->stream != nullptr, TRN_ERR_ARG,
                "graph capture needs a non-null stream (trn_ctx_use_own_stream)");
    build_items(b, sweeps);
    uint64_t key[12];
    key[0] = (uint64_t)xy_in_d; key[1] = (uint64_t)xy_d; key[2] = (uint64_t)solut_d;
    key[3] = (uint64_t)tensor_d; key[4] = (uint64_t)qualit_d; key[5] = (uint64_t)sweeps;
    key[6] = (uint64_t)(uint32_t)lp_norm;
    std::memcpy(&key[7], &target_factor, 8);
    std::memcpy(&key[8], &h_min, 8);
    std::memcpy(&key[9], &h_max, 8);
    key[10] = (uint64_t)ctx->stream;
    key[11] = (uint64_t)b->items_h.size();
    std::vector<uint64_t> k(key, key + 12);
    if (!b->exec || b->graph_key != k) {
      drop_graph(b);
      BatchArgs a = make_args(b, xy_in_d, xy_d, solut_d, target_factor, lp_norm, h_min, h_max,
                              tensor_d, qualit_d);
      const bool timing = ctx->timing;
      ctx->timing = false;  // no event records inside the capture
      TRN_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      try {
        enqueue_pass(ctx, b, a, sweeps);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        ctx->timing = timing;
        throw;
      }
      ctx->timing = timing;
      TRN_HIP(hipStreamEndCapture(ctx->stream, &b->graph));
      TRN_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_key = k;
    }
    if (ctx->timing) {
      const int st = stat_index(ctx, "batch_graph");
      hipEvent_t e0, e1;
      timing_begin(ctx, st, &e0, &e1);
      TRN_HIP(hipGraphLaunch(b->exec, ctx->stream));
      timing_end(ctx, st, e0, e1, (368.0 + 196.0 * sweeps) * b->nv);
    } else {
      TRN_HIP(hipGraphLaunch(b->exec, ctx->stream));
    }
  });
}

int trn_batch_check(trn_ctx* ctx, trn_batch* b) {
  return guard_impl("trn_batch_check", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(b != nullptr, TRN_ERR_ARG, "null batch");
    TRN_REQUIRE(b->ctrl != nullptr, TRN_ERR_STATE, "no pass has run");
    check_pass(ctx, b);
  });
}

int trn_batch_autotune(trn_ctx* ctx, trn_batch* b, const double* xy_in_d, double* xy_d,
                       const double* solut_d, double target_factor, int32_t lp_norm,
                       double h_min, double h_max, double* tensor_d, double* qualit_d,
                       int32_t sweeps, int32_t* best, double* best_ms) {
  return guard_impl("trn_batch_autotune", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(b != nullptr && xy_in_d != nullptr, TRN_ERR_ARG, "null argument");
    check_params(xy_d, solut_d, tensor_d, qualit_d, h_min, h_max, sweeps);
    TRN_REQUIRE(kBatchPassEnabled, TRN_ERR_STATE, kBatchOff);
    const int blocks[3] = {64, 128, 256};
    const int scales[3] = {1, 2, 4};
    const int occ[3] = {2, 4, 8};
    double best_t = 1e30;
    int cfg[4] = {b->block, b->tile_v, b->tile_t, b->blocks_per_cu};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TRN_HIP(hipEventCreate(&e0));
    TRN_HIP(hipEventCreate(&e1));
    const bool timing = ctx->timing;
    ctx->timing = false;
    try {
      for (int bi = 0; bi < 3; ++bi)
        for (int si = 0; si < 3; ++si)
          for (int oi = 0; oi < 3; ++oi) {
            const int blk = blocks[bi], tv = blk * scales[si], tt = 4 * tv;
            if (trn_batch_set_config(b, blk, tv, tt, occ[oi]) != TRN_OK) continue;
            build_items(b, sweeps);
            BatchArgs a = make_args(b, xy_in_d, xy_d, solut_d, target_factor, lp_norm, h_min,
                                    h_max, tensor_d, qualit_d);
            enqueue_pass(ctx, b, a, sweeps);  // warm-up
            TRN_HIP(hipEventRecord(e0, ctx->stream));
            enqueue_pass(ctx, b, a, sweeps);
            TRN_HIP(hipEventRecord(e1, ctx->stream));
            TRN_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            TRN_HIP(hipEventElapsedTime(&ms, e0, e1));
            check_pass(ctx, b);
            if (ms < best_t) {
              best_t = ms;
              cfg[0] = blk; cfg[1] = tv; cfg[2] = tt; cfg[3] = occ[oi];
            }
          }
    } catch (...) {
      ctx->timing = timing;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    ctx->timing = timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int rc = trn_batch_set_config(b, cfg[0], cfg[1], cfg[2], cfg[3]);
    if (rc != TRN_OK) fail(rc, trn_last_error());
    build_items(b, sweeps);
    if (best) std::copy(cfg, cfg + 4, best);
    if (best_ms) *best_ms = best_t;
  });
}

// The single-launch pipelined pass.  Switched off: on MI355X it did not
// complete within minutes on an 11-mesh batch and the cause has not been
// shown; the call reports that instead of launching it.
int trn_batch_pass_pipelined(trn_ctx* ctx, trn_batch* b, const double* xy_in_d, double* xy_d,
                             const double* solut_d, double target_factor, int32_t lp_norm,
                             double h_min, double h_max, double* tensor_d, double* qualit_d,
                             int32_t sweeps) {
  return guard_impl("trn_batch_pass_pipelined", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(b != nullptr, TRN_ERR_ARG, "null batch");
    check_params(xy_d, solut_d, tensor_d, qualit_d, h_min, h_max, sweeps);
    TRN_REQUIRE(kPipelinedEnabled && kBatchPassEnabled, TRN_ERR_STATE, kBatchOff);
    build_items(b, sweeps);
    BatchArgs a = make_args(b, xy_in_d, xy_d, solut_d, target_factor, lp_norm, h_min, h_max,
                            tensor_d, qualit_d);
    const int32_t nctrl = 1 + 2 * b->nstage;
    const int32_t ncount = 2 * kSlots * (sweeps + 1);
    const int32_t ndone = b->nmesh * b->nstage;
    launch(ctx, "batch_reset", 0.0, k_batch_reset,
           dim3(nblk(std::max(std::max(nctrl, ncount), ndone), 256)), dim3(256), 0, b->ctrl,
           nctrl, b->counters, ncount, b->done, ndone);
    launch(ctx, "batch_pipelined", (368.0 + 196.0 * sweeps) * b->nv, k_batch_pipelined,
           dim3(b->num_cus * b->blocks_per_cu), dim3(b->block),
           std::max(patch_shmem(b->mesh, b->block), phase_shmem(b->mesh, b->block)), a);
  });
}

}  // extern "C"
