// trn_runtime.hpp -- context, error handling, launch timing and scratch
// management shared by every translation unit of libtrinity_amd.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/trinity_amd.h"

namespace trn {

// Internal errors travel as exceptions and are turned into status codes at
// the C boundary (trn_guard); nothing throws across the ABI.
struct Error {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg);
void set_last_error(const std::string& msg);

#define TRN_HIP(expr)                                                        \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess)                                                    \
      ::trn::fail(TRN_ERR_HIP, std::string(#expr) + ": " +                   \
                                   hipGetErrorString(_e));                   \
  } while (0)

#define TRN_REQUIRE(cond, code, msg)                                         \
  do {                                                                       \
    if (!(cond)) ::trn::fail((code), (msg));                                 \
  } while (0)

// Per-kernel timing (opt-in): hipEvents bracket each launch on the stream
// the kernel runs on; durations are folded into per-name totals when the
// caller asks for them.
struct KernelStat {
  std::string name;
  double total_ms = 0.0;
  int64_t launches = 0;
  double bytes = 0.0;  // algorithmic bytes per launch x launches (if known)
  double bytes_per_unit = 0.0;  // for work counted on the device (timing_units)
};

struct PendingTiming {
  int stat;
  hipEvent_t start, stop;
};

// Workspace: one growing device buffer, carved into aligned slices per call.
struct Workspace {
  char* base = nullptr;
  size_t size = 0;
  int live = 0;  // Scratch objects carving it (at most one)
};

}  // namespace trn

struct trn_ctx {
  int device = 0;
  hipStream_t stream = nullptr;      // stream every launch goes to
  hipStream_t own_stream = nullptr;  // the context's own (default priority)
  trn::Workspace ws;
  trn::Workspace aux;                // temporaries of mesh creation (grows)
  bool timing = false;
  std::string timing_only;           // non-empty: time only launches of this name
  std::vector<trn::KernelStat> stats;
  std::vector<trn::PendingTiming> pending;
  std::vector<hipEvent_t> event_pool;
  int32_t* pinned_i32 = nullptr;     // small pinned host staging buffers
  double* pinned_f64 = nullptr;
  unsigned long long* unit_dev = nullptr;  // [kTimingUnitSlots] device work counters
};

namespace trn {
// Compact topology of one vertex for the smoothing sweeps (56 bytes): the
// 1-ring and incident triangles of a vertex with at most kVtopoDeg of each,
// in CSR order, and per incident triangle a star byte r | a << 2 | b << 5
// (tri[t][r] = i, tri[t][(r+1)%3] = ring[a], tri[t][(r+2)%3] = ring[b]).
// meta = deg | nb << 4, or kVtopoGeneric when the vertex has more than
// kVtopoDeg incident triangles or neighbours (then the CSR arrays are used).
constexpr int kVtopoDeg = 6;
constexpr uint8_t kVtopoGeneric = 0xFF;
// Padded to 64 bytes, so an entry never straddles two 128-byte lines.
struct alignas(64) VTopo {
  int32_t ring[kVtopoDeg];
  int32_t tris[kVtopoDeg];
  uint8_t star[kVtopoDeg];
  uint8_t meta;
  uint8_t tag;
};
static_assert(sizeof(VTopo) == 64, "VTopo is 64 bytes");

// What a move reads of a neighbour (metric log, coordinates, metric) in
// the smoothing sweeps' working copy: 64 bytes per vertex, two vertices per
// 128-byte line (a miss in L2 brings a whole line), with the metric's
// third entry in a dense array of its own (eight vertices per 64 bytes).
// Filled by each smoothing call from the caller's arrays; a committed move
// writes its new data there and to those arrays.
struct alignas(64) VData {
  double4 logm;   // matrix log of the metric (numeric.cpp:115-154)
  double2 xy;
  double t01[2];  // metric entries m11, m12
};
static_assert(sizeof(VData) == 64, "VData is 64 bytes");

constexpr int kPatchSlab = 64;  // vertices per slab of the stored patches
constexpr int kUnionKeys = 4096;  // largest slab (rows x kPatchSlab) given a union
constexpr int kUnionCap = 2048;   // most triangles in one slab's union
}  // namespace trn

struct trn_mesh {
  trn_ctx* ctx = nullptr;
  int32_t nv = 0, nt = 0;
  int64_t nnz_v2v = 0;
  int32_t max_deg_t = 0;     // max incident triangles per vertex
  int32_t max_deg_v = 0;     // max 1-ring size
  int32_t max_patch = 0;     // max 2-ring element patch size
  int32_t patch_cap = 0;     // LDS slots per thread used by the gathers
  int32_t* tri = nullptr;    // [3nt]
  uint8_t* tags = nullptr;   // [nv]
  int32_t* v2t_off = nullptr;
  int32_t* v2t = nullptr;
  int32_t* v2v_off = nullptr;
  int32_t* v2v = nullptr;
  // Per incident triangle (parallel to v2t): where the vertex sits in the
  // triangle and which of its 1-ring entries are the two other corners,
  // packed r | a << 2 | b << 10 with tri[t][r] = i, tri[t][(r+1)%3] =
  // v2v[v2v_off[i]+a], tri[t][(r+2)%3] = v2v[v2v_off[i]+b].
  int32_t* star = nullptr;
  // Vertices along a Z-curve of the coordinates given at creation (NULL:
  // none were given).
  int32_t* zorder = nullptr;
  // Vertex at each position of the patch slabs (slab b holds positions
  // b * kPatchSlab ..): the Z-curve order when there is one, so that a slab
  // is a compact patch of the mesh (NULL: index order).  Not owned.
  const int32_t* slab_v = nullptr;
  // Visiting order of the vertices of a colour class: NULL (ascending, the
  // default) or zorder (trn_mesh_set_class_order).  Not owned.
  const int32_t* class_v = nullptr;
  // [nv] working copy of the smoothing sweeps (vdata, vt2: metric m22)
  // and the vertices' compact topology (filled at creation); smoothing
  // calls on one mesh must not overlap.
  trn::VData* vdata = nullptr;
  double* vt2 = nullptr;
  trn::VTopo* vtopo = nullptr;
  // 2-ring element patches (Mesh::getVicinity(i, 2), ascending) in sliced
  // ELL form: vertices in slabs of kPatchSlab; entry e of the patch of
  // vertex i = slab_v[b * kPatchSlab + l] is patch_ell[(patch_row[b] + e) *
  // kPatchSlab + l], e < patch_n[i]; slab b has patch_row[b+1] -
  // patch_row[b] rows (its largest patch), shorter patches padded with -1.
  int32_t* patch_n = nullptr;    // [nv]
  int32_t* patch_row = nullptr;  // [nslab + 1]
  int32_t* patch_ell = nullptr;  // [patch_row[nslab] * kPatchSlab]
  // Per slab, the ascending union of its vertices' patches (union_n[b]
  // triangles at union_ids + union_off[b], union_off[b] = patch_row[b] *
  // kPatchSlab: the slab's ELL block, which can hold any union of its
  // entries; union_n[b] < 0: no union, the slab gathers through
  // patch_ell), and patch_local, shaped like patch_ell, the position of
  // each patch entry in its slab's union.
  int32_t* union_n = nullptr;        // [nslab]
  int32_t* union_off = nullptr;      // [nslab + 1]
  int32_t* union_ids = nullptr;      // [patch_row[nslab] * kPatchSlab]
  uint16_t* patch_local = nullptr;   // [patch_row[nslab] * kPatchSlab]
  // The same unions packed one after the other (slab b's entries from
  // union_cu[b]), as the three corners of each entry's triangle: the
  // gathers stage a union entry's terms from its corners directly.
  int32_t* union_cu = nullptr;       // [nslab + 1]
  int32_t* union_tri = nullptr;      // [3 * union_cu[nslab]]
  int32_t max_union = 0;             // largest union_n
  bool union_all = false;            // every slab has a union
  // Launch shapes of the plain kernels (trn_mesh_autotune measures them on
  // the mesh): slabs per block of the gradient and Hessian gathers, and the
  // waves per SIMD the smoothing phase kernel is compiled for (its register
  // budget).  None of them changes a result.
  struct {
    int32_t grad_waves = 1, hess_waves = 1, phase_minw = 3;
    bool tuned = false;
  } tune;
  int32_t color_tile = 0;            // colouring tile size (0: chosen from bandwidth)
  int32_t bandwidth = 0;             // largest |u - v| over the edges
};

namespace trn {

// Launch helpers ---------------------------------------------------------
int stat_index(trn_ctx* ctx, const char* name);
// Work of a launch that only the device knows (e.g. the length of a list it
// reads): with timing on, returns a device counter the kernel adds its work
// units to; they are turned into bytes (units x bytes_per_unit) when the
// timings are collected.  nullptr when launches of `name` are not timed.
constexpr int kTimingUnitSlots = 64;
unsigned long long* timing_units(trn_ctx* ctx, const char* name, double bytes_per_unit);
void timing_begin(trn_ctx* ctx, int stat, hipEvent_t* start, hipEvent_t* stop);
void timing_end(trn_ctx* ctx, int stat, hipEvent_t start, hipEvent_t stop,
                double bytes);

// Whether launches of the statistic `name` are timed: timing on, and no
// name filter or this name.
inline bool timed(const trn_ctx* ctx, const char* name) {
  return ctx->timing && (ctx->timing_only.empty() || ctx->timing_only == name);
}

// Launches `kernel` on ctx->stream; when its launches are timed, brackets it
// with events and credits `bytes` algorithmic bytes to the named statistic.
template <typename K, typename... Args>
inline void launch(trn_ctx* ctx, const char* name, double bytes, K kernel,
                   dim3 grid, dim3 block, size_t shmem, Args... args) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = -1;
  const bool t = timed(ctx, name);
  if (t) {
    st = stat_index(ctx, name);
    timing_begin(ctx, st, &e0, &e1);
  }
  hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, args...);
  TRN_HIP(hipGetLastError());
  if (t) timing_end(ctx, st, e0, e1, bytes);
}

// Scratch slices from the context workspace.  An entry point sizes all of
// A second growing device buffer of the context for the temporaries of mesh
// creation (the stored patches before their ELL copy), held until the
// context is destroyed.  Queued work reading an earlier use must have ended
// (it synchronises before growing).
void* aux_workspace(trn_ctx* ctx, size_t bytes);

// its scratch up front (Scratch(ctx, total)), which grows the workspace if
// needed, then carves slices in order.  Slices live until the next entry
// point on the same context; every entry point is stream-ordered on
// ctx->stream, so reuse across calls is safe.
inline size_t scratch_bytes(size_t bytes) { return (bytes + 255) & ~size_t(255); }

// Slices of the context workspace for one entry point's launches.  Every
// Scratch carves from the start of the workspace (and may reallocate it),
// so at most one may be live per context: a second one while another is
// live is refused (TRN_ERR_STATE) instead of aliasing the first's buffers.
class Scratch {
 public:
  Scratch(trn_ctx* ctx, size_t total);
  ~Scratch();
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  template <typename T>
  T* carve(size_t count) {
    size_t bytes = scratch_bytes(count * sizeof(T));
    TRN_REQUIRE(off_ + bytes <= ctx_->ws.size, TRN_ERR_STATE, "scratch overflow");
    T* p = reinterpret_cast<T*>(ctx_->ws.base + off_);
    off_ += bytes;
    return p;
  }

 private:
  trn_ctx* ctx_;
  size_t off_ = 0;
};

// Slices of a caller-provided device buffer, carved in order (the same
// alignment as Scratch), for internal launch sequences whose buffers live
// in an object of their own (a batch) rather than the context workspace.
class Carve {
 public:
  Carve(void* base, size_t bytes) : base_((char*)base), cap_(bytes) {}
  template <typename T>
  T* carve(size_t count) {
    size_t bytes = scratch_bytes(count * sizeof(T));
    TRN_REQUIRE(off_ + bytes <= cap_, TRN_ERR_STATE, "carve overflow");
    T* p = reinterpret_cast<T*>(base_ + off_);
    off_ += bytes;
    return p;
  }

 private:
  char* base_;
  size_t cap_;
  size_t off_ = 0;
};

void sync(trn_ctx* ctx);
void memset_async(trn_ctx* ctx, void* p, int v, size_t bytes);

// XCD-aware block index: the hardware hands workgroups to the 8 XCDs round
// robin (block b -> XCD b % 8).  Remapping gives each XCD one contiguous
// range of logical blocks, so neighbouring vertex blocks, which share
// stencil data, run on the same XCD and hit the same L2.
__device__ __forceinline__ uint32_t xcd_block_of(const uint32_t b, const uint32_t n) {
  const uint32_t q = n >> 3, r = n & 7u, x = b & 7u, k = b >> 3;
  return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ uint32_t xcd_block() { return xcd_block_of(blockIdx.x, gridDim.x); }

}  // namespace trn
