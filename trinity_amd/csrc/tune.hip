// tune.hip -- the autotuner of the plain kernels' launch shapes.
//
// For one mesh it measures, with HIP events on the context's stream,
//   * the slabs per block (1, 2, 4) of the gradient and of the Hessian
//     gathers (metric.hip), on the caller's coordinates and field;
//   * the register shape (2, 3, 4 waves per SIMD) of the smoothing phase
//     kernel (smoothing.hip), on the largest colour class of the mesh's own
//     colouring, first sweep, with the metric of the caller's field;
// and keeps the fastest of each in the mesh.  The vertex records and the
// quality cache the phase runs modify are saved before and restored after,
// so the tuner leaves no trace but the shapes.  No shape changes a result:
// the gathers sum every patch in the same order whatever the block, and a
// phase kernel's arithmetic is the same at every register budget.
#include <algorithm>
#include <string>
#include <vector>

#include "trn_api.hpp"

namespace trn {
namespace {

// Device buffers of one tuning run, freed on every exit path.
struct TuneBuffers {
  std::vector<void*> ptrs;
  template <class T>
  T* alloc(size_t n) {
    void* p = nullptr;
    TRN_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    ptrs.push_back(p);
    return static_cast<T*>(p);
  }
  ~TuneBuffers() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

struct Events {
  hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
  Events() {
    for (auto& x : e) TRN_HIP(hipEventCreate(&x));
  }
  ~Events() {
    for (auto& x : e)
      if (x) (void)hipEventDestroy(x);
  }
  float ms(int a, int b) const {
    float t = 0.f;
    TRN_HIP(hipEventElapsedTime(&t, e[a], e[b]));
    return t;
  }
};

constexpr int kReps = 3;  // the first of each candidate is a warm-up

}  // namespace
}  // namespace trn

using namespace trn;

extern "C" {

int trn_mesh_autotune(trn_ctx* ctx, trn_mesh* mesh, const double* xy_d, const double* solut_d,
                      int32_t lp_norm, int32_t* shapes, double* times_ms) {
  return guard_impl("trn_mesh_autotune", [&] {
    DeviceScope ds(ctx);
    TRN_REQUIRE(mesh && xy_d && solut_d, TRN_ERR_ARG, "null argument");
    TRN_REQUIRE(!ctx->timing, TRN_ERR_STATE, "per-kernel timing must be off while tuning");
    trn_mesh* m = mesh;
    TuneBuffers buf;
    Events ev;
    const int cand[3] = {1, 2, 4};

    // 1. gathers: each candidate kReps times, the fastest after a warm-up
    const size_t gbytes = tune_gather_scratch(m);
    void* gs = buf.alloc<char>(gbytes);
    float grad_t[3], hess_t[3];
    for (int c = 0; c < 3; ++c) {
      grad_t[c] = hess_t[c] = 1e30f;
      for (int r = 0; r < kReps; ++r) {
        tune_gathers(ctx, m, xy_d, solut_d, lp_norm, gs, gbytes, cand[c], cand[c], ev.e);
        TRN_HIP(hipEventSynchronize(ev.e[3]));
        if (r == 0) continue;
        grad_t[c] = std::min(grad_t[c], ev.ms(0, 1));
        hess_t[c] = std::min(hess_t[c], ev.ms(2, 3));
      }
    }
    const int gb = (int)(std::min_element(grad_t, grad_t + 3) - grad_t);
    const int hb = (int)(std::min_element(hess_t, hess_t + 3) - hess_t);

    // 2. smoothing phase on the largest class of the mesh's colouring
    double* tensor = buf.alloc<double>(3 * (size_t)m->nv);
    double* xy = buf.alloc<double>(2 * (size_t)m->nv);
    double* qualit = buf.alloc<double>((size_t)m->nt);
    int32_t* color = buf.alloc<int32_t>((size_t)m->nv);
    int32_t* order = buf.alloc<int32_t>((size_t)m->nv);
    uint8_t* stale = buf.alloc<uint8_t>((size_t)m->nv);
    int32_t* counters = buf.alloc<int32_t>(256);
    // the working copy (vdata, vt2) as found, and each run's starting point
    const size_t nvd = sizeof(VData) * (size_t)m->nv, nt2 = sizeof(double) * (size_t)m->nv;
    VData* vd_caller = buf.alloc<VData>((size_t)m->nv);
    double* t2_caller = buf.alloc<double>((size_t)m->nv);
    VData* vd_prepared = buf.alloc<VData>((size_t)m->nv);
    double* t2_prepared = buf.alloc<double>((size_t)m->nv);
    double* q_prepared = buf.alloc<double>((size_t)m->nt);
    copy_async(ctx, vd_caller, m->vdata, nvd);
    copy_async(ctx, t2_caller, m->vt2, nt2);
    copy_async(ctx, xy, xy_d, sizeof(double) * 2 * (size_t)m->nv);
    int rc = trn_metric_run(ctx, m, xy, solut_d, 0.8, lp_norm, 1e-9, 0.4, tensor, nullptr,
                            nullptr, nullptr);
    TRN_REQUIRE(rc == TRN_OK, rc, std::string("metric: ") + trn_last_error());
    rc = trn_smooth_prepare(ctx, m, xy, tensor, qualit, 1);
    TRN_REQUIRE(rc == TRN_OK, rc, std::string("prepare: ") + trn_last_error());
    std::vector<int32_t> off(66, 0);
    int32_t nc = 0;
    rc = trn_color(ctx, m, 1, 64, color, order, off.data(), &nc);
    TRN_REQUIRE(rc == TRN_OK, rc, std::string("colouring: ") + trn_last_error());
    int big = 0;
    for (int c = 1; c < nc; ++c)
      if (off[c + 1] - off[c] > off[big + 1] - off[big]) big = c;
    const int32_t count = nc ? off[big + 1] - off[big] : 0;
    copy_async(ctx, vd_prepared, m->vdata, nvd);
    copy_async(ctx, t2_prepared, m->vt2, nt2);
    copy_async(ctx, q_prepared, qualit, sizeof(double) * (size_t)m->nt);
    float phase_t[3] = {0.f, 0.f, 0.f};
    const int minw[3] = {2, 3, 4};
    if (count > 0) {
      for (int c = 0; c < 3; ++c) {
        phase_t[c] = 1e30f;
        for (int r = 0; r < kReps; ++r) {
          copy_async(ctx, m->vdata, vd_prepared, nvd);
          copy_async(ctx, m->vt2, t2_prepared, nt2);
          copy_async(ctx, qualit, q_prepared, sizeof(double) * (size_t)m->nt);
          memset_async(ctx, stale, 0, (size_t)m->nv);
          memset_async(ctx, counters, 0, sizeof(int32_t) * 256);
          tune_phase(ctx, m, count, order + off[big], qualit, stale, counters, minw[c], ev.e);
          TRN_HIP(hipEventSynchronize(ev.e[1]));
          if (r == 0) continue;
          phase_t[c] = std::min(phase_t[c], ev.ms(0, 1));
        }
      }
    }
    copy_async(ctx, m->vdata, vd_caller, nvd);
    copy_async(ctx, m->vt2, t2_caller, nt2);
    sync(ctx);
    const int pb = count > 0 ? (int)(std::min_element(phase_t, phase_t + 3) - phase_t) : 1;

    m->tune.grad_waves = cand[gb];
    m->tune.hess_waves = cand[hb];
    m->tune.phase_minw = minw[pb];
    m->tune.tuned = true;
    if (shapes) {
      shapes[0] = m->tune.grad_waves;
      shapes[1] = m->tune.hess_waves;
      shapes[2] = m->tune.phase_minw;
    }
    if (times_ms) {  // the default shape's time, then the chosen one's
      times_ms[0] = grad_t[0];
      times_ms[1] = grad_t[gb];
      times_ms[2] = hess_t[0];
      times_ms[3] = hess_t[hb];
      times_ms[4] = phase_t[1];
      times_ms[5] = phase_t[pb];
    }
  });
}

int trn_mesh_shapes(const trn_mesh* mesh, int32_t* shapes) {
  return guard_impl("trn_mesh_shapes", [&] {
    TRN_REQUIRE(mesh && shapes, TRN_ERR_ARG, "null argument");
    shapes[0] = mesh->tune.grad_waves;
    shapes[1] = mesh->tune.hess_waves;
    shapes[2] = mesh->tune.phase_minw;
  });
}

int trn_mesh_set_shapes(trn_mesh* mesh, const int32_t* shapes) {
  return guard_impl("trn_mesh_set_shapes", [&] {
    TRN_REQUIRE(mesh && shapes, TRN_ERR_ARG, "null argument");
    auto pow2 = [](int32_t w) { return w == 1 || w == 2 || w == 4; };
    TRN_REQUIRE(pow2(shapes[0]) && pow2(shapes[1]), TRN_ERR_ARG,
                "slabs per gather block must be 1, 2 or 4");
    TRN_REQUIRE(shapes[2] >= 2 && shapes[2] <= 4, TRN_ERR_ARG,
                "phase register shape must be 2, 3 or 4 waves per SIMD");
    mesh->tune.grad_waves = shapes[0];
    mesh->tune.hess_waves = shapes[1];
    mesh->tune.phase_minw = shapes[2];
  });
}

}  // extern "C"
