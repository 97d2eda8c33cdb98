// metric.cpp (drop-in) -- trinity::Metrics on MI355X.
//
// Replaces src/metric.cpp and src/hessian.cpp of hobywan/trinity: the
// public interface is the reference's own (include/trinity/metric.h:38-44,
// compiled against that header), and run() performs Metrics::run
// (src/metric.cpp:59-70) with one trn_metric_run call: gradient and Hessian
// recovery over the 2-ring patches, local L^p normalisation, complexity and
// global normalisation, all on the GPU.  Side effects match the reference:
// the metric lands in Mesh::geom.tensor, the nodal field is released
// (:92-96), Stats receives nb_nodes evaluations and tasks (:275-279), the
// shared progress counter mesh->param.iter advances (:282-286) and the
// console output of every verbosity level is the reference's (:249-255,
// :281-302).
#include "trinity/metric.h"

#include <algorithm>
#include <cmath>
#include <string>

#include "trn_dropin.hpp"

namespace trinity {

using trinity_amd_dropin::check;
using trinity_amd_dropin::context;
using trinity_amd_dropin::DeviceMesh;

Metrics::Metrics(Mesh* input_mesh, double target_factor, int Lp_norm, double h_min,
                 double h_max)
    : mesh(input_mesh),
      nb_nodes(mesh->nb.nodes),
      nb_elems(mesh->nb.elems),
      nb_cores(mesh->nb.cores),
      verbose(mesh->param.verb),
      iter(mesh->param.iter),
      rounds(mesh->param.rounds) {
  // same parameters as metric.cpp:38-46
  param.chunk = mesh->nb.nodes / mesh->nb.cores;
  param.target = (int) std::floor(mesh->nb.nodes * target_factor);
  param.norm = Lp_norm;
  param.h_min = h_min;
  param.h_max = h_max;
  param.scale.fact = 0;
  param.scale.exp = Lp_norm ? -1. / (2 * Lp_norm + 2) : 0.5;
  param.eigen.min = 1 / std::pow(h_min, 2);
  param.eigen.max = 1 / std::pow(h_max, 2);
  field.complexity = 0;
  field.solut = mesh->geom.solut.data();
  field.tensor = mesh->geom.tensor.data();
  field.stencil = nullptr;   // the 2-ring patches live on the device
  field.gradient = nullptr;
}

Metrics::~Metrics() {}

// Console header of the pass (metric.cpp:242-259).
void Metrics::initialize() {
  if (not verbose) {
    std::printf("\n\r= Remeshing  ... %3d %% =", 0);
  } else if (verbose == 1) {
    std::printf("%-18s%s", "= metric field", "...");
  } else if (verbose == 2) {
    std::printf("Computing multi-scale metric field ... \n");
  }
  std::fflush(stdout);
  time.start = timer::now();
}

// Counters, progress and summary of the pass (metric.cpp:262-303).  The
// rate divides by at least 1 ms: a pass can finish within a millisecond.
void Metrics::recap(Stats* total) {
  const int end = timer::elapsed_ms(time.start);
  if (total != nullptr) {
    total->eval += nb_nodes;
    total->task += nb_nodes;
    total->elap += end;
  }
  if (not verbose) {
    std::printf("\r= Remeshing  ... %3d %% =", (int) std::floor(100 * (++iter) / (4 * rounds + 1)));
  } else if (verbose == 1) {
    std::printf("%10d task/sec \e[32m(%4.2f s)\e[0m\n",
                (int) std::floor(nb_nodes / (std::max(end, 1) * 1e-3)), (float) end / 1e3);
  } else if (verbose == 2) {
    std::printf("= norm L^%s\n", param.norm <= 0 ? "inf" : std::to_string(param.norm).data());
    std::printf("= target : %.2e\n", (float) param.target);
    std::printf("= complex: %.1f\n", (float) std::floor(field.complexity));
    std::printf("= scale_f: %.1f\n", param.scale.fact);
    std::printf("done. \e[32m(%d ms)\e[0m\n\n", end);
  }
  std::fflush(stdout);
}

void Metrics::run(Stats* total) {
  initialize();
  trn_ctx* ctx = context();
  DeviceMesh& dev = DeviceMesh::get();
  const int nv = nb_nodes, nt = nb_elems;
  trn_mesh* topo = dev.topology(nv, nt, mesh->topo.elems.data(), mesh->sync.tags);
  const double* xy = dev.xy.upload(mesh->geom.points.data(), 2 * (size_t) nv);
  const double* f = dev.solut.upload(mesh->geom.solut.data(), (size_t) nv);
  double* tensor = dev.tensor.reserve(3 * (size_t) nv);

  // trn_metric_run takes the target factor and floors nv * factor
  // (metric.cpp:39); (target + 1/2) / nv floors back to param.target for
  // any nv below 2^52, whatever nb.nodes was when the object was built.
  const double factor = (param.target + 0.5) / nv;
  double complexity = 0;
  check(trn_metric_run(ctx, topo, xy, f, factor, param.norm, param.h_min, param.h_max, tensor,
                       &complexity, nullptr, nullptr),
        "trn_metric_run");
  field.complexity = complexity;
  param.scale.fact = param.target / field.complexity;  // metric.cpp:197-198
  dev.tensor.download(mesh->geom.tensor.data(), 3 * (size_t) nv);

  mesh->geom.solut.clear();  // metric.cpp:92-96
  mesh->geom.solut.shrink_to_fit();
  recap(total);
}

}  // namespace trinity
