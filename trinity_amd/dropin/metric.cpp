// metric.cpp (drop-in) -- trinity::Metrics on MI355X.
//
// Replaces src/metric.cpp and src/hessian.cpp of hobywan/trinity: the
// public interface is the reference's own (include/trinity/metric.h:38-44,
// compiled against that header), and run() performs Metrics::run
// (src/metric.cpp:59-70) through the C ABI: gradient and Hessian recovery
// over the 2-ring patches, local L^p normalisation, complexity and global
// normalisation, all on the GPU.  Side effects match the reference: the
// metric lands in Mesh::geom.tensor, the nodal field is released
// (:92-96), and Stats receives nb_nodes evaluations and tasks (:275-279).
#include "trinity/metric.h"

#include <algorithm>
#include <cmath>

#include "trn_dropin.hpp"

namespace trinity {

using trinity_amd_dropin::check;
using trinity_amd_dropin::context;
using trinity_amd_dropin::DeviceArray;
using trinity_amd_dropin::DeviceTopology;

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

void Metrics::run(Stats* total) {
  auto start = timer::now();
  trn_ctx* ctx = context();
  const int nv = nb_nodes, nt = nb_elems;
  DeviceTopology topo(nv, nt, mesh->topo.elems.data(), mesh->sync.tags);
  DeviceArray<double> xy(2 * (size_t) nv), f(nv), grad(2 * (size_t) nv), tensor(3 * (size_t) nv);
  xy.upload(mesh->geom.points.data());
  f.upload(mesh->geom.solut.data());

  check(trn_gradient(ctx, topo.get(), xy.get(), f.get(), grad.get()), "trn_gradient");
  check(trn_hessian(ctx, topo.get(), xy.get(), grad.get(), tensor.get()), "trn_hessian");
  check(trn_normalize_local(ctx, nv, tensor.get(), param.norm), "trn_normalize_local");
  double complexity = 0;
  check(trn_complexity(ctx, topo.get(), xy.get(), tensor.get(), &complexity), "trn_complexity");
  field.complexity = complexity;
  param.scale.fact = param.target / field.complexity;  // metric.cpp:197-198
  check(trn_normalize_global(ctx, nv, tensor.get(), param.scale.fact, param.h_min, param.h_max),
        "trn_normalize_global");
  check(trn_ctx_synchronize(ctx), "trn_ctx_synchronize");
  tensor.download(mesh->geom.tensor.data());

  mesh->geom.solut.clear();  // metric.cpp:92-96
  mesh->geom.solut.shrink_to_fit();

  const int end = timer::elapsed_ms(start);
  if (total != nullptr) {
    total->eval += nb_nodes;
    total->task += nb_nodes;
    total->elap += end;
  }
  if (verbose == 1) {
    std::printf("%-18s%s%10d task/sec \e[32m(%4.2f s)\e[0m\n", "= metric field", "...",
                (int) std::floor(nb_nodes / (std::max(end, 1) * 1e-3)), (float) end / 1e3);
  } else if (verbose == 2) {
    std::printf("Computing multi-scale metric field on GPU ... done. (%d ms)\n", end);
    std::printf("= complex: %.1f\n= scale_f: %.1f\n", (float) std::floor(field.complexity),
                param.scale.fact);
  }
  std::fflush(stdout);
}

}  // namespace trinity
