// smoothing.cpp (drop-in) -- trinity::Smooth on MI355X.
//
// Replaces src/smoothing.cpp of hobywan/trinity behind the reference's own
// interface (include/trinity/smoothing.h:39-42).  run() performs
// Smooth::run (src/smoothing.cpp:274-302) through the C ABI: the vertex
// colouring of Partit::extractColoring (src/partition.cpp:79-197) on the GPU,
// written back into the Partit's colour classes (task.subset / cardin /
// nb.parts), then `level` sweeps of moveSmartLaplacian over the classes in
// ascending colour order.  Side effects match the reference: points,
// tensors and element qualities of the mesh are updated, sync.activ marks
// the movable vertices and the primal graph is rebuilt (:129-142), and
// Stats receives the task and commit counts (:229-241).
#include "trinity/smoothing.h"

#include <algorithm>

#include "trn_dropin.hpp"

namespace trinity {

using trinity_amd_dropin::check;
using trinity_amd_dropin::context;
using trinity_amd_dropin::DeviceArray;
using trinity_amd_dropin::DeviceTopology;

Smooth::Smooth(Mesh* input, Partit* algo, int level)
    : mesh(input),
      heuris(algo),
      cores(mesh->nb.cores),
      nb_nodes(mesh->nb.nodes),
      nb_elems(mesh->nb.elems),
      verbose(mesh->param.verb),
      iter(mesh->param.iter),
      rounds(mesh->param.rounds) {
  sync.activ = mesh->sync.activ;
  geom.qualit = mesh->geom.qualit.data();
  task.depth = level;
}

void Smooth::run(Stats* total) {
  auto start = timer::now();
  trn_ctx* ctx = context();
  const int nv = nb_nodes, nt = nb_elems;

  // Smooth::preProcess side effects (smoothing.cpp:129-142)
#pragma omp parallel
  {
#pragma omp for
    for (int i = 0; i < nv; ++i)
      if (!mesh->topo.stenc[i].empty()) sync.activ[i] = char(mesh->isBoundary(i) ? 0 : 1);
    mesh->extractPrimalGraph();
  }

  DeviceTopology topo(nv, nt, mesh->topo.elems.data(), mesh->sync.tags);
  DeviceArray<double> xy(2 * (size_t) nv), tensor(3 * (size_t) nv), qualit(nt);
  DeviceArray<int32_t> color(nv), order(nv);
  xy.upload(mesh->geom.points.data());
  tensor.upload(mesh->geom.tensor.data());

  // colouring with the Partit's palette: colours 1..max.part-1
  const int palette = heuris->max.part - 1;
  std::vector<int32_t> class_off(palette + 2, 0);
  int32_t ncolors = 0;
  check(trn_color(ctx, topo.get(), 1, palette, color.get(), order.get(), class_off.data(),
                  &ncolors),
        "trn_color");
  std::vector<int32_t> host_order(nv);
  order.download(host_order.data());
  std::fill(heuris->task.cardin, heuris->task.cardin + heuris->max.part, 0);
  for (int c = 0; c < ncolors; ++c) {
    heuris->task.cardin[c] = class_off[c + 1] - class_off[c];
    std::copy(host_order.begin() + class_off[c], host_order.begin() + class_off[c + 1],
              heuris->task.subset[c]);
  }
  heuris->nb.parts = ncolors;

  std::vector<int32_t> commits(std::max(task.depth, 1)), tasks(std::max(task.depth, 1));
  check(trn_smooth_run(ctx, topo.get(), xy.get(), tensor.get(), qualit.get(), order.get(),
                       class_off.data(), ncolors, task.depth, 1, commits.data(), tasks.data()),
        "trn_smooth_run");
  xy.download(mesh->geom.points.data());
  tensor.download(mesh->geom.tensor.data());
  qualit.download(mesh->geom.qualit.data());

  int stat[2] = {0, 0};
  for (int s = 0; s < task.depth; ++s) {
    stat[0] += tasks[s];
    stat[1] += commits[s];
  }
  nb.tasks = task.depth ? tasks[task.depth - 1] : 0;
  nb.commit = task.depth ? commits[task.depth - 1] : 0;
  const int end = std::max(timer::elapsed_ms(start), 1);
  if (total != nullptr) {  // smoothing.cpp:234-240
    total->eval += stat[0];
    total->task += stat[1];
    total->elap += end;
  }
  if (verbose == 1) {
    std::printf("%-18s%s%10d task/sec \e[32m(%4.2f s)\e[0m\n", "= smoothing", "...",
                (int) std::floor(stat[0] / (end * 1e-3)), (float) end / 1e3);
  } else if (verbose == 2) {
    std::printf("Process smoothing on GPU ... %d colours, %d tasks, %d commits (%d ms)\n",
                ncolors, stat[0], stat[1], end);
  }
  std::fflush(stdout);
}

}  // namespace trinity
