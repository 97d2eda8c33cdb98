// smoothing.cpp (drop-in) -- trinity::Smooth on MI355X.
//
// Replaces src/smoothing.cpp of hobywan/trinity behind the reference's own
// interface (include/trinity/smoothing.h:39-42).  run() performs
// Smooth::run (src/smoothing.cpp:274-302) through the C ABI, in the
// reference's four timed steps (Stats::step[0..3], :239-240):
//   primal  -- Smooth::preProcess (:129-142): sync.activ marks the movable
//              vertices and the primal graph is rebuilt on the host, as the
//              reference's later passes read them; the device topology is
//              rebuilt only if the elements or tags changed;
//   color   -- Partit::extractColoring (src/partition.cpp:79-197) on the GPU
//              with the Partit's palette 1..max.part-1; a vertex whose
//              neighbours hold every colour stays at colour 0, in no class,
//              and is never moved (partition.cpp:104-110, smoothing.cpp:170);
//              the classes land in the Partit (task.subset / cardin /
//              nb.parts, task.mapping first holding the colours);
//   qualit  -- Smooth::cacheQuality (:145-155);
//   kernel  -- `level` sweeps of moveSmartLaplacian over the classes in
//              ascending colour order (:158-180), points, tensors and
//              element qualities written back to the mesh.
// Stats receives the task and commit counts and the step times
// (:229-241), mesh->param.iter advances (:243-246), and the console output
// of every verbosity level is the reference's (:182-220, :243-270); at
// level 2 a line reports the vertices left uncoloured, if any.
#include "trinity/smoothing.h"

#include <algorithm>
#include <cmath>

#include "trn_dropin.hpp"

namespace trinity {

using trinity_amd_dropin::check;
using trinity_amd_dropin::context;
using trinity_amd_dropin::DeviceMesh;

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

// Console header and clocks of the pass (smoothing.cpp:183-196).
void Smooth::initialize() {
  if (verbose == 1) {
    std::printf("%-18s%s", "= smoothing", "...");
  } else if (verbose == 2) {
    std::printf("Process smoothing ... ");
  }
  std::fflush(stdout);
  time.start = time.iter = time.tic = timer::now();
}

// Per-sweep totals; the field widths come from the first sweep (:199-209).
void Smooth::saveStat(int level, int* stat, int* form) {
  stat[0] += nb.tasks;
  stat[1] += nb.commit;
  if (!level) {
    form[0] = tools::format(nb.tasks);
    form[1] = tools::format(nb.commit);
  }
}

// Per-sweep line at verbosity 2 (:212-226); a sweep without tasks shows 0 %.
void Smooth::showStat(int level, int* form) {
  if (verbose == 2) {
    const int round = level + 1;
    const int percent = nb.tasks ? nb.commit * 100 / nb.tasks : 0;
    const int secs = timer::round(time.iter);
    std::printf("\n= round %2d. %*d tasks \e[0m(100 %%)\e[0m, "
                "%*d comm. \e[0m(%2d %%) \e[32m(%d ms)\e[0m",
                round, form[0], nb.tasks, form[1], nb.commit, percent, secs);
    std::fflush(stdout);
  }
}

// Counters, step times, progress and summary of the pass (:229-271).
void Smooth::recap(int* elap, int* stat, int* form, Stats* total) {
  const int end = std::max(timer::elapsed_ms(time.start), 1);
  if (total != nullptr) {
    total->eval += stat[0];
    total->task += stat[1];
    total->elap += end;
    for (int i = 0; i < 4; ++i) total->step[i] += elap[i];
  }
  *form = tools::format(elap[3]);
  if (not verbose) {
    std::printf("\r= Remeshing  ... %3d %% =", (int) std::floor(100 * (++iter) / (4 * rounds + 1)));
  } else if (verbose == 1) {
    std::printf("%10d task/sec \e[32m(%4.2f s)\e[0m\n", (int) std::floor(stat[0] / (end * 1e-3)),
                (float) end / 1E3);
  } else if (verbose == 2) {
    std::printf("\n\n");
    std::printf("= rate : %d move/sec (%d tasks) \n", (int) std::floor(stat[0] / (end * 1e-3)),
                stat[0]);
    int step[4];
    for (int i = 0; i < 4; ++i) step[i] = elap[i] * 100 / end;
    std::printf("= time per step\n");
    std::printf("  %2d %% primal \e[32m(%*d ms)\e[0m\n", step[0], *form, elap[0]);
    std::printf("  %2d %% color  \e[32m(%*d ms)\e[0m\n", step[1], *form, elap[1]);
    std::printf("  %2d %% qualit \e[32m(%*d ms)\e[0m\n", step[2], *form, elap[2]);
    std::printf("  %2d %% kernel \e[32m(%*d ms)\e[0m\n", step[3], *form, elap[3]);
    std::printf("done. \e[32m(%d ms)\e[0m\n", end);
    tools::separator();
  }
  std::fflush(stdout);
}

void Smooth::run(Stats* total) {
  initialize();
  trn_ctx* ctx = context();
  DeviceMesh& dev = DeviceMesh::get();
  const int nv = nb_nodes, nt = nb_elems;
  int elap[] = {0, 0, 0, 0};
  int form[] = {0, 0};
  int stat[] = {0, 0};

  // 1. primal: Smooth::preProcess side effects (smoothing.cpp:129-142)
#pragma omp parallel
  {
#pragma omp for
    for (int i = 0; i < nv; ++i)
      if (!mesh->topo.stenc[i].empty()) sync.activ[i] = char(mesh->isBoundary(i) ? 0 : 1);
    mesh->extractPrimalGraph();
  }
  trn_mesh* topo = dev.topology(nv, nt, mesh->topo.elems.data(), mesh->sync.tags);
  double* xy = dev.xy.upload(mesh->geom.points.data(), 2 * (size_t) nv);
  double* tensor = dev.tensor.upload(mesh->geom.tensor.data(), 3 * (size_t) nv);
  double* qualit = dev.qualit.reserve(std::max(nt, 1));
  timer::save(time.tic, elap);

  // 2. color: the Partit's palette, overflow left uncoloured
  const int palette = heuris->max.part - 1;
  int32_t* color = dev.color.reserve(std::max(nv, 1));
  int32_t* order = dev.order.reserve(std::max(nv, 1));
  std::vector<int32_t> class_off(palette + 2, 0);
  int32_t ncolors = 0;
  int64_t uncoloured = 0;
  check(trn_color_partial(ctx, topo, 1, palette, color, order, class_off.data(), &ncolors,
                          &uncoloured),
        "trn_color_partial");
  heuris->reset();
  dev.color.download(heuris->task.mapping, (size_t) nv);
  std::vector<int32_t> host_order(class_off[ncolors]);
  dev.order.download(host_order.data(), host_order.size());
  for (int c = 0; c < ncolors; ++c) {
    heuris->task.cardin[c] = class_off[c + 1] - class_off[c];
    std::copy(host_order.begin() + class_off[c], host_order.begin() + class_off[c + 1],
              heuris->task.subset[c]);
  }
  heuris->nb.parts = ncolors;
  timer::save(time.tic, elap + 1);

  // 3. qualit: the element quality cache
  check(trn_smooth_prepare(ctx, topo, xy, tensor, qualit, 1), "trn_smooth_prepare");
  check(trn_ctx_synchronize(ctx), "trn_ctx_synchronize");
  timer::save(time.tic, elap + 2);

  // 4. kernel: the sweeps; one call per sweep when each sweep is shown
  const int depth = task.depth;
  std::vector<int32_t> commits(std::max(depth, 1)), tasks(std::max(depth, 1));
  const int per_call = verbose == 2 ? 1 : std::max(depth, 1);
  for (int level = 0; level < depth; level += per_call) {
    const int n = std::min(per_call, depth - level);
    check(trn_smooth_run_prepared(ctx, topo, xy, tensor, qualit, order, class_off.data(),
                                  ncolors, n, commits.data() + level, tasks.data() + level),
          "trn_smooth_run_prepared");
    for (int s = level; s < level + n; ++s) {
      nb.tasks = tasks[s];
      nb.commit = commits[s];
      saveStat(s, stat, form);
      showStat(s, form);
    }
  }
  if (depth == 0) nb.tasks = nb.commit = 0;
  dev.xy.download(mesh->geom.points.data(), 2 * (size_t) nv);
  dev.tensor.download(mesh->geom.tensor.data(), 3 * (size_t) nv);
  dev.qualit.download(mesh->geom.qualit.data(), (size_t) nt);
  timer::save(time.tic, elap + 3);

  if (verbose == 2 && uncoloured > 0)
    std::printf("\n= %lld vertices left uncoloured (palette of %d colours)",
                (long long) uncoloured, palette);
  recap(elap, stat, form, total);
}

}  // namespace trinity
