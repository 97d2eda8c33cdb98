/*
 * ref_harness.cpp -- drives the UNMODIFIED hobywan/trinity sources
 * (compiled from /root/reference/src by oracle/Makefile into oracle/_ref/)
 * to produce per-stage golden dumps and CPU timings of the reference path.
 *
 * TEST INFRASTRUCTURE ONLY.  It is the checker's checker: it validates the
 * C restatement (trinity_oracle.c) and serves as the CPU baseline.
 *
 * Stage methods are private in the reference; this translation unit alone
 * maps `private` to `public` after every standard header the reference
 * includes has been included (a global -Dprivate=public breaks libstdc++).
 *
 * Input: a TRNM binary mesh (see trinity_amd/meshio.py):
 *   char[4] "TRNM"; int32 version(=1); int32 nv; int32 nt;
 *   f64 xy[2nv]; int32 tri[3nt] (0-based); u8 tags[nv]; f64 solut[nv]
 *
 * Usage:
 *   ref_harness dump  MESH OUTDIR tf p hmin hmax sweeps [COLOR]
 *   ref_harness bench MESH JSON   tf p hmin hmax sweeps reps
 *   ref_harness load  MEDIT BB OUTDIR [JSON]   (Mesh::load of the text files:
 *                     xy, tri (0-based), tags, solut dumps; JSON: its time)
 */
#include <stack>
#include <set>
#include <map>
#include <vector>
#include <string>
#include <algorithm>
#include <omp.h>
#include <thread>
#include <atomic>
#include <memory>
#include <cstring>
#include <sstream>
#include <fstream>
#include <iostream>
#include <iomanip>
#include <cmath>
#include <random>
#include <limits>
#include <chrono>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <csignal>
#include <cassert>
#include <functional>
#include <unordered_map>
#include <numeric>
#include <utility>
#include <unistd.h>

#define private public
#include "trinity.h"
#undef private

using namespace trinity;

struct MeshData {
  int nv = 0, nt = 0;
  std::vector<double> xy, solut;
  std::vector<int32_t> tri;
  std::vector<uint8_t> tags;
};

static void die(const char* msg) {
  std::fprintf(stderr, "ref_harness: %s\n", msg);
  std::exit(2);
}

static MeshData readMesh(const char* path) {
  MeshData m;
  FILE* f = std::fopen(path, "rb");
  if (!f) die("cannot open mesh");
  char magic[4];
  int32_t hdr[3];
  if (std::fread(magic, 1, 4, f) != 4 || std::memcmp(magic, "TRNM", 4))
    die("bad magic");
  if (std::fread(hdr, 4, 3, f) != 3 || hdr[0] != 1) die("bad header");
  m.nv = hdr[1];
  m.nt = hdr[2];
  m.xy.resize(2 * (size_t)m.nv);
  m.tri.resize(3 * (size_t)m.nt);
  m.tags.resize(m.nv);
  m.solut.resize(m.nv);
  if (std::fread(m.xy.data(), 8, m.xy.size(), f) != m.xy.size() ||
      std::fread(m.tri.data(), 4, m.tri.size(), f) != m.tri.size() ||
      std::fread(m.tags.data(), 1, m.tags.size(), f) != m.tags.size() ||
      std::fread(m.solut.data(), 8, m.solut.size(), f) != m.solut.size())
    die("short mesh file");
  std::fclose(f);
  return m;
}

/* Equivalent of Mesh::load (io.cpp:53-161) from binary arrays: points,
 * 0-based elements, tags (bound = 4, corner = 8) and the nodal field, then
 * rebuildTopology exactly as load does (io.cpp:128-129). */
static void fillMesh(Mesh& mesh, const MeshData& d) {
  for (size_t i = 0; i < d.xy.size(); ++i) mesh.geom.points[i] = d.xy[i];
  for (size_t i = 0; i < d.tri.size(); ++i) mesh.topo.elems[i] = d.tri[i];
  for (int i = 0; i < d.nv; ++i) mesh.sync.tags[i] = d.tags[i];
  for (int i = 0; i < d.nv; ++i) mesh.geom.solut[i] = d.solut[i];
#pragma omp parallel
  mesh.rebuildTopology();
}

static void writeBin(const std::string& path, const void* p, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) die("cannot write output");
  if (bytes && std::fwrite(p, 1, bytes, f) != bytes) die("short write");
  std::fclose(f);
}

static int dumpMode(int argc, char** argv) {
  if (argc < 9) die("dump needs MESH OUTDIR tf p hmin hmax sweeps [COLOR]");
  MeshData d = readMesh(argv[2]);
  std::string out = argv[3];
  double tf = std::atof(argv[4]);
  int p = std::atoi(argv[5]);
  double hmin = std::atof(argv[6]);
  double hmax = std::atof(argv[7]);
  int sweeps = std::atoi(argv[8]);
  const char* colorPath = argc > 9 ? argv[9] : nullptr;

  int size[2] = {d.nv, d.nt};
  Mesh mesh(size, 64, 2, 0, 1);
  fillMesh(mesh, d);

  Metrics metric(&mesh, tf, p, hmin, hmax);
#pragma omp parallel
  {
    metric.initialize();
    metric.recoverHessianField();
  }
  writeBin(out + "/gradient.bin", metric.field.gradient, sizeof(double) * 2 * d.nv);
  writeBin(out + "/hessian.bin", mesh.geom.tensor.data(), sizeof(double) * 3 * d.nv);
#pragma omp parallel
  metric.normalizeLocally();
  writeBin(out + "/metric_local.bin", mesh.geom.tensor.data(), sizeof(double) * 3 * d.nv);
#pragma omp parallel
  {
    metric.computeComplexity();
    metric.normalizeGlobally();
  }
  writeBin(out + "/complexity.bin", &metric.field.complexity, sizeof(double));
  writeBin(out + "/metric.bin", mesh.geom.tensor.data(), sizeof(double) * 3 * d.nv);
  metric.clear();

  Partit partit(&mesh, 8);
  Smooth smooth(&mesh, &partit, sweeps);
  bool external = colorPath != nullptr;
  if (external) {
    std::vector<int32_t> color(d.nv);
    FILE* f = std::fopen(colorPath, "rb");
    if (!f || std::fread(color.data(), 4, d.nv, f) != (size_t)d.nv)
      die("bad colour file");
    std::fclose(f);
    int nparts = 0;
    for (int i = 0; i < d.nv; ++i) nparts = std::max(nparts, color[i]);
    if (nparts > partit.max.part) die("too many colours for Partit(mesh, 8)");
    for (int c = 0; c < partit.max.part; ++c) partit.task.cardin[c] = 0;
    for (int i = 0; i < d.nv; ++i)
      if (color[i] > 0)
        partit.task.subset[color[i] - 1][partit.task.cardin[color[i] - 1]++] = i;
    partit.nb.parts = nparts;
  }
  std::vector<int32_t> commits(sweeps), tasks(sweeps);
#pragma omp parallel
  {
    smooth.preProcess();
    if (!external) partit.extractColoring(&mesh);
#pragma omp barrier
    smooth.cacheQuality();
#pragma omp barrier
#pragma omp master
    writeBin(out + "/qualit0.bin", mesh.geom.qualit.data(), sizeof(double) * d.nt);
#pragma omp barrier
    for (int level = 0; level < sweeps; ++level) {
      smooth.movePoints();
#pragma omp master
      {
        commits[level] = smooth.nb.commit;
        tasks[level] = smooth.nb.tasks;
      }
#pragma omp barrier
    }
  }
  std::vector<int32_t> color(d.nv, 0);
  for (int c = 0; c < partit.nb.parts; ++c)
    for (int k = 0; k < partit.task.cardin[c]; ++k)
      color[partit.task.subset[c][k]] = c + 1;
  writeBin(out + "/color.bin", color.data(), sizeof(int32_t) * d.nv);
  writeBin(out + "/xy.bin", mesh.geom.points.data(), sizeof(double) * 2 * d.nv);
  writeBin(out + "/tensor.bin", mesh.geom.tensor.data(), sizeof(double) * 3 * d.nv);
  writeBin(out + "/qualit.bin", mesh.geom.qualit.data(), sizeof(double) * d.nt);
  writeBin(out + "/commits.bin", commits.data(), sizeof(int32_t) * sweeps);
  writeBin(out + "/tasks.bin", tasks.data(), sizeof(int32_t) * sweeps);
  return 0;
}

/* Times the public API exactly as src/main.cpp:40-56 calls it: a fresh Mesh
 * per repetition (Metrics is single-use and Smooth mutates the mesh), then
 * Metrics::run and Smooth::run, each bracketed by a steady clock. */
static int benchMode(int argc, char** argv) {
  if (argc < 10) die("bench needs MESH JSON tf p hmin hmax sweeps reps");
  MeshData d = readMesh(argv[2]);
  const char* json = argv[3];
  double tf = std::atof(argv[4]);
  int p = std::atoi(argv[5]);
  double hmin = std::atof(argv[6]);
  double hmax = std::atof(argv[7]);
  int sweeps = std::atoi(argv[8]);
  int reps = std::atoi(argv[9]);
  std::vector<double> tm, ts;
  int commit = 0;
  for (int r = 0; r < reps; ++r) {
    int size[2] = {d.nv, d.nt};
    Mesh mesh(size, 64, 2, 0, 1);
    fillMesh(mesh, d);
    Metrics metric(&mesh, tf, p, hmin, hmax);
    Partit partit(&mesh, 8);
    Smooth smooth(&mesh, &partit, sweeps);
    Stats stat[2];
    auto t0 = std::chrono::steady_clock::now();
    metric.run(stat);
    auto t1 = std::chrono::steady_clock::now();
    smooth.run(stat + 1);
    auto t2 = std::chrono::steady_clock::now();
    tm.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    ts.push_back(std::chrono::duration<double, std::milli>(t2 - t1).count());
    commit = stat[1].task;
  }
  std::sort(tm.begin(), tm.end());
  std::sort(ts.begin(), ts.end());
  FILE* f = std::fopen(json, "w");
  if (!f) die("cannot write json");
  std::fprintf(f,
               "{\"nv\": %d, \"nt\": %d, \"threads\": %d, \"reps\": %d, "
               "\"metric_ms\": %.3f, \"smooth_ms\": %.3f, \"commits\": %d}\n",
               d.nv, d.nt, omp_get_max_threads(), reps, tm[tm.size() / 2],
               ts[ts.size() / 2], commit);
  std::fclose(f);
  return 0;
}

// Mesh::load (src/io.cpp:53-161) of a .mesh/.bb pair, as src/main.cpp
// sizes it (io::find of the two counts), dumping what it parsed.
static int loadMode(int argc, char** argv) {
  if (argc < 5) die("usage: ref_harness load MEDIT BB OUTDIR [JSON]");
  std::ifstream file(argv[2], std::ios::in);
  int size[] = {io::find("Vertices", file), io::find("Triangles", file)};
  file.close();
  if (size[0] <= 0 || size[1] <= 0) die("no Vertices or Triangles");
  Mesh mesh(size, 64, 2, 0, 1);  // depth 2: node/element capacities of 3x
  const auto t0 = std::chrono::steady_clock::now();
  mesh.load(argv[2], argv[3]);
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  const std::string out = argv[4];
  const int nv = mesh.nb.nodes, nt = mesh.nb.elems;
  writeBin(out + "/xy.bin", mesh.geom.points.data(), sizeof(double) * 2 * nv);
  writeBin(out + "/tri.bin", mesh.topo.elems.data(), sizeof(int) * 3 * nt);
  writeBin(out + "/tags.bin", mesh.sync.tags, (size_t)nv);
  writeBin(out + "/solut.bin", mesh.geom.solut.data(), sizeof(double) * nv);
  if (argc > 5) {
    FILE* f = std::fopen(argv[5], "w");
    if (!f) die("cannot write json");
    std::fprintf(f, "{\"nv\": %d, \"nt\": %d, \"threads\": %d, \"load_ms\": %.3f}\n", nv, nt,
                 omp_get_max_threads(), ms);
    std::fclose(f);
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) die("usage: ref_harness dump|bench|load ...");
  if (!std::strcmp(argv[1], "load")) return loadMode(argc, argv);
  if (!std::strcmp(argv[1], "dump")) return dumpMode(argc, argv);
  if (!std::strcmp(argv[1], "bench")) return benchMode(argc, argv);
  die("unknown mode");
  return 2;
}
