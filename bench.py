#!/usr/bin/env python3
"""Benchmark of the MI355X metric-and-smoothing path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--sweeps S] [--p P] [--no-cpu-baseline] [--no-extras]

One step is one pass of the path over one synthetic mesh already resident in
HBM: L2-projection gradient/Hessian recovery over the 2-ring patches, metric
construction (eigen clamp, L^p local normalisation, complexity, target
scaling), vertex colouring, and S coloured smoothing sweeps -- what
Metrics::run + Smooth::run do in the reference (src/metric.cpp:59-70,
src/smoothing.cpp:274-302).  The mesh adjacency (v2t/v2v, stored 2-ring
patches, vertex records) is built once before timing, as Mesh::load does
before Metrics::run; the line also reports the per-call figure with that
construction included.  Each step restores the input coordinates (a
16 B/vertex device copy) because smoothing moves them.

Default workload (the stress configuration the target is quoted on):
4096^2 jittered grid (16,785,409 vertices, 33,554,432 triangles),
boundary-layer field exp(-(y-0.3)/1e-3) + 0.1 sin(8 pi x), L^4 metric with
target factor 0.8, h_min=1e-9, h_max=0.4, GPU colouring, 5 sweeps.  The CPU
baseline is the reference's OpenMP Metrics::run + Smooth::run on the same
mesh on this host's cores.  Unless --no-extras, the JSON line also carries
(under "extras") the 2048^2 three-fronts pass at L^1, L^2 and L^4; the
batched pass over 512 uneven meshes against looping the plain pass over
them; the adjoint forward/backward on the 1024^2 grid and on the default
mesh; the drop-in's per-call sequence from host arrays; the refinement
filter with Steiner points and the coarsening's independent set; and
gradient-based against Laplacian smoothing scored by the same energy.

Inside the timed region only the dominant kernel's launches (smooth_phase)
carry HIP events, for the roofline figures; the per-kernel table and the
stage breakdown come from a few further passes with every launch bracketed.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E, published
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector, published

# Algorithmic HBM bytes per vertex (each array read / written once per stage;
# int32 indices, FP64 values, CSR lists 4 B offset + 4 B per entry, T = 2V):
BYTES_METRIC_PER_V = 120 + 136 + 64 + 48   # gradient, Hessian+local, complexity, global
BYTES_COLOR_PER_V = 32
BYTES_SWEEP_PER_V = 196


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def cpu_arithmetic_only(args, m, threads) -> dict:
    """The same sample through the oracle (the C restatement, OpenMP, -O2):
    2-ring patches merged from the CSR per vertex instead of the reference's
    std::set stencils, which are most of Metrics::run on the CPU.  Topology
    and colouring are built before timing; median of args.cpu_reps."""
    import numpy as np
    import oracle
    oracle.set_threads(threads)
    topo = oracle.Topology(m.nv, m.tri)
    col, nc = oracle.color_firstfit(topo, 8)
    tm, tsm = [], []
    for _ in range(args.cpu_reps):
        t0 = time.time()
        st = oracle.metric_stages(m.xy, m.tri, m.solut, 0.8, args.p, topo=topo)
        t1 = time.time()
        oracle.smooth_run(m.xy, m.tri, m.tags, st["metric"], col, nc, args.sweeps, topo=topo)
        t2 = time.time()
        tm.append((t1 - t0) * 1e3)
        tsm.append((t2 - t1) * 1e3)
    mm, sm = float(np.median(tm)), float(np.median(tsm))
    log(f"cpu arithmetic-only (oracle, {threads} threads): metric {mm:.1f} ms, "
        f"smooth {sm:.1f} ms")
    return {"value": round(m.nv / ((mm + sm) * 1e-3) / 1e6, 4), "unit": "Mvertices/s",
            "cores": threads, "kind": "port", "metric_ms": round(mm, 3),
            "smooth_ms": round(sm, 3),
            "note": "oracle C restatement; CSR-merged 2-ring patches instead of std::set "
                    "stencils; topology and colouring untimed"}


def reference_load(mp, bp, threads) -> dict:
    """The reference's Mesh::load (text parse + rebuildTopology) of the
    ingest files (oracle/_ref/ref_harness `load`)."""
    import oracle
    with tempfile.TemporaryDirectory() as td:
        js = os.path.join(td, "r.json")
        subprocess.run([oracle.REF_HARNESS, "load", mp, bp, td, js], check=True,
                       env=dict(os.environ, OMP_NUM_THREADS=str(threads)),
                       stdout=subprocess.DEVNULL)
        return json.load(open(js))


def cpu_baseline(args, ingest_files=None) -> dict | None:
    """The reference OpenMP path (oracle/_ref/ref_bench: the unmodified
    hobywan/trinity sources, -O3 -funroll-loops) timed on this host on a
    bounded sample of the same workload family; falls back to the oracle
    (the C restatement) when the reference build is absent."""
    import numpy as np
    import oracle
    from trinity_amd import meshgen as mg

    wl = args.cpu_workload or args.workload
    m = mg.workload(wl, seed=args.seed)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or cpu_cores()
    sample = (f"{wl} (V={m.nv:,}, the {'same mesh as the headline' if wl == args.workload else 'a smaller mesh of the family'}), "
              f"p={args.p}, {args.sweeps} sweeps, Metrics::run + Smooth::run wall time, "
              f"{'median of ' + str(args.cpu_reps) if args.cpu_reps > 1 else 'one run'}; "
              f"mesh loading untimed")
    if os.path.exists(oracle.REF_BENCH):
        with tempfile.TemporaryDirectory() as td:
            mp = os.path.join(td, "m.trnm")
            mg.write_trnm(mp, m)
            js = os.path.join(td, "r.json")
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores",
                       OMP_PROC_BIND="close")
            t0 = time.time()
            subprocess.run([oracle.REF_BENCH, "bench", mp, js, "0.8", str(args.p), "1e-9", "0.4",
                            str(args.sweeps), str(args.cpu_reps)], check=True, env=env,
                           stdout=subprocess.DEVNULL)
            r = json.load(open(js))
            log(f"cpu baseline (reference, {threads} threads): {r} in {time.time() - t0:.1f}s")
        t = (r["metric_ms"] + r["smooth_ms"]) * 1e-3
        out = {"value": round(m.nv / t / 1e6, 4), "unit": "Mvertices/s", "cores": threads,
               "kind": "reference", "sample": sample, "workload": wl, "vertices": m.nv,
               "metric_ms": r["metric_ms"], "smooth_ms": r["smooth_ms"]}
        try:
            out["arithmetic_only"] = cpu_arithmetic_only(args, m, threads)
        except Exception as e:  # the reference figure stands on its own
            log(f"arithmetic-only cpu split failed: {e!r}")
        if ingest_files is not None:
            try:
                rl = reference_load(*ingest_files, threads)
                log(f"reference Mesh::load: {rl}")
                out["reference_text_load"] = {
                    "ms": rl["load_ms"], "threads": rl["threads"],
                    "note": "Mesh::load of the ingest .mesh/.bb files of the same mesh: "
                            "iostream parse and rebuildTopology (compare "
                            "extras.ingest.text_to_device_mesh_ms)"}
            except Exception as e:
                log(f"reference load timing failed: {e!r}")
        return out
    oracle.set_threads(threads)
    ts = []
    for _ in range(args.cpu_reps):
        t0 = time.time()
        st = oracle.metric_stages(m.xy, m.tri, m.solut, 0.8, args.p)
        col, nc = oracle.color_firstfit(st["topo"], 8)
        oracle.smooth_run(m.xy, m.tri, m.tags, st["metric"], col, nc, args.sweeps, topo=st["topo"])
        ts.append(time.time() - t0)
    t = sorted(ts)[len(ts) // 2]
    return {"value": round(m.nv / t / 1e6, 4), "unit": "Mvertices/s", "cores": threads,
            "kind": "port", "sample": sample}


def free_gpu():
    import gc
    import torch
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def run_plain(T, mg, ctx, workload, p, sweeps, steps, warmup, seed, color_seed,
              locality=False, cctx=None, cstream=None, stat_steps=3,
              timed_kernel="smooth_phase", create_reps=0, autotune=False):
    """Times `steps` passes (coordinates restored, metric, colouring, sweeps)
    and returns the wall time per pass and the per-kernel statistics.

    Inside the timed region only the launches of `timed_kernel` are
    bracketed with HIP events (each event pair costs a few microseconds of
    the pass); the statistics of every kernel come from `stat_steps` further
    passes with every launch bracketed, run after the timed region.

    With cctx (a second context on the non-blocking stream cstream) the
    colouring, which reads only the topology, runs on cstream while the
    recovery and metric kernels and then the smoothing preparation (record
    packing, quality cache) run on ctx's stream; cstream first waits for all
    earlier work of ctx's stream, so the colouring never overwrites buffers
    the previous pass still reads, and the sweeps start after the colouring
    has returned (trn_color synchronises its stream)."""
    import numpy as np
    import torch
    m = mg.workload(workload, seed=seed)
    xy0 = torch.from_numpy(m.xy).cuda()
    # coordinates at creation: the 2-ring patches are stored in Z-curve slabs
    dm = T.DeviceMesh(ctx, m.nv, m.tri, m.tags, xy=xy0)
    if locality:
        T.set_class_order(dm, True)
    tuned = None
    if autotune:
        t0 = time.perf_counter()
        tuned = T.autotune(ctx, dm, xy0, torch.from_numpy(m.solut).cuda(), p)
        tuned["tuner_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        log(f"{workload}: autotuned shapes {tuned}")
    f = torch.from_numpy(m.solut).cuda()
    xy = torch.empty_like(xy0)
    tensor = torch.empty((m.nv, 3), dtype=torch.float64, device="cuda")
    qualit = torch.empty(m.nt, dtype=torch.float64, device="cuda")
    log(f"{workload}: V={m.nv:,} T={m.nt:,} max 2-ring patch {dm.max_patch}, p={p}, "
        f"{sweeps} sweeps")
    ncolors = []
    create_ms = None
    create_kernels = None
    if create_reps:
        # the topology the reference builds inside each call (2-ring patches
        # in Metrics::run, mesh.cpp:390-434; the primal graph in Smooth::run,
        # mesh.cpp:297-330): trn_mesh_create from device-resident triangles
        # and tags (CSR, stored patches and unions, star codes, records)
        tri_d = torch.from_numpy(m.tri.reshape(-1)).cuda()
        tags_d = torch.from_numpy(m.tags).cuda()
        T.DeviceMesh(ctx, m.nv, tri_d, tags_d, xy=xy0).close()
        times = []
        for _ in range(create_reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            d2 = T.DeviceMesh(ctx, m.nv, tri_d, tags_d, xy=xy0)
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
            d2.close()
        create_ms = float(np.median(times))
        # one more creation with every launch bracketed: its kernel breakdown
        ctx.set_timing(True, only=None)
        ctx.reset_timing()
        T.DeviceMesh(ctx, m.nv, tri_d, tags_d, xy=xy0).close()
        create_kernels = {n: round(v["ms"], 3) for n, v in
                          sorted(ctx.timings().items(), key=lambda kv: -kv[1]["ms"])[:10]}
        ctx.set_timing(False)
        ctx.reset_timing()
        del tri_d, tags_d
        log(f"{workload}: device mesh creation {create_ms:.3f} ms (median of {create_reps}); "
            f"kernels {create_kernels}")

    def step():
        xy.copy_(xy0)
        if cctx is not None:
            cstream.wait_stream(torch.cuda.current_stream())
        T.metric_run(ctx, dm, xy, f, 0.8, p, 1e-9, 0.4, tensor=tensor, want_complexity=False)
        if cctx is not None:
            # the record packing and the quality cache need the metric, not
            # the colouring: queue them before waiting for the colouring
            T.smooth_prepare(ctx, dm, xy, tensor, qualit, recompute_qualit=True)
        col = T.color(cctx if cctx is not None else ctx, dm, seed=color_seed, max_colors=64)
        ncolors.append(col.ncolors)
        if cctx is not None:
            T.smooth_run_prepared(ctx, dm, xy, tensor, qualit, col, sweeps, counters=False)
        else:
            T.smooth_run(ctx, dm, xy, tensor, qualit, col, sweeps, recompute_qualit=True,
                         counters=False)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ctxs = [ctx] + ([cctx] if cctx is not None else [])
    for c in ctxs:
        c.set_timing(False)
        c.reset_timing()
    ctx.set_timing(True, only=timed_kernel)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    timed_stats = ctx.timings()
    ms = (t1 - t0) * 1e3 / steps
    for c in ctxs:
        c.set_timing(True, only=None)
        c.reset_timing()
    for _ in range(stat_steps):
        step()
    torch.cuda.synchronize()
    stats = {}
    for c in ctxs:
        stats.update(c.timings())
        c.set_timing(False)
    nv, nt = m.nv, m.nt
    dm.close()
    del m, xy0, dm, f, xy, tensor, qualit
    free_gpu()
    return {"nv": nv, "nt": nt, "ms": ms, "stats": stats, "timed_stats": timed_stats,
            "colours": int(np.median(ncolors)), "steps": steps, "stat_steps": stat_steps,
            "colouring_overlapped": cctx is not None, "create_ms": create_ms,
            "create_kernels": create_kernels, "autotune": tuned}


def summarize(r, sweeps, label=None):
    """Per-kernel averages and per-stage rates from the HIP-event timings of
    run_plain's instrumented passes (every launch bracketed)."""
    steps, stats, nv = r["stat_steps"], r["stats"], r["nv"]
    kern = {}
    for name, s in stats.items():
        kern[name] = {"ms_per_step": s["ms"] / steps, "launches": s["launches"],
                      "avg_us": 1e3 * s["ms"] / max(s["launches"], 1),
                      "alg_GBps": (s["bytes"] / (s["ms"] * 1e-3) / 1e9) if s["ms"] > 0 else 0.0}
    if label:
        log(f"-- {label}: {r['ms']:.3f} ms/pass, {nv / (r['ms'] * 1e-3) / 1e6:.1f} Mvertices/s")
    for name, k in sorted(kern.items(), key=lambda kv: -kv[1]["ms_per_step"])[:12]:
        log(f"  {name:18s} {k['ms_per_step']:8.3f} ms/step  x{k['launches']:5d}  "
            f"avg {k['avg_us']:9.1f} us  alg {k['alg_GBps']:8.1f} GB/s")
    metric_names = ["tri_grad", "vtx_grad", "tri_hess", "vtx_hess_norm", "tri_cplx",
                    "cplx_final", "vtx_global"]
    metric_ms = sum(kern[n]["ms_per_step"] for n in metric_names if n in kern)
    color_ms = sum(v["ms_per_step"] for n, v in kern.items()
                   if n.startswith("color") or n.startswith("class") or n.startswith("scan")
                   or n == "reduce_max")
    smooth_ms = sum(kern[n]["ms_per_step"] for n in ("tensor_log", "pack_records",
                                                     "quality_cache", "class_live",
                                                     "smooth_phase",
                                                     "counter_reduce")
                    if n in kern)
    stages = {
        "recovery_metric": {"ms": round(metric_ms, 4),
                            "alg_GBps": round(BYTES_METRIC_PER_V * nv / (metric_ms * 1e-3) / 1e9, 1),
                            "hbm_frac": round(BYTES_METRIC_PER_V * nv / (metric_ms * 1e-3) / 1e9
                                              / HBM_PEAK_GBS, 4),
                            "bytes_per_vertex": BYTES_METRIC_PER_V},
        "colouring": {"ms": round(color_ms, 4), "colours": r["colours"],
                      "overlapped_with_metric": bool(r.get("colouring_overlapped", False))},
        "smoothing": {"ms": round(smooth_ms, 4), "sweeps": sweeps,
                      "Mvertex_moves_per_s": round(nv * sweeps / (smooth_ms * 1e-3) / 1e6, 1),
                      "alg_GBps": round(BYTES_SWEEP_PER_V * nv * sweeps / (smooth_ms * 1e-3) / 1e9, 1)},
    }
    return kern, stages


def read_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 PMC summary (profiles/traffic.txt: lines 'kernel workload
    fetch_bytes write_bytes'), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 4 and not line.startswith("#") and parts[0] == kernel \
                and parts[1] == workload:
            return float(parts[2]) + float(parts[3])
    return None


def read_valu(kernel, workload):
    """(FP64 flops issued, VALU wave instructions, active-lane fraction or
    None) per launch of `kernel` on `workload` from the committed rocprofv3
    summary (profiles/valu.txt: 'kernel workload fp64_flops valu_wave_instr
    waves [active_lane_fraction]'), or None."""
    path = os.path.join(ROOT, "profiles", "valu.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 4 and not line.startswith("#") and parts[0] == kernel \
                and parts[1] == workload:
            active = float(parts[5]) if len(parts) >= 6 else None
            return float(parts[2]), float(parts[3]), active
    return None


def run_host_boundary(T, mg, ctx, workload, p, sweeps, reps=3, seed=1234, color_seed=1):
    """The drop-in shims' per-call sequence from host arrays: triangles,
    tags, coordinates and field copied to the device, the device mesh built
    (CSR, stored patches, unions, records), the pass run on one stream, and
    coordinates and metric copied back.  Median wall time of `reps` calls
    after one untimed call."""
    import numpy as np
    import torch
    m = mg.workload(workload, seed=seed)
    tri = np.ascontiguousarray(m.tri.reshape(-1), dtype=np.int32)
    tags = np.ascontiguousarray(m.tags, dtype=np.uint8)
    xy_h = np.ascontiguousarray(m.xy, dtype=np.float64)
    f_h = np.ascontiguousarray(m.solut, dtype=np.float64)
    # results go back into existing host arrays, as the shims write the
    # reference's own (already touched) coordinate and metric arrays
    xy_back = torch.from_numpy(np.zeros_like(xy_h))
    t_back = torch.from_numpy(np.zeros((m.nv, 3), dtype=np.float64))
    times, parts = [], []
    for r in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xy = torch.from_numpy(xy_h).cuda()
        f = torch.from_numpy(f_h).cuda()
        tri_d, tags_d = torch.from_numpy(tri).cuda(), torch.from_numpy(tags).cuda()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        dm = T.DeviceMesh(ctx, m.nv, tri_d, tags_d, xy=xy)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        tensor = torch.empty((m.nv, 3), dtype=torch.float64, device="cuda")
        qualit = torch.empty(m.nt, dtype=torch.float64, device="cuda")
        T.metric_run(ctx, dm, xy, f, 0.8, p, 1e-9, 0.4, tensor=tensor, want_complexity=False)
        col = T.color(ctx, dm, seed=color_seed, max_colors=64)
        T.smooth_run(ctx, dm, xy, tensor, qualit, col, sweeps, recompute_qualit=True,
                     counters=False)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        xy_back.copy_(xy)
        t_back.copy_(tensor)
        t4 = time.perf_counter()
        dm.close()
        del xy, f, tri_d, tags_d, tensor, qualit, col
        if r:
            times.append((t4 - t0) * 1e3)
            parts.append([(t1 - t0) * 1e3, (t2 - t1) * 1e3, (t3 - t2) * 1e3, (t4 - t3) * 1e3])
    ms = float(np.median(times))
    pm = np.median(np.array(parts), axis=0)
    nv = m.nv
    del m
    free_gpu()
    log(f"host boundary {workload}: {ms:.2f} ms per call, {nv / (ms * 1e-3) / 1e6:.1f} "
        f"Mvertices/s (copy in {pm[0]:.2f}, mesh {pm[1]:.2f}, pass {pm[2]:.2f}, "
        f"copy out {pm[3]:.2f} ms)")
    return {"workload": workload, "vertices": nv, "ms_per_call": round(ms, 3),
            "Mvertices_per_s": round(nv / (ms * 1e-3) / 1e6, 2),
            "ms_copy_in": round(float(pm[0]), 3), "ms_mesh_create": round(float(pm[1]), 3),
            "ms_pass": round(float(pm[2]), 3), "ms_copy_out": round(float(pm[3]), 3),
            "includes": "host-to-device copies of triangles, tags, xy and f (pageable), device "
                        "mesh creation, the pass on one stream, device-to-host copies of xy and "
                        "the metric into existing host arrays"}


def run_refine(T, mg, ctx, workload, p, reps=5, seed=1234):
    """Refine::filterElems + Steiner points (trn_refine_filter) on the metric
    of `workload`: median wall time of `reps` calls (each synchronises once,
    for the list sizes)."""
    import numpy as np
    import torch
    m = mg.workload(workload, seed=seed)
    xy = torch.from_numpy(m.xy).cuda()
    dm = T.DeviceMesh(ctx, m.nv, m.tri, m.tags, xy=xy)
    tensor, _ = T.metric_run(ctx, dm, xy, torch.from_numpy(m.solut).cuda(), 0.8, p, 1e-9, 0.4,
                             want_complexity=False)
    r = T.refine_filter(ctx, dm, xy, tensor)
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = T.refine_filter(ctx, dm, xy, tensor)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.median(times))
    nt = m.nt
    dm.close()
    del m, xy, tensor
    free_gpu()
    log(f"refine filter {workload}: {ms:.3f} ms, {r['split']:,} edges to split, "
        f"{r['tasks']:,} triangles to cut")
    return {"workload": workload, "triangles": nt, "ms": round(ms, 4),
            "Mtriangles_per_s": round(nt / (ms * 1e-3) / 1e6, 1), "edges_to_split": r["split"],
            "triangles_to_cut": r["tasks"]}


def run_indep_set(T, mg, ctx, workload, frac=0.5, reps=5, seed=1234):
    """Partit::extractIndepSet (trn_indep_set) on the subgraph of a seeded
    `frac` of the vertices of `workload` (ascending ids, as the coarsening's
    filter lists them): median wall time of `reps` calls."""
    import numpy as np
    import torch
    m = mg.workload(workload, seed=seed)
    dm = T.DeviceMesh(ctx, m.nv, m.tri, m.tags)
    rng = np.random.default_rng(seed)
    nodes = np.flatnonzero(rng.random(m.nv) < frac).astype(np.int32)
    nodes_t = torch.from_numpy(nodes).cuda()
    s = T.indep_set(ctx, dm, nodes_t, seed=1)
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s = T.indep_set(ctx, dm, nodes_t, seed=1)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.median(times))
    n, k = int(nodes.size), int(s.numel())
    dm.close()
    del m, nodes_t, s
    free_gpu()
    log(f"independent set {workload}: {n:,} nodes -> {k:,} in {ms:.3f} ms")
    return {"workload": workload, "nodes": n, "set_size": k, "ms": round(ms, 4),
            "Mnodes_per_s": round(n / (ms * 1e-3) / 1e6, 1)}


def run_ingest(T, mg, ctx, workload, seed, td):
    """Input to a device mesh, end to end, from the text files the reference
    reads (Mesh::load, src/io.cpp:53-161) and from the binary form: file ->
    host parse (text) or direct read (binary) -> device arrays ->
    trn_mesh_create.  The files are written to td first (untimed) and read
    twice; the second, page-cached read is the one reported."""
    import numpy as np
    import torch
    m = mg.workload(workload, seed=seed)
    mp, bp, tp = (os.path.join(td, n) for n in ("m.mesh", "m.bb", "m.trnm"))
    t0 = time.perf_counter()
    mg.write_medit(mp, bp, m)
    mg.write_trnm(tp, m)
    log(f"ingest: wrote {os.path.getsize(mp) / 1e6:.0f} MB text, "
        f"{os.path.getsize(tp) / 1e6:.0f} MB binary in {time.perf_counter() - t0:.1f} s")
    nv, nt = m.nv, m.nt
    del m
    out = {"workload": workload, "vertices": nv,
           "text_bytes": os.path.getsize(mp) + os.path.getsize(bp),
           "binary_bytes": os.path.getsize(tp)}
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xy, tri, tags, f = T.read_medit(mp, bp)
        t1 = time.perf_counter()
        xy_d = torch.from_numpy(xy).cuda()
        f_d = torch.from_numpy(f).cuda()
        dm = T.DeviceMesh(ctx, nv, torch.from_numpy(tri.reshape(-1)).cuda(),
                          torch.from_numpy(tags).cuda(), xy=xy_d)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        dm.close()
        del xy, tri, tags, f, xy_d, f_d
        t3 = time.perf_counter()
        xy_d, tri_d, tags_d, f_d = T.read_trnm(tp, ctx)
        t4 = time.perf_counter()
        dm = T.DeviceMesh(ctx, nv, tri_d.reshape(-1), tags_d, xy=xy_d)
        torch.cuda.synchronize()
        t5 = time.perf_counter()
        dm.close()
        del xy_d, tri_d, tags_d, f_d
        free_gpu()
    out.update({"text_parse_ms": round((t1 - t0) * 1e3, 1),
                "text_to_device_mesh_ms": round((t2 - t0) * 1e3, 1),
                "binary_read_ms": round((t4 - t3) * 1e3, 1),
                "binary_to_device_mesh_ms": round((t5 - t3) * 1e3, 1),
                "io_threads": int(os.environ.get("TRN_IO_THREADS", "0")) or min(16, os.cpu_count() or 1),
                "note": "trn_medit_read (multi-threaded parse with Mesh::load's semantics) or "
                        "trn_trnm_read (pinned, overlapped copies), then trn_mesh_create "
                        "(CSR, 2-ring patches, records); second of two reads (page cache warm)"})
    log(f"ingest: {out}")
    return out, (mp, bp)


def run_adjoint(T, mg, ctx, workload, steps, p=2):
    """Forward (metric + qualities + energy) and backward (dE/dxy, dE/df)
    of the energy sum_t (1 - q_M(t))^2."""
    import torch
    m = mg.workload(workload)
    xy = torch.from_numpy(m.xy).cuda()
    dm = T.DeviceMesh(ctx, m.nv, m.tri, m.tags, xy=xy)
    adj = T.Adjoint(ctx, dm)
    f = torch.from_numpy(m.solut).cuda()
    gx = torch.empty_like(xy)
    gf = torch.empty_like(f)
    adj.forward(xy, f, 0.8, p)
    adj.backward(gx, gf)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        E = adj.forward(xy, f, 0.8, p)
    t1 = time.perf_counter()
    for _ in range(steps):
        adj.backward(gx, gf)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd, bwd = (t1 - t0) * 1e3 / steps, (t2 - t1) * 1e3 / steps
    log(f"adjoint {workload}: forward {fwd:.3f} ms, backward {bwd:.3f} ms, E={E:.6g}")
    finite = bool(torch.isfinite(gx).all().item() and torch.isfinite(gf).all().item())
    out = {"workload": workload, "vertices": m.nv, "triangles": m.nt, "lp_norm": p,
           "forward_ms": round(fwd, 4), "backward_ms": round(bwd, 4),
           "Mvertices_per_s_forward_backward": round(m.nv / ((fwd + bwd) * 1e-3) / 1e6, 1),
           "energy": E, "gradient_finite": finite}
    adj.close()
    dm.close()
    del adj, dm, xy, f, gx, gf
    free_gpu()
    return out


def compare_smoothing(T, mg, ctx, workload, iters):
    """Gradient-based smoothing (preconditioned descent on the energy)
    against the Laplacian sweeps, from the same mesh, scored by the same
    energy.  The sweeps keep the metric of the starting positions while the
    energy rebuilds it from the moved ones, so their energy is not monotone
    in the number of sweeps; the descent's is."""
    import numpy as np
    import torch
    m = mg.workload(workload)
    dm = T.DeviceMesh(ctx, m.nv, m.tri, m.tags, xy=torch.from_numpy(m.xy).cuda())
    adj = T.Adjoint(ctx, dm)
    f = torch.from_numpy(m.solut).cuda()
    xy = torch.from_numpy(m.xy).cuda()
    E0 = adj.forward(xy, f, 0.8, 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    en, acc = adj.smooth(xy, f, 0.8, 2, iters=iters, step=0.25 / math.sqrt(m.nv))
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    xl = torch.from_numpy(m.xy).cuda()
    ten = torch.empty((m.nv, 3), dtype=torch.float64, device="cuda")
    q = torch.empty(m.nt, dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    T.metric_run(ctx, dm, xl, f, 0.8, 2, 1e-9, 0.4, tensor=ten, want_complexity=False)
    col = T.color(ctx, dm, seed=1)
    T.smooth_run(ctx, dm, xl, ten, q, col, iters, counters=False)
    torch.cuda.synchronize()
    tl = time.perf_counter() - t0
    El = adj.forward(xl, f, 0.8, 2)
    log(f"smoothing modes {workload}: E0={E0:.6g}  gradient({acc} steps)={en[-1]:.6g} "
        f"in {tg * 1e3:.1f} ms  laplacian({iters} sweeps)={El:.6g} in {tl * 1e3:.1f} ms")
    out = {"workload": workload, "iterations": iters, "energy_initial": E0,
           "gradient": {"energy": en[-1], "accepted_steps": acc, "ms": round(tg * 1e3, 2)},
           "laplacian": {"energy": El, "sweeps": iters, "ms": round(tl * 1e3, 2)}}
    adj.close()
    dm.close()
    del adj, dm, xy, xl, f, ten, q
    free_gpu()
    return out


def run_batched(T, mg, count=512, p=2, sweeps=2, steps=5, warmup=2, loop_steps=2, seed=99,
                color_seed=1):
    """The uneven batch: `count` jittered grids with per-mesh vertex counts
    log-uniform in [2e3, 2e5], each with its own seeded fronts field.  One
    batched step = the batch colouring (trn_batch_color) + the pass replayed
    from its hipGraph (coordinates restored, metric with per-mesh complexity
    and target, quality cache, sweeps) -- every launch covering every mesh.
    One plain step = the plain pass (coordinates restored, trn_metric_run,
    trn_color, trn_smooth_run) looped over the meshes, each already resident
    as its own device mesh.  Both are whole-batch wall times."""
    import numpy as np
    import torch
    t0 = time.perf_counter()
    ms = mg.batch_meshes(count, seed=seed)
    nv = sum(m.nv for m in ms)
    nt = sum(m.nt for m in ms)
    log(f"batch: {count} meshes, {nv:,} vertices, {nt:,} triangles "
        f"(generated in {time.perf_counter() - t0:.1f} s)")
    # batched, on a context with its own stream (graph capture needs one)
    gctx = T.Context(0, use_torch_stream=False)
    gctx.use_own_stream()
    xy_host = np.concatenate([m.xy for m in ms])
    b = T.Batch(gctx, [(m.nv, m.tri, m.tags) for m in ms], xy=xy_host)
    xy_in = torch.from_numpy(xy_host).cuda()
    solut = torch.from_numpy(np.concatenate([m.solut for m in ms])).cuda()
    xy = torch.empty_like(xy_in)
    ten = torch.empty((nv, 3), dtype=torch.float64, device="cuda")
    q = torch.empty(nt, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ncol = b.color(seed=color_seed)
    b.run(xy_in, xy, solut, ten, q, 0.8, p, 1e-9, 0.4, sweeps=sweeps)
    gctx.synchronize()
    # per-kernel view of one eager pass
    gctx.set_timing(True)
    gctx.reset_timing()
    b.run(xy_in, xy, solut, ten, q, 0.8, p, 1e-9, 0.4, sweeps=sweeps)
    kstats = gctx.timings()
    gctx.set_timing(False)
    for _ in range(warmup):
        b.color(seed=color_seed)
        b.run_graph(xy_in, xy, solut, ten, q, 0.8, p, 1e-9, 0.4, sweeps=sweeps)
    gctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.color(seed=color_seed)
        b.run_graph(xy_in, xy, solut, ten, q, 0.8, p, 1e-9, 0.4, sweeps=sweeps)
    gctx.synchronize()
    bms = (time.perf_counter() - t0) * 1e3 / steps
    b.close()
    del xy_in, solut, xy, ten, q
    gctx.close()
    free_gpu()
    log(f"batched: {bms:.3f} ms per pass, {nv / (bms * 1e-3) / 1e6:.1f} Mvertices/s, "
        f"{ncol} colours")
    # plain pass looped over the meshes
    ctx = T.Context(0)
    dms, arrs = [], []
    for m in ms:
        x0 = torch.from_numpy(m.xy).cuda()
        dms.append(T.DeviceMesh(ctx, m.nv, m.tri, m.tags, xy=x0))
        arrs.append((x0, torch.from_numpy(m.solut).cuda(),
                     torch.empty((m.nv, 2), dtype=torch.float64, device="cuda"),
                     torch.empty((m.nv, 3), dtype=torch.float64, device="cuda"),
                     torch.empty(m.nt, dtype=torch.float64, device="cuda")))

    def loop():
        for dm, (x0, f, x, t, qq) in zip(dms, arrs):
            x.copy_(x0)
            T.metric_run(ctx, dm, x, f, 0.8, p, 1e-9, 0.4, tensor=t, want_complexity=False)
            col = T.color(ctx, dm, seed=color_seed, max_colors=64)
            T.smooth_run(ctx, dm, x, t, qq, col, sweeps, recompute_qualit=True, counters=False)

    loop()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(loop_steps):
        loop()
    torch.cuda.synchronize()
    pms = (time.perf_counter() - t0) * 1e3 / loop_steps
    for dm in dms:
        dm.close()
    del dms, arrs
    ctx.close()
    free_gpu()
    log(f"plain loop over {count} meshes: {pms:.3f} ms per pass, "
        f"{nv / (pms * 1e-3) / 1e6:.1f} Mvertices/s")
    bpv = BYTES_METRIC_PER_V + BYTES_COLOR_PER_V + BYTES_SWEEP_PER_V * sweeps
    kern = {n: {"ms": round(v["ms"], 4), "launches": v["launches"]}
            for n, v in sorted(kstats.items(), key=lambda kv: -kv[1]["ms"])[:10]}
    dominant = max(kstats, key=lambda n: kstats[n]["ms"])
    roof = roofline_for(dominant, kstats[dominant], f"batch{count}", kstats)
    roof["note"] = ("dominant kernel of one eager batched pass (HIP events on every launch); " +
                    roof["note"])
    return {"meshes": count, "vertices": nv, "triangles": nt, "lp_norm": p, "sweeps": sweeps,
            "colours": ncol, "roofline": roof,
            "batched": {"ms_per_pass": round(bms, 3),
                        "Mvertices_per_s": round(nv / (bms * 1e-3) / 1e6, 2),
                        "alg_GBps": round(bpv * nv / (bms * 1e-3) / 1e9, 1),
                        "hbm_frac": round(bpv * nv / (bms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "how": "trn_batch_color + trn_batch_pass_graph (hipGraph replay)"},
            "plain_loop": {"ms_per_pass": round(pms, 3),
                           "Mvertices_per_s": round(nv / (pms * 1e-3) / 1e6, 2),
                           "how": "per mesh: coordinates restored, trn_metric_run, trn_color, "
                                  "trn_smooth_run, each mesh already a device mesh"},
            "speedup_batched_vs_plain": round(pms / bms, 2),
            "eager_pass_kernels": kern,
            "bytes_per_vertex": bpv}


def roofline_for(kname, dstat, workload, stats_all):
    """Roofline of the dominant kernel from its HIP-event launch times:
    compute-bound (FP64) when the committed rocprofv3 summary holds its FP64
    work on this workload, else HBM-bound on its algorithmic bytes."""
    bytes_per_launch = dstat["bytes"] / max(dstat["launches"], 1)
    avg_s = dstat["ms"] * 1e-3 / max(dstat["launches"], 1)
    alg_gbps = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    traffic = read_traffic(kname, workload)
    valu = read_valu(kname, workload)
    hbm = {"achieved": round(alg_gbps, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(alg_gbps / HBM_PEAK_GBS, 5),
           "bytes_per_launch": round(bytes_per_launch, 1)}
    if valu and avg_s > 0:
        flops, winstr, active = valu
        useful = flops * active if active else flops
        tf = useful / avg_s / 1e12
        out = {"kernel": kname, "bound": "mfma", "compute": "fp64",
               "achieved": round(tf, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
               "frac": round(tf / FP64_PEAK_TFLOPS, 4), "traffic": traffic,
               "avg_launch_us": round(avg_s * 1e6, 2),
               "fp64_flops_per_launch": useful,
               "lane_utilisation": active,
               "issued_fp64_flops_per_launch": flops,
               "valu_issue_frac": round(winstr * 4 / (1024 * avg_s * 2.4e9), 3),
               "hbm": hbm,
               "note": ("compute roof: FP64 vector arithmetic (the kernel issues no MFMA "
                        "instruction; MI355X's FP64 vector and FP64 matrix peaks are both "
                        "78.6 TFLOP/s).  achieved = FP64 flops per launch with masked lanes "
                        "excluded (rocprofv3 F64 VALU instruction counters x 64 lanes x the "
                        "active-lane fraction, profiles/valu.txt) / HIP-event launch time; "
                        "hbm = algorithmic bytes per launch / the same time; traffic = HBM "
                        "bytes per launch from FETCH_SIZE / WRITE_SIZE (profiles/traffic.txt)")}
        return out
    return {"kernel": kname, "bound": "hbm", "achieved": hbm["achieved"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": hbm["frac"], "traffic": traffic,
            "avg_launch_us": round(avg_s * 1e6, 2),
            "note": ("achieved = algorithmic bytes per launch / HIP-event launch time (events on "
                     "this kernel's launches inside the timed region); traffic = HBM bytes per "
                     "launch from rocprofv3 FETCH_SIZE and WRITE_SIZE (profiles/traffic.txt)")}


def pass_stage(nv, ms, sweeps):
    bpv = BYTES_METRIC_PER_V + BYTES_COLOR_PER_V + BYTES_SWEEP_PER_V * sweeps
    gbps = bpv * nv / (ms * 1e-3) / 1e9
    return {"ms": round(ms, 4), "bytes_per_vertex": bpv, "alg_GBps": round(gbps, 1),
            "hbm_frac": round(gbps / HBM_PEAK_GBS, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="jitter4096:layer")
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--sweeps", type=int, default=5)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--color-seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline pass (no 2048^2 passes, batch, adjoint, ...)")
    ap.add_argument("--serial-colouring", action="store_true",
                    help="colour on the pass's stream after the metric instead of on a "
                         "second stream beside it")
    ap.add_argument("--locality-order", action="store_true",
                    help="visit colour classes along a Z-curve of the coordinates "
                         "(the patch slabs follow it in any case)")
    ap.add_argument("--cpu-workload", default=None,
                    help="mesh of the CPU baseline (default: the headline's)")
    ap.add_argument("--cpu-reps", type=int, default=1)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of tuning them on the mesh")
    ap.add_argument("--ingest", action="store_true",
                    help="also time file input to a device mesh (text vs binary) on the "
                         "workload, and the reference's Mesh::load beside it")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 512 MiB device-to-device 3 times (scales the rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE counters against known traffic)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 26
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "copies": 3}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; the multi-GPU sharding is a design "
              "(see the sharding section of the design notes) and is not measured", file=sys.stderr)
        sys.exit(2)

    import trinity_amd as T
    from trinity_amd import meshgen as mg

    import torch
    ctx = T.Context(0)
    cctx = cstream = None
    if not args.serial_colouring:
        cstream = torch.cuda.Stream()  # non-blocking, default priority
        cctx = T.Context(0, use_torch_stream=False)
        cctx.set_stream(cstream.cuda_stream)
    r = run_plain(T, mg, ctx, args.workload, args.p, args.sweeps, args.steps, args.warmup,
                  args.seed, args.color_seed, args.locality_order, cctx, cstream, create_reps=3,
                  autotune=not args.no_autotune)
    ms = r["ms"]
    value = r["nv"] / (ms * 1e-3) / 1e6
    kern, stages = summarize(r, args.sweeps, f"{args.workload} p={args.p}")
    stages["pass"] = pass_stage(r["nv"], ms, args.sweeps)
    if r["create_ms"] is not None:
        stages["per_call_with_topology"] = {
            "ms": round(ms + r["create_ms"], 4), "mesh_create_ms": round(r["create_ms"], 4),
            "Mvertices_per_s": round(r["nv"] / ((ms + r["create_ms"]) * 1e-3) / 1e6, 2),
            "mesh_create_kernels_ms": r["create_kernels"],
            "note": "the pass plus trn_mesh_create from device-resident triangles and tags "
                    "(v2t/v2v CSR, stored 2-ring patches and their unions, star codes, vertex "
                    "records), which the reference rebuilds inside Metrics::run (getVicinity) "
                    "and Smooth::run (extractPrimalGraph)"}
    stats = r["stats"]
    dname = max(kern.items(), key=lambda kv: kv[1]["ms_per_step"])[0]
    # the dominant kernel's launches carry HIP events inside the timed region
    dstat = r["timed_stats"].get(dname, stats[dname])
    roof = roofline_for(dname, dstat, args.workload, stats)

    extras = {}
    ingest_files = None
    ingest_dir = tempfile.TemporaryDirectory() if args.ingest else None
    if args.ingest:
        extras["ingest"], ingest_files = run_ingest(T, mg, ctx, args.workload, args.seed,
                                                    ingest_dir.name)
    if not args.no_extras:
        wl2 = "jitter2048:fronts"
        for p in (2, 1, 4):
            rp = run_plain(T, mg, ctx, wl2, p, 3, 10 if p == 2 else 3, 2 if p == 2 else 1,
                           args.seed, args.color_seed, False, cctx, cstream,
                           stat_steps=3 if p == 2 else 2)
            kp, st = summarize(rp, 3, f"{wl2} p={p}")
            st["pass"] = pass_stage(rp["nv"], rp["ms"], 3)
            e = {"workload": wl2, "vertices": rp["nv"], "lp_norm": p, "sweeps": 3,
                 "ms_per_pass": round(rp["ms"], 4),
                 "Mvertices_per_s": round(rp["nv"] / (rp["ms"] * 1e-3) / 1e6, 2), "stages": st}
            if p == 2:
                dn = max(kp.items(), key=lambda kv: kv[1]["ms_per_step"])[0]
                e["roofline"] = roofline_for(dn, rp["timed_stats"].get(dn, rp["stats"][dn]), wl2,
                                             rp["stats"])
            extras[f"fronts2048_p{p}"] = e
        try:
            extras["batched"] = run_batched(T, mg)
        except Exception as e:  # the headline stands on its own
            log(f"batched run failed: {e!r}")
            extras["batched"] = {"error": repr(e)}
        extras["host_boundary"] = run_host_boundary(T, mg, ctx, wl2, 2, 3, seed=args.seed,
                                                    color_seed=args.color_seed)
        extras["refine_filter"] = run_refine(T, mg, ctx, wl2, 2, seed=args.seed)
        extras["indep_set"] = run_indep_set(T, mg, ctx, wl2, seed=args.seed)
        extras["adjoint"] = run_adjoint(T, mg, ctx, "jitter1024:tanh", 5)
        extras["adjoint_headline_mesh"] = run_adjoint(T, mg, ctx, args.workload, 2, p=args.p)
        extras["smoothing_modes"] = compare_smoothing(T, mg, ctx, "jitter256:fronts+quad", 5)
        extras["smoothing_modes_20"] = compare_smoothing(T, mg, ctx, "jitter256:fronts+quad", 20)

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args, ingest_files)
        except Exception as e:  # the GPU number stands on its own
            log(f"cpu baseline failed: {e!r}")
    if ingest_dir is not None:
        ingest_dir.cleanup()

    out = {
        "metric": "Mvertices/s per recovery+metric+smooth pass",
        "value": round(value, 3),
        "unit": "Mvertices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "vertices": r["nv"], "triangles": r["nt"],
                   "lp_norm": args.p, "target_factor": 0.8, "h_min": 1e-9, "h_max": 0.4,
                   "sweeps": args.sweeps,
                   "colouring": "gpu tile-sequential first fit, seams repaired by seeded "
                                "speculative rounds" +
                   ("" if args.serial_colouring else ", on a second stream beside the metric"),
                   "patch_slabs": "z-curve", "parallelism": "single GPU"},
        "roofline": roof,
        "autotune": r["autotune"],
        "stages": stages,
        "extras": extras,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
        out["speedup_vs_cpu_note"] = ("GPU pass rate / the CPU's Metrics::run + Smooth::run "
                                      "rate on " + ("the same mesh" if cpu.get("workload") ==
                                                    args.workload else cpu.get("workload", "?")))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
