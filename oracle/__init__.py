"""ctypes binding of the C restatement (liboracle.so) and of the compiled
reference harness (oracle/_ref/).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke``
(as the checker) and by the CPU-baseline leg of ``bench.py``.  The product
package ``trinity_amd`` never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")
REF_BENCH = os.path.join(HERE, "_ref", "ref_bench")

_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int32)
_u8 = ctypes.POINTER(ctypes.c_uint8)


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_build_v2t.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _i]
        L.orc_build_v2v.argtypes = [ctypes.c_int, _i, _i, _i, _i, _i, ctypes.c_int64]
        L.orc_build_v2v.restype = ctypes.c_int64
        L.orc_patch2.argtypes = [ctypes.c_int, _i, _i, _i, _i, _i, ctypes.c_int]
        L.orc_gradient.argtypes = [ctypes.c_int, _d, _i, _d, _i, _i, _i, _i, _d]
        L.orc_hessian.argtypes = [ctypes.c_int, _d, _i, _d, _i, _i, _i, _i, _d]
        L.orc_eigen.argtypes = [_d, _d, _d, _d]
        L.orc_interpolate_tensor.argtypes = [_d, _d, ctypes.c_int]
        L.orc_quality.argtypes = [_d, _d, _d, _d, _d, _d]
        L.orc_quality.restype = ctypes.c_double
        L.orc_riemann_dist.argtypes = [_d, _d, _d, _d]
        L.orc_riemann_dist.restype = ctypes.c_double
        L.orc_ccw.argtypes = [_d, _d, _d]
        L.orc_metric_exponent.argtypes = [ctypes.c_int]
        L.orc_metric_exponent.restype = ctypes.c_double
        L.orc_metric_target.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_normalize_local.argtypes = [ctypes.c_int, _d, ctypes.c_int]
        L.orc_complexity.argtypes = [ctypes.c_int, _d, _i, _d]
        L.orc_complexity.restype = ctypes.c_double
        L.orc_normalize_global.argtypes = [ctypes.c_int, _d, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double]
        L.orc_color_firstfit.argtypes = [ctypes.c_int, _i, _i, ctypes.c_int, _i]
        L.orc_color_valid.argtypes = [ctypes.c_int, _i, _i, _i, ctypes.c_int]
        L.orc_quality_cache.argtypes = [ctypes.c_int, _d, _i, _d, _d]
        L.orc_move_vertex.argtypes = [ctypes.c_int, _d, _d, _d, _i, _i, _i, _i, _i]
        L.orc_smooth_run.argtypes = [ctypes.c_int, ctypes.c_int, _d, _d, _d, _i, _u8,
                                     _i, _i, _i, _i, _i, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, _i, _i]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_energy.argtypes = [ctypes.c_int, ctypes.c_int, _d, _i, _d, _i, _i, _i, _i,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 _d, _d]
        L.orc_energy.restype = ctypes.c_double
        L.orc_energy_grad.argtypes = [ctypes.c_int, ctypes.c_int, _d, _i, _d, _i, _i, _i, _i,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, _d, _d]
        L.orc_energy_grad.restype = ctypes.c_double
        L.orc_libm.argtypes = [ctypes.c_int, ctypes.c_int64, _d, _d, _d]
        L.orc_steiner_point.argtypes = [_d, _d, _d, _d, _d, _d]
        L.orc_color_priority.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
        L.orc_color_priority.restype = ctypes.c_uint32
        L.orc_indep_set.argtypes = [ctypes.c_int, _i, _i, _i, ctypes.c_int, ctypes.c_uint64, _i]
        L.orc_refine_filter.argtypes = [ctypes.c_int, _d, _i, _d, _i, _i, ctypes.c_double, _i, _i,
                                        _i, ctypes.c_int, _i]
        _lib = L
    return _lib


def _p(a: np.ndarray, kind):
    return a.ctypes.data_as(kind)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


class Topology:
    """CSR adjacency with the semantics of Mesh::rebuildTopology
    (src/mesh.cpp:191-249)."""

    def __init__(self, nv: int, tri: np.ndarray):
        tri = np.ascontiguousarray(tri, dtype=np.int32)
        nt = tri.shape[0]
        self.nv, self.nt = nv, nt
        self.tri = tri
        self.v2t_off = np.zeros(nv + 1, dtype=np.int32)
        self.v2t = np.zeros(max(3 * nt, 1), dtype=np.int32)
        lib().orc_build_v2t(nv, nt, _p(tri, _i), _p(self.v2t_off, _i), _p(self.v2t, _i))
        cap = max(6 * nt, 1)
        self.v2v_off = np.zeros(nv + 1, dtype=np.int32)
        v2v = np.zeros(cap, dtype=np.int32)
        nnz = lib().orc_build_v2v(nv, _p(tri, _i), _p(self.v2t_off, _i), _p(self.v2t, _i),
                                  _p(self.v2v_off, _i), _p(v2v, _i), cap)
        if nnz < 0:
            raise RuntimeError("v2v capacity exceeded")
        self.v2v = v2v[:nnz].copy()
        if self.v2v.size == 0:
            self.v2v = np.zeros(1, dtype=np.int32)

    def args(self):
        return (_p(self.v2t_off, _i), _p(self.v2t, _i), _p(self.v2v_off, _i), _p(self.v2v, _i))

    def patch(self, i: int) -> np.ndarray:
        out = np.zeros(4096, dtype=np.int32)
        k = lib().orc_patch2(i, *self.args(), _p(out, _i), 4096)
        if k < 0:
            raise RuntimeError("patch too large")
        return out[:k].copy()


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def metric_stages(xy, tri, solut, tf=0.8, p=2, hmin=1e-9, hmax=0.4, topo=None) -> dict:
    """Metrics::run stage by stage (src/metric.cpp:59-70)."""
    L = lib()
    xy = f64(xy).reshape(-1, 2)
    nv = xy.shape[0]
    tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
    nt = tri.shape[0]
    solut = f64(solut)
    topo = topo or Topology(nv, tri)
    grad = np.zeros((nv, 2))
    L.orc_gradient(nv, _p(xy, _d), _p(tri, _i), _p(solut, _d), *topo.args(), _p(grad, _d))
    hess = np.zeros((nv, 3))
    L.orc_hessian(nv, _p(xy, _d), _p(tri, _i), _p(grad, _d), *topo.args(), _p(hess, _d))
    mloc = hess.copy()
    L.orc_normalize_local(nv, _p(mloc, _d), int(p))
    cplx = L.orc_complexity(nt, _p(xy, _d), _p(tri, _i), _p(mloc, _d))
    target = L.orc_metric_target(nv, float(tf))
    metric = mloc.copy()
    L.orc_normalize_global(nv, _p(metric, _d), target / cplx, float(hmin), float(hmax))
    return {"gradient": grad, "hessian": hess, "metric_local": mloc,
            "complexity": cplx, "metric": metric, "target": target, "topo": topo}


def energy(xy, tri, f, tf=0.8, p=2, hmin=1e-9, hmax=0.4, topo=None, want_fields=False):
    """E = sum_t (1 - q_M(t))^2 with M built from (xy, f) as Metrics::run
    does (trinity_adjoint.c).  With want_fields, also (metric, qualities)."""
    xy = f64(xy).reshape(-1, 2)
    nv = xy.shape[0]
    tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
    topo = topo or Topology(nv, tri)
    ten = np.zeros((nv, 3)) if want_fields else None
    q = np.zeros(tri.shape[0]) if want_fields else None
    E = lib().orc_energy(nv, tri.shape[0], _p(xy, _d), _p(tri, _i), _p(f64(f), _d), *topo.args(),
                         float(tf), int(p), float(hmin), float(hmax),
                         _p(ten, _d) if want_fields else None, _p(q, _d) if want_fields else None)
    return (E, ten, q) if want_fields else E


def energy_grad(xy, tri, f, tf=0.8, p=2, hmin=1e-9, hmax=0.4, topo=None):
    """(E, dE/dxy [nv,2], dE/df [nv]) -- the FP64 reverse-mode gradient."""
    xy = f64(xy).reshape(-1, 2)
    nv = xy.shape[0]
    tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
    topo = topo or Topology(nv, tri)
    dxy = np.zeros((nv, 2))
    df = np.zeros(nv)
    E = lib().orc_energy_grad(nv, tri.shape[0], _p(xy, _d), _p(tri, _i), _p(f64(f), _d),
                              *topo.args(), float(tf), int(p), float(hmin), float(hmax),
                              _p(dxy, _d), _p(df, _d))
    return E, dxy, df


def eigen(m) -> tuple[np.ndarray, np.ndarray]:
    m = f64(m)
    val = np.zeros(2)
    vec = np.zeros(4)
    lib().orc_eigen(_p(m, _d), _p(val, _d), _p(vec, _d), vec[2:].ctypes.data_as(_d))
    return val, vec


def interpolate_tensor(ms) -> np.ndarray:
    ms = f64(ms).reshape(-1)
    out = np.zeros(3)
    lib().orc_interpolate_tensor(_p(ms, _d), _p(out, _d), ms.size // 3)
    return out


def quality(pa, pb, pc, ta, tb, tc) -> float:
    a = [f64(v) for v in (pa, pb, pc, ta, tb, tc)]
    return lib().orc_quality(*[_p(v, _d) for v in a])


def riemann_dist(pa, pb, ta, tb) -> float:
    a = [f64(v) for v in (pa, pb, ta, tb)]
    return lib().orc_riemann_dist(*[_p(v, _d) for v in a])


def libm(op: int, x, y=None) -> np.ndarray:
    """glibc log/exp/pow/pow3/sqrt (op 0..4) of x (and y for pow)."""
    x = f64(x)
    y = f64(y) if y is not None else x
    out = np.zeros_like(x)
    lib().orc_libm(int(op), x.size, _p(x, _d), _p(y, _d), _p(out, _d))
    return out


def color_firstfit(topo: Topology, max_part: int = 8) -> tuple[np.ndarray, int]:
    color = np.zeros(topo.nv, dtype=np.int32)
    nc = lib().orc_color_firstfit(topo.nv, _p(topo.v2v_off, _i), _p(topo.v2v, _i),
                                  int(max_part), _p(color, _i))
    return color, nc


def color_valid(topo: Topology, color: np.ndarray, ncolors: int) -> bool:
    color = np.ascontiguousarray(color, dtype=np.int32)
    return bool(lib().orc_color_valid(topo.nv, _p(topo.v2v_off, _i), _p(topo.v2v, _i),
                                      _p(color, _i), int(ncolors)))


def quality_cache(xy, tri, tensor) -> np.ndarray:
    xy = f64(xy)
    tri = np.ascontiguousarray(tri, dtype=np.int32)
    tensor = f64(tensor)
    q = np.zeros(tri.reshape(-1, 3).shape[0])
    lib().orc_quality_cache(q.size, _p(xy, _d), _p(tri, _i), _p(tensor, _d), _p(q, _d))
    return q


def indep_set(topo, nodes, seed=1) -> np.ndarray:
    """Partit::extractIndepSet on the subgraph of `nodes` (greedy in
    decreasing seeded-priority order); returns in_set per node (0/1)."""
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    out = np.zeros(max(nodes.size, 1), dtype=np.int32)
    lib().orc_indep_set(topo.nv, _p(topo.v2v_off, _i), _p(topo.v2v, _i), _p(nodes, _i),
                        nodes.size, int(seed) & ((1 << 64) - 1), _p(out, _i))
    return out[:nodes.size].copy()


def color_priority(seed, v) -> int:
    return int(lib().orc_color_priority(int(seed) & ((1 << 64) - 1), int(v)))


def steiner_point(pa, pb, ta, tb) -> tuple[np.ndarray, np.ndarray]:
    """numeric::computeSteinerPoint (src/numeric.cpp:387-445)."""
    pa, pb, ta, tb = f64(pa), f64(pb), f64(ta), f64(tb)
    pt, tn = np.zeros(2), np.zeros(3)
    lib().orc_steiner_point(_p(pa, _d), _p(pb, _d), _p(ta, _d), _p(tb, _d), _p(pt, _d),
                            _p(tn, _d))
    return pt, tn


def refine_filter(xy, tri, tensor, topo=None, l_max=1.414213562373095) -> dict:
    """Refine::filterElems (src/refinement.cpp:216-265) run by one thread:
    per-triangle long-edge counts, the listed edges (v1, v2, opposite) in
    (triangle, edge) order, the triangles to cut and the four counters."""
    xy = f64(xy)
    tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
    tensor = f64(tensor)
    nt = tri.shape[0]
    topo = topo or Topology(xy.reshape(-1, 2).shape[0], tri)
    pattern = np.zeros(nt, dtype=np.int32)
    cap = 3 * nt
    edges = np.zeros(3 * max(cap, 1), dtype=np.int32)
    elems = np.zeros(max(nt, 1), dtype=np.int32)
    counts = np.zeros(4, dtype=np.int32)
    rc = lib().orc_refine_filter(nt, _p(xy, _d), _p(tri, _i), _p(tensor, _d),
                                 _p(topo.v2t_off, _i), _p(topo.v2t, _i), float(l_max),
                                 _p(pattern, _i), _p(edges, _i), _p(elems, _i), cap,
                                 _p(counts, _i))
    if rc < 0:
        raise RuntimeError("edge capacity exceeded")
    ev, adds, split, tasks = (int(c) for c in counts)
    return {"pattern": pattern, "edges": edges[:3 * split].reshape(-1, 3).copy(),
            "elems": elems[:tasks].copy(), "eval": ev, "adds": adds, "split": split,
            "tasks": tasks}


def smooth_run(xy, tri, tags, tensor, color, ncolors, sweeps, topo=None, qualit=None):
    """Smooth::run under a given colouring (src/smoothing.cpp:274-302).
    Returns (xy, tensor, qualit, commits, tasks) as new arrays."""
    xy = f64(xy).reshape(-1, 2).copy()
    nv = xy.shape[0]
    tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
    nt = tri.shape[0]
    tags = np.ascontiguousarray(tags, dtype=np.uint8)
    tensor = f64(tensor).reshape(-1, 3).copy()
    topo = topo or Topology(nv, tri)
    recompute = qualit is None
    q = np.zeros(nt) if qualit is None else f64(qualit).copy()
    color = np.ascontiguousarray(color, dtype=np.int32)
    commits = np.zeros(max(sweeps, 1), dtype=np.int32)
    tasks = np.zeros(max(sweeps, 1), dtype=np.int32)
    lib().orc_smooth_run(nv, nt, _p(xy, _d), _p(tensor, _d), _p(q, _d), _p(tri, _i),
                         _p(tags, _u8), *topo.args(), _p(color, _i), int(ncolors),
                         int(sweeps), int(recompute), _p(commits, _i), _p(tasks, _i))
    return xy, tensor, q, commits[:sweeps], tasks[:sweeps]


# --------------------------------------------------------------------------
# compiled reference (oracle/_ref)

def ref_available() -> bool:
    return os.path.exists(REF_HARNESS)


def _trnm_bytes(xy, tri, tags, solut) -> bytes:
    xy = f64(xy).reshape(-1, 2)
    tri = np.ascontiguousarray(tri, dtype="<i4").reshape(-1, 3)
    head = b"TRNM" + np.array([1, xy.shape[0], tri.shape[0]], dtype="<i4").tobytes()
    return (head + xy.astype("<f8").tobytes() + tri.tobytes()
            + np.ascontiguousarray(tags, dtype="u1").tobytes() + f64(solut).astype("<f8").tobytes())


def ref_dump(xy, tri, tags, solut, tf=0.8, p=2, hmin=1e-9, hmax=0.4, sweeps=1,
             color=None, threads=1) -> dict:
    """Run the compiled reference on the input and return every stage."""
    xy = f64(xy).reshape(-1, 2)
    nv = xy.shape[0]
    nt = np.asarray(tri).reshape(-1, 3).shape[0]
    with tempfile.TemporaryDirectory() as td:
        mp = os.path.join(td, "in.trnm")
        with open(mp, "wb") as fh:
            fh.write(_trnm_bytes(xy, tri, tags, solut))
        cmd = [REF_HARNESS, "dump", mp, td, repr(float(tf)), str(int(p)), repr(float(hmin)),
               repr(float(hmax)), str(int(sweeps))]
        if color is not None:
            cp = os.path.join(td, "color.bin")
            np.ascontiguousarray(color, dtype="<i4").tofile(cp)
            cmd.append(cp)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        subprocess.run(cmd, check=True, env=env, stdout=subprocess.DEVNULL)

        def rd(name, dt, shape):
            return np.fromfile(os.path.join(td, name), dtype=dt).reshape(shape)
        return {
            "gradient": rd("gradient.bin", "<f8", (nv, 2)),
            "hessian": rd("hessian.bin", "<f8", (nv, 3)),
            "metric_local": rd("metric_local.bin", "<f8", (nv, 3)),
            "complexity": float(rd("complexity.bin", "<f8", (1,))[0]),
            "metric": rd("metric.bin", "<f8", (nv, 3)),
            "qualit0": rd("qualit0.bin", "<f8", (nt,)),
            "color": rd("color.bin", "<i4", (nv,)),
            "xy": rd("xy.bin", "<f8", (nv, 2)),
            "tensor": rd("tensor.bin", "<f8", (nv, 3)),
            "qualit": rd("qualit.bin", "<f8", (nt,)),
            "commits": rd("commits.bin", "<i4", (sweeps,)),
            "tasks": rd("tasks.bin", "<i4", (sweeps,)),
        }
