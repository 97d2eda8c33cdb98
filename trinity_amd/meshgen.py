"""Synthetic meshes and nodal fields for the metric-and-smoothing path.

Meshes are the unit square split into an n x n grid of cells, each cell cut
along one diagonal into two counter-clockwise triangles
    (i,j),(i+1,j),(i+1,j+1)   and   (i,j),(i+1,j+1),(i,j+1),
vertex (i,j) numbered j*(n+1)+i.  Boundary vertices carry tag 4 (the
reference's ``mask::bound``, include/trinity/header.h:57) and the four
corners 4|8 (``mask::corner``), as ``Mesh::load`` sets them from the
``Edges``/``Corners`` sections (src/io.cpp:80-104).  Interior vertices may be
jittered by a seeded uniform offset of +-jitter*h per coordinate; the
connectivity is unchanged.

The ``.mesh``/``.bb`` writers produce the text formats that
``Mesh::load`` parses (src/io.cpp:53-161) so that the same input can be fed
to the reference CLI.  The TRNM binary format is this package's own
(header ``b"TRNM"``, int32 version 1, int32 nv, int32 nt, then f64 xy[2nv],
int32 tri[3nt] 0-based, u8 tags[nv], f64 solut[nv]).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

TAG_BOUND = 4
TAG_CORNER = 8


@dataclass
class Mesh:
    xy: np.ndarray              # (nv, 2) float64
    tri: np.ndarray             # (nt, 3) int32, 0-based, CCW
    tags: np.ndarray            # (nv,) uint8
    solut: np.ndarray | None = None   # (nv,) float64 nodal field
    meta: dict = field(default_factory=dict)

    @property
    def nv(self) -> int:
        return int(self.xy.shape[0])

    @property
    def nt(self) -> int:
        return int(self.tri.shape[0])


def grid_mesh(n: int, jitter: float = 0.0, seed: int = 1234, flip: float = 0.0) -> Mesh:
    """n x n cell grid on [0,1]^2 (V=(n+1)^2, T=2n^2), optional jitter.
    flip: fraction of cells (random by seed) split along the other diagonal,
    which gives interior vertex degrees 4..8 instead of 6 everywhere."""
    if n < 1:
        raise ValueError("n must be >= 1")
    h = 1.0 / n
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    x = (ii * h).astype(np.float64).ravel()
    y = (jj * h).astype(np.float64).ravel()
    # exact end points
    x[ii.ravel() == n] = 1.0
    y[jj.ravel() == n] = 1.0
    bnd = (ii.ravel() == 0) | (ii.ravel() == n) | (jj.ravel() == 0) | (jj.ravel() == n)
    tags = np.zeros((n + 1) * (n + 1), dtype=np.uint8)
    tags[bnd] = TAG_BOUND
    corners = [0, n, n * (n + 1), (n + 1) * (n + 1) - 1]
    tags[corners] = TAG_BOUND | TAG_CORNER
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        interior = np.flatnonzero(~bnd)
        off = rng.uniform(-jitter * h, jitter * h, size=(interior.size, 2))
        x[interior] += off[:, 0]
        y[interior] += off[:, 1]
    ci, cj = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    v00 = (cj * (n + 1) + ci).ravel()
    v10 = v00 + 1
    v01 = v00 + (n + 1)
    v11 = v01 + 1
    tri = np.empty((2 * n * n, 3), dtype=np.int32)
    tri[0::2, 0] = v00
    tri[0::2, 1] = v10
    tri[0::2, 2] = v11
    tri[1::2, 0] = v00
    tri[1::2, 1] = v11
    tri[1::2, 2] = v01
    if flip > 0.0:
        rng = np.random.default_rng(seed + 7)
        sel = np.flatnonzero(rng.random(n * n) < flip)
        tri[2 * sel, 0] = v00[sel]
        tri[2 * sel, 1] = v10[sel]
        tri[2 * sel, 2] = v01[sel]
        tri[2 * sel + 1, 0] = v10[sel]
        tri[2 * sel + 1, 1] = v11[sel]
        tri[2 * sel + 1, 2] = v01[sel]
    xy = np.stack([x, y], axis=1)
    return Mesh(xy=np.ascontiguousarray(xy), tri=tri, tags=tags,
                meta={"n": n, "jitter": jitter, "seed": seed, "flip": flip})


def _interior_edges(tri: np.ndarray, nv: int):
    """Edges shared by two triangles: (a, b, c, d, t1, t2) with t1 = (a, b, c)
    and t2 = (b, a, d) up to rotation, both counter-clockwise."""
    t = np.asarray(tri, dtype=np.int64)
    nt = t.shape[0]
    a = np.concatenate([t[:, 0], t[:, 1], t[:, 2]])
    b = np.concatenate([t[:, 1], t[:, 2], t[:, 0]])
    c = np.concatenate([t[:, 2], t[:, 0], t[:, 1]])
    tid = np.tile(np.arange(nt), 3)
    key = np.minimum(a, b) * nv + np.maximum(a, b)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    pair = np.flatnonzero(ks[1:] == ks[:-1])
    h1, h2 = order[pair], order[pair + 1]
    return a[h1], b[h1], c[h1], c[h2], tid[h1], tid[h2], np.unique(ks)


def raise_degrees(m: Mesh, max_deg: int = 12, rounds: int = 40, seed: int = 5,
                  min_area: float = 0.12) -> Mesh:
    """Random edge flips that spread the vertex degrees of a triangulation up
    to max_deg, the way anisotropic adaptation leaves them.  Each round flips
    a vertex-disjoint set of interior edges (a, b) -> (c, d), favouring flips
    onto vertices of high degree, where the quad a, d, b, c is convex, both
    new triangles keep at least min_area of the pair's mean area, (c, d) is
    not an edge yet, deg(c), deg(d) stay <= max_deg and deg(a), deg(b) >= 4.
    Positions and tags are kept; the triangles stay counter-clockwise."""
    rng = np.random.default_rng(seed)
    tri = np.array(m.tri, dtype=np.int32, copy=True)
    nv = m.nv
    x, y = m.xy[:, 0], m.xy[:, 1]

    def orient(p, q, r):
        return (x[q] - x[p]) * (y[r] - y[p]) - (y[q] - y[p]) * (x[r] - x[p])

    for _ in range(rounds):
        a, b, c, d, t1, t2, keys = _interior_edges(tri, nv)
        deg = np.bincount(np.concatenate([a, b]), minlength=nv)  # interior edges
        bnd_e = boundary_edges(tri)
        if bnd_e:
            be = np.asarray(bnd_e, dtype=np.int64).ravel()
            deg += np.bincount(be, minlength=nv)
        ok = (deg[c] < max_deg) & (deg[d] < max_deg) & (deg[a] > 4) & (deg[b] > 4)
        o1, o2 = orient(c, a, d), orient(c, d, b)
        mean = 0.5 * (orient(a, b, c) + orient(b, a, d))
        ok &= (o1 > min_area * mean) & (o2 > min_area * mean)
        new_key = np.minimum(c, d) * nv + np.maximum(c, d)
        pos = np.searchsorted(keys, new_key)
        pos = np.minimum(pos, keys.size - 1)
        ok &= keys[pos] != new_key
        cand = np.flatnonzero(ok)
        if cand.size == 0:
            break
        pr = rng.random(cand.size) + 0.05 * (deg[c[cand]] + deg[d[cand]])
        best = np.full(nv, -1.0)
        for v in (a, b, c, d):
            np.maximum.at(best, v[cand], pr)
        win = np.ones(cand.size, dtype=bool)
        for v in (a, b, c, d):
            win &= best[v[cand]] == pr
        e = cand[win]
        tri[t1[e]] = np.stack([c[e], a[e], d[e]], axis=1)
        tri[t2[e]] = np.stack([c[e], d[e], b[e]], axis=1)
    out = Mesh(xy=m.xy.copy(), tri=tri, tags=m.tags.copy(),
               solut=None if m.solut is None else m.solut.copy(), meta=dict(m.meta))
    out.meta["max_deg"] = max_deg
    return out


def vertex_degrees(tri: np.ndarray, nv: int) -> np.ndarray:
    """Number of distinct neighbours of every vertex."""
    t = np.asarray(tri, dtype=np.int64)
    a = np.concatenate([t[:, 0], t[:, 1], t[:, 2]])
    b = np.concatenate([t[:, 1], t[:, 2], t[:, 0]])
    key = np.unique(np.minimum(a, b) * nv + np.maximum(a, b))
    return np.bincount(np.concatenate([key // nv, key % nv]), minlength=nv)


def renumber(m: Mesh, perm: np.ndarray) -> Mesh:
    """The same mesh with vertex perm[k] renamed k."""
    perm = np.asarray(perm, dtype=np.int64)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    return Mesh(xy=np.ascontiguousarray(m.xy[perm]), tri=inv[m.tri].astype(np.int32),
                tags=m.tags[perm].copy(),
                solut=None if m.solut is None else np.ascontiguousarray(m.solut[perm]),
                meta=dict(m.meta))


# --------------------------------------------------------------------------
# nodal fields

def field_tanh_front(xy: np.ndarray) -> np.ndarray:
    """f = tanh(50 (y - 0.5 - 0.1 sin(2 pi x)))  (256^2 CPU case)."""
    x, y = xy[:, 0], xy[:, 1]
    return np.tanh(50.0 * (y - 0.5 - 0.1 * np.sin(2.0 * math.pi * x)))


def field_fronts(xy: np.ndarray, seed: int = 7, count: int = 3) -> tuple[np.ndarray, dict]:
    """Sum of `count` tanh fronts with random orientation theta~U[0,2pi),
    steepness s~U[20,100] and offset o~U[0.2,0.8]."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.0, 2.0 * math.pi, size=count)
    s = rng.uniform(20.0, 100.0, size=count)
    o = rng.uniform(0.2, 0.8, size=count)
    x, y = xy[:, 0], xy[:, 1]
    f = np.zeros(xy.shape[0], dtype=np.float64)
    for k in range(count):
        f += np.tanh(s[k] * (x * math.cos(theta[k]) + y * math.sin(theta[k]) - o[k]))
    return f, {"theta": theta.tolist(), "s": s.tolist(), "o": o.tolist()}


def field_boundary_layer(xy: np.ndarray) -> np.ndarray:
    """f = exp(-(y-0.3)/1e-3) + 0.1 sin(8 pi x)  (anisotropy up to 1e3)."""
    x, y = xy[:, 0], xy[:, 1]
    return np.exp(-(y - 0.3) / 1e-3) + 0.1 * np.sin(8.0 * math.pi * x)


def field_fronts_quad(xy: np.ndarray, seed: int = 7) -> tuple[np.ndarray, dict]:
    """Fronts plus the anisotropic quadratic 0.3x^2 + 0.2xy + 0.5y^2, whose
    Hessian [[0.6, 0.2], [0.2, 1.0]] keeps the recovered Hessians away from
    the eigenvalue clamp and from the diagonal shortcut of the eigen solver
    (off-diagonal products below 1e-9), where the metric is not smooth: the
    energy of this field can be checked by central differences."""
    f, info = field_fronts(xy, seed=seed)
    x, y = xy[:, 0], xy[:, 1]
    return f + 0.3 * x * x + 0.2 * x * y + 0.5 * y * y, info


def field_linear(xy: np.ndarray, a: float = 2.0, b: float = -5.0) -> np.ndarray:
    return a * xy[:, 0] + b * xy[:, 1]


def field_quadratic(xy: np.ndarray) -> np.ndarray:
    """f = x^2 + 3xy - 2y^2 (Hessian [[2,3],[3,-4]])."""
    x, y = xy[:, 0], xy[:, 1]
    return x * x + 3.0 * x * y - 2.0 * y * y


# --------------------------------------------------------------------------
# named workloads (sizes as in the benchmark description)

def workload(name: str, seed: int = 1234) -> Mesh:
    """Build one of the standard workloads:
    uniform256   : 256^2 uniform grid, tanh front           (66,049 V)
    jitter2048   : 2048^2 jittered grid, 3 random fronts    (4,198,401 V)
    jitter1024   : 1024^2 jittered grid, tanh front         (1,050,625 V)
    jitter4096   : 4096^2 jittered grid, boundary layer     (16,785,409 V)
    flipN        : N^2 jittered grid, half the cells split along the other
                   diagonal (vertex degrees 4..8), 3 random fronts
    hubN         : N^2 jittered grid after 10 rounds of random edge flips
                   (vertex degrees 4..12), 3 random fronts
    Other sizes: 'jitterN' / 'uniformN' with the field picked as above
    ('fronts' for jittered, 'tanh' for uniform) or suffixes ':tanh',
    ':fronts', ':fronts+quad', ':layer', ':linear', ':quadratic'.
    """
    base, _, fld = name.partition(":")
    if base.startswith("uniform"):
        n = int(base[len("uniform"):])
        m = grid_mesh(n, 0.0, seed)
        default = "tanh"
    elif base.startswith("jitter"):
        n = int(base[len("jitter"):])
        m = grid_mesh(n, 0.25, seed)
        default = {1024: "tanh", 4096: "layer"}.get(n, "fronts")
    elif base.startswith("flip"):
        n = int(base[len("flip"):])
        m = grid_mesh(n, 0.25, seed, flip=0.5)
        default = "fronts"
    elif base.startswith("hub"):
        n = int(base[len("hub"):])
        m = raise_degrees(grid_mesh(n, 0.25, seed), 12, rounds=10, seed=seed + 3)
        default = "fronts"
    else:
        raise ValueError(f"unknown workload {name!r}")
    fld = fld or default
    if fld == "tanh":
        m.solut = field_tanh_front(m.xy)
    elif fld == "fronts":
        m.solut, info = field_fronts(m.xy, seed=seed + 1)
        m.meta["fronts"] = info
    elif fld == "fronts+quad":
        m.solut, info = field_fronts_quad(m.xy, seed=seed + 1)
        m.meta["fronts"] = info
    elif fld == "layer":
        m.solut = field_boundary_layer(m.xy)
    elif fld == "linear":
        m.solut = field_linear(m.xy)
    elif fld == "quadratic":
        m.solut = field_quadratic(m.xy)
    else:
        raise ValueError(f"unknown field {fld!r}")
    m.meta["workload"] = name
    m.meta["field"] = fld
    return m


def batch_sizes(count: int = 512, lo: float = 2e3, hi: float = 2e5, seed: int = 99) -> list[int]:
    """Per-mesh grid sides for the uneven batch: n_i = round(exp(U[ln lo,
    ln hi])) vertices, side = round(sqrt(n_i)) cells."""
    rng = np.random.default_rng(seed)
    n = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), size=count)))
    return [max(2, int(round(math.sqrt(v)))) for v in n]


def batch_meshes(count: int = 512, seed: int = 99) -> list[Mesh]:
    out = []
    for k, side in enumerate(batch_sizes(count, seed=seed)):
        m = grid_mesh(side, 0.25, seed + 1000 + k)
        m.solut, info = field_fronts(m.xy, seed=seed + 5000 + k)
        m.meta["fronts"] = info
        out.append(m)
    return out


# --------------------------------------------------------------------------
# I/O

def write_trnm(path: str, m: Mesh) -> None:
    if m.solut is None:
        raise ValueError("mesh has no nodal field")
    with open(path, "wb") as fh:
        fh.write(b"TRNM")
        fh.write(np.array([1, m.nv, m.nt], dtype="<i4").tobytes())
        fh.write(np.ascontiguousarray(m.xy, dtype="<f8").tobytes())
        fh.write(np.ascontiguousarray(m.tri, dtype="<i4").tobytes())
        fh.write(np.ascontiguousarray(m.tags, dtype="u1").tobytes())
        fh.write(np.ascontiguousarray(m.solut, dtype="<f8").tobytes())


def read_trnm(path: str) -> Mesh:
    with open(path, "rb") as fh:
        if fh.read(4) != b"TRNM":
            raise ValueError("not a TRNM file")
        ver, nv, nt = np.frombuffer(fh.read(12), dtype="<i4")
        if ver != 1:
            raise ValueError("unsupported TRNM version")
        xy = np.frombuffer(fh.read(16 * nv), dtype="<f8").reshape(nv, 2).copy()
        tri = np.frombuffer(fh.read(12 * nt), dtype="<i4").reshape(nt, 3).copy()
        tags = np.frombuffer(fh.read(nv), dtype="u1").copy()
        solut = np.frombuffer(fh.read(8 * nv), dtype="<f8").copy()
    return Mesh(xy=xy, tri=tri, tags=tags, solut=solut)


def boundary_edges(tri: np.ndarray) -> list[tuple[int, int]]:
    """Edges that belong to exactly one triangle, in triangle orientation."""
    t = np.asarray(tri, dtype=np.int64)
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    key = np.minimum(e[:, 0], e[:, 1]) * (int(t.max()) + 1) + np.maximum(e[:, 0], e[:, 1])
    uniq, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    once = cnt[inv] == 1
    return [(int(a), int(b)) for a, b in e[once]]


def _write_rows(fh, fmt: str, rows: np.ndarray, chunk: int = 1 << 16) -> None:
    """fmt % row for every row, formatted a chunk at a time in one call."""
    for s in range(0, rows.shape[0], chunk):
        part = rows[s:s + chunk]
        fh.write((fmt * part.shape[0]) % tuple(part.reshape(-1).tolist()))


def write_medit(mesh_path: str, bb_path: str, m: Mesh) -> None:
    """Write the .mesh/.bb pair parsed by Mesh::load (src/io.cpp:53-161)."""
    with open(mesh_path, "w") as fh:
        fh.write("MeshVersionFormatted 1\nDimension\n2\nVertices\n%d\n" % m.nv)
        _write_rows(fh, "%.17g %.17g 1\n", np.asarray(m.xy, dtype=np.float64))
        # boundary edges: those of a single triangle
        edges = np.asarray(boundary_edges(m.tri), dtype=np.int64).reshape(-1, 2)
        fh.write("Edges\n%d\n" % edges.shape[0])
        _write_rows(fh, "%d %d 1\n", edges + 1)
        fh.write("Triangles\n%d\n" % m.nt)
        _write_rows(fh, "%d %d %d 1\n", np.asarray(m.tri, dtype=np.int64) + 1)
        corners = np.flatnonzero(m.tags & TAG_CORNER)
        fh.write("Corners\n%d\n" % corners.size)
        _write_rows(fh, "%d\n", corners + 1)
        fh.write("End\n")
    with open(bb_path, "w") as fh:
        fh.write("2 1 %d 2\n" % m.nv)
        _write_rows(fh, "%.17g\n", np.asarray(m.solut, dtype=np.float64))


def read_medit(mesh_path: str, bb_path: str) -> Mesh:
    """Parse a .mesh/.bb pair with the semantics of Mesh::load
    (src/io.cpp:53-161): 1-based indices, Edges -> bound tag, Corners ->
    bound|corner, .bb values from line 2."""
    with open(mesh_path) as fh:
        tokens = fh.read().split()
    pos = {}
    for k, tok in enumerate(tokens):
        if tok in ("Vertices", "Edges", "Triangles", "Corners") and tok not in pos:
            pos[tok] = k
    k = pos["Vertices"] + 1
    nv = int(tokens[k])
    vals = np.array(tokens[k + 1:k + 1 + 3 * nv], dtype=np.float64).reshape(nv, 3)
    xy = np.ascontiguousarray(vals[:, :2])
    tags = np.zeros(nv, dtype=np.uint8)
    if "Edges" in pos:
        k = pos["Edges"] + 1
        ne = int(tokens[k])
        e = np.array(tokens[k + 1:k + 1 + 3 * ne], dtype=np.int64).reshape(ne, 3)
        tags[e[:, 0] - 1] = TAG_BOUND
        tags[e[:, 1] - 1] = TAG_BOUND
    k = pos["Triangles"] + 1
    nt = int(tokens[k])
    t = np.array(tokens[k + 1:k + 1 + 4 * nt], dtype=np.int64).reshape(nt, 4)
    tri = np.ascontiguousarray((t[:, :3] - 1).astype(np.int32))
    if "Corners" in pos:
        k = pos["Corners"] + 1
        nc = int(tokens[k])
        c = np.array(tokens[k + 1:k + 1 + nc], dtype=np.int64)
        tags[c - 1] = TAG_BOUND | TAG_CORNER
    with open(bb_path) as fh:
        fh.readline()
        solut = np.array(fh.read().split()[:nv], dtype=np.float64)
    return Mesh(xy=xy, tri=tri, tags=tags, solut=solut)
