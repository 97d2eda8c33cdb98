"""GPU parity tests of the HIP path, called through the C ABI, against the
reference's golden fixtures and the oracle on the same seeded inputs.

Tolerances (FP64):
  gradient, raw Hessian                    bit-exact (same operation order,
                                           no contraction, IEEE div/sqrt)
  locally normalised metric                bit-exact (the device pow is
                                           glibc's, trn_libm.hpp)
  complexity                               1e-12 relative (tree-ordered sum)
  final metric                             1e-10 per-vertex norm-relative
  element quality                          1e-10 relative
  smoothed coordinates (same colouring)    1e-12 per-vertex norm-relative,
                                           commit and task counts exact
"""
import numpy as np
import pytest

import oracle
from conftest import golden_cases, load_golden, normrel
from trinity_amd import meshgen as mg

pytestmark = pytest.mark.gpu

CASES = golden_cases()


def T():
    import trinity_amd
    return trinity_amd


def dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("wl", ["uniform16:tanh", "jitter33:fronts", "jitter64:layer"])
def test_topology_matches_rebuild_topology(ctx, wl):
    m = mg.workload(wl)
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    got = dm.csr()
    topo = oracle.Topology(m.nv, m.tri)
    for k in ("v2t_off", "v2t", "v2v_off", "v2v"):
        ref = getattr(topo, k)
        assert np.array_equal(got[k], ref[:got[k].size]), k
    assert dm.max_patch == max(len(topo.patch(i)) for i in range(m.nv))


def _metric_on_gpu(ctx, g):
    import torch
    tf, p, hmin, hmax, sweeps = g["params"]
    nv = g["xy_in"].shape[0]
    dm = T().DeviceMesh(ctx, nv, g["tri"], g["tags"])
    xy = dev(g["xy_in"])
    f = dev(g["solut"])
    grad = torch.empty((nv, 2), dtype=torch.float64, device="cuda")
    hess = torch.empty((nv, 3), dtype=torch.float64, device="cuda")
    ten, c = T().metric_run(ctx, dm, xy, f, tf, int(p), hmin, hmax, gradient=grad, hessian=hess)
    mloc = T().normalize_local(ctx, hess.clone(), int(p))
    return dm, xy, {"gradient": grad.cpu().numpy(), "hessian": hess.cpu().numpy(),
                    "metric_local": mloc.cpu().numpy(), "metric": ten.cpu().numpy(),
                    "complexity": c}


@pytest.mark.parametrize("name", CASES)
def test_metric_matches_reference_golden(ctx, name):
    g = load_golden(name)
    dm, xy, r = _metric_on_gpu(ctx, g)
    assert abs(r["complexity"] - g["complexity"][0]) <= 1e-12 * abs(g["complexity"][0])
    if "sample_v" in g:
        iv = g["sample_v"]
        assert bits_equal(r["gradient"][iv], g["s_gradient"])
        assert bits_equal(r["hessian"][iv], g["s_hessian"])
        assert bits_equal(r["metric_local"][iv], g["s_metric_local"])
        assert normrel(r["metric"][iv], g["s_metric"]).max() <= 1e-10
    else:
        assert bits_equal(r["gradient"], g["gradient"])
        assert bits_equal(r["hessian"], g["hessian"])
        assert bits_equal(r["metric_local"], g["metric_local"])
        assert normrel(r["metric"], g["metric"]).max() <= 1e-10


def test_stage_entry_points_compose_to_metric_run(ctx):
    import torch
    m = mg.workload("jitter48:fronts")
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    xy, f = dev(m.xy), dev(m.solut)
    full, c = T().metric_run(ctx, dm, xy, f, 0.8, 2, 1e-9, 0.4)
    g = T().gradient(ctx, dm, xy, f)
    h = T().hessian(ctx, dm, xy, g)
    T().normalize_local(ctx, h, 2)
    c2 = T().complexity(ctx, dm, xy, h)
    assert c2 == c
    T().normalize_global(ctx, h, int(np.floor(m.nv * 0.8)) / c2, 1e-9, 0.4)
    assert torch.equal(h, full)


def test_metric_full_mesh_matches_oracle(ctx):
    m = mg.workload("jitter300:layer")
    st = oracle.metric_stages(m.xy, m.tri, m.solut, 0.8, 4, 1e-9, 0.4)
    g = {"params": np.array([0.8, 4, 1e-9, 0.4, 1]), "xy_in": m.xy, "tri": m.tri,
         "tags": m.tags, "solut": m.solut}
    dm, xy, r = _metric_on_gpu(ctx, g)
    assert bits_equal(r["gradient"], st["gradient"])
    assert bits_equal(r["hessian"], st["hessian"])
    assert abs(r["complexity"] - st["complexity"]) <= 1e-12 * st["complexity"]
    assert normrel(r["metric"], st["metric"]).max() <= 1e-10


@pytest.mark.parametrize("wl", ["jitter300:layer", "hub160:fronts"])
def test_metric_with_zcurve_slabs_is_bit_identical(ctx, wl):
    """A mesh created with coordinates stores its 2-ring patches in slabs of
    Z-curve-consecutive vertices instead of index-consecutive ones: the
    gathers then run over other slabs and unions, and every stage output is
    bit-identical to the index-order slabs'."""
    import torch
    m = mg.workload(wl)
    xy, f = dev(m.xy), dev(m.solut)
    out = []
    for coords in (None, m.xy):
        dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags, xy=coords)
        grad = torch.empty((m.nv, 2), dtype=torch.float64, device="cuda")
        hess = torch.empty((m.nv, 3), dtype=torch.float64, device="cuda")
        ten, c = T().metric_run(ctx, dm, xy, f, 0.8, 4, 1e-9, 0.4, gradient=grad, hessian=hess)
        out.append((grad.cpu(), hess.cpu(), ten.cpu(), c))
        dm.close()
    (g0, h0, t0, c0), (g1, h1, t1, c1) = out
    assert torch.equal(g0, g1) and torch.equal(h0, h1) and torch.equal(t0, t1) and c0 == c1


def _wheel(n=60):
    """A hub joined to a rim of n vertices, and an outer ring of n boundary
    vertices joined to the rim: the hub has n incident triangles (generic
    move path), every rim vertex's 2-ring holds all n fan triangles and
    more (a slab too large for an LDS union), the outer slab is small."""
    ang = 2 * np.pi * np.arange(n) / n
    rim = np.stack([0.5 + 0.25 * np.cos(ang), 0.5 + 0.25 * np.sin(ang)], 1)
    out = np.stack([0.5 + 0.45 * np.cos(ang + 0.01), 0.5 + 0.45 * np.sin(ang + 0.01)], 1)
    xy = np.vstack([[0.5 + 0.013, 0.5 - 0.007], rim, out])  # off-centre hub
    r = lambda k: 1 + k % n
    o = lambda k: 1 + n + k % n
    tri = []
    for k in range(n):
        tri += [[0, r(k), r(k + 1)], [r(k), o(k), o(k + 1)], [r(k), o(k + 1), r(k + 1)]]
    tri = np.array(tri, dtype=np.int32)
    tags = np.zeros(2 * n + 1, dtype=np.uint8)
    tags[1 + n:] = mg.TAG_BOUND
    solut = np.tanh(8 * (xy[:, 0] - 0.45)) + xy[:, 1] ** 2
    return xy, tri, tags, solut


def _renumbered(wl, seed=3):
    """A workload with its vertices and triangles in random order, so no
    64-vertex slab or vertex block is spatially compact."""
    m = mg.workload(wl)
    rng = np.random.default_rng(seed)
    pv = rng.permutation(m.nv)            # old -> new vertex id
    pt = rng.permutation(m.nt)            # new triangle k is old triangle pt[k]
    xy = np.empty_like(m.xy)
    xy[pv] = m.xy
    tags = np.empty_like(m.tags)
    tags[pv] = m.tags
    solut = np.empty_like(m.solut)
    solut[pv] = m.solut
    tri = pv[m.tri[pt]].astype(np.int32)
    return xy, tri, tags, solut


@pytest.mark.parametrize("case", ["wheel60", "renumbered:jitter80:fronts+quad",
                                  "renumbered:flip64:fronts"])
def test_irregular_meshes_match_oracle(ctx, case):
    """Metric (gradient and Hessian bit-exact) and smoothing (same colouring,
    counters exact) on a hub of degree 60 (generic move path, patches too
    large for a slab union) and on randomly renumbered grids (scattered
    slabs, records and live lists)."""
    import torch
    if case == "wheel60":
        xy, tri, tags, solut = _wheel(60)
    else:
        xy, tri, tags, solut = _renumbered(case.split(":", 1)[1])
    st = oracle.metric_stages(xy, tri, solut, 0.8, 2, 1e-9, 0.4)
    g = {"params": np.array([0.8, 2, 1e-9, 0.4, 1]), "xy_in": xy, "tri": tri, "tags": tags,
         "solut": solut}
    dm, _, r = _metric_on_gpu(ctx, g)
    assert bits_equal(r["gradient"], st["gradient"])
    assert bits_equal(r["hessian"], st["hessian"])
    assert normrel(r["metric"], st["metric"]).max() <= 1e-10
    col = T().color(ctx, dm, seed=2, max_colors=64)
    cnp = col.color.cpu().numpy()
    sweeps = 4
    xo, to, qo, co, tao = oracle.smooth_run(xy, tri, tags, st["metric"], cnp, col.ncolors, sweeps,
                                            topo=st["topo"])
    xg, tg, qg, com, tas = _smooth_gpu(ctx, dm, xy, st["metric"], cnp, col.ncolors, sweeps)
    assert tas == tao.tolist()
    assert com == co.tolist()
    assert normrel(xg, xo).max() <= 1e-12
    assert normrel(tg, to).max() <= 1e-10
    if case == "wheel60":
        assert tas[0] == 61 and com[0] >= 1  # hub and rim are tasks; some moved


@pytest.mark.parametrize("wl", ["uniform40:tanh", "jitter100:fronts"])
def test_colouring_valid_deterministic_seeded(ctx, wl):
    m = mg.workload(wl)
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    topo = oracle.Topology(m.nv, m.tri)
    a = T().color(ctx, dm, seed=3, max_colors=7)
    b = T().color(ctx, dm, seed=3, max_colors=7)
    T().set_color_tile(dm, 256)  # more seams, so the seeded repair has work
    c = T().color(ctx, dm, seed=4, max_colors=7)
    d = T().color(ctx, dm, seed=4, max_colors=7)
    ca, cb = a.color.cpu().numpy(), b.color.cpu().numpy()
    assert np.array_equal(ca, cb)
    assert np.array_equal(c.color.cpu().numpy(), d.color.cpu().numpy())
    for col in (a, c):
        assert T().color_check(ctx, dm, col.color, col.ncolors) == (0, 0)
        assert oracle.color_valid(topo, col.color.cpu().numpy(), col.ncolors)
        assert 1 <= col.ncolors <= 7
        # classes: colour c lists its vertices in ascending order (no
        # coordinates were given, so no locality order)
        order = col.order.cpu().numpy()
        cv = col.color.cpu().numpy()
        off = col.class_off
        assert off[0] == 0 and off[-1] == m.nv
        for k in range(col.ncolors):
            seg = order[off[k]:off[k + 1]]
            assert (cv[seg] == k + 1).all() and (np.diff(seg) > 0).all()


@pytest.mark.parametrize("wl", ["jitter128:fronts", "jitter300:layer", "flip200:fronts",
                                "renumbered:jitter120:fronts"])
def test_colouring_tiles_few_colours(ctx, wl):
    """The tile-sequential first fit (index order inside tiles, seams
    repaired) colours row-numbered grids like the reference's one-thread
    colouring inside each tile (4 colours on a triangulated grid; the seams
    between tiles allow up to 6), and stays
    valid on randomly flipped diagonals and on randomly numbered meshes
    (where almost every edge is a seam)."""
    if wl.startswith("renumbered:"):
        xy, tri, tags, _ = _renumbered(wl.split(":", 1)[1])
        nv = xy.shape[0]
    else:
        m = mg.workload(wl)
        nv, tri, tags = m.nv, m.tri, m.tags
    dm = T().DeviceMesh(ctx, nv, tri, tags)
    topo = oracle.Topology(nv, tri)
    for tile in (0, 1024, 4096):
        T().set_color_tile(dm, tile)
        col = T().color(ctx, dm, seed=1, max_colors=64)
        assert T().color_check(ctx, dm, col.color, col.ncolors) == (0, 0)
        assert oracle.color_valid(topo, col.color.cpu().numpy(), col.ncolors)
        if wl.startswith("jitter"):
            assert col.ncolors <= 6, (tile, col.ncolors)
    dm.close()


def _tile_first_fit(nv, off, adj, tile, max_colors):
    """The tile-sequential first fit in plain Python: even tiles first, each
    vertex in index order taking the smallest colour absent among its
    earlier neighbours in its tile; then odd tiles, which also avoid their
    neighbours in even tiles.  -1 marks a vertex past the palette, 0 an
    isolated one."""
    col = np.zeros(nv, dtype=np.int64)
    ntiles = (nv + tile - 1) // tile
    for parity in (0, 1):
        for tb in range(parity, ntiles, 2):
            t0, t1 = tb * tile, min(tb * tile + tile, nv)
            for v in range(t0, t1):
                nb = adj[off[v]:off[v + 1]]
                if nb.size == 0:
                    continue
                used = set()
                for u in nb:
                    if t0 <= u < v or (parity and not t0 <= u < t1 and (u // tile) % 2 == 0):
                        if col[u] > 0:
                            used.add(int(col[u]))
                c = 1
                while c in used:
                    c += 1
                col[v] = c if c <= max_colors else -1
    return col


@pytest.mark.parametrize("wl", ["jitter40:fronts", "jitter150:fronts", "flip120:fronts",
                                "hub96:fronts"])
def test_colouring_tiles_equal_sequential_model(ctx, wl):
    """With tiles longer than every edge of the numbering (no seams), the
    colouring is exactly the tile-sequential first fit: pinned against a
    plain sequential restatement, for rows shorter than a 64-vertex chunk
    (chunks with several earlier in-chunk neighbours: the scalar walk) and
    longer (one earlier in-chunk neighbour: the pointer-jumping chains),
    on grids, flipped diagonals and degree-12 hubs, with a palette that
    fits and one that does not (partial colouring)."""
    m = mg.workload(wl)
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    topo = oracle.Topology(m.nv, m.tri)
    span = max(int(np.abs(topo.v2v[topo.v2v_off[v]:topo.v2v_off[v + 1]] - v).max(initial=0))
               for v in range(m.nv))
    for tile in (0, (span // 64 + 1) * 64, (span // 64 + 3) * 64):
        T().set_color_tile(dm, tile)
        t = tile if tile else max(1024, (span // 64 + 1) * 64)
        want = _tile_first_fit(m.nv, topo.v2v_off, topo.v2v, t, 64)
        col = T().color(ctx, dm, seed=1, max_colors=64)
        assert np.array_equal(col.color.cpu().numpy(), want), (wl, tile)
        want3 = _tile_first_fit(m.nv, topo.v2v_off, topo.v2v, t, 3)
        part, unc = T().color_partial(ctx, dm, seed=1, max_colors=3)
        got3 = part.color.cpu().numpy()
        assert np.array_equal(got3, np.where(want3 < 0, 0, want3)), (wl, tile)
        assert unc == int((want3 < 0).sum())
    dm.close()


def test_colour_partial_leaves_vertices_uncoloured(ctx):
    """With a palette too small for the 60-spoke hub, the reference leaves a
    vertex uncoloured (partition.cpp:104-110) and never smooths it
    (smoothing.cpp:170): trn_color_partial returns a proper partial
    colouring, every uncoloured vertex out of the classes and counted, and
    with a palette large enough the same call colours everything."""
    import torch
    xy, tri, tags, _ = _wheel(60)
    nv = xy.shape[0]
    dm = T().DeviceMesh(ctx, nv, tri, tags)
    col, unc = T().color_partial(ctx, dm, seed=1, max_colors=3)
    c = col.color.cpu().numpy()
    assert unc >= 1 and (c == 0).sum() == unc
    assert c.min() >= 0 and c.max() <= 3
    for a, b in ((0, 1), (1, 2), (2, 0)):
        ca, cb = c[tri[:, a]], c[tri[:, b]]
        assert not ((ca == cb) & (ca > 0)).any()
    order = col.order.cpu().numpy()
    assert col.class_off[-1] == nv - unc
    assert (c[order[:col.class_off[-1]]] > 0).all()
    with pytest.raises(T().TrinityError):
        T().color(ctx, dm, seed=1, max_colors=3)
    full, unc2 = T().color_partial(ctx, dm, seed=1, max_colors=7)
    assert unc2 == 0
    ref = T().color(ctx, dm, seed=1, max_colors=7)
    assert np.array_equal(full.color.cpu().numpy(), ref.color.cpu().numpy())
    # the sweeps leave the uncoloured vertex where it is
    metric = oracle.metric_stages(xy, tri, _wheel(60)[3], 0.8, 2)["metric"]
    xg = torch.from_numpy(xy.copy()).cuda()
    tg = torch.from_numpy(metric.copy()).cuda()
    q = torch.empty(tri.shape[0], dtype=torch.float64, device="cuda")
    T().smooth_run(ctx, dm, xg, tg, q, col, 2)
    moved = (xg.cpu().numpy() != xy).any(axis=1)
    assert not moved[c == 0].any()
    dm.close()


def test_locality_order_is_a_permutation_of_each_class(ctx):
    m = mg.workload("jitter90:fronts")
    plain = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    local = T().DeviceMesh(ctx, m.nv, m.tri, m.tags, xy=m.xy)
    with pytest.raises(T().TrinityError):
        T().set_class_order(plain, True)          # no coordinates to order by
    same = T().color(ctx, local, seed=5)          # ascending unless asked
    T().set_class_order(local, True)
    a = T().color(ctx, plain, seed=5)
    b = T().color(ctx, local, seed=5)
    assert np.array_equal(same.order.cpu().numpy(), a.order.cpu().numpy())
    assert np.array_equal(a.color.cpu().numpy(), b.color.cpu().numpy())
    assert np.array_equal(a.class_off, b.class_off)
    oa, ob = a.order.cpu().numpy(), b.order.cpu().numpy()
    assert not np.array_equal(oa, ob)
    for k in range(a.ncolors):
        s = slice(a.class_off[k], a.class_off[k + 1])
        assert np.array_equal(np.sort(ob[s]), oa[s])


def test_colouring_reports_palette_overflow(ctx):
    m = mg.workload("uniform8:tanh")
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    with pytest.raises(T().TrinityError) as e:
        T().color(ctx, dm, seed=1, max_colors=2)
    assert e.value.code == T().TRN_ERR_PALETTE


def test_colouring_overwrites_earlier_colours(ctx):
    """The first round writes every vertex's colour, so colour and order
    buffers holding an earlier call's (or any) values give the same result
    as fresh ones."""
    import ctypes
    import torch
    m = mg.workload("jitter40:fronts")
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    fresh = T().color(ctx, dm, seed=9)
    col = torch.full((m.nv,), 5, dtype=torch.int32, device="cuda")
    order = torch.full((m.nv,), -1, dtype=torch.int32, device="cuda")
    off = (ctypes.c_int32 * 80)()
    nc = ctypes.c_int32()
    rc = T().lib().trn_color(ctx.handle, dm.handle, 9, 64, ctypes.c_void_p(col.data_ptr()),
                             ctypes.c_void_p(order.data_ptr()), off, ctypes.byref(nc))
    assert rc == 0
    assert nc.value == fresh.ncolors
    assert list(off[:nc.value + 1]) == list(fresh.class_off)
    assert torch.equal(col, fresh.color) and torch.equal(order, fresh.order)


def test_colouring_leaves_isolated_vertices_out(ctx):
    """Vertices in no triangle keep colour 0 and appear in no class (the
    reference's preProcess skips them; partition.cpp colours only vertices
    with neighbours)."""
    m = mg.workload("uniform12:tanh")
    extra = 7
    nv = m.nv + extra
    tags = np.concatenate([m.tags, np.zeros(extra, dtype=m.tags.dtype)])
    dm = T().DeviceMesh(ctx, nv, m.tri, tags)
    col = T().color(ctx, dm, seed=2)
    cv = col.color.cpu().numpy()
    assert (cv[m.nv:] == 0).all() and (cv[:m.nv] >= 1).all()
    assert col.class_off[0] == 0 and col.class_off[-1] == m.nv
    listed = np.sort(col.order.cpu().numpy()[:m.nv])
    assert np.array_equal(listed, np.arange(m.nv))
    assert T().color_check(ctx, dm, col.color, col.ncolors) == (0, 0)


def test_timing_filter_brackets_only_the_named_kernel(ctx):
    import torch
    m = mg.workload("jitter40:fronts")
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    xy = dev(m.xy)
    f = dev(m.solut)
    ten = torch.empty((m.nv, 3), dtype=torch.float64, device="cuda")
    q = torch.empty(m.nt, dtype=torch.float64, device="cuda")
    T().metric_run(ctx, dm, xy, f, 0.8, 2, 1e-9, 0.4, tensor=ten, want_complexity=False)
    col = T().color(ctx, dm, seed=1)
    try:
        ctx.set_timing(True, only="smooth_phase")
        ctx.reset_timing()
        T().smooth_run(ctx, dm, xy.clone(), ten.clone(), q, col, 2)
        only = ctx.timings()
        ctx.set_timing(True)
        ctx.reset_timing()
        T().smooth_run(ctx, dm, xy.clone(), ten.clone(), q, col, 2)
        every = ctx.timings()
    finally:
        ctx.set_timing(False)
    assert set(only) == {"smooth_phase"}
    assert only["smooth_phase"]["launches"] == 2 * col.ncolors
    assert only["smooth_phase"]["bytes"] > 0
    assert {"smooth_phase", "pack_records", "quality_cache"} <= set(every)
    assert every["smooth_phase"]["launches"] == only["smooth_phase"]["launches"]


def test_quality_stats_match_mesh_compute_quality(ctx):
    """Mesh::computeQuality(double[3]) (mesh.cpp:647-681): qualities of the
    active triangles (a deleted one is skipped), folded into the caller's
    array as min, max and (initial + sum) / count."""
    import torch
    m = mg.workload("jitter64:fronts")
    tri = m.tri.copy()
    tri[17] = -1  # deleted triangle (Mesh::isActiveElem)
    active = tri[:, 0] >= 0
    dm = T().DeviceMesh(ctx, m.nv, tri, m.tags)
    metric = oracle.metric_stages(m.xy, m.tri, m.solut)["metric"]
    xy, ten = dev(m.xy), dev(metric)
    (qmin, qmax, qmean), n, qc = T().quality_stats(ctx, dm, xy, ten)
    cache = T().quality_cache(ctx, dm, xy, ten)
    assert n == int(active.sum())
    assert bits_equal(qc.cpu().numpy()[active], cache.cpu().numpy()[active])
    ref = oracle.quality_cache(m.xy, m.tri, metric)[active]
    assert abs(qmin - ref.min()) <= 1e-10 * ref.min()
    assert abs(qmax - ref.max()) <= 1e-10 * ref.max()
    assert abs(qmean - ref.mean()) <= 1e-12 * ref.mean()
    # folding into existing values
    (a, b, c), n2, _ = T().quality_stats(ctx, dm, xy, ten, quality=(0.25, 2.0, 10.0))
    assert n2 == n and a == min(0.25, qmin) and b == 2.0
    assert abs(c - (10.0 + qmean * n) / n) <= 1e-12 * c


@pytest.mark.parametrize("wl,lmax", [("jitter64:fronts", 1.414213562373095),
                                     ("jitter40:layer", 1.0)])
def test_refine_filter_matches_oracle(ctx, wl, lmax):
    """Refine::filterElems + Steiner points against the oracle (pinned by
    closed forms in test_oracle_kat.py), one triangle deleted.  Lengths,
    lists and Steiner coordinates use only IEEE arithmetic and the Riemann
    distance (exact here); the Steiner tensor goes through log/exp of
    metric eigenvalues, where the device's equal glibc's: bit-exact."""
    m = mg.workload(wl)
    tri = m.tri.copy()
    tri[11] = -1
    dm = T().DeviceMesh(ctx, m.nv, tri, m.tags)
    metric = oracle.metric_stages(m.xy, m.tri, m.solut)["metric"]
    r = T().refine_filter(ctx, dm, dev(m.xy), dev(metric), l_max=lmax)
    ref = oracle.refine_filter(m.xy, tri, metric, l_max=lmax)
    for k in ("eval", "adds", "split", "tasks"):
        assert r[k] == ref[k], k
    assert ref["split"] > 0 and ref["eval"] == len(tri) - 1
    assert np.array_equal(r["pattern"].cpu().numpy(), ref["pattern"])
    assert np.array_equal(r["edges"].cpu().numpy(), ref["edges"])
    assert np.array_equal(r["elems"].cpu().numpy(), ref["elems"])
    sxy, sten = r["steiner_xy"].cpu().numpy(), r["steiner_tensor"].cpu().numpy()
    for e, (a, b, _) in enumerate(ref["edges"]):
        pt, tn = oracle.steiner_point(m.xy[a], m.xy[b], metric[a], metric[b])
        assert np.array_equal(sxy[e], pt)
        assert np.array_equal(sten[e], tn)


def test_refine_filter_reports_capacity(ctx):
    import torch
    m = mg.grid_mesh(10)
    c = (1.5 * 10) ** 2  # every edge long: 3 n^2 + 2 n = 320 edges
    ten = torch.tensor(np.tile([c, 0.0, c], (m.nv, 1)), device="cuda")
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    with pytest.raises(T().TrinityError) as e:
        T().refine_filter(ctx, dm, dev(m.xy), ten, edge_cap=100)
    assert e.value.code == T().TRN_ERR_CAPACITY
    r = T().refine_filter(ctx, dm, dev(m.xy), ten)
    assert r["split"] == 320 and r["tasks"] == 200 and r["adds"] == 600


@pytest.mark.parametrize("wl,frac,seed", [("jitter64:fronts", 0.5, 3), ("uniform40:tanh", 1.0, 8)])
def test_indep_set_matches_oracle(ctx, wl, frac, seed):
    """Partit::extractIndepSet on the subgraph of a random vertex subset
    (shuffled): the device set equals the oracle's greedy set (pinned by
    its characterisation in test_oracle_kat.py), in input order."""
    import torch
    m = mg.workload(wl)
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    rng = np.random.default_rng(seed)
    nodes = rng.permutation(m.nv)[:int(frac * m.nv)].astype(np.int32)
    got = T().indep_set(ctx, dm, dev(nodes, torch.int32), seed=seed).cpu().numpy()
    ins = oracle.indep_set(oracle.Topology(m.nv, m.tri), nodes, seed=seed)
    assert np.array_equal(got, nodes[ins == 1])


def test_indep_set_rejects_bad_ids_and_takes_empty_input(ctx):
    import torch
    m = mg.workload("uniform8:tanh")
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    with pytest.raises(T().TrinityError) as e:
        T().indep_set(ctx, dm, dev(np.array([0, 3, m.nv + 2], np.int32), torch.int32))
    assert e.value.code == T().TRN_ERR_ARG
    assert T().indep_set(ctx, dm, torch.empty(0, dtype=torch.int32, device="cuda")).numel() == 0


def test_colour_checker_flags_conflicts(ctx):
    m = mg.workload("uniform8:tanh")
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    col = T().color(ctx, dm, seed=1)
    bad = col.color.clone()
    bad[10] = bad[11]
    assert T().color_check(ctx, dm, bad, col.ncolors)[0] >= 1


def test_invalid_triangles_are_rejected(ctx):
    m = mg.workload("uniform4:tanh")
    tri = m.tri.copy()
    tri[3, 1] = m.nv + 5
    with pytest.raises(T().TrinityError) as e:
        T().DeviceMesh(ctx, m.nv, tri, m.tags)
    assert e.value.code == T().TRN_ERR_ARG


def _smooth_gpu(ctx, dm, xy0, metric, color_np, ncolors, sweeps):
    import torch
    col = T().color_classes(ctx, dev(color_np, torch.int32), ncolors, mesh=dm)
    xy = dev(xy0)
    ten = dev(metric)
    q = torch.empty(dm.nt, dtype=torch.float64, device="cuda")
    com, tas = T().smooth_run(ctx, dm, xy, ten, q, col, sweeps)
    return xy.cpu().numpy(), ten.cpu().numpy(), q.cpu().numpy(), com, tas


@pytest.mark.parametrize("name", CASES)
def test_smoothing_matches_reference_golden(ctx, name):
    """Fed the reference's colouring and metric, the sweeps reproduce the
    reference's coordinates, qualities and accepted-move counts."""
    g = load_golden(name)
    tf, p, hmin, hmax, sweeps = g["params"]
    nv = g["xy_in"].shape[0]
    dm = T().DeviceMesh(ctx, nv, g["tri"], g["tags"])
    if "sample_v" in g:
        # metric and colouring from the bit-pinned oracle
        st = oracle.metric_stages(g["xy_in"], g["tri"], g["solut"], tf, int(p), hmin, hmax)
        metric = st["metric"]
        color, nc = oracle.color_firstfit(st["topo"], 8)
        assert np.array_equal(color[g["sample_v"]], g["s_color"])
    else:
        metric, color = g["metric"], g["color"]
        nc = int(color.max())
    xy, ten, q, com, tas = _smooth_gpu(ctx, dm, g["xy_in"], metric, color, nc, int(sweeps))
    assert com == g["commits"].tolist()
    assert tas == g["tasks"].tolist()
    if "sample_v" in g:
        iv, it = g["sample_v"], g["sample_t"]
        assert normrel(xy[iv], g["s_xy"]).max() <= 1e-12
        assert normrel(ten[iv], g["s_tensor"]).max() <= 1e-10
        assert normrel(q[it], g["s_qualit"]).max() <= 1e-10
    else:
        assert normrel(xy, g["xy"]).max() <= 1e-12
        assert normrel(ten, g["tensor"]).max() <= 1e-10
        assert normrel(q, g["qualit"]).max() <= 1e-10


def test_quality_cache_matches_oracle(ctx):
    import torch
    m = mg.workload("jitter80:fronts")
    st = oracle.metric_stages(m.xy, m.tri, m.solut)
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    q = T().quality_cache(ctx, dm, dev(m.xy), dev(st["metric"]))
    q0 = oracle.quality_cache(m.xy, m.tri, st["metric"])
    assert normrel(q.cpu().numpy(), q0).max() <= 1e-10


@pytest.mark.parametrize("wl,p,sweeps", [("jitter200:fronts", 2, 3), ("uniform256:tanh", 2, 1),
                                         ("jitter160:layer", 4, 2), ("flip160:fronts", 2, 3),
                                         ("jitter128:fronts+quad", 2, 8)])
def test_smoothing_with_gpu_colouring_matches_oracle(ctx, wl, p, sweeps):
    m = mg.workload(wl)
    st = oracle.metric_stages(m.xy, m.tri, m.solut, 0.8, p)
    # with coordinates: Z-curve patch slabs, classes visited in Z-curve
    # order (same results)
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags, xy=m.xy)
    T().set_class_order(dm, True)
    col = T().color(ctx, dm, seed=11, max_colors=8)
    cnp = col.color.cpu().numpy()
    xo, to, qo, co, tao = oracle.smooth_run(m.xy, m.tri, m.tags, st["metric"], cnp, col.ncolors,
                                            sweeps, topo=st["topo"])
    xy, ten, q, com, tas = _smooth_gpu(ctx, dm, m.xy, st["metric"], cnp, col.ncolors, sweeps)
    assert tas == tao.tolist()
    assert com == co.tolist()
    assert normrel(xy, xo).max() <= 1e-12
    assert normrel(ten, to).max() <= 1e-10


def test_prepare_then_sweeps_equals_smooth_run(ctx):
    """trn_smooth_prepare followed by trn_smooth_run_prepared is
    trn_smooth_run: identical coordinates, metric, qualities and counters."""
    import torch
    m = mg.workload("flip96:fronts+quad")
    st = oracle.metric_stages(m.xy, m.tri, m.solut, 0.8, 2)
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    col = T().color(ctx, dm, seed=5, max_colors=16)
    xa, ta = dev(m.xy), dev(st["metric"])
    qa = torch.empty(dm.nt, dtype=torch.float64, device="cuda")
    ca = T().smooth_run(ctx, dm, xa, ta, qa, col, 4)
    xb, tb = dev(m.xy), dev(st["metric"])
    qb = torch.empty(dm.nt, dtype=torch.float64, device="cuda")
    T().smooth_prepare(ctx, dm, xb, tb, qb)
    cb = T().smooth_run_prepared(ctx, dm, xb, tb, qb, col, 4)
    assert ca == cb
    for a, b in ((xa, xb), (ta, tb), (qa, qb)):
        assert np.array_equal(a.cpu().numpy().view(np.uint64), b.cpu().numpy().view(np.uint64))


def test_device_exp_pow_bit_identical_to_glibc_log_correctly_rounded(ctx):
    """exp and pow (every exponent the path uses, and pow(x, 3) of the
    quality) equal the host libm bit for bit on the device; log is correctly
    rounded, hence equal to the host's on metric eigenvalues and within one
    ulp (host misrounding, about 2e-3 of arguments) near 1."""
    import mpmath
    mpmath.mp.prec = 300
    rng = np.random.default_rng(0)
    n = 1 << 18
    yl = -1.0 / (2 * rng.integers(0, 5, n) + 2)
    yl[yl == -0.5] = 0.5                       # the p = 0 exponent +1/2
    cases = [
        (1, rng.uniform(-45, 45, n), None),     # logs of metric eigenvalues
        (1, rng.uniform(-740, 705, n), None),   # scaled and subnormal results
        (2, np.exp(rng.uniform(-90, 90, n)), yl),
        (2, np.exp(rng.uniform(-20, 20, n)), rng.uniform(-30, 30, n)),
        (3, rng.uniform(0, 1, n), None),        # (m (2 - m))^3 with m in (0, 1]
        (4, np.exp(rng.uniform(-30, 30, n)), None),
    ]
    for op, x, y in cases:
        ref = oracle.libm(op, x, y)
        out = T().debug_math(ctx, op, dev(x), dev(y) if y is not None else None).cpu().numpy()
        assert bits_equal(out, ref), (op, int((out != ref).sum()))
    eig = np.exp(rng.uniform(np.log(6.25), np.log(1e18), n))
    out = T().debug_math(ctx, 0, dev(eig)).cpu().numpy()
    assert bits_equal(out, oracle.libm(0, eig))
    ratio = 1 + rng.uniform(0, 1, n)           # approxRiemannDist's l_max / l_min
    out = T().debug_math(ctx, 0, dev(ratio)).cpu().numpy()
    ref = oracle.libm(0, ratio)
    diff = out.view(np.int64) - ref.view(np.int64)
    assert np.abs(diff).max() <= 1 and np.count_nonzero(diff) <= 3e-3 * n
    sample = np.concatenate([ratio[:2000], ratio[diff != 0][:500]])
    got = T().debug_math(ctx, 0, dev(sample)).cpu().numpy()
    want = np.array([float(mpmath.log(mpmath.mpf(float(v)))) for v in sample])
    assert bits_equal(got, want)


@pytest.mark.parametrize("wl", ["hub256:fronts", "renumbered:hub160:fronts"])
def test_colouring_degree12_palette13(ctx, wl):
    """A vertex graph of degree up to 12 (random edge flips of a jittered
    grid, as adaptation leaves it) with the palette 1..13 that Partit(mesh,
    14) gives: valid, deterministic, within the palette, also when the
    numbering is random."""
    if wl.startswith("renumbered:"):
        xy, tri, tags, _ = _renumbered(wl.split(":", 1)[1])
        nv = xy.shape[0]
    else:
        m = mg.workload(wl)
        nv, tri, tags = m.nv, m.tri, m.tags
    assert mg.vertex_degrees(tri, nv).max() == 12
    dm = T().DeviceMesh(ctx, nv, tri, tags)
    topo = oracle.Topology(nv, tri)
    for tile in (0, 256):
        T().set_color_tile(dm, tile)
        a = T().color(ctx, dm, seed=5, max_colors=13)
        b = T().color(ctx, dm, seed=5, max_colors=13)
        assert np.array_equal(a.color.cpu().numpy(), b.color.cpu().numpy())
        assert T().color_check(ctx, dm, a.color, a.ncolors) == (0, 0)
        assert oracle.color_valid(topo, a.color.cpu().numpy(), a.ncolors)
        assert 1 <= a.ncolors <= 13
        print(wl, "tile", tile, "colours", a.ncolors)
    dm.close()


@pytest.mark.parametrize("wl", ["jitter256:layer", "hub128:fronts"])
def test_launch_shapes_do_not_change_results(ctx, wl):
    """Every launch shape the autotuner chooses from (gather slabs per block,
    phase register shape) gives bit-identical metric, coordinates, qualities
    and counters; the tuner keeps valid shapes and restores the records."""
    import torch
    m = mg.workload(wl)
    xy0, f = dev(m.xy), dev(m.solut)
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags, xy=xy0)
    col = T().color(ctx, dm, seed=3, max_colors=16)

    def run():
        xy = xy0.clone()
        ten, c = T().metric_run(ctx, dm, xy, f, 0.8, 2, 1e-9, 0.4)
        q = torch.empty(m.nt, dtype=torch.float64, device="cuda")
        com, tas = T().smooth_run(ctx, dm, xy, ten, q, col, 2)
        return ten.cpu(), c, xy.cpu(), q.cpu(), com, tas

    assert T().mesh_shapes(dm) == (1, 1, 3)
    base = run()
    for shapes in ((2, 4, 2), (4, 2, 4), (1, 1, 3)):
        T().set_shapes(dm, shapes)
        assert T().mesh_shapes(dm) == shapes
        r = run()
        for a, b in zip(base, r):
            assert (torch.equal(a, b) if isinstance(a, torch.Tensor) else a == b), shapes
    tuned = T().autotune(ctx, dm, xy0, f, 2)
    sh = tuple(tuned["shapes"].values())
    assert sh == T().mesh_shapes(dm)
    assert sh[0] in (1, 2, 4) and sh[1] in (1, 2, 4) and sh[2] in (2, 3, 4)
    for k, t in tuned["ms"].items():
        assert 0 < t["tuned"] <= t["default"], (k, t)
    r = run()
    for a, b in zip(base, r):
        assert (torch.equal(a, b) if isinstance(a, torch.Tensor) else a == b)
    with pytest.raises(T().TrinityError):
        T().set_shapes(dm, (3, 1, 3))
    dm.close()
