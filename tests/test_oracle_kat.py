"""Known-answer and invariant tests that pin the oracle by mathematics, so a
dropped term, a swapped index or a transposed operand fails independently
of the fixtures."""
import math

import numpy as np
import pytest

import oracle
from trinity_amd import meshgen as mg


# ---- recovery ---------------------------------------------------------------

def test_linear_field_gradient_is_4x_exact_gradient():
    # norm = 2/area (hessian.cpp:52) is 4x the P1 basis gradient, so a
    # linear field recovers 4*grad f at every vertex, boundary included
    m = mg.grid_mesh(30, 0.25, seed=3)
    f = 2.0 * m.xy[:, 0] - 5.0 * m.xy[:, 1]
    st = oracle.metric_stages(m.xy, m.tri, f)
    np.testing.assert_allclose(st["gradient"][:, 0], 8.0, rtol=0, atol=1e-11)
    np.testing.assert_allclose(st["gradient"][:, 1], -20.0, rtol=0, atol=1e-11)
    # the element Hessians vanish: sum_j deriv_j = 0 for a constant gradient
    assert np.abs(st["hessian"]).max() < 1e-9


def test_quadratic_field_hessian_is_16x_exact_hessian():
    # two 4x factors (gradient and Hessian); exact where the 2-ring patch of
    # every patch vertex lies inside the uniform grid
    n = 24
    m = mg.grid_mesh(n, 0.0)
    f = mg.field_quadratic(m.xy)
    st = oracle.metric_stages(m.xy, m.tri, f)
    i = np.arange(m.nv) % (n + 1)
    j = np.arange(m.nv) // (n + 1)
    inner = (i >= 4) & (i <= n - 4) & (j >= 4) & (j <= n - 4)
    h = st["hessian"][inner]
    np.testing.assert_allclose(h[:, 0], 16 * 2.0, rtol=1e-12)
    np.testing.assert_allclose(h[:, 1], 16 * 3.0, rtol=1e-12)
    np.testing.assert_allclose(h[:, 2], 16 * -4.0, rtol=1e-12)
    # and the gradient is 4x exact at ring >= 2
    inner2 = (i >= 2) & (i <= n - 2) & (j >= 2) & (j <= n - 2)
    x, y = m.xy[inner2, 0], m.xy[inner2, 1]
    np.testing.assert_allclose(st["gradient"][inner2, 0], 4 * (2 * x + 3 * y), rtol=0, atol=1e-11)
    np.testing.assert_allclose(st["gradient"][inner2, 1], 4 * (3 * x - 4 * y), rtol=0, atol=1e-11)


def test_patch_is_union_of_incident_triangles_of_closed_1ring():
    m = mg.grid_mesh(6, 0.2, seed=1)
    topo = oracle.Topology(m.nv, m.tri)
    for v in range(m.nv):
        ring = set(topo.v2v[topo.v2v_off[v]:topo.v2v_off[v + 1]].tolist()) | {v}
        brute = sorted(t for t in range(m.nt) if ring & set(m.tri[t].tolist()))
        assert topo.patch(v).tolist() == brute
    # 24 elements for interior vertices of the diagonal-split grid
    assert len(topo.patch(3 * 7 + 3)) == 24


# ---- eigen decomposition ------------------------------------------------------

def test_eigen_fast_path_on_diagonal():
    val, vec = oracle.eigen([3.0, 0.0, 0.0, 7.0])
    assert val.tolist() == [7.0, 3.0]
    assert vec.tolist() == [0.0, 1.0, 1.0, 0.0]


@pytest.mark.parametrize("theta", [0.3, 1.1, 2.0, 2.9])
@pytest.mark.parametrize("lam", [(1.0, 5.0), (40.0, 2.0), (-3.0, 8.0)])
def test_eigen_general_path_reconstructs(theta, lam):
    c, s = math.cos(theta), math.sin(theta)
    R = np.array([[c, -s], [s, c]])
    M = R @ np.diag(lam) @ R.T
    M = 0.5 * (M + M.T)
    val, vec = oracle.eigen([M[0, 0], M[0, 1], M[1, 0], M[1, 1]])
    assert val[0] <= val[1]
    np.testing.assert_allclose(sorted(val), sorted(lam), rtol=1e-13, atol=1e-13)
    P = np.array([[vec[0], vec[1]], [vec[2], vec[3]]])  # columns = eigenvectors
    np.testing.assert_allclose(P @ np.diag(val) @ P.T, M, rtol=0, atol=1e-13 * max(map(abs, lam)))


# ---- tensor interpolation, quality, length --------------------------------------

def test_log_euclidean_mean_of_equal_tensors_is_identity_map():
    t = [4.0, 1.5, 2.0]
    r = oracle.interpolate_tensor([t, t, t, t])
    np.testing.assert_allclose(r, t, rtol=1e-14)


def test_log_euclidean_mean_of_diagonal_tensors_is_geometric_mean():
    r = oracle.interpolate_tensor([[2.0, 0.0, 9.0], [8.0, 0.0, 1.0]])
    np.testing.assert_allclose(r, [4.0, 0.0, 3.0], rtol=1e-14, atol=1e-300)


def test_quality_of_unit_equilateral_triangle_is_one():
    L = 0.01
    m = 1.0 / L ** 2
    pa, pb, pc = [0.0, 0.0], [L, 0.0], [L / 2, L * math.sqrt(3) / 2]
    T = [m, 0.0, m]
    assert oracle.quality(pa, pb, pc, T, T, T) == pytest.approx(1.0, rel=1e-13)
    # a 2x too coarse metric: perimeter 3/2, mean = min(P/3, 3/P) = 1/2, and
    # 12 sqrt3 * area * sqrt(det) / P^2 = 1 for any equilateral triangle, so
    # q = (mean (2 - mean))^3
    T2 = [m / 4, 0.0, m / 4]
    q2 = oracle.quality(pa, pb, pc, T2, T2, T2)
    assert q2 == pytest.approx((0.5 * 1.5) ** 3, rel=1e-12)


def test_riemann_distance_isotropic_and_geometric():
    pa, pb = [0.1, 0.2], [0.4, 0.6]
    e = math.hypot(0.3, 0.4)
    assert oracle.riemann_dist(pa, pb, [4.0, 0, 4.0], [4.0, 0, 4.0]) == pytest.approx(2 * e, rel=1e-15)
    la, lb = 2 * e, 5 * e
    r = lb / la
    expect = lb * (r - 1) / (r * math.log(r))
    assert oracle.riemann_dist(pa, pb, [4.0, 0, 4.0], [25.0, 0, 25.0]) == pytest.approx(expect, rel=1e-14)


def test_complexity_of_constant_metric_is_area_times_sqrt_det():
    m = mg.grid_mesh(12, 0.2, seed=4)
    T = np.tile([9.0, 1.0, 4.0], (m.nv, 1))
    import ctypes
    L = oracle.lib()
    xy = np.ascontiguousarray(m.xy)
    c = L.orc_complexity(m.nt, xy.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                         m.tri.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                         T.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert c == pytest.approx(1.0 * math.sqrt(9 * 4 - 1), rel=1e-12)


def test_metric_eigenvalues_are_clamped():
    m = mg.workload("jitter40:fronts")
    st = oracle.metric_stages(m.xy, m.tri, m.solut, 0.8, 2, 1e-9, 0.4)
    M = st["metric"]
    tr = M[:, 0] + M[:, 2]
    det = M[:, 0] * M[:, 2] - M[:, 1] ** 2
    disc = np.sqrt(np.maximum(tr * tr / 4 - det, 0))
    lo, hi = tr / 2 - disc, tr / 2 + disc
    assert lo.min() >= 6.25 * (1 - 1e-9)           # 1/h_max^2
    assert hi.max() <= 1e18 * (1 + 1e-9)           # 1/h_min^2
    # the target complexity is met before clamping: sum A*sqrt(det M)
    assert st["target"] == int(math.floor(m.nv * 0.8))


# ---- colouring -----------------------------------------------------------------

def test_firstfit_colouring_is_sequential_greedy():
    m = mg.grid_mesh(16, 0.0)
    topo = oracle.Topology(m.nv, m.tri)
    col, nc = oracle.color_firstfit(topo, 8)
    assert oracle.color_valid(topo, col, nc)
    for v in range(m.nv):   # smallest colour absent from lower neighbours
        nb = topo.v2v[topo.v2v_off[v]:topo.v2v_off[v + 1]]
        used = set(col[nb[nb < v]].tolist())
        assert col[v] == min(c for c in range(1, 8) if c not in used)
    assert nc == 4


def test_colour_validity_checker_rejects_conflicts():
    m = mg.grid_mesh(8, 0.0)
    topo = oracle.Topology(m.nv, m.tri)
    col, nc = oracle.color_firstfit(topo, 8)
    bad = col.copy()
    bad[10] = bad[11]
    assert not oracle.color_valid(topo, bad, nc)
    bad = col.copy()
    bad[5] = 0
    assert not oracle.color_valid(topo, bad, nc)


# ---- smoothing invariants ------------------------------------------------------------

def test_uniform_grid_constant_metric_does_not_move():
    m = mg.grid_mesh(16, 0.0)
    T = np.tile([100.0, 0.0, 100.0], (m.nv, 1))
    topo = oracle.Topology(m.nv, m.tri)
    col, nc = oracle.color_firstfit(topo, 8)
    xy, ten, q, com, tas = oracle.smooth_run(m.xy, m.tri, m.tags, T, col, nc, 2, topo=topo)
    np.testing.assert_allclose(xy, m.xy, rtol=0, atol=1e-15)
    assert tas.tolist() == [(16 - 1) ** 2] * 2


def test_smoothing_keeps_boundary_and_orientation():
    m = mg.workload("jitter24:fronts")
    st = oracle.metric_stages(m.xy, m.tri, m.solut)
    col, nc = oracle.color_firstfit(st["topo"], 8)
    q0 = oracle.quality_cache(m.xy, m.tri, st["metric"])
    xy, ten, q, com, tas = oracle.smooth_run(m.xy, m.tri, m.tags, st["metric"], col, nc, 3,
                                             topo=st["topo"])
    bnd = (m.tags & 4) != 0
    assert np.array_equal(xy[bnd], m.xy[bnd])
    assert np.array_equal(ten[bnd], st["metric"][bnd])
    a, b, c = xy[m.tri[:, 0]], xy[m.tri[:, 1]], xy[m.tri[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1])
    assert (0.5 * cross > 1e-9).all()
    assert com.min() > 0 and (com <= tas).all()
    # accepted moves only raise the worst quality around each moved vertex
    assert q.min() >= q0.min()


# ---- refinement filter and Steiner points (src/refinement.cpp:216-296,
# numeric.cpp:387-445), pinned by closed forms -------------------------------

def test_steiner_point_in_one_isotropic_metric_is_the_midpoint():
    pa, pb = np.array([0.1, 0.2]), np.array([0.4, 0.65])
    t = np.array([5.0, 0.0, 5.0])
    pt, tn = oracle.steiner_point(pa, pb, t, t)
    assert np.array_equal(pt, 0.5 * pa + 0.5 * pb)
    assert np.allclose(tn, t, rtol=1e-14, atol=0)


def test_steiner_point_between_two_isotropic_metrics():
    # lengths 2L and 3L: weight (sqrt(6) L - 3L) / (2L - 3L) = 3 - sqrt(6);
    # log-Euclidean mean of 4 I and 9 I is 6 I
    pa, pb = np.array([0.0, 0.0]), np.array([0.3, 0.4])
    pt, tn = oracle.steiner_point(pa, pb, [4.0, 0.0, 4.0], [9.0, 0.0, 9.0])
    w = 3.0 - np.sqrt(6.0)
    assert np.allclose(pt, pa + w * (pb - pa), rtol=1e-14, atol=1e-16)
    assert np.allclose(tn, [6.0, 0.0, 6.0], rtol=1e-14, atol=1e-15)


def _grid_with_metric(n, edge_len):
    m = mg.grid_mesh(n)
    c = (edge_len * n) ** 2  # axis edges of length 1/n measure edge_len in c I
    ten = np.tile([c, 0.0, c], (m.nv, 1))
    return m, ten


def test_refine_filter_lists_each_long_diagonal_once():
    n = 12
    m, ten = _grid_with_metric(n, 1.2)  # axis 1.2 < sqrt 2 < diagonal 1.2 sqrt 2
    r = oracle.refine_filter(m.xy, m.tri, ten)
    nt = 2 * n * n
    assert r["eval"] == nt and r["tasks"] == nt and r["adds"] == nt
    assert (r["pattern"] == 1).all()
    assert r["split"] == n * n
    e = r["edges"]
    assert (e[:, 0] < e[:, 1]).all() and (e[:, 2] >= 0).all()
    assert len({(a, b) for a, b, _ in e}) == n * n
    # each listed edge is a diagonal of its cell, held by the opposite triangle
    for a, b, opp in e:
        assert b - a == n + 2
        assert {int(a), int(b)} <= set(m.tri[opp].tolist())


def test_refine_filter_lists_boundary_edges_from_their_only_triangle():
    n = 9
    m, ten = _grid_with_metric(n, 1.5)  # every edge is long
    r = oracle.refine_filter(m.xy, m.tri, ten)
    assert (r["pattern"] == 3).all()
    assert r["split"] == 3 * n * n + 2 * n  # every edge of the grid once
    e = r["edges"]
    assert len({(min(a, b), max(a, b)) for a, b, _ in e}) == r["split"]
    assert int((e[:, 2] < 0).sum()) == 4 * n


# ---- independent set for coarsening (src/partition.cpp:200-306) ----------

@pytest.mark.parametrize("wl,frac,seed", [("jitter40:fronts", 0.6, 1), ("uniform24:tanh", 1.0, 7),
                                          ("jitter30:layer", 0.2, 3)])
def test_indep_set_is_the_greedy_maximal_independent_set(wl, frac, seed):
    """A vertex is in the set iff no neighbour in the subgraph with a larger
    (priority, id) key is: this characterises the greedy set uniquely, and
    makes it independent and maximal."""
    m = mg.workload(wl)
    topo = oracle.Topology(m.nv, m.tri)
    rng = np.random.default_rng(seed)
    nodes = np.sort(rng.choice(m.nv, size=int(frac * m.nv), replace=False)).astype(np.int32)
    ins = oracle.indep_set(topo, nodes, seed=seed)
    sub = np.zeros(m.nv, bool)
    sub[nodes] = True
    inset = np.zeros(m.nv, bool)
    inset[nodes[ins == 1]] = True
    key = {int(v): (oracle.color_priority(seed, int(v)), int(v)) for v in nodes}
    for v in nodes:
        ring = topo.v2v[topo.v2v_off[v]:topo.v2v_off[v + 1]]
        ring = [int(u) for u in ring if sub[u]]
        blocked = any(inset[u] and key[u] > key[int(v)] for u in ring)
        assert inset[v] == (not blocked)
        if inset[v]:
            assert not any(inset[u] for u in ring)        # independent
        else:
            assert any(inset[u] for u in ring)            # maximal
