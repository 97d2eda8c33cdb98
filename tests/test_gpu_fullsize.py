"""GPU parity at the benchmark's full sizes, against the oracle run on the
host (OpenMP C, up to 16 threads):

  jitter2048:fronts  4,198,401 vertices, L^1, L^2 and L^4, 3 sweeps
  jitter4096:layer  16,785,409 vertices, L^4, 5 sweeps (the stress pass)
  flip1024:fronts    1,050,625 vertices with randomly flipped diagonals
                     (1 to 8 incident triangles; 31 % of the vertices have
                     more than 6 and take the generic move), L^2, 3 sweeps

Checked, at the tolerances of the small sizes:
  gradient, raw Hessian        bit-exact
  locally normalised metric    bit-exact (the device pow is glibc's)
  final metric                 1e-10 per vertex (norm-relative) against the
                               oracle's own; also bit-identical to the
                               oracle's global normalisation of the device's
                               local metric with the device complexity
  sweeps (device colouring,    per-sweep commit and task counts exact,
  oracle's metric fed to both) coordinates 1e-12, smoothed metric 1e-10
Size-independent properties beside them: the colouring valid, boundary
vertices unmoved, every cached quality positive.

One bound is wider than at the small sizes, for reasons of size: the
complexity is a sum of nt positive terms that the oracle adds in `threads`
sequential chunks, whose rounding error is bounded by
(nt / threads + threads) * 2^-52 relative (the device's tree is far
tighter); 1.9e-12 was measured on the 33.5 M-triangle grid, against 1e-12
used at the small sizes.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import normrel
from trinity_amd import meshgen as mg

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("wl,p,sweeps", [("jitter2048:fronts", 2, 3),
                                         ("jitter2048:fronts", 1, 3),
                                         ("jitter2048:fronts", 4, 3),
                                         ("jitter4096:layer", 4, 5),
                                         ("flip1024:fronts", 2, 3)])
def test_full_size_pass_matches_oracle(ctx, wl, p, sweeps):
    import torch
    oracle.set_threads(min(16, os.cpu_count() or 1))  # the same count bounds the sum below
    m = mg.workload(wl)
    print(f"{wl}: V={m.nv:,} T={m.nt:,}", flush=True)
    st = oracle.metric_stages(m.xy, m.tri, m.solut, 0.8, p)
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    xy, f = dev(m.xy), dev(m.solut)
    grad = torch.empty((m.nv, 2), dtype=torch.float64, device="cuda")
    hess = torch.empty((m.nv, 3), dtype=torch.float64, device="cuda")
    ten, c = T().metric_run(ctx, dm, xy, f, 0.8, p, 1e-9, 0.4, gradient=grad, hessian=hess)
    assert bits_equal(grad.cpu().numpy(), st["gradient"])
    assert bits_equal(hess.cpu().numpy(), st["hessian"])
    threads = min(16, os.cpu_count() or 1)
    tol = (m.nt / threads + threads) * 2.0 ** -52
    assert abs(c - st["complexity"]) <= tol * st["complexity"]
    # local normalisation: pow is the only step that is not an IEEE operation
    mloc = T().normalize_local(ctx, hess.clone(), p).cpu().numpy()
    rl = normrel(mloc, st["metric_local"])
    # global normalisation of the device's own local metric with the device
    # complexity, by the oracle: IEEE operations only, so bit-identical
    ref = np.ascontiguousarray(mloc.copy())
    oracle.lib().orc_normalize_global(m.nv, oracle._p(ref, oracle._d), st["target"] / c, 1e-9, 0.4)
    tn = ten.cpu().numpy()
    rm = normrel(tn, st["metric"])
    print(f"local metric: max {rl.max():.3e}; final metric vs the oracle's own: max {rm.max():.3e}, "
          f"above 1e-10: {(rm > 1e-10).sum()}", flush=True)
    assert bits_equal(mloc, st["metric_local"])
    assert bits_equal(tn, ref)
    assert rm.max() <= 1e-10
    del grad, hess, ten

    col = T().color(ctx, dm, seed=1, max_colors=64)
    assert T().color_check(ctx, dm, col.color, col.ncolors) == (0, 0)
    print(f"colours: {col.ncolors}", flush=True)
    if wl.startswith("jitter"):
        assert col.ncolors <= 6  # the tile first fit on a row-numbered grid
    cnp = col.color.cpu().numpy()
    xo, to, qo, co, tao = oracle.smooth_run(m.xy, m.tri, m.tags, st["metric"], cnp,
                                            col.ncolors, sweeps, topo=st["topo"])
    xg, tg = dev(m.xy), dev(st["metric"])
    q = torch.empty(dm.nt, dtype=torch.float64, device="cuda")
    com, tas = T().smooth_run(ctx, dm, xg, tg, q, col, sweeps)
    assert tas == tao.tolist()
    assert com == co.tolist()
    xgn, tgn, qn = xg.cpu().numpy(), tg.cpu().numpy(), q.cpu().numpy()
    assert normrel(xgn, xo).max() <= 1e-12
    r = normrel(tgn, to)
    top = np.argsort(r)[-5:]
    print(f"smoothed metric: max {r.max():.3e}, above 1e-12: {(r > 1e-12).sum()}, above 1e-10: "
          f"{(r > 1e-10).sum()}; worst vertices {top.tolist()} {r[top].tolist()}", flush=True)
    for v in top[-2:]:
        ev = np.linalg.eigvalsh(np.array([[to[v, 0], to[v, 1]], [to[v, 1], to[v, 2]]]))
        print(f"  vertex {v}: eigenvalues {ev.tolist()}, moved {bool((xgn[v] != m.xy[v]).any())}",
              flush=True)
    assert r.max() <= 1e-10
    boundary = (m.tags & 4) != 0
    assert np.array_equal(xgn[boundary], m.xy[boundary])
    assert (qn > 0).all()
    dm.close()
