"""GPU tests of the batched form (trn_batch_*): many meshes of uneven sizes
in one persistent launch per pass.

Each mesh of a batch must come out as a plain pass over that mesh alone:
  metric                 1e-12 per-vertex norm-relative vs the plain path
                         (per-mesh complexity summed in another fixed order),
                         1e-10 vs the oracle
  complexity             1e-12 relative vs the plain path and the oracle
  element qualities      bit-exact vs the plain quality cache fed the same
                         metric
  smoothed coordinates,  bit-exact vs the plain sweeps fed the same metric
  tensors, qualities     and the union colouring restricted to the mesh;
                         1e-12 (coordinates) vs the oracle fed the same
  accepted-move counts   exact
A pass replayed from its hipGraph equals the eager pass bit for bit on
every replay, and no result depends on the launch shape the autotuner picks.

SWITCHED OFF: on MI355X the batched pass did not complete on the 11-mesh
batch of the first test below (three runs: twice the single-launch pipelined
schedule, once the per-stage schedule; 120 s and 60 s test limits), while
the same sources complete and match the plain pass in a CPU emulation of
the HIP runtime.  The GPU cause has not been shown, so the pass entry points
refuse without launching; the parity tests are kept, skipped, and the
entry points must refuse.
"""
import numpy as np
import pytest

import oracle
from conftest import normrel
from trinity_amd import meshgen as mg

pytestmark = pytest.mark.gpu

SIDES = [3, 7, 12, 20, 5, 33, 16, 9, 41, 2, 27]
BATCH_OFF = ("the batched pass stalled on MI355X (no completion within the test limit on "
             "an 11-mesh batch, three runs, both schedules) and its cause has not been shown; "
             "its entry points are switched off")
TF, P, HMIN, HMAX = 0.8, 2, 1e-9, 0.4


def T():
    import trinity_amd
    return trinity_amd


def dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def _meshes(sides=SIDES, seed=5):
    out = []
    for k, side in enumerate(sides):
        m = mg.grid_mesh(side, 0.25, seed + k)
        m.solut, _ = mg.field_fronts(m.xy, seed=seed + 100 + k)
        out.append(m)
    return out


def _fields(ms):
    import torch
    xy_in = dev(np.concatenate([m.xy for m in ms]))
    solut = dev(np.concatenate([m.solut for m in ms]))
    nv = xy_in.shape[0]
    nt = sum(m.nt for m in ms)
    return (xy_in, torch.empty_like(xy_in), solut,
            torch.empty((nv, 3), dtype=torch.float64, device="cuda"),
            torch.empty(nt, dtype=torch.float64, device="cuda"))


def _batch(ctx, ms):
    return T().Batch(ctx, [(m.nv, m.tri, m.tags) for m in ms])


@pytest.mark.skip(reason=BATCH_OFF)
def test_batch_metric_matches_plain_and_oracle(ctx):
    import torch
    ms = _meshes()
    b = _batch(ctx, ms)
    xy_in, xy, solut, ten, q = _fields(ms)
    cplx = torch.empty(len(ms), dtype=torch.float64, device="cuda")
    b.run(xy_in, xy, solut, ten, q, TF, P, HMIN, HMAX, sweeps=0, complexity=cplx)
    b.check()
    ten_b, q_b, c_b = ten.cpu().numpy(), q.cpu().numpy(), cplx.cpu().numpy()
    assert np.array_equal(xy.cpu().numpy(), xy_in.cpu().numpy())
    for k, m in enumerate(ms):
        v0, v1 = b.voff[k], b.voff[k + 1]
        t0, t1 = b.toff[k], b.toff[k + 1]
        dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
        tp, cp = T().metric_run(ctx, dm, dev(m.xy), dev(m.solut), TF, P, HMIN, HMAX)
        assert abs(c_b[k] - cp) <= 1e-12 * abs(cp), k
        assert normrel(ten_b[v0:v1], tp.cpu().numpy()).max() <= 1e-12, k
        st = oracle.metric_stages(m.xy, m.tri, m.solut, TF, P, HMIN, HMAX)
        assert abs(c_b[k] - st["complexity"]) <= 1e-12 * abs(st["complexity"]), k
        assert normrel(ten_b[v0:v1], st["metric"]).max() <= 1e-10, k
        qp = T().quality_cache(ctx, dm, dev(m.xy), dev(ten_b[v0:v1]))
        assert np.array_equal(q_b[t0:t1], qp.cpu().numpy()), k


@pytest.mark.skip(reason=BATCH_OFF)
def test_batch_smoothing_matches_plain_and_oracle(ctx):
    import torch
    ms = _meshes(seed=17)
    sweeps = 2
    b = _batch(ctx, ms)
    nc = b.color(seed=5, max_colors=64)
    xy_in, xy, solut, ten, q = _fields(ms)
    b.run(xy_in, xy, solut, ten, q, TF, P, HMIN, HMAX, sweeps=0)
    metric = ten.cpu().numpy()
    com_b, tas_b = b.run(xy_in, xy, solut, ten, q, TF, P, HMIN, HMAX, sweeps=sweeps,
                         counters=True)
    xy_b, ten_b, q_b = xy.cpu().numpy(), ten.cpu().numpy(), q.cpu().numpy()
    # the batch colouring is trn_color on the union mesh with the same seed
    tri_u = np.concatenate([m.tri + b.voff[k] for k, m in enumerate(ms)])
    tags_u = np.concatenate([m.tags for m in ms])
    du = T().DeviceMesh(ctx, b.nv, tri_u, tags_u)
    col = T().color(ctx, du, seed=5, max_colors=64)
    assert col.ncolors == nc
    assert T().color_check(ctx, du, col.color, nc) == (0, 0)
    cu = col.color.cpu().numpy()
    com_sum, tas_sum = np.zeros(sweeps, int), np.zeros(sweeps, int)
    for k, m in enumerate(ms):
        v0, v1 = b.voff[k], b.voff[k + 1]
        t0, t1 = b.toff[k], b.toff[k + 1]
        dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
        cm = T().color_classes(ctx, dev(cu[v0:v1], torch.int32), nc, mesh=dm)
        xp, tp = dev(m.xy), dev(metric[v0:v1])
        qp = torch.empty(m.nt, dtype=torch.float64, device="cuda")
        com, tas = T().smooth_run(ctx, dm, xp, tp, qp, cm, sweeps)
        com_sum += com
        tas_sum += tas
        assert np.array_equal(xy_b[v0:v1], xp.cpu().numpy()), k
        assert np.array_equal(ten_b[v0:v1], tp.cpu().numpy()), k
        assert np.array_equal(q_b[t0:t1], qp.cpu().numpy()), k
        xo, to, qo, co, tao = oracle.smooth_run(m.xy, m.tri, m.tags, metric[v0:v1], cu[v0:v1],
                                                nc, sweeps)
        assert com == co.tolist() and tas == tao.tolist(), k
        assert normrel(xy_b[v0:v1], xo).max() <= 1e-12, k
        assert normrel(ten_b[v0:v1], to).max() <= 1e-10, k
    assert com_b == com_sum.tolist()
    assert tas_b == tas_sum.tolist()


@pytest.mark.skip(reason=BATCH_OFF)
def test_batch_graph_replays_equal_eager_pass(ctx):
    import torch
    import trinity_amd
    ms = _meshes(sides=[9, 30, 4, 18, 25], seed=3)
    c2 = trinity_amd.Context(0, use_torch_stream=False)
    try:
        c2.use_own_stream()
        b = _batch(c2, ms)
        b.color(seed=2)
        xy_in, xy, solut, ten, q = _fields(ms)
        torch.cuda.synchronize()
        b.run(xy_in, xy, solut, ten, q, 0.7, 4, HMIN, HMAX, sweeps=3)
        c2.synchronize()
        b.check()
        ref = [xy.cpu().numpy(), ten.cpu().numpy(), q.cpu().numpy()]
        for rep in range(4):
            # eager work between replays: clobber the outputs on torch's stream
            xy.fill_(float(rep) - 7.5)
            ten.fill_(3.25 * rep)
            q.zero_()
            torch.cuda.synchronize()
            b.run_graph(xy_in, xy, solut, ten, q, 0.7, 4, HMIN, HMAX, sweeps=3)
            c2.synchronize()
            b.check()
            out = [xy.cpu().numpy(), ten.cpu().numpy(), q.cpu().numpy()]
            for a, r in zip(out, ref):
                assert np.array_equal(a, r), rep
        b.close()
    finally:
        c2.close()


@pytest.mark.skip(reason=BATCH_OFF)
def test_batch_autotune_changes_only_the_time(ctx):
    ms = _meshes(sides=[40, 6, 22, 31, 13], seed=9)
    b = _batch(ctx, ms)
    b.color(seed=1)
    xy_in, xy, solut, ten, q = _fields(ms)
    b.set_config(64, 256, 1024, 4)
    b.run(xy_in, xy, solut, ten, q, sweeps=2)
    b.check()
    ref = [xy.cpu().numpy(), ten.cpu().numpy(), q.cpu().numpy()]
    best, ms_best = b.autotune(xy_in, xy, solut, ten, q, sweeps=2)
    assert best[0] in (64, 128, 256) and best[3] in (2, 4, 8)
    assert best[1] in (best[0], 2 * best[0], 4 * best[0]) and best[2] == 4 * best[1]
    assert b.config() == best and ms_best > 0
    for cfg in [(128, 128, 512, 2), (256, 1024, 4096, 8), best]:
        b.set_config(*cfg)
        b.run(xy_in, xy, solut, ten, q, sweeps=2)
        b.check()
        for a, r in zip([xy.cpu().numpy(), ten.cpu().numpy(), q.cpu().numpy()], ref):
            assert np.array_equal(a, r), cfg


def test_batch_rejects_bad_input(ctx):
    m = mg.grid_mesh(4)
    bad = m.tri.copy()
    bad[0, 0] = m.nv  # an index outside its own mesh would join two meshes
    with pytest.raises(T().TrinityError):
        T().Batch(ctx, [(m.nv, m.tri, m.tags), (m.nv, bad, m.tags)])
    b = _batch(ctx, [m])
    with pytest.raises(T().TrinityError):
        b.set_config(96, 256, 1024, 4)
    xy_in, xy, solut, ten, q = _fields([mg.workload("uniform4:tanh")])
    with pytest.raises(T().TrinityError):  # sweeps need a colouring
        b.run(xy_in, xy, solut, ten, q, sweeps=1)


@pytest.mark.skip(reason=BATCH_OFF)
def test_batch_pipelined_matches_staged(ctx):
    ms = _meshes()
    b = _batch(ctx, ms)
    b.color(seed=5)
    xy_in, xy, solut, ten, q = _fields(ms)
    b.run(xy_in, xy, solut, ten, q, TF, P, HMIN, HMAX, sweeps=2)
    ref = [xy.cpu().numpy(), ten.cpu().numpy(), q.cpu().numpy()]
    b.run_pipelined(xy_in, xy, solut, ten, q, TF, P, HMIN, HMAX, sweeps=2)
    b.check()
    for a, r in zip([xy.cpu().numpy(), ten.cpu().numpy(), q.cpu().numpy()], ref):
        assert np.array_equal(a, r)


def test_batch_pass_entry_points_refuse_without_launching(ctx):
    ms = _meshes(sides=[5, 8])
    b = _batch(ctx, ms)
    assert b.nv == sum(m.nv for m in ms)
    xy_in, xy, solut, ten, q = _fields(ms)
    for call in (b.run, b.run_graph, b.run_pipelined, b.autotune):
        with pytest.raises(T().TrinityError, match="switched off"):
            call(xy_in, xy, solut, ten, q, sweeps=0)
