"""The allocation-free sweeps (trn_smooth_run_prepared_noalloc) against
trn_smooth_run_prepared, through the C ABI: bit-identical coordinates,
metric and qualities and equal counters, on two calls in a row, the second
on a smaller mesh (the fixed scratch keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest

from trinity_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def T():
    import trinity_amd
    return trinity_amd


def _run(ctx, dm, m, metric, col, sweeps, noalloc):
    import torch
    L = T().lib()
    xy = torch.from_numpy(m.xy.copy()).cuda()
    ten = torch.from_numpy(metric.copy()).cuda()
    q = torch.empty(m.nt, dtype=torch.float64, device="cuda")
    T().smooth_prepare(ctx, dm, xy, ten, q)
    off = (ctypes.c_int32 * (col.ncolors + 1))(*[int(v) for v in col.class_off])
    commits = (ctypes.c_int32 * sweeps)()
    tasks = (ctypes.c_int32 * sweeps)()
    fn = L.trn_smooth_run_prepared_noalloc if noalloc else L.trn_smooth_run_prepared
    rc = fn(ctx.handle, dm.handle, ctypes.c_void_p(xy.data_ptr()), ctypes.c_void_p(ten.data_ptr()),
            ctypes.c_void_p(q.data_ptr()), ctypes.c_void_p(col.order.data_ptr()), off,
            int(col.ncolors), int(sweeps), commits, tasks)
    assert rc == 0, T().lib().trn_last_error()
    return (xy.cpu().numpy(), ten.cpu().numpy(), q.cpu().numpy(), list(commits), list(tasks))


def _case(ctx, workload, p, sweeps):
    import torch
    m = mg.workload(workload)
    dm = T().DeviceMesh(ctx, m.nv, m.tri, m.tags)
    metric, _ = T().metric_run(ctx, dm, torch.from_numpy(m.xy).cuda(),
                               torch.from_numpy(m.solut).cuda(), 0.8, p, 1e-9, 0.4)
    metric = metric.cpu().numpy()
    col = T().color(ctx, dm, seed=7, max_colors=64)
    ref = _run(ctx, dm, m, metric, col, sweeps, False)
    got = _run(ctx, dm, m, metric, col, sweeps, True)
    for a, b in zip(ref[:3], got[:3]):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert ref[3:] == got[3:]
    assert sum(got[3]) > 0
    dm.close()


def test_noalloc_sweeps_equal_allocating_sweeps(ctx):
    _case(ctx, "jitter300:fronts", 2, 3)     # first call on this GPU
    _case(ctx, "flip90:fronts", 4, 2)        # smaller, generic moves too
