"""trinity_amd -- MI355X (gfx950) implementation of hobywan/trinity's
metric-and-smoothing path, driven through the C ABI of libtrinity_amd.so
(include/trinity_amd.h).

This module is a thin ctypes binding used by the tests and the benchmark;
PyTorch only supplies device memory and the stream.  There is no CPU
fallback: every call goes to the HIP kernels, and loading fails loudly when
the library has not been built (``python -m trinity_amd.build``).

The Python names mirror the reference passes:
    Metrics.run            -> metric_run      (src/metric.cpp:59-70)
    Partit.extractColoring -> color           (src/partition.cpp:79-197)
    Smooth.run             -> smooth_run      (src/smoothing.cpp:274-302)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtrinity_amd.so")

TRN_OK = 0
TRN_ERR_ARG = -1
TRN_ERR_HIP = -2
TRN_ERR_CAPACITY = -3
TRN_ERR_PALETTE = -4
TRN_ERR_STATE = -5

# every symbol declared in include/trinity_amd.h
EXPORTS = [
    "trn_version", "trn_last_error", "trn_ctx_create", "trn_ctx_destroy",
    "trn_ctx_set_stream", "trn_ctx_use_own_stream", "trn_ctx_synchronize", "trn_ctx_set_timing",
    "trn_ctx_set_timing_only",
    "trn_ctx_reset_timing", "trn_ctx_timing_count", "trn_ctx_timing_get",
    "trn_mesh_create", "trn_mesh_destroy", "trn_mesh_info", "trn_mesh_csr",
    "trn_metric_run", "trn_gradient", "trn_hessian", "trn_normalize_local",
    "trn_complexity", "trn_normalize_global", "trn_color", "trn_color_partial",
    "trn_mesh_set_color_tile", "trn_mesh_set_class_order", "trn_color_classes",
    "trn_medit_info", "trn_medit_read", "trn_trnm_info", "trn_trnm_read", "trn_trnm_write",
    "trn_mesh_autotune", "trn_mesh_shapes", "trn_mesh_set_shapes",
    "trn_color_check", "trn_indep_set", "trn_quality_cache", "trn_quality_stats", "trn_refine_filter",
    "trn_smooth_run",
    "trn_smooth_prepare",
    "trn_smooth_run_prepared",
    "trn_smooth_run_prepared_noalloc",
    "trn_batch_create", "trn_batch_create_xy", "trn_batch_destroy", "trn_batch_info", "trn_batch_color",
    "trn_batch_coloring", "trn_batch_pass", "trn_batch_pass_graph",
    "trn_adjoint_create", "trn_adjoint_destroy", "trn_adjoint_forward", "trn_adjoint_backward",
    "trn_adjoint_smooth", "trn_debug_math",
]


class TrinityError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_P = ctypes.POINTER


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m trinity_amd.build` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "trn_version": ([], ctypes.c_char_p),
        "trn_last_error": ([], ctypes.c_char_p),
        "trn_ctx_create": ([ctypes.c_int, _P(_vp)], ctypes.c_int),
        "trn_ctx_destroy": ([_vp], ctypes.c_int),
        "trn_ctx_set_stream": ([_vp, _vp], ctypes.c_int),
        "trn_ctx_use_own_stream": ([_vp], ctypes.c_int),
        "trn_ctx_synchronize": ([_vp], ctypes.c_int),
        "trn_ctx_set_timing": ([_vp, ctypes.c_int], ctypes.c_int),
        "trn_ctx_set_timing_only": ([_vp, ctypes.c_char_p], ctypes.c_int),
        "trn_ctx_reset_timing": ([_vp], ctypes.c_int),
        "trn_ctx_timing_count": ([_vp, _P(_i32)], ctypes.c_int),
        "trn_ctx_timing_get": ([_vp, _i32, _P(ctypes.c_char_p), _P(_f64), _P(_i64), _P(_f64)],
                               ctypes.c_int),
        "trn_mesh_create": ([_vp, _i32, _i32, _vp, _vp, _vp, _P(_vp)], ctypes.c_int),
        "trn_mesh_destroy": ([_vp], ctypes.c_int),
        "trn_mesh_info": ([_vp, _P(_i64), _P(_i32), _P(_i32), _P(_i32)], ctypes.c_int),
        "trn_mesh_csr": ([_vp, _P(_vp), _P(_vp), _P(_vp), _P(_vp)], ctypes.c_int),
        "trn_metric_run": ([_vp, _vp, _vp, _vp, _f64, _i32, _f64, _f64, _vp, _P(_f64), _vp, _vp],
                           ctypes.c_int),
        "trn_gradient": ([_vp, _vp, _vp, _vp, _vp], ctypes.c_int),
        "trn_hessian": ([_vp, _vp, _vp, _vp, _vp], ctypes.c_int),
        "trn_normalize_local": ([_vp, _i32, _vp, _i32], ctypes.c_int),
        "trn_complexity": ([_vp, _vp, _vp, _vp, _P(_f64)], ctypes.c_int),
        "trn_normalize_global": ([_vp, _i32, _vp, _f64, _f64, _f64], ctypes.c_int),
        "trn_color": ([_vp, _vp, ctypes.c_uint64, _i32, _vp, _vp, _P(_i32), _P(_i32)],
                      ctypes.c_int),
        "trn_color_partial": ([_vp, _vp, ctypes.c_uint64, _i32, _vp, _vp, _P(_i32), _P(_i32),
                               _P(_i64)], ctypes.c_int),
        "trn_mesh_set_color_tile": ([_vp, _i32], ctypes.c_int),
        "trn_mesh_set_class_order": ([_vp, _i32], ctypes.c_int),
        "trn_mesh_autotune": ([_vp, _vp, _vp, _vp, _i32, _P(_i32), _P(_f64)], ctypes.c_int),
        "trn_mesh_shapes": ([_vp, _P(_i32)], ctypes.c_int),
        "trn_mesh_set_shapes": ([_vp, _P(_i32)], ctypes.c_int),
        "trn_medit_info": ([ctypes.c_char_p, _P(_i32), _P(_i32)], ctypes.c_int),
        "trn_medit_read": ([ctypes.c_char_p, ctypes.c_char_p, _i32, _i32, _vp, _vp, _vp, _vp],
                           ctypes.c_int),
        "trn_trnm_info": ([ctypes.c_char_p, _P(_i32), _P(_i32)], ctypes.c_int),
        "trn_trnm_read": ([_vp, ctypes.c_char_p, _i32, _i32, _vp, _vp, _vp, _vp], ctypes.c_int),
        "trn_trnm_write": ([ctypes.c_char_p, _i32, _i32, _vp, _vp, _vp, _vp], ctypes.c_int),
        "trn_color_classes": ([_vp, _vp, _i32, _vp, _i32, _vp, _P(_i32)], ctypes.c_int),
        "trn_color_check": ([_vp, _vp, _vp, _i32, _P(_i64), _P(_i64)], ctypes.c_int),
        "trn_quality_cache": ([_vp, _vp, _vp, _vp, _vp], ctypes.c_int),
        "trn_indep_set": ([_vp, _vp, _vp, _i32, ctypes.c_uint64, _vp, _P(_i32)], ctypes.c_int),
        "trn_quality_stats": ([_vp, _vp, _vp, _vp, _vp, _P(_f64), _P(_i64)], ctypes.c_int),
        "trn_refine_filter": ([_vp, _vp, _vp, _vp, _f64, _vp, _vp, _vp, _vp, _vp, _i32,
                               _P(_i32)], ctypes.c_int),
        "trn_smooth_run": ([_vp, _vp, _vp, _vp, _vp, _vp, _P(_i32), _i32, _i32, _i32, _P(_i32),
                            _P(_i32)], ctypes.c_int),
        "trn_smooth_prepare": ([_vp, _vp, _vp, _vp, _vp, _i32], ctypes.c_int),
        "trn_smooth_run_prepared": ([_vp, _vp, _vp, _vp, _vp, _vp, _P(_i32), _i32, _i32,
                                     _P(_i32), _P(_i32)], ctypes.c_int),
        "trn_smooth_run_prepared_noalloc": ([_vp, _vp, _vp, _vp, _vp, _vp, _P(_i32), _i32, _i32,
                                             _P(_i32), _P(_i32)], ctypes.c_int),
        "trn_batch_create": ([_vp, _i32, _P(_i32), _P(_i32), _vp, _vp, _P(_vp)], ctypes.c_int),
        "trn_batch_create_xy": ([_vp, _i32, _P(_i32), _P(_i32), _vp, _vp, _vp, _P(_vp)],
                                ctypes.c_int),
        "trn_batch_destroy": ([_vp], ctypes.c_int),
        "trn_batch_info": ([_vp, _P(_i32), _P(_i32), _P(_vp)], ctypes.c_int),
        "trn_batch_color": ([_vp, _vp, ctypes.c_uint64, _i32, _P(_i32)], ctypes.c_int),
        "trn_batch_coloring": ([_vp, _P(_vp), _P(_vp), _P(_i32), _P(_i32)], ctypes.c_int),
        "trn_batch_pass": ([_vp, _vp, _vp, _vp, _vp, _f64, _i32, _f64, _f64, _vp, _vp, _i32, _vp,
                            _P(_i32), _P(_i32)], ctypes.c_int),
        "trn_batch_pass_graph": ([_vp, _vp, _vp, _vp, _vp, _f64, _i32, _f64, _f64, _vp, _vp,
                                  _i32], ctypes.c_int),
        "trn_adjoint_create": ([_vp, _vp, _P(_vp)], ctypes.c_int),
        "trn_adjoint_destroy": ([_vp], ctypes.c_int),
        "trn_adjoint_forward": ([_vp, _vp, _vp, _vp, _f64, _i32, _f64, _f64, _P(_f64), _vp, _vp],
                                ctypes.c_int),
        "trn_adjoint_backward": ([_vp, _vp, _vp, _vp], ctypes.c_int),
        "trn_adjoint_smooth": ([_vp, _vp, _vp, _vp, _f64, _i32, _f64, _f64, _i32, _f64, _P(_f64),
                               _P(_i32)], ctypes.c_int),
        "trn_debug_math": ([_vp, _i32, _i64, _vp, _vp, _vp], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def lib():
    return _load()


def _check(code: int) -> None:
    if code != TRN_OK:
        msg = lib().trn_last_error().decode(errors="replace")
        raise TrinityError(code, msg)


def version() -> str:
    return lib().trn_version().decode()


def _torch():
    import torch
    return torch


def _ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def to_device(a, dtype=None, device="cuda"):
    """numpy array or torch tensor -> contiguous torch tensor on `device`."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        t = a
        if dtype is not None:
            t = t.to(dtype)
        return t.to(device).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype).contiguous()


class Context:
    """A device, a stream and the scratch workspace (trn_ctx)."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        torch = _torch()
        h = _vp()
        _check(lib().trn_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device
        if use_torch_stream:
            # launch on exactly the stream torch allocates on (its default is
            # the null stream, pointer 0), so torch reads see finished results
            s = torch.cuda.current_stream(device)
            _check(lib().trn_ctx_set_stream(h, _vp(s.cuda_stream) if s.cuda_stream else None))

    def set_stream(self, stream_ptr: int | None) -> None:
        """Route work to a hipStream_t (None/0 = the null stream)."""
        _check(lib().trn_ctx_set_stream(self.handle, _vp(stream_ptr) if stream_ptr else None))

    def use_own_stream(self) -> None:
        _check(lib().trn_ctx_use_own_stream(self.handle))

    def synchronize(self) -> None:
        _check(lib().trn_ctx_synchronize(self.handle))

    def set_timing(self, on: bool, only: str | None = None) -> None:
        """Per-kernel HIP-event timing; `only` restricts it to the launches
        of one statistic (e.g. "smooth_phase")."""
        _check(lib().trn_ctx_set_timing_only(self.handle, only.encode() if only else None))
        _check(lib().trn_ctx_set_timing(self.handle, int(bool(on))))

    def reset_timing(self) -> None:
        _check(lib().trn_ctx_reset_timing(self.handle))

    def timings(self) -> dict:
        n = _i32()
        _check(lib().trn_ctx_timing_count(self.handle, ctypes.byref(n)))
        out = {}
        for k in range(n.value):
            name = ctypes.c_char_p()
            ms = _f64()
            cnt = _i64()
            by = _f64()
            _check(lib().trn_ctx_timing_get(self.handle, k, ctypes.byref(name), ctypes.byref(ms),
                                            ctypes.byref(cnt), ctypes.byref(by)))
            out[name.value.decode()] = {"ms": ms.value, "launches": cnt.value, "bytes": by.value}
        return out

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().trn_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMesh:
    """Device topology of a triangle mesh (trn_mesh): v2t/v2v CSR built on
    the GPU with the semantics of Mesh::rebuildTopology."""

    def __init__(self, ctx: Context, nv: int, tri, tags, xy=None):
        torch = _torch()
        self.ctx = ctx
        self.nv = int(nv)
        tri_t = to_device(np.asarray(tri, dtype=np.int32).reshape(-1) if not isinstance(tri, torch.Tensor)
                          else tri.reshape(-1), torch.int32)
        tags_t = to_device(np.asarray(tags, dtype=np.uint8) if not isinstance(tags, torch.Tensor)
                           else tags, torch.uint8)
        self.nt = tri_t.numel() // 3
        xy_t = None
        if xy is not None:
            xy_t = to_device(xy if isinstance(xy, torch.Tensor) else np.asarray(xy, dtype=np.float64),
                             torch.float64)
        h = _vp()
        _check(lib().trn_mesh_create(ctx.handle, self.nv, self.nt, _vp(tri_t.data_ptr()),
                                     _vp(tags_t.data_ptr()),
                                     _vp(xy_t.data_ptr()) if xy_t is not None else None,
                                     ctypes.byref(h)))
        self.handle = h
        nnz, mdt, mdv, mp = _i64(), _i32(), _i32(), _i32()
        _check(lib().trn_mesh_info(h, ctypes.byref(nnz), ctypes.byref(mdt), ctypes.byref(mdv),
                                   ctypes.byref(mp)))
        self.nnz_v2v, self.max_deg_t, self.max_deg_v, self.max_patch = (
            nnz.value, mdt.value, mdv.value, mp.value)

    def csr(self) -> dict:
        """Copies of the device CSR arrays as numpy (for tests)."""
        a, b, c, d = _vp(), _vp(), _vp(), _vp()
        _check(lib().trn_mesh_csr(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                  ctypes.byref(d)))
        self.ctx.synchronize()
        v2t_off = _d2h_i32(a.value, self.nv + 1)
        v2v_off = _d2h_i32(c.value, self.nv + 1)
        return {"v2t_off": v2t_off, "v2t": _d2h_i32(b.value, int(v2t_off[-1])),
                "v2v_off": v2v_off, "v2v": _d2h_i32(d.value, int(v2v_off[-1]))}

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().trn_mesh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_hip = None


def _d2h_i32(ptr: int, n: int) -> np.ndarray:
    """Copies n int32 from library-owned device memory to a numpy array."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    out = np.empty(max(n, 0), dtype=np.int32)
    if n > 0:
        rc = _hip.hipMemcpy(out.ctypes.data_as(_vp), _vp(ptr), 4 * n, 2)  # DeviceToHost
        if rc != 0:
            raise TrinityError(TRN_ERR_HIP, f"hipMemcpy failed ({rc})")
    return out


# ---------------------------------------------------------------------------
# input

def _np_ptr(a):
    return _vp(a.ctypes.data) if a is not None else None


def read_medit(mesh_path: str, bb_path: str | None = None):
    """Mesh::load's parse of a .mesh (and .bb) file on the host: returns
    (xy [nv,2] f64, tri [nt,3] int32 0-based, tags [nv] u8, solut [nv] f64 or
    None) as numpy arrays."""
    nv, nt = _i32(), _i32()
    _check(lib().trn_medit_info(mesh_path.encode(), ctypes.byref(nv), ctypes.byref(nt)))
    xy = np.empty((nv.value, 2), dtype=np.float64)
    tri = np.empty((nt.value, 3), dtype=np.int32)
    tags = np.empty(nv.value, dtype=np.uint8)
    solut = np.empty(nv.value, dtype=np.float64) if bb_path else None
    _check(lib().trn_medit_read(mesh_path.encode(), bb_path.encode() if bb_path else None,
                                nv.value, nt.value, _np_ptr(xy), _np_ptr(tri), _np_ptr(tags),
                                _np_ptr(solut)))
    return xy, tri, tags, solut


def trnm_info(path: str) -> tuple[int, int]:
    nv, nt = _i32(), _i32()
    _check(lib().trn_trnm_info(path.encode(), ctypes.byref(nv), ctypes.byref(nt)))
    return nv.value, nt.value


def read_trnm(path: str, ctx: "Context | None" = None, with_solut: bool = True):
    """A TRNM file into host numpy arrays (ctx None) or device torch tensors
    (ctx given): (xy, tri, tags, solut or None)."""
    nv, nt = trnm_info(path)
    if ctx is None:
        xy = np.empty((nv, 2), dtype=np.float64)
        tri = np.empty((nt, 3), dtype=np.int32)
        tags = np.empty(nv, dtype=np.uint8)
        solut = np.empty(nv, dtype=np.float64) if with_solut else None
        _check(lib().trn_trnm_read(None, path.encode(), nv, nt, _np_ptr(xy), _np_ptr(tri),
                                   _np_ptr(tags), _np_ptr(solut)))
        return xy, tri, tags, solut
    torch = _torch()
    xy = torch.empty((nv, 2), dtype=torch.float64, device="cuda")
    tri = torch.empty((nt, 3), dtype=torch.int32, device="cuda")
    tags = torch.empty(nv, dtype=torch.uint8, device="cuda")
    solut = torch.empty(nv, dtype=torch.float64, device="cuda") if with_solut else None
    _check(lib().trn_trnm_read(ctx.handle, path.encode(), nv, nt, _vp(xy.data_ptr()),
                               _vp(tri.data_ptr()), _vp(tags.data_ptr()),
                               _vp(solut.data_ptr()) if solut is not None else None))
    return xy, tri, tags, solut


def write_trnm(path: str, xy, tri, tags, solut) -> None:
    """Host arrays into a TRNM file."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    tri = np.ascontiguousarray(tri, dtype=np.int32)
    tags = np.ascontiguousarray(tags, dtype=np.uint8)
    solut = np.ascontiguousarray(solut, dtype=np.float64)
    _check(lib().trn_trnm_write(path.encode(), xy.shape[0], tri.size // 3, _np_ptr(xy),
                                _np_ptr(tri), _np_ptr(tags), _np_ptr(solut)))


# ---------------------------------------------------------------------------
# passes

def metric_run(ctx: Context, mesh: DeviceMesh, xy, solut, target_factor=0.8, lp_norm=2,
               h_min=1e-9, h_max=0.4, tensor=None, gradient=None, hessian=None,
               want_complexity=True):
    """Metrics::run on the device; returns (tensor, complexity or None)."""
    torch = _torch()
    if tensor is None:
        tensor = torch.empty((mesh.nv, 3), dtype=torch.float64, device=xy.device)
    c = _f64()
    _check(lib().trn_metric_run(ctx.handle, mesh.handle, _vp(xy.data_ptr()),
                                _vp(solut.data_ptr()), float(target_factor), int(lp_norm),
                                float(h_min), float(h_max), _vp(tensor.data_ptr()),
                                ctypes.byref(c) if want_complexity else None,
                                _vp(_ptr(gradient)) if gradient is not None else None,
                                _vp(_ptr(hessian)) if hessian is not None else None))
    return tensor, (c.value if want_complexity else None)


def gradient(ctx, mesh, xy, solut, out=None):
    torch = _torch()
    out = out if out is not None else torch.empty((mesh.nv, 2), dtype=torch.float64, device=xy.device)
    _check(lib().trn_gradient(ctx.handle, mesh.handle, _vp(xy.data_ptr()), _vp(solut.data_ptr()),
                              _vp(out.data_ptr())))
    return out


def hessian(ctx, mesh, xy, grad, out=None):
    torch = _torch()
    out = out if out is not None else torch.empty((mesh.nv, 3), dtype=torch.float64, device=xy.device)
    _check(lib().trn_hessian(ctx.handle, mesh.handle, _vp(xy.data_ptr()), _vp(grad.data_ptr()),
                             _vp(out.data_ptr())))
    return out


def normalize_local(ctx, tensor, lp_norm):
    _check(lib().trn_normalize_local(ctx.handle, tensor.shape[0], _vp(tensor.data_ptr()),
                                     int(lp_norm)))
    return tensor


def complexity(ctx, mesh, xy, tensor) -> float:
    c = _f64()
    _check(lib().trn_complexity(ctx.handle, mesh.handle, _vp(xy.data_ptr()),
                                _vp(tensor.data_ptr()), ctypes.byref(c)))
    return c.value


def normalize_global(ctx, tensor, scale_fact, h_min=1e-9, h_max=0.4):
    _check(lib().trn_normalize_global(ctx.handle, tensor.shape[0], _vp(tensor.data_ptr()),
                                      float(scale_fact), float(h_min), float(h_max)))
    return tensor


class Coloring:
    def __init__(self, color, order, class_off, ncolors):
        self.color = color          # device int32 [nv]
        self.order = order          # device int32 [nv]
        self.class_off = class_off  # numpy int32 [ncolors+1]
        self.ncolors = ncolors


def color(ctx, mesh, seed=1, max_colors=64) -> Coloring:
    """Partit::extractColoring on the device (seeded, deterministic)."""
    torch = _torch()
    dev = f"cuda:{ctx.device}"
    col = torch.empty(mesh.nv, dtype=torch.int32, device=dev)
    order = torch.empty(mesh.nv, dtype=torch.int32, device=dev)
    off = (ctypes.c_int32 * 80)()
    nc = _i32()
    _check(lib().trn_color(ctx.handle, mesh.handle, int(seed) & ((1 << 64) - 1), int(max_colors),
                           _vp(col.data_ptr()), _vp(order.data_ptr()), off, ctypes.byref(nc)))
    return Coloring(col, order, np.array(off[:nc.value + 1], dtype=np.int32), nc.value)


def color_partial(ctx, mesh, seed=1, max_colors=7) -> tuple[Coloring, int]:
    """Colouring with the reference's palette behaviour: vertices that find
    every colour 1..max_colors taken stay uncoloured (0, in no class).
    Returns (coloring, number of uncoloured vertices)."""
    torch = _torch()
    dev = f"cuda:{ctx.device}"
    col = torch.empty(mesh.nv, dtype=torch.int32, device=dev)
    order = torch.empty(mesh.nv, dtype=torch.int32, device=dev)
    off = (ctypes.c_int32 * 80)()
    nc = _i32()
    unc = _i64()
    _check(lib().trn_color_partial(ctx.handle, mesh.handle, int(seed) & ((1 << 64) - 1),
                                   int(max_colors), _vp(col.data_ptr()), _vp(order.data_ptr()),
                                   off, ctypes.byref(nc), ctypes.byref(unc)))
    return Coloring(col, order, np.array(off[:nc.value + 1], dtype=np.int32), nc.value), unc.value


def set_color_tile(mesh, tile: int) -> None:
    """Tile size of the colouring's first step (0 = chosen from nv)."""
    _check(lib().trn_mesh_set_color_tile(mesh.handle, int(tile)))


SHAPE_NAMES = ("gradient_slabs_per_block", "hessian_slabs_per_block", "phase_waves_per_simd")


def autotune(ctx, mesh, xy, solut, lp_norm=2) -> dict:
    """trn_mesh_autotune: the launch shapes it keeps, and per tuned kernel
    the default shape's time and the chosen one's (ms)."""
    sh = (_i32 * 3)()
    tm = (_f64 * 6)()
    _check(lib().trn_mesh_autotune(ctx.handle, mesh.handle, _vp(xy.data_ptr()),
                                   _vp(solut.data_ptr()), int(lp_norm), sh, tm))
    kernels = ("vtx_grad", "vtx_hess_norm", "smooth_phase")
    return {"shapes": dict(zip(SHAPE_NAMES, list(sh))),
            "ms": {k: {"default": round(tm[2 * i], 4), "tuned": round(tm[2 * i + 1], 4)}
                   for i, k in enumerate(kernels)}}


def mesh_shapes(mesh) -> tuple[int, int, int]:
    sh = (_i32 * 3)()
    _check(lib().trn_mesh_shapes(mesh.handle, sh))
    return tuple(sh)


def set_shapes(mesh, shapes) -> None:
    sh = (_i32 * 3)(*[int(x) for x in shapes])
    _check(lib().trn_mesh_set_shapes(mesh.handle, sh))


def set_class_order(mesh, zcurve: bool) -> None:
    """Colour classes in ascending order (False) or along the Z-curve of the
    coordinates the mesh was created with (True)."""
    _check(lib().trn_mesh_set_class_order(mesh.handle, int(bool(zcurve))))


def color_classes(ctx, color_t, ncolors, mesh=None) -> Coloring:
    order = _torch().empty_like(color_t)
    off = (ctypes.c_int32 * 80)()
    _check(lib().trn_color_classes(ctx.handle, mesh.handle if mesh is not None else None,
                                   color_t.numel(), _vp(color_t.data_ptr()), int(ncolors),
                                   _vp(order.data_ptr()), off))
    return Coloring(color_t, order, np.array(off[:ncolors + 1], dtype=np.int32), int(ncolors))


def indep_set(ctx, mesh, nodes_t, seed=1):
    """Partit::extractIndepSet on the subgraph of the vertices nodes_t
    (device int32): returns the independent vertices (device, input order)."""
    torch = _torch()
    out = torch.empty(max(nodes_t.numel(), 1), dtype=torch.int32, device=nodes_t.device)
    n = _i32()
    _check(lib().trn_indep_set(ctx.handle, mesh.handle, _vp(nodes_t.data_ptr()),
                               int(nodes_t.numel()), int(seed) & ((1 << 64) - 1),
                               _vp(out.data_ptr()), ctypes.byref(n)))
    return out[:n.value]


def color_check(ctx, mesh, color_t, ncolors) -> tuple[int, int]:
    be, bv = _i64(), _i64()
    _check(lib().trn_color_check(ctx.handle, mesh.handle, _vp(color_t.data_ptr()), int(ncolors),
                                 ctypes.byref(be), ctypes.byref(bv)))
    return be.value, bv.value


def quality_cache(ctx, mesh, xy, tensor, out=None):
    torch = _torch()
    out = out if out is not None else torch.empty(mesh.nt, dtype=torch.float64, device=xy.device)
    _check(lib().trn_quality_cache(ctx.handle, mesh.handle, _vp(xy.data_ptr()),
                                   _vp(tensor.data_ptr()), _vp(out.data_ptr())))
    return out


def quality_stats(ctx, mesh, xy, tensor, out=None, quality=(3.0, 0.0, 0.0)):
    """Mesh::computeQuality(double[3]): rebuilds the quality cache and
    returns ((min, max, mean) folded into `quality` as the reference folds
    into its array, active triangle count, cache)."""
    torch = _torch()
    out = out if out is not None else torch.empty(mesh.nt, dtype=torch.float64, device=xy.device)
    q = (_f64 * 3)(*[float(v) for v in quality])
    n = _i64()
    _check(lib().trn_quality_stats(ctx.handle, mesh.handle, _vp(xy.data_ptr()),
                                   _vp(tensor.data_ptr()), _vp(out.data_ptr()), q,
                                   ctypes.byref(n)))
    return (q[0], q[1], q[2]), n.value, out


def refine_filter(ctx, mesh, xy, tensor, l_max=1.414213562373095, steiner=True,
                  edge_cap=None) -> dict:
    """Refine::filterElems + the Steiner points of computeSteinerPoints on the
    device.  Returns pattern, edges (split x 3: v1, v2, opposite), elems,
    Steiner points and tensors (when steiner) and the four counters."""
    torch = _torch()
    dev = xy.device
    cap = int(edge_cap) if edge_cap is not None else 3 * mesh.nt
    pattern = torch.empty(max(mesh.nt, 1), dtype=torch.int32, device=dev)
    edges = torch.empty(3 * max(cap, 1), dtype=torch.int32, device=dev)
    elems = torch.empty(max(mesh.nt, 1), dtype=torch.int32, device=dev)
    sxy = torch.empty(2 * max(cap, 1), dtype=torch.float64, device=dev) if steiner else None
    sten = torch.empty(3 * max(cap, 1), dtype=torch.float64, device=dev) if steiner else None
    counts = (_i32 * 4)()
    _check(lib().trn_refine_filter(ctx.handle, mesh.handle, _vp(xy.data_ptr()),
                                   _vp(tensor.data_ptr()), float(l_max),
                                   _vp(pattern.data_ptr()), _vp(edges.data_ptr()),
                                   _vp(elems.data_ptr()),
                                   _vp(sxy.data_ptr()) if steiner else None,
                                   _vp(sten.data_ptr()) if steiner else None, cap, counts))
    ev, adds, split, tasks = (int(c) for c in counts)
    out = {"pattern": pattern[:mesh.nt], "edges": edges[:3 * split].view(-1, 3),
           "elems": elems[:tasks], "eval": ev, "adds": adds, "split": split, "tasks": tasks}
    if steiner:
        out["steiner_xy"] = sxy[:2 * split].view(-1, 2)
        out["steiner_tensor"] = sten[:3 * split].view(-1, 3)
    return out


def smooth_run(ctx, mesh, xy, tensor, qualit, coloring: Coloring, sweeps: int,
               recompute_qualit=True, counters=True):
    """Smooth::run (after colouring) on the device, in place on xy/tensor/
    qualit.  Returns (commits, tasks) per sweep when counters=True."""
    off = (ctypes.c_int32 * (coloring.ncolors + 1))(*[int(v) for v in coloring.class_off])
    commits = (ctypes.c_int32 * max(sweeps, 1))()
    tasks = (ctypes.c_int32 * max(sweeps, 1))()
    _check(lib().trn_smooth_run(ctx.handle, mesh.handle, _vp(xy.data_ptr()),
                                _vp(tensor.data_ptr()), _vp(qualit.data_ptr()),
                                _vp(coloring.order.data_ptr()), off, int(coloring.ncolors),
                                int(sweeps), int(bool(recompute_qualit)),
                                commits if counters else None, tasks if counters else None))
    if counters:
        return list(commits[:sweeps]), list(tasks[:sweeps])
    return None


def smooth_prepare(ctx, mesh, xy, tensor, qualit, recompute_qualit=True):
    """First part of smooth_run: loads xy/tensor into the mesh's vertex
    records and rebuilds the quality cache; needs no colouring."""
    _check(lib().trn_smooth_prepare(ctx.handle, mesh.handle, _vp(xy.data_ptr()),
                                    _vp(tensor.data_ptr()), _vp(qualit.data_ptr()),
                                    int(bool(recompute_qualit))))


def smooth_run_prepared(ctx, mesh, xy, tensor, qualit, coloring: Coloring, sweeps: int,
                        counters=True):
    """Second part of smooth_run: the sweeps on the prepared records, then
    xy/tensor written back.  Returns (commits, tasks) when counters=True."""
    off = (ctypes.c_int32 * (coloring.ncolors + 1))(*[int(v) for v in coloring.class_off])
    commits = (ctypes.c_int32 * max(sweeps, 1))()
    tasks = (ctypes.c_int32 * max(sweeps, 1))()
    _check(lib().trn_smooth_run_prepared(ctx.handle, mesh.handle, _vp(xy.data_ptr()),
                                         _vp(tensor.data_ptr()), _vp(qualit.data_ptr()),
                                         _vp(coloring.order.data_ptr()), off,
                                         int(coloring.ncolors), int(sweeps),
                                         commits if counters else None,
                                         tasks if counters else None))
    if counters:
        return list(commits[:sweeps]), list(tasks[:sweeps])
    return None


class Batch:
    """Many meshes of uneven sizes processed together (trn_batch): the
    disjoint union of the meshes runs through the plain kernels, every launch
    covering every mesh, with per-mesh complexity and target.  No
    counterpart in the reference, which runs Metrics::run and Smooth::run on
    one mesh at a time.

    meshes: list of (nv, tri[nt,3] 0-based per mesh, tags[nv]).  Field
    arrays passed to `run` are the per-mesh arrays concatenated in list
    order; `voff` / `toff` give each mesh's first vertex / triangle.  xy
    (optional, the meshes' coordinates concatenated, host or device) only
    orders the stored patches along each mesh's Z-curve
    (trn_batch_create_xy)."""

    def __init__(self, ctx: Context, meshes, xy=None):
        self.ctx = ctx
        self.nmesh = len(meshes)
        nv = np.array([int(m[0]) for m in meshes], dtype=np.int32)
        tris = [np.asarray(m[1], dtype=np.int32).reshape(-1, 3) for m in meshes]
        nt = np.array([t.shape[0] for t in tris], dtype=np.int32)
        tri = np.ascontiguousarray(np.concatenate(tris).reshape(-1))
        tags = np.ascontiguousarray(np.concatenate([np.asarray(m[2], dtype=np.uint8)
                                                    for m in meshes]))
        self.voff = np.concatenate([[0], np.cumsum(nv, dtype=np.int64)])
        self.toff = np.concatenate([[0], np.cumsum(nt, dtype=np.int64)])
        self.nv, self.nt = int(self.voff[-1]), int(self.toff[-1])
        h = _vp()
        if xy is None:
            _check(lib().trn_batch_create(ctx.handle, self.nmesh,
                                          nv.ctypes.data_as(_P(_i32)),
                                          nt.ctypes.data_as(_P(_i32)), tri.ctypes.data_as(_vp),
                                          tags.ctypes.data_as(_vp), ctypes.byref(h)))
        else:
            if isinstance(xy, np.ndarray):
                xy_h = np.ascontiguousarray(xy, dtype=np.float64)
                xp = xy_h.ctypes.data_as(_vp)
            else:
                xy_h = xy.contiguous()
                xp = _ptr(xy_h)
            if int(np.prod(xy_h.shape)) != 2 * self.nv:
                raise ValueError("xy must hold 2 values per vertex of the batch")
            _check(lib().trn_batch_create_xy(ctx.handle, self.nmesh,
                                             nv.ctypes.data_as(_P(_i32)),
                                             nt.ctypes.data_as(_P(_i32)),
                                             tri.ctypes.data_as(_vp), tags.ctypes.data_as(_vp),
                                             xp, ctypes.byref(h)))
        self.handle = h
        self.ncolors = 0

    def color(self, seed=1, max_colors=64) -> int:
        nc = _i32()
        _check(lib().trn_batch_color(self.ctx.handle, self.handle, int(seed) & ((1 << 64) - 1),
                                     int(max_colors), ctypes.byref(nc)))
        self.ncolors = nc.value
        return nc.value

    def coloring(self) -> np.ndarray:
        """The union's colouring (host copy, [nv], colours 1..ncolors)."""
        cptr = _vp()
        _check(lib().trn_batch_coloring(self.handle, ctypes.byref(cptr), None, None, None))
        return _d2h_i32(cptr.value, self.nv)

    def run(self, xy_in, xy, solut, tensor, qualit, target_factor=0.8, lp_norm=2, h_min=1e-9,
            h_max=0.4, sweeps=3, complexity=None, counters=False):
        """One pass (metric + quality cache + sweeps) over every mesh.
        Returns (commits, tasks) per sweep, summed over the meshes, when
        counters=True."""
        commits = (_i32 * max(sweeps, 1))()
        tasks = (_i32 * max(sweeps, 1))()
        _check(lib().trn_batch_pass(
            self.ctx.handle, self.handle, _vp(_ptr(xy_in)) if xy_in is not None else None,
            _vp(xy.data_ptr()), _vp(solut.data_ptr()), float(target_factor), int(lp_norm),
            float(h_min), float(h_max), _vp(tensor.data_ptr()), _vp(qualit.data_ptr()),
            int(sweeps), _vp(complexity.data_ptr()) if complexity is not None else None,
            commits if counters else None, tasks if counters else None))
        if counters:
            return list(commits[:sweeps]), list(tasks[:sweeps])
        return None

    def run_graph(self, xy_in, xy, solut, tensor, qualit, target_factor=0.8, lp_norm=2,
                  h_min=1e-9, h_max=0.4, sweeps=3):
        """The same pass replayed from a hipGraph captured on the first call
        with these arguments (needs a created stream, e.g. the context's own:
        Context.use_own_stream)."""
        _check(lib().trn_batch_pass_graph(
            self.ctx.handle, self.handle, _vp(_ptr(xy_in)) if xy_in is not None else None,
            _vp(xy.data_ptr()), _vp(solut.data_ptr()), float(target_factor), int(lp_norm),
            float(h_min), float(h_max), _vp(tensor.data_ptr()), _vp(qualit.data_ptr()),
            int(sweeps)))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().trn_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Adjoint:
    """Energy E = sum_t (1 - q_M(t))^2 of the metric built from (xy, f) and
    its gradient (trn_adjoint).  No counterpart in the reference."""

    def __init__(self, ctx: Context, mesh: DeviceMesh):
        self.ctx, self.mesh = ctx, mesh
        h = _vp()
        _check(lib().trn_adjoint_create(ctx.handle, mesh.handle, ctypes.byref(h)))
        self.handle = h
        self._keep = None

    def forward(self, xy, f, target_factor=0.8, lp_norm=2, h_min=1e-9, h_max=0.4,
                tensor=None, qualit=None) -> float:
        e = _f64()
        self._keep = (xy, f)  # the backward reads them
        _check(lib().trn_adjoint_forward(
            self.ctx.handle, self.handle, _vp(xy.data_ptr()), _vp(f.data_ptr()),
            float(target_factor), int(lp_norm), float(h_min), float(h_max), ctypes.byref(e),
            _vp(tensor.data_ptr()) if tensor is not None else None,
            _vp(qualit.data_ptr()) if qualit is not None else None))
        return e.value

    def backward(self, dxy=None, df=None):
        torch = _torch()
        dev = f"cuda:{self.ctx.device}"
        dxy = dxy if dxy is not None else torch.empty((self.mesh.nv, 2), dtype=torch.float64,
                                                      device=dev)
        df = df if df is not None else torch.empty(self.mesh.nv, dtype=torch.float64, device=dev)
        _check(lib().trn_adjoint_backward(self.ctx.handle, self.handle, _vp(dxy.data_ptr()),
                                          _vp(df.data_ptr())))
        return dxy, df

    def smooth(self, xy, f, target_factor=0.8, lp_norm=2, h_min=1e-9, h_max=0.4, iters=5,
               step=1e-3):
        """Gradient-based smoothing in place on xy; returns (energies, accepted)."""
        en = (_f64 * (iters + 1))()
        acc = _i32()
        self._keep = (xy, f)
        _check(lib().trn_adjoint_smooth(
            self.ctx.handle, self.handle, _vp(xy.data_ptr()), _vp(f.data_ptr()),
            float(target_factor), int(lp_norm), float(h_min), float(h_max), int(iters),
            float(step), en, ctypes.byref(acc)))
        return list(en), acc.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().trn_adjoint_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def debug_math(ctx, op: int, x, y=None):
    torch = _torch()
    out = torch.empty_like(x)
    _check(lib().trn_debug_math(ctx.handle, int(op), x.numel(), _vp(x.data_ptr()),
                                _vp(y.data_ptr()) if y is not None else None, _vp(out.data_ptr())))
    return out
