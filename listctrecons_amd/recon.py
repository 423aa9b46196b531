"""Python host API over the C ABI (include/lct.h).

Mirrors the reference's plug-in structure (Include/Rebiner.h, Filter.h,
Backprojector.h, Reconstructor.h): stage calls with the reference's per-block
semantics, and ``Context.reconstruct`` for the whole FDK loop of
Src/Reconstructor.cpp:30-49.  Device memory is torch CUDA tensors (plumbing);
every computation runs in liblct.so.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import sys

import numpy as np

from . import _lib
from ._lib import check, lib


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _check_buf(t, n: int, name: str, dtype=None) -> None:
    """A contiguous CUDA tensor of `dtype` (float32 by default) holding at
    least n elements (what the kernel will index)."""
    import torch
    dtype = torch.float32 if dtype is None else dtype
    if not isinstance(t, torch.Tensor) or t.dtype != dtype or t.numel() < int(n):
        got = f"{t.dtype} x {t.numel()}" if isinstance(t, torch.Tensor) else type(t).__name__
        raise ValueError(f"{name}: expected {dtype} with at least {int(n)} elements, got {got}")


def _storage_dtype(storage: int):
    import torch
    return {_lib.F32: torch.float32, _lib.F16: torch.float16, _lib.BF16: torch.bfloat16}[storage]


def _check_f32(t, shape, name: str) -> None:
    """A float32 tensor of exactly `shape` (what the kernel will index)."""
    import torch
    if t.dtype != torch.float32 or tuple(t.shape) != tuple(int(x) for x in shape):
        raise ValueError(f"{name}: expected float32 of shape {tuple(shape)}, "
                         f"got {t.dtype} {tuple(t.shape)}")


def _dev(t) -> C.c_void_p:
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous():
        raise ValueError("expected a contiguous CUDA tensor")
    return C.c_void_p(t.data_ptr())


# ------------------------------------------------------------ host planner

def plan_helical(prm: _lib.Prm, angles: np.ndarray, couch: np.ndarray) -> _lib.Plan:
    """ReconData::convert_from_file_params (Src/DataStructs.cpp:121-250)."""
    a = np.ascontiguousarray(angles, np.float32)
    c = np.ascontiguousarray(couch, np.float32)
    if a.size < prm.readings or c.size < prm.readings:
        raise ValueError("angle / couch arrays shorter than Readings")
    out = _lib.Plan()
    check(lib().lct_plan_helical(C.byref(prm), _fptr(a), _fptr(c), C.byref(out)))
    return out


def block_info(plan: _lib.Plan, couch: np.ndarray, block: int):
    """BlockInfo::update (Src/DataStructs.cpp:77-108) ->
    (proj_start, proj_end, block_start, block_end, ref_view)."""
    c = np.ascontiguousarray(couch, np.float32)
    s, e, r = C.c_int(), C.c_int(), C.c_int()
    bs, be = C.c_float(), C.c_float()
    check(lib().lct_block_info(C.byref(plan), _fptr(c), block, C.byref(s), C.byref(e),
                               C.byref(bs), C.byref(be), C.byref(r)))
    return s.value, e.value, bs.value, be.value, r.value


def slab_views(plan: _lib.Plan, couch: np.ndarray, block_begin: int, block_end: int):
    """Projection-file views [begin, end) that blocks [block_begin, block_end)
    read (the union of their BlockInfo windows, DataOffset applied)."""
    c = np.ascontiguousarray(couch, np.float32)
    vb, ve = C.c_long(), C.c_long()
    check(lib().lct_slab_views(C.byref(plan), _fptr(c), block_begin, block_end, C.byref(vb),
                               C.byref(ve)))
    return vb.value, ve.value


def ramp_taps(window: int, ns: int, ds: float) -> np.ndarray:
    """RampFilterGPU::calculate_filter (Src/Filter.cpp:8-20) + apodisation."""
    h = np.empty(2 * ns, np.float32)
    check(lib().lct_ramp_taps(window, ns, C.c_float(ds), _fptr(h)))
    return h


def block_betas(plan: _lib.Plan, angles: np.ndarray, proj_start: int) -> np.ndarray:
    """betas of a block (CUDA/Backprojector.cu:145) padded to n_angle + npt/2 + 1."""
    g = plan.ctg
    n = g.n_angle + g.n_proj_turn // 2 + 1
    out = np.zeros(n, np.float32)
    src = proj_start + g.add_projections
    avail = max(0, min(n, angles.size - src))
    out[:avail] = angles[src:src + avail]
    return out


# ------------------------------------------------------------ options

@dataclasses.dataclass
class Options:
    pipeline: int = _lib.PIPE_BLOCKS
    rebin_kind: int = _lib.REBIN_FAN2PAR
    weight_kind: int = _lib.WEIGHT_NONE
    window: int = _lib.WINDOW_NONE
    filter_algo: int = _lib.FILTER_DIRECT_F32
    storage: int = _lib.F32
    bp_kind: int = _lib.BP_HELICAL
    interp: int = _lib.INTERP_LDS

    def to_c(self) -> _lib.ReconOpts:
        o = _lib.ReconOpts()
        for f in dataclasses.fields(self):
            setattr(o, f.name, int(getattr(self, f.name)))
        return o


def preset(name: str, window: int = _lib.WINDOW_NONE) -> Options:
    """Named paths: "plain" is the reference's per-block FP32 loop on the GPU;
    "fast_f32" / "mixed_f16" / "mixed_bf16" are the HBM-resident paths with
    f32 / f16 / bf16 filtered-projection storage (FP32 accumulation);
    "fast_f32_x6" is "fast_f32" with the six-product BF16 Toeplitz MFMA filter
    (FP32-level accuracy) instead of the FP32 FMA one."""
    if name == "plain":
        return Options(window=window)
    if name == "fast_f32":
        return Options(pipeline=_lib.PIPE_RESIDENT, window=window)
    if name == "mixed_f16":
        return Options(pipeline=_lib.PIPE_RESIDENT, window=window, storage=_lib.F16,
                       filter_algo=_lib.FILTER_TOEPLITZ_BF16X3)
    if name == "mixed_bf16":
        return Options(pipeline=_lib.PIPE_RESIDENT, window=window, storage=_lib.BF16,
                       filter_algo=_lib.FILTER_TOEPLITZ_BF16X3)
    if name == "resident_direct_f16":
        return Options(pipeline=_lib.PIPE_RESIDENT, window=window, storage=_lib.F16)
    if name == "fast_f32_x6":
        return Options(pipeline=_lib.PIPE_RESIDENT, window=window,
                       filter_algo=_lib.FILTER_TOEPLITZ_BF16X6)
    if name == "fast_f32_toeplitz":
        return Options(pipeline=_lib.PIPE_RESIDENT, window=window,
                       filter_algo=_lib.FILTER_TOEPLITZ_BF16X3)
    if name.endswith("_pi_ori"):
        # the PI-original stage set (Rebiner.cu:156-233, PreWeighting.cu:8-44,
        # Backprojector.cu:171-307) on the named path: "plain_pi_ori" is the
        # reference's per-block loop, the resident ones fold its view pairs
        o = preset(name[: -len("_pi_ori")], window)
        o.rebin_kind, o.weight_kind, o.bp_kind = _lib.REBIN_PI_ORI, _lib.WEIGHT_PI_ORI, _lib.BP_PI_ORI
        return o
    if name.endswith("_gather"):
        o = preset(name[: -len("_gather")], window)
        o.interp = _lib.INTERP_GATHER
        return o
    raise KeyError(name)


# ------------------------------------------------------------ context

class Context:
    """One lct_ctx on one device (CTDataGPU's role, Src/DataStructs.cpp:57-66)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().lct_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def close(self) -> None:
        if self.h:
            lib().lct_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the HIP runtime may already be torn down: the
        # process releases device memory itself (the default argument keeps
        # the check reachable after module globals are cleared)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self) -> None:
        check(lib().lct_synchronize(self.h))

    def set_stream(self, stream: int) -> None:
        """Run this context's work on a HIP stream handle (e.g. a torch
        stream's `cuda_stream`); 0 is the default stream."""
        check(lib().lct_ctx_set_stream(self.h, C.c_void_p(stream)))

    def set_profiling(self, on: bool) -> None:
        check(lib().lct_set_profiling(self.h, 1 if on else 0))

    def timings(self) -> dict:
        t = _lib.Timings()
        check(lib().lct_get_timings(self.h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in _lib.Timings._fields_}

    # -------------------------------------------------------- stages
    def rebin(self, kind: int, weight: int, g: _lib.CTGeom, fan, par) -> None:
        _check_buf(fan, g.n_angle_add * g.nt * g.ns, "fan")
        _check_buf(par, g.n_angle * g.nt * g.ns, "par")
        check(lib().lct_rebin(self.h, kind, weight, C.byref(g), _dev(fan), _dev(par)))

    def weighting(self, kind: int, g: _lib.CTGeom, par) -> None:
        _check_buf(par, g.n_angle * g.nt * g.ns, "par")
        check(lib().lct_weighting(self.h, kind, C.byref(g), _dev(par)))

    def filter(self, g: _lib.CTGeom, taps: np.ndarray, rows, n_rows: int) -> None:
        t = np.ascontiguousarray(taps, np.float32)
        if t.size < 2 * g.ns:
            raise ValueError("taps needs 2 ns values")
        _check_buf(rows, n_rows * g.ns, "rows")
        check(lib().lct_filter(self.h, C.byref(g), _fptr(t), _dev(rows), n_rows))

    def filter_to_bp_layout(self, algo: int, storage: int, ns: int, nt: int, taps: np.ndarray,
                            rows, views: int, q) -> None:
        """Ramp filter of rows [views][nt][ns] into Q [views][ns+3][nt+8] (q pre-zeroed)."""
        t = np.ascontiguousarray(taps, np.float32)
        if t.size < 2 * ns:
            raise ValueError("taps needs 2 ns values")
        _check_buf(rows, views * nt * ns, "rows")
        _check_buf(q, views * (ns + 3) * (nt + 8), "q", _storage_dtype(storage))
        check(lib().lct_filter_to_bp_layout(self.h, algo, storage, ns, nt, _fptr(t), _dev(rows),
                                            views, _dev(q)))

    def backproject(self, kind: int, g: _lib.CTGeom, ig: _lib.ImgGeom, q, betas, img) -> None:
        b = np.ascontiguousarray(betas, np.float32)
        if b.size < g.n_angle + g.n_proj_turn // 2 + 1:
            raise ValueError("betas needs n_angle + n_proj_turn/2 + 1 values")
        _check_buf(q, g.n_angle * g.nt * g.ns, "q")
        _check_buf(img, ig.nz * ig.ny * ig.nx, "img")
        check(lib().lct_backproject(self.h, kind, C.byref(g), C.byref(ig), _dev(q), _fptr(b),
                                    _dev(img)))

    def backproject_noalloc(self, kind: int, g: _lib.CTGeom, ig: _lib.ImgGeom, q, betas,
                            img) -> None:
        """lct_backproject_noalloc: backproject() without any allocation (its
        tables in the library's fixed per-GPU array)."""
        b = np.ascontiguousarray(betas, np.float32)
        if b.size < g.n_angle + g.n_proj_turn // 2 + 1:
            raise ValueError("betas needs n_angle + n_proj_turn/2 + 1 values")
        _check_buf(q, g.n_angle * g.nt * g.ns, "q")
        _check_buf(img, ig.nz * ig.ny * ig.nx, "img")
        check(lib().lct_backproject_noalloc(self.h, kind, C.byref(g), C.byref(ig), _dev(q),
                                            _fptr(b), _dev(img)))

    # -------------------------------------------------------- pipeline
    def reconstruct(self, plan: _lib.Plan, opts: Options, angles: np.ndarray,
                    couch: np.ndarray, fan, vol=None):
        """FDKReconstructor::reconstruct over all planned blocks; fan is the
        whole projection file on the device ([views][nt][ns] float32)."""
        import torch
        a = np.ascontiguousarray(angles, np.float32)
        c = np.ascontiguousarray(couch, np.float32)
        g = plan.ctg
        _check_tables(plan, a, c)
        _check_buf(fan, g.nt * g.ns, "fan")
        views = fan.numel() // (g.nt * g.ns)
        if vol is None:
            vol = torch.empty((plan.n_slices_recon, plan.ig.ny, plan.ig.nx), dtype=torch.float32,
                              device=fan.device)
        _check_f32(vol, (plan.n_slices_recon, plan.ig.ny, plan.ig.nx), "vol")
        oc = opts.to_c()
        check(lib().lct_reconstruct(self.h, C.byref(plan), C.byref(oc), _fptr(a), _fptr(c),
                                    _dev(fan), views, _dev(vol)))
        return vol

    def last_block_filtered(self, plan: _lib.Plan, out=None):
        """Filtered views [n_angle][nt][ns] of the last reconstructed block
        (what the reference writes to rebin.raw)."""
        import torch
        g = plan.ctg
        if out is None:
            out = torch.empty((g.n_angle, g.nt, g.ns), dtype=torch.float32, device="cuda")
        _check_f32(out, (g.n_angle, g.nt, g.ns), "out")
        check(lib().lct_last_block_filtered(self.h, _dev(out)))
        return out

    def reconstruct_slabs(self, plan: _lib.Plan, opts: Options, angles: np.ndarray,
                          couch: np.ndarray, fan, fan_view0: int, block_begin: int,
                          block_end: int, vol=None):
        """Blocks [block_begin, block_end) only; fan holds projection-file
        views starting at fan_view0.  Returns (block_end - block_begin) * 32
        slices."""
        import torch
        a = np.ascontiguousarray(angles, np.float32)
        c = np.ascontiguousarray(couch, np.float32)
        g = plan.ctg
        _check_tables(plan, a, c)
        _check_buf(fan, g.nt * g.ns, "fan")
        views = fan.numel() // (g.nt * g.ns)
        if vol is None:
            vol = torch.empty(((block_end - block_begin) * 32, plan.ig.ny, plan.ig.nx),
                              dtype=torch.float32, device=fan.device)
        _check_f32(vol, ((block_end - block_begin) * 32, plan.ig.ny, plan.ig.nx), "vol")
        oc = opts.to_c()
        check(lib().lct_reconstruct_slabs(self.h, C.byref(plan), C.byref(oc), _fptr(a), _fptr(c),
                                          _dev(fan), fan_view0, views, block_begin, block_end,
                                          _dev(vol)))
        return vol

    # -------------------------------------------------------- synthetic data
    def project_ellipsoids(self, g: _lib.CTGeom, angles, couch, ell: np.ndarray, out,
                           view0: int = 0, n_views: int | None = None) -> None:
        a = np.ascontiguousarray(angles, np.float32)
        c = np.ascontiguousarray(couch, np.float32)
        e = np.ascontiguousarray(ell, np.float64).reshape(-1, 16)
        if n_views is None:
            n_views = a.size - view0
        if view0 < 0 or n_views < 0 or view0 + n_views > min(a.size, c.size):
            raise ValueError("views outside the angle / couch tables")
        _check_buf(out, n_views * g.nt * g.ns, "out")
        check(lib().lct_project_ellipsoids(self.h, C.byref(g), _fptr(a), _fptr(c), view0,
                                           n_views, e.ctypes.data_as(C.POINTER(C.c_double)),
                                           e.shape[0], _dev(out)))

    def forward_project(self, g: _lib.CTGeom, angles, couch, vol, ig: _lib.ImgGeom,
                        z0: float, out, view0: int = 0, n_views: int | None = None,
                        step: float = 0.25) -> None:
        """Reprojection of vol [nz][ny][nx] (slice k at z0 + k dz) into out [views][nt][ns]."""
        a = np.ascontiguousarray(angles, np.float32)
        c = np.ascontiguousarray(couch, np.float32)
        if n_views is None:
            n_views = a.size - view0
        _check_f32(vol, (vol.shape[0], ig.ny, ig.nx), "vol")
        _check_f32(out, (n_views, g.nt, g.ns), "out")
        if view0 < 0 or view0 + n_views > min(a.size, c.size):
            raise ValueError("views outside the angle / couch tables")
        check(lib().lct_forward_project(self.h, C.byref(g), _fptr(a), _fptr(c), view0, n_views,
                                        _dev(vol), C.byref(ig), C.c_float(z0), vol.shape[0],
                                        C.c_float(step), _dev(out)))

    def forward_project_adjoint(self, g: _lib.CTGeom, angles, couch, proj, ig: _lib.ImgGeom,
                                z0: float, out, view0: int = 0, step: float = 0.25) -> None:
        """Transpose of forward_project: proj [views][nt][ns] -> out [nz][ny][nx]."""
        a = np.ascontiguousarray(angles, np.float32)
        c = np.ascontiguousarray(couch, np.float32)
        _check_f32(proj, (proj.shape[0], g.nt, g.ns), "proj")
        _check_f32(out, (out.shape[0], ig.ny, ig.nx), "out")
        if view0 < 0 or view0 + proj.shape[0] > min(a.size, c.size):
            raise ValueError("views outside the angle / couch tables")
        check(lib().lct_forward_project_adjoint(self.h, C.byref(g), _fptr(a), _fptr(c), view0,
                                                proj.shape[0], _dev(proj), C.byref(ig),
                                                C.c_float(z0), out.shape[0], C.c_float(step),
                                                _dev(out)))

    def voxelize(self, ig: _lib.ImgGeom, z0: float, nz: int, ell: np.ndarray, out,
                 supersample: int = 2) -> None:
        e = np.ascontiguousarray(ell, np.float64).reshape(-1, 16)
        _check_buf(out, nz * ig.ny * ig.nx, "out")
        check(lib().lct_voxelize_ellipsoids(self.h, C.byref(ig), float(z0), nz,
                                            e.ctypes.data_as(C.POINTER(C.c_double)), e.shape[0],
                                            supersample, _dev(out)))

    def poisson_noise(self, proj, i0: float, seed: int) -> None:
        _check_buf(proj, proj.numel(), "proj")
        check(lib().lct_poisson_noise(self.h, _dev(proj), proj.numel(), C.c_float(i0),
                                      C.c_uint64(seed)))


def _check_tables(plan: _lib.Plan, angles: np.ndarray, couch: np.ndarray) -> None:
    """The planner read plan.n_readings angles and couch positions; the
    pipeline reads the same."""
    if angles.size < plan.n_readings or couch.size < plan.n_readings:
        raise ValueError(f"angle / couch tables shorter than the plan's {plan.n_readings} readings")


def scan_geometry(w) -> _lib.CTGeom:
    """CTGeom fields the projector needs for a workload (ws, wt, ds, dt, dso)."""
    g = _lib.CTGeom()
    g.ns, g.nt = w.ns, w.nt
    g.ws, g.wt = w.central_channel, (w.nt - 1) / 2.0
    g.ds, g.dt, g.dso = w.ds, w.dt, w.dso
    return g


def synth_scan(ctx: Context, w, noise: bool = True):
    """Projection file of workload `w` on the device: [data_offset + readings][nt][ns]."""
    import torch
    angles, couch = w.angles_couch()
    g = scan_geometry(w)
    fan = torch.zeros((w.data_offset + w.readings, w.nt, w.ns), dtype=torch.float32,
                      device=f"cuda:{ctx.device}")
    body = fan[w.data_offset:]
    ctx.project_ellipsoids(g, angles, couch, w.ellipsoids(), body)
    if noise:
        ctx.poisson_noise(body, w.i0, w.noise_seed)
    ctx.synchronize()
    return fan, angles, couch


class Cnn:
    """The learned residual correction y = x + net(x) (lct_cnn_* in lct.h)."""

    def __init__(self, ctx: Context, seed: int = 0):
        h = C.c_void_p()
        check(lib().lct_cnn_create(ctx.h, C.c_uint64(seed), C.byref(h)))
        self.h = h
        self.ctx = ctx
        self.n_params = int(lib().lct_cnn_param_count())

    def close(self) -> None:
        if self.h:
            lib().lct_cnn_destroy(self.h)
            self.h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the HIP runtime may already be torn down: the
        # process releases device memory itself (the default argument keeps
        # the check reachable after module globals are cleared)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        check(lib().lct_cnn_get_params(self.h, _fptr(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        q = np.ascontiguousarray(p, np.float32)
        assert q.size == self.n_params
        check(lib().lct_cnn_set_params(self.h, _fptr(q)))

    def grads(self) -> np.ndarray:
        g = np.empty(self.n_params, np.float32)
        check(lib().lct_cnn_get_grads(self.h, _fptr(g)))
        return g

    def forward(self, x, y=None):
        """x, y: CUDA float32 tensors [B][D][H][W]."""
        import torch
        if y is None:
            y = torch.empty_like(x)
        B, D, H, W = x.shape
        _check_f32(x, (B, D, H, W), "x")
        _check_f32(y, (B, D, H, W), "y")
        check(lib().lct_cnn_forward(self.h, _dev(x), B, D, H, W, _dev(y)))
        return y

    def loss_grad(self, x, t) -> float:
        B, D, H, W = x.shape
        _check_f32(x, (B, D, H, W), "x")
        _check_f32(t, (B, D, H, W), "t")
        loss = np.zeros(1, np.float32)
        check(lib().lct_cnn_loss_grad(self.h, _dev(x), _dev(t), B, D, H, W, _fptr(loss)))
        return float(loss[0])

    def adam_step(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
                  eps: float = 1e-8) -> None:
        check(lib().lct_cnn_adam_step(self.h, C.c_float(lr), C.c_float(beta1),
                                      C.c_float(beta2), C.c_float(eps)))
