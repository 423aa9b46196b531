"""ctypes binding of the FP64 CPU oracle (oracle/liblct_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Imported by tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg, never by the
product package.  Every function restates the reference file:line named in
``lct_oracle.c``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblct_oracle.so")


class CTGeom(C.Structure):
    """CTGeom of Include/DataStructs.h:9-100 (same field order as lct_ct_geom)."""

    _fields_ = [
        ("orbit", C.c_float), ("orbit_start", C.c_float),
        ("ns", C.c_int), ("nt", C.c_int),
        ("ws", C.c_float), ("wt", C.c_float),
        ("n_angle", C.c_int), ("n_angle_add", C.c_int),
        ("ds", C.c_float), ("dt", C.c_float), ("pitch", C.c_float),
        ("source_z0", C.c_float), ("dso", C.c_float),
        ("n_proj_turn", C.c_int), ("z_dir", C.c_int), ("add_projections", C.c_int),
    ]


class ImgGeom(C.Structure):
    """ImgGeom of Include/DataStructs.h:102-156."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("wx", C.c_float), ("wy", C.c_float), ("wz", C.c_float),
        ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
    ]


class Prm(C.Structure):
    """Numeric keys of the .prm file (Src/FileParams.cpp:10)."""

    _fields_ = [
        ("ns", C.c_int), ("nt", C.c_int),
        ("ds", C.c_float), ("dt", C.c_float), ("pitch_value", C.c_float),
        ("nx", C.c_int), ("ny", C.c_int),
        ("dx", C.c_float), ("dz", C.c_float), ("offset_x", C.c_float), ("offset_y", C.c_float),
        ("readings", C.c_int), ("view_offset", C.c_int), ("data_offset", C.c_int),
        ("r_src_to_iso", C.c_float), ("central_channel", C.c_float),
        ("n_proj_turn", C.c_int),
    ]


class Plan(C.Structure):
    """ReconInfo (Include/DataStructs.h:225-273) plus the derived geometry."""

    _fields_ = [
        ("ctg", CTGeom), ("ig", ImgGeom),
        ("n_readings", C.c_int),
        ("data_begin_pos", C.c_float), ("data_end_pos", C.c_float),
        ("allow_begin", C.c_float), ("allow_end", C.c_float),
        ("start_pos", C.c_float), ("end_pos", C.c_float),
        ("recon_start_pos", C.c_float), ("recon_end_pos", C.c_float),
        ("n_slices_recon", C.c_int), ("n_slices_block", C.c_int),
        ("n_blocks", C.c_int), ("data_offset", C.c_int),
        ("recon_direction", C.c_int),
    ]


ELL_STRIDE = 16

_lib = None


def build(force: bool = False) -> str:
    """Compile liblct_oracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        f32p, f64p, i32p = P(C.c_float), P(C.c_double), P(C.c_int)
        L.lco_array_search.argtypes = [C.c_float, f32p, C.c_int, C.c_int]
        L.lco_plan_helical.argtypes = [P(Prm), f32p, f32p, P(Plan)]
        L.lco_block_info.argtypes = [P(Plan), f32p, C.c_int, i32p, i32p,
                                     P(C.c_float), P(C.c_float)]
        L.lco_ramp_taps.argtypes = [C.c_int, C.c_double, C.c_int, f64p]
        for name in ("lco_rebin_fan2par", "lco_rebin_null", "lco_rebin_pi_ori"):
            getattr(L, name).argtypes = [P(CTGeom), f32p, f64p]
        L.lco_weight_pi_ori.argtypes = [P(CTGeom), f64p]
        L.lco_filter_rows.argtypes = [C.c_int, f64p, f64p, f64p, C.c_long]
        for name in ("lco_bp_helical", "lco_bp_pi_ori"):
            getattr(L, name).argtypes = [P(CTGeom), P(ImgGeom), f64p, f32p, i32p, C.c_int,
                                         C.c_int, C.c_int, f64p]
        L.lco_reconstruct.argtypes = [P(Plan), f32p, C.c_long, f32p, f32p, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_int, i32p, C.c_int,
                                      C.c_int, C.c_int, f64p]
        L.lco_project_ellipsoids.argtypes = [P(CTGeom), f32p, f32p, C.c_long, C.c_long,
                                             f64p, C.c_int, f64p]
        L.lco_voxelize.argtypes = [P(ImgGeom), C.c_double, C.c_int, f64p, C.c_int, C.c_int,
                                   f64p]
        L.lco_forward_project.argtypes = [P(CTGeom), f32p, f32p, C.c_long, C.c_long, f64p,
                                          P(ImgGeom), C.c_double, C.c_int, C.c_double, f64p]
        L.lco_forward_project_adjoint.argtypes = [P(CTGeom), f32p, f32p, C.c_long, C.c_long,
                                                  f64p, P(ImgGeom), C.c_double, C.c_int,
                                                  C.c_double, f64p]
        L.lco_num_threads.restype = C.c_int
        L.lco_set_num_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def num_threads() -> int:
    return lib().lco_num_threads()


def set_num_threads(n: int) -> None:
    lib().lco_set_num_threads(int(n))


def array_search(key: float, array: np.ndarray, search_type: int) -> int:
    a = _f32(array)
    return lib().lco_array_search(C.c_float(key), _p(a, C.c_float), a.size, search_type)


def plan_helical(prm: Prm, angles: np.ndarray, couch: np.ndarray) -> Plan:
    """Src/DataStructs.cpp:121-250.  Raises ValueError on the reference's exit(1) paths."""
    a, c = _f32(angles), _f32(couch)
    out = Plan()
    rc = lib().lco_plan_helical(C.byref(prm), _p(a, C.c_float), _p(c, C.c_float), C.byref(out))
    if rc == -2:
        raise ValueError("Requested reconstruction is outside of allowed data range")
    if rc != 0:
        raise ValueError("invalid planner arguments")
    return out


def block_info(plan: Plan, couch: np.ndarray, block: int):
    """Src/DataStructs.cpp:77-108 -> (proj_start, proj_end, block_start, block_end)."""
    c = _f32(couch)
    s, e = C.c_int(), C.c_int()
    bs, be = C.c_float(), C.c_float()
    lib().lco_block_info(C.byref(plan), _p(c, C.c_float), block, C.byref(s), C.byref(e),
                         C.byref(bs), C.byref(be))
    return s.value, e.value, bs.value, be.value


def ramp_taps(ns: int, ds: float, window: int = 0) -> np.ndarray:
    h = np.empty(2 * ns, dtype=np.float64)
    lib().lco_ramp_taps(ns, float(ds), window, _p(h, C.c_double))
    return h


def rebin(g: CTGeom, fan: np.ndarray, kind: int = 1) -> np.ndarray:
    """kind 0 null, 1 fan2par, 2 pi_ori.  fan [NA+][nt][ns] -> par [NA][nt][ns] float64."""
    f = _f32(fan)
    assert f.size == g.n_angle_add * g.nt * g.ns
    out = np.empty((g.n_angle, g.nt, g.ns), dtype=np.float64)
    fn = {0: lib().lco_rebin_null, 1: lib().lco_rebin_fan2par, 2: lib().lco_rebin_pi_ori}[kind]
    fn(C.byref(g), _p(f, C.c_float), _p(out, C.c_double))
    return out


def weight_pi_ori(g: CTGeom, par: np.ndarray) -> np.ndarray:
    out = _f64(par).copy()
    lib().lco_weight_pi_ori(C.byref(g), _p(out, C.c_double))
    return out


def filter_rows(h: np.ndarray, rows: np.ndarray) -> np.ndarray:
    x = _f64(rows)
    ns = x.shape[-1]
    hh = _f64(h)
    assert hh.size == 2 * ns
    out = np.empty_like(x)
    lib().lco_filter_rows(ns, _p(hh, C.c_double), _p(x, C.c_double), _p(out, C.c_double),
                          x.size // ns)
    return out


def backproject(g: CTGeom, ig: ImgGeom, q: np.ndarray, betas: np.ndarray, kind: int = 0,
                iz_list=None, iy_range=None) -> np.ndarray:
    """kind 0: CUDA/Backprojector.cu:12-99, kind 1: :171-250.

    betas must hold n_angle + n_proj_turn//2 + 1 angles (conjugate indices of
    clipped windows reach past n_angle).  Returns img [nz][ny][nx] (NaN where
    not requested)."""
    qq = _f64(q)
    assert qq.size == g.n_angle * g.nt * g.ns
    b = _f32(betas)
    if b.size < g.n_angle + g.n_proj_turn // 2 + 1:
        b = np.concatenate([b, np.zeros(g.n_angle + g.n_proj_turn // 2 + 1 - b.size, np.float32)])
    iz = np.arange(ig.nz, dtype=np.int32) if iz_list is None else np.asarray(iz_list, np.int32)
    iy0, iy1 = (0, ig.ny) if iy_range is None else iy_range
    img = np.full((ig.nz, ig.ny, ig.nx), np.nan, dtype=np.float64)
    fn = lib().lco_bp_pi_ori if kind == 1 else lib().lco_bp_helical
    fn(C.byref(g), C.byref(ig), _p(qq, C.c_double), _p(b, C.c_float), _p(iz, C.c_int),
       iz.size, iy0, iy1, _p(img, C.c_double))
    return img


def reconstruct(plan: Plan, fan_file: np.ndarray, angles: np.ndarray, couch: np.ndarray,
                rebin_kind: int = 1, weight_kind: int = 0, window: int = 0, bp_kind: int = 0,
                blocks=None, iz_list=None, iy_range=None) -> np.ndarray:
    """Src/Reconstructor.cpp:33-48 over blocks [b0, b1).  Returns vol
    [n_slices_recon][ny][nx] float64 (NaN where not computed)."""
    f = _f32(fan_file)
    a, c = _f32(angles), _f32(couch)
    ns, nt = plan.ctg.ns, plan.ctg.nt
    views = f.size // (ns * nt)
    b0, b1 = (0, plan.n_blocks) if blocks is None else blocks
    iz = np.arange(32, dtype=np.int32) if iz_list is None else np.asarray(iz_list, np.int32)
    iy0, iy1 = (0, plan.ig.ny) if iy_range is None else iy_range
    vol = np.full((plan.n_slices_recon, plan.ig.ny, plan.ig.nx), np.nan, dtype=np.float64)
    rc = lib().lco_reconstruct(C.byref(plan), _p(f, C.c_float), views, _p(a, C.c_float),
                               _p(c, C.c_float), rebin_kind, weight_kind, window, bp_kind,
                               b0, b1, _p(iz, C.c_int), iz.size, iy0, iy1,
                               _p(vol, C.c_double))
    if rc != 0:
        raise ValueError(f"oracle reconstruct failed ({rc}): projection window outside data")
    return vol


def project_ellipsoids(g: CTGeom, angles: np.ndarray, couch: np.ndarray, ell: np.ndarray,
                       view0: int = 0, n_views: int | None = None) -> np.ndarray:
    a, c = _f32(angles), _f32(couch)
    e = _f64(ell).reshape(-1, ELL_STRIDE)
    if n_views is None:
        n_views = a.size - view0
    out = np.empty((n_views, g.nt, g.ns), dtype=np.float64)
    lib().lco_project_ellipsoids(C.byref(g), _p(a, C.c_float), _p(c, C.c_float), view0,
                                 n_views, _p(e, C.c_double), e.shape[0], _p(out, C.c_double))
    return out


def forward_project(g: CTGeom, angles: np.ndarray, couch: np.ndarray, vol: np.ndarray,
                    ig: ImgGeom, z0: float, step: float, view0: int = 0,
                    n_views: int | None = None) -> np.ndarray:
    """Reprojection of vol [nz][ny][nx] (trilinear, midpoint rule)."""
    a, c = _f32(angles), _f32(couch)
    v = _f64(vol)
    if n_views is None:
        n_views = a.size - view0
    out = np.empty((n_views, g.nt, g.ns), dtype=np.float64)
    lib().lco_forward_project(C.byref(g), _p(a, C.c_float), _p(c, C.c_float), view0, n_views,
                              _p(v, C.c_double), C.byref(ig), float(z0), v.shape[0],
                              float(step), _p(out, C.c_double))
    return out


def forward_project_adjoint(g: CTGeom, angles: np.ndarray, couch: np.ndarray,
                            proj: np.ndarray, ig: ImgGeom, z0: float, nz: int, step: float,
                            view0: int = 0) -> np.ndarray:
    """Transpose of forward_project: proj [n_views][nt][ns] -> vol [nz][ny][nx]."""
    a, c = _f32(angles), _f32(couch)
    p = _f64(proj)
    vol = np.empty((nz, ig.ny, ig.nx), dtype=np.float64)
    lib().lco_forward_project_adjoint(C.byref(g), _p(a, C.c_float), _p(c, C.c_float), view0,
                                      p.shape[0], _p(p, C.c_double), C.byref(ig), float(z0), nz,
                                      float(step), _p(vol, C.c_double))
    return vol


def voxelize(ig: ImgGeom, z0: float, nz: int, ell: np.ndarray, supersample: int = 1):
    e = _f64(ell).reshape(-1, ELL_STRIDE)
    vol = np.empty((nz, ig.ny, ig.nx), dtype=np.float64)
    lib().lco_voxelize(C.byref(ig), float(z0), nz, _p(e, C.c_double), e.shape[0],
                       supersample, _p(vol, C.c_double))
    return vol
