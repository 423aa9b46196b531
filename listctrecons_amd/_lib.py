"""ctypes binding of liblct.so (include/lct.h).

The library is built in tree by ``python -m listctrecons_amd.build``.  There
is no fallback: if the library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblct.so")

LCT_OK = 0
LCT_ERR_INVALID = -1
LCT_ERR_HIP = -2
LCT_ERR_NOMEM = -3
LCT_ERR_RANGE = -4
LCT_ERR_UNSUPPORTED = -5

F32, F16, BF16 = 0, 1, 2
REBIN_NULL, REBIN_FAN2PAR, REBIN_PI_ORI = 0, 1, 2
WEIGHT_NONE, WEIGHT_PI_ORI = 0, 1
WINDOW_NONE, WINDOW_HANN, WINDOW_SHEPP_LOGAN = 0, 1, 2
FILTER_DIRECT_F32, FILTER_TOEPLITZ_BF16X3, FILTER_TOEPLITZ_BF16X6 = 0, 1, 2
BP_HELICAL, BP_PI_ORI = 0, 1
INTERP_LDS, INTERP_GATHER = 0, 1
PIPE_BLOCKS, PIPE_RESIDENT = 0, 1


class LctError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[lct {code}] {msg}")
        self.code = code


class CTGeom(C.Structure):
    """lct_ct_geom == CTGeom (Include/DataStructs.h:9-100)."""

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
    """lct_img_geom == ImgGeom (Include/DataStructs.h:102-156)."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("wx", C.c_float), ("wy", C.c_float), ("wz", C.c_float),
        ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
    ]


class Prm(C.Structure):
    """lct_prm: the numeric keys of a .prm file (Src/FileParams.cpp:10)."""

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
    """lct_plan: ReconInfo (Include/DataStructs.h:225-273) + geometry."""

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


class ReconOpts(C.Structure):
    _fields_ = [
        ("pipeline", C.c_int), ("rebin_kind", C.c_int), ("weight_kind", C.c_int),
        ("window", C.c_int), ("filter_algo", C.c_int), ("storage", C.c_int),
        ("bp_kind", C.c_int), ("interp", C.c_int),
    ]


class Timings(C.Structure):
    _fields_ = [
        ("rebin_ms", C.c_float), ("filter_ms", C.c_float), ("backproject_ms", C.c_float),
        ("total_ms", C.c_float), ("n_backproject_launches", C.c_int),
        ("backproject_updates", C.c_longlong), ("backproject_kernel_ms", C.c_float),
    ]


# name -> (restype, argtypes)
_P = C.POINTER
_vp, _fp, _dp, _ip = C.c_void_p, _P(C.c_float), _P(C.c_double), _P(C.c_int)
SIGNATURES = {
    "lct_ctx_create": (C.c_int, [C.c_int, _P(C.c_void_p)]),
    "lct_ctx_destroy": (C.c_int, [_vp]),
    "lct_ctx_set_stream": (C.c_int, [_vp, _vp]),
    "lct_synchronize": (C.c_int, [_vp]),
    "lct_last_error": (C.c_char_p, []),
    "lct_version": (C.c_char_p, []),
    "lct_device_malloc": (C.c_int, [_vp, _P(C.c_void_p), C.c_size_t]),
    "lct_device_free": (C.c_int, [_vp, _vp]),
    "lct_memcpy_h2d": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "lct_memcpy_d2h": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "lct_memcpy_d2d": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "lct_memset_zero": (C.c_int, [_vp, _vp, C.c_size_t]),
    "lct_plan_helical": (C.c_int, [_P(Prm), _fp, _fp, _P(Plan)]),
    "lct_block_info": (C.c_int, [_P(Plan), _fp, C.c_int, _ip, _ip, _fp, _fp, _ip]),
    "lct_slab_views": (C.c_int, [_P(Plan), _fp, C.c_int, C.c_int, _P(C.c_long), _P(C.c_long)]),
    "lct_ramp_taps": (C.c_int, [C.c_int, C.c_int, C.c_float, _fp]),
    "lct_rebin": (C.c_int, [_vp, C.c_int, C.c_int, _P(CTGeom), _vp, _vp]),
    "lct_weighting": (C.c_int, [_vp, C.c_int, _P(CTGeom), _vp]),
    "lct_filter": (C.c_int, [_vp, _P(CTGeom), _fp, _vp, C.c_long]),
    "lct_backproject": (C.c_int, [_vp, C.c_int, _P(CTGeom), _P(ImgGeom), _vp, _fp, _vp]),
    "lct_backproject_noalloc": (C.c_int, [_vp, C.c_int, _P(CTGeom), _P(ImgGeom), _vp, _fp, _vp]),
    "lct_filter_to_bp_layout": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _fp, _vp,
                                          C.c_long, _vp]),
    "lct_reconstruct": (C.c_int, [_vp, _P(Plan), _P(ReconOpts), _fp, _fp, _vp, C.c_long, _vp]),
    "lct_reconstruct_slabs": (C.c_int, [_vp, _P(Plan), _P(ReconOpts), _fp, _fp, _vp, C.c_long,
                                        C.c_long, C.c_int, C.c_int, _vp]),
    "lct_last_block_filtered": (C.c_int, [_vp, _vp]),
    "lct_get_timings": (C.c_int, [_vp, _P(Timings)]),
    "lct_set_profiling": (C.c_int, [_vp, C.c_int]),
    "lct_project_ellipsoids": (C.c_int, [_vp, _P(CTGeom), _fp, _fp, C.c_long, C.c_long, _dp,
                                         C.c_int, _vp]),
    "lct_poisson_noise": (C.c_int, [_vp, _vp, C.c_size_t, C.c_float, C.c_uint64]),
    "lct_forward_project": (C.c_int, [_vp, _P(CTGeom), _fp, _fp, C.c_long, C.c_long, _vp,
                                      _P(ImgGeom), C.c_float, C.c_int, C.c_float, _vp]),
    "lct_forward_project_adjoint": (C.c_int, [_vp, _P(CTGeom), _fp, _fp, C.c_long, C.c_long,
                                              _vp, _P(ImgGeom), C.c_float, C.c_int, C.c_float,
                                              _vp]),
    "lct_voxelize_ellipsoids": (C.c_int, [_vp, _P(ImgGeom), C.c_double, C.c_int, _dp, C.c_int,
                                          C.c_int, _vp]),
    "lct_cnn_param_count": (C.c_long, []),
    "lct_cnn_create": (C.c_int, [_vp, C.c_uint64, _P(C.c_void_p)]),
    "lct_cnn_destroy": (C.c_int, [_vp]),
    "lct_cnn_get_params": (C.c_int, [_vp, _fp]),
    "lct_cnn_set_params": (C.c_int, [_vp, _fp]),
    "lct_cnn_get_grads": (C.c_int, [_vp, _fp]),
    "lct_cnn_forward": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp]),
    "lct_cnn_loss_grad": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _fp]),
    "lct_cnn_adam_step": (C.c_int, [_vp, C.c_float, C.c_float, C.c_float, C.c_float]),
}

_lib = None


def lib():
    """Loads liblct.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m listctrecons_amd.build` "
                              "(the HIP path has no fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != LCT_OK:
        msg = lib().lct_last_error()
        raise LctError(rc, msg.decode() if msg else "")
