"""ctypes binding of libstfwi.so (include/stfwi.h).

The library is built in-tree by ``stfwi_amd.build``.  There is no CPU
fallback: if the library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstfwi.so")

ABI_VERSION = 2
F32, F64 = 0, 1
SRC_FORCE, SRC_MOMENT = 0, 1
IMAGING = {"none": 0, "zhou": 1, "exact": 2}
GRAD = {"vs": 1, "vs_vp": 2, "all": 3}
CH_VX, CH_VZ, CH_EX, CH_EZ = 1, 2, 4, 8
SNAP = {"vx": 1, "vz": 2, "ex": 4, "ez": 8, "es": 16}

# symbol -> (restype, argtypes)
_VP = C.c_void_p
SIGNATURES = {
    "stfwi_abi_version": (C.c_int, []),
    "stfwi_last_error": (C.c_char_p, []),
    "stfwi_plan_create": (C.c_int, [_VP, C.POINTER(_VP)]),
    "stfwi_plan_destroy": (C.c_int, [_VP]),
    "stfwi_plan_bytes": (C.c_longlong, [_VP]),
    "stfwi_plan_footprint": (C.c_int, [_VP, C.POINTER(C.c_longlong)]),
    "stfwi_plan_choose_ckpt": (C.c_int, [_VP, C.c_longlong, C.POINTER(C.c_int)]),
    "stfwi_set_model": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "stfwi_set_shots": (C.c_int, [_VP, _VP, C.c_int, _VP, _VP, _VP, _VP]),
    "stfwi_forward": (C.c_int, [_VP, _VP, C.c_int, _VP, C.c_int, _VP, C.c_int, _VP]),
    "stfwi_adjoint": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int, _VP]),
    "stfwi_gradient_finalize": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int, _VP, _VP]),
    "stfwi_time_kernel": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, _VP, C.POINTER(C.c_double)]),
    "stfwi_autotune": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP]),
    "stfwi_set_tiles": (C.c_int, [_VP, _VP]),
    "stfwi_set_groups": (C.c_int, [_VP, C.c_int]),
    "stfwi_set_lean": (C.c_int, [_VP, C.c_int]),
    "stfwi_set_graphs": (C.c_int, [_VP, C.c_int]),
    "stfwi_smooth2d": (C.c_int, [_VP, _VP, _VP, C.c_int, C.c_int, C.c_double, C.c_int, _VP]),
    "stfwi_illumination": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_double, C.c_int, _VP, _VP]),
    "stfwi_eprecond3": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _VP, _VP,
                                  C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _VP, _VP, _VP]),
    "stfwi_eprecond1": (C.c_int, [_VP, _VP, C.c_longlong, C.c_double, C.c_int, _VP, _VP]),
    "stfwi_eprecond1_noalloc": (C.c_int, [_VP, _VP, C.c_longlong, C.c_double, C.c_int, _VP, _VP]),
    "stfwi_set_adjoint_energy": (C.c_int, [_VP, _VP, C.c_int]),
    "stfwi_taper_grad": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                   C.c_int, _VP, _VP]),
    "stfwi_source_taper": (C.c_int, [C.c_int, C.c_int, C.c_double, _VP, C.c_int, C.c_double, C.c_int, _VP, _VP]),
    "stfwi_sosfilt": (C.c_int, [_VP, C.c_int, C.c_int, _VP, C.c_int, C.c_int, C.c_int, C.c_int, _VP]),
    "stfwi_zhou_imaging": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.c_int, C.c_int, C.c_int,
                                     C.c_double, C.c_double, C.c_double, C.c_int, _VP, _VP]),
    "stfwi_residual": (C.c_int, [_VP, _VP, _VP, C.c_int, _VP, C.c_int, C.c_int, C.c_double, C.c_int,
                                 _VP, _VP, _VP]),
    "stfwi_lbfgs_direction": (C.c_int, [_VP, _VP, _VP, _VP, _VP, C.c_int, C.c_longlong, C.c_longlong,
                                        C.c_double, _VP, _VP, _VP]),
    "stfwi_lbfgs_pair": (C.c_int, [_VP, _VP, _VP, _VP, C.c_longlong, _VP, _VP, _VP, _VP]),
    "stfwi_dot64": (C.c_int, [_VP, _VP, C.c_longlong, _VP, _VP]),
    "stfwi_step_clip": (C.c_int, [_VP, _VP, C.c_double, _VP, _VP, C.c_longlong, _VP, _VP]),
    "stfwi_bound_mask": (C.c_int, [_VP, _VP, _VP, _VP, C.c_longlong, _VP, _VP]),
    "stfwi_conv2d_bf16": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, _VP, _VP, C.c_int, C.c_int,
                                    C.c_int, _VP, _VP]),
    "stfwi_conv2d_bf16_dgrad": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, _VP, C.c_int, C.c_int, _VP,
                                          _VP]),
    "stfwi_conv2d_bf16_wgrad_work": (C.c_longlong, [C.c_int, C.c_int, C.c_int]),
    "stfwi_conv2d_bf16_wgrad": (C.c_int, [_VP, _VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _VP, _VP,
                                          C.c_int, _VP, C.c_longlong, _VP]),
    "stfwi_relu_bwd": (C.c_int, [_VP, _VP, C.c_longlong, _VP]),
    "stfwi_pool2": (C.c_int, [_VP, C.c_longlong, C.c_int, C.c_int, _VP, C.c_int, _VP]),
    "stfwi_up2": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, _VP, C.c_longlong, C.c_longlong, C.c_longlong,
                            C.c_int, _VP]),
    "stfwi_adam": (C.c_int, [_VP, _VP, _VP, _VP, C.c_longlong, C.c_double, C.c_double, C.c_double, C.c_double,
                             C.c_int, _VP]),
    "stfwi_mse": (C.c_int, [_VP, _VP, _VP, C.c_longlong, C.c_double, _VP, _VP, _VP, _VP]),
}


class Config(C.Structure):
    _fields_ = [("nz_dom", C.c_int), ("nx_dom", C.c_int), ("pad", C.c_int), ("free_surface", C.c_int),
                ("fd_order", C.c_int), ("dx", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
                ("fc", C.c_double), ("dtype", C.c_int), ("n_shots_max", C.c_int), ("nt_max", C.c_int),
                ("nrec_max", C.c_int), ("imaging", C.c_int), ("grad_params", C.c_int),
                ("store_history", C.c_int), ("ckpt_every", C.c_int), ("hist_frames", C.c_longlong)]


class Shot(C.Structure):
    _fields_ = [("src_x", C.c_int), ("src_z", C.c_int), ("src_type", C.c_int), ("nt", C.c_int),
                ("rec_offset", C.c_int), ("nrec", C.c_int)]


class StfwiError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libstfwi.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise StfwiError(f"{LIB_PATH} is missing: build it with `python -m stfwi_amd.build` "
                             "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.stfwi_abi_version() != ABI_VERSION:
            raise StfwiError("libstfwi.so ABI version mismatch")
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().stfwi_last_error().decode(errors="replace")
        if rc == -2:
            raise NotImplementedError(msg)
        if rc == -1:
            raise ValueError(msg)
        raise StfwiError(f"stfwi error {rc}: {msg}")
    return rc
