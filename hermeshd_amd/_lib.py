"""ctypes binding of libhermeshd_amd.so (the C ABI declared in
include/hermeshd_amd.h).  Loading fails loudly when the library is missing:
there is no CPU fallback for the product path."""
from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhermeshd_amd.so")

HD_OK = 0
HD_ERR_INVALID = -1
HD_ERR_HIP = -2
HD_ERR_NONFINITE = -3
HD_ERR_UNSUPPORTED = -4

# every extern "C" symbol declared in include/hermeshd_amd.h
EXPORTED = [
    "hd_create", "hd_destroy", "hd_limit", "hd_rhs", "hd_euler_step", "hd_min_dt", "hd_min_dt_static", "hd_update",
    "hd_step", "hd_stage", "hd_rhs_vjp", "hd_stage_vjp", "hd_set_step", "hd_set_inflow", "hd_set_mask", "hd_apply_mask", "hd_sample", "hd_check", "hd_timing_enable", "hd_timing_read", "hd_device_bytes",
    "hd_last_error",
    "hd_ens_create", "hd_ens_destroy", "hd_ens_state_size", "hd_ens_offsets", "hd_ens_device_bytes",
    "hd_ens_limit", "hd_ens_rhs", "hd_ens_update", "hd_ens_run", "hd_ens_set_step", "hd_ens_check",
    "hd_ens_launch_count", "hd_ens_launch_info", "hd_ens_set_launch", "hd_ens_get_launch", "hd_ens_tune",
]
# entry points whose return value is not an error code
NON_RC = ("hd_device_bytes", "hd_last_error", "hd_ens_state_size", "hd_ens_device_bytes", "hd_ens_launch_count",
          "hd_ens_get_launch")


class HDConfigC(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("nbasis", C.c_int32), ("iquad", C.c_int32), ("nQ", C.c_int32),
        ("ivis", C.c_int32), ("ieos", C.c_int32), ("riemann", C.c_int32), ("limiter", C.c_int32),
        ("integrator", C.c_int32), ("bc", C.c_int32 * 6), ("noise", C.c_int32), ("reserved0", C.c_int32 * 2),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double), ("gamma", C.c_double),
        ("rh_floor", C.c_double), ("T_floor", C.c_double), ("coll", C.c_double), ("colvis", C.c_double),
        ("cflm", C.c_double), ("body_force", C.c_double * 3), ("eta_sd", C.c_double), ("zeta_sd", C.c_double),
        ("seed", C.c_uint64), ("realization", C.c_uint32), ("reserved1", C.c_uint32),
        ("eos_p1", C.c_double), ("eos_pbase", C.c_double),
        ("coll_q", C.c_double), ("col_q", C.c_double),
    ]


class StageVjpArgs(C.Structure):
    _fields_ = [
        ("a", C.c_double), ("b", C.c_double), ("acc0", C.c_void_p), ("acc1", C.c_void_p),
        ("w0", C.c_double), ("w1", C.c_double), ("q_raw", C.c_void_p), ("out", C.c_void_p), ("qn", C.c_void_p),
        ("dcoll_dev", C.c_void_p),
    ]


class HDError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class NonFiniteError(HDError):
    def __init__(self, code: int, msg: str, cell: int):
        super().__init__(code, msg)
        self.cell = cell


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m hermeshd_amd.build` "
                          "(or __graft_entry__.build()); the product path has no CPU fallback")
    L = C.CDLL(LIB_PATH)
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    L.hd_create.argtypes = [C.POINTER(HDConfigC), C.POINTER(C.c_void_p)]
    L.hd_destroy.argtypes = [vp]
    L.hd_limit.argtypes = [vp, vp, vp]
    L.hd_rhs.argtypes = [vp, vp, vp, vp]
    L.hd_euler_step.argtypes = [vp, vp, vp, vp, vp]
    L.hd_min_dt.argtypes = [vp, vp, vp, vp]
    L.hd_min_dt_static.argtypes = [C.POINTER(HDConfigC), vp, vp, vp]
    L.hd_update.argtypes = [vp, vp, vp, vp, vp, vp]
    L.hd_step.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.hd_set_step.argtypes = [vp, C.c_uint64, vp]
    L.hd_set_inflow.argtypes = [vp, vp, vp]
    L.hd_set_mask.argtypes = [vp, vp, C.c_double, vp]
    L.hd_apply_mask.argtypes = [vp, vp, vp]
    L.hd_sample.argtypes = [vp, vp, C.c_int32, C.c_uint32, dp, vp, vp]
    L.hd_stage.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp]
    L.hd_rhs_vjp.argtypes = [vp, vp, vp, vp, vp]
    L.hd_stage_vjp.argtypes = [vp, vp, vp, vp, vp, C.POINTER(StageVjpArgs), vp]
    L.hd_check.argtypes = [vp, vp, C.POINTER(C.c_int64)]
    L.hd_timing_enable.argtypes = [vp, C.c_int]
    L.hd_timing_read.argtypes = [vp, dp, C.POINTER(C.c_int64)]
    L.hd_device_bytes.argtypes = [vp]
    L.hd_device_bytes.restype = C.c_int64
    L.hd_last_error.restype = C.c_char_p
    i64p = C.POINTER(C.c_int64)
    L.hd_ens_create.argtypes = [C.POINTER(HDConfigC), C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                C.POINTER(C.c_void_p)]
    L.hd_ens_destroy.argtypes = [vp]
    L.hd_ens_state_size.argtypes = [vp]
    L.hd_ens_state_size.restype = C.c_int64
    L.hd_ens_offsets.argtypes = [vp, i64p]
    L.hd_ens_device_bytes.argtypes = [vp]
    L.hd_ens_device_bytes.restype = C.c_int64
    L.hd_ens_limit.argtypes = [vp, vp, vp]
    L.hd_ens_rhs.argtypes = [vp, vp, vp, vp]
    L.hd_ens_update.argtypes = [vp, vp, vp, vp, vp, vp]
    L.hd_ens_run.argtypes = [vp, vp, vp, vp, vp, C.c_int64, vp]
    L.hd_ens_set_step.argtypes = [vp, C.c_uint64, vp]
    L.hd_ens_check.argtypes = [vp, vp, i64p, i64p]
    L.hd_ens_launch_count.argtypes = []
    L.hd_ens_launch_count.restype = C.c_int32
    L.hd_ens_launch_info.argtypes = [C.c_int32, C.POINTER(C.c_int32)]
    L.hd_ens_set_launch.argtypes = [vp, C.c_int32]
    L.hd_ens_get_launch.argtypes = [vp]
    L.hd_ens_get_launch.restype = C.c_int
    L.hd_ens_tune.argtypes = [vp, vp, vp, vp, C.c_int32, vp, C.POINTER(C.c_int32), dp]
    for name in EXPORTED:
        if name not in NON_RC:
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc: int):
    if rc != HD_OK:
        raise HDError(rc, lib().hd_last_error().decode())


def config_struct(d: dict) -> HDConfigC:
    s = HDConfigC()
    for k, v in d.items():
        if k in ("bc", "body_force"):
            getattr(s, k)[:] = v
        elif hasattr(s, k):
            setattr(s, k, v)
    return s
