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
    "hd_create", "hd_destroy", "hd_limit", "hd_rhs", "hd_euler_step", "hd_min_dt", "hd_update",
    "hd_step", "hd_set_step", "hd_check", "hd_timing_enable", "hd_timing_read", "hd_device_bytes",
    "hd_last_error",
]


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
    L.hd_update.argtypes = [vp, vp, vp, vp, vp, vp]
    L.hd_step.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.hd_set_step.argtypes = [vp, C.c_uint64, vp]
    L.hd_check.argtypes = [vp, vp, C.POINTER(C.c_int64)]
    L.hd_timing_enable.argtypes = [vp, C.c_int]
    L.hd_timing_read.argtypes = [vp, dp, C.POINTER(C.c_int64)]
    L.hd_device_bytes.argtypes = [vp]
    L.hd_device_bytes.restype = C.c_int64
    L.hd_last_error.restype = C.c_char_p
    for name in EXPORTED:
        if name not in ("hd_device_bytes", "hd_last_error"):
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
