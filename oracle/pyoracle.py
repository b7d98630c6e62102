"""ctypes binding of the C oracle (hd_oracle.c).

TEST INFRASTRUCTURE: the checker for the HIP product path.  Only tests/, the
smoke check and bench.py's CPU-baseline leg may import this module.

Arrays are numpy float64, C-contiguous with shape (nbasis, nQ, nz, ny, nx),
which is the memory layout of the reference's Fortran Q(nx,ny,nz,nQ,nbasis).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhd_oracle.so")


class OrcConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("nbasis", C.c_int), ("iquad", C.c_int), ("nQ", C.c_int),
        ("ivis", C.c_int), ("ieos", C.c_int), ("riemann", C.c_int), ("limiter", C.c_int),
        ("integrator", C.c_int), ("bc", C.c_int * 6), ("noise", C.c_int), ("pad_", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("gamma", C.c_double), ("rh_floor", C.c_double), ("T_floor", C.c_double),
        ("coll", C.c_double), ("colvis", C.c_double), ("cflm", C.c_double),
        ("body_force", C.c_double * 3), ("eta_sd", C.c_double), ("zeta_sd", C.c_double),
        ("seed", C.c_uint64), ("realization", C.c_uint32), ("pad2_", C.c_uint32),
        ("eos_p1", C.c_double), ("eos_pbase", C.c_double),
        ("coll_q", C.c_double), ("col_q", C.c_double),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (no FP contraction, like the Fortran build)."""
    src = os.path.join(HERE, "hd_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-o", LIB_PATH, src,
                               "-lm"])
    return LIB_PATH


_lib = None
_libs = {}


def _configure(L):
    dp = C.POINTER(C.c_double)
    L.orc_create.restype = C.c_void_p
    L.orc_create.argtypes = [C.POINTER(OrcConfig)]
    L.orc_destroy.argtypes = [C.c_void_p]
    for name in ("orc_limiter",):
        getattr(L, name).argtypes = [C.c_void_p, dp]
    for name in ("orc_glflux", "orc_source", "orc_rhs"):
        getattr(L, name).argtypes = [C.c_void_p, dp, dp]
    L.orc_euler_step.argtypes = [C.c_void_p, dp, dp, C.c_double]
    L.orc_min_dt.argtypes = [C.c_void_p, dp]
    L.orc_min_dt.restype = C.c_double
    L.orc_update.argtypes = [C.c_void_p, dp, dp, dp, C.c_double]
    L.orc_step.argtypes = [C.c_void_p, dp, dp, dp, dp, dp]
    L.orc_flux_points.argtypes = [C.c_void_p, dp, dp, C.c_int, C.c_int]
    L.orc_riemann_face.argtypes = [C.c_void_p, dp, dp, dp, C.c_int, C.c_int]
    L.orc_cfcal.argtypes = [C.c_void_p, dp, C.c_int]
    L.orc_cfcal.restype = C.c_double
    L.orc_tables.argtypes = [C.c_void_p] + [dp] * 13
    L.orc_set_step_index.argtypes = [C.c_void_p, C.c_uint64]
    L.orc_set_noise_dt.argtypes = [C.c_void_p, C.c_double]
    L.orc_set_noise_impulse.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_int, C.c_double]
    L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.orc_noise_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint64, dp]
    L.orc_noise_stress.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64, C.c_double, C.c_double, dp]
    L.orc_set_inflow.argtypes = [C.c_void_p, dp]
    L.orc_set_mask.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
    L.orc_apply_mask.argtypes = [C.c_void_p, dp]
    L.orc_sample.argtypes = [C.c_void_p, dp, C.c_int, C.c_uint32, C.c_void_p, C.c_void_p]
    L.orc_last_error.restype = C.c_char_p
    L.orc_last_bad_cell.restype = C.c_int64
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads.restype = C.c_int
    # one thread unless a caller asks for more (set_threads): the tests run
    # several oracle processes side by side
    L.orc_set_threads(1)
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _configure(C.CDLL(LIB_PATH))
    return _lib


def load_library(path: str):
    """A separately built oracle library (the mutation tests build variants of
    hd_oracle.c with one planted mistake each and check that the pins fail)."""
    if path not in _libs:
        _libs[path] = _configure(C.CDLL(path))
    return _libs[path]


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleError(RuntimeError):
    pass


class Oracle:
    """One oracle context (tables + scratch) for a configuration dict as
    produced by ``HDConfig.as_ints()``."""

    def __init__(self, cfg: dict, library=None):
        self.L = library if library is not None else lib()
        self.cfg = dict(cfg)
        s = OrcConfig()
        for k, v in cfg.items():
            if k in ("bc", "body_force"):
                getattr(s, k)[:] = v
            else:
                setattr(s, k, v)
        self._s = s
        self.h = self.L.orc_create(C.byref(s))
        if not self.h:
            raise OracleError(self.L.orc_last_error().decode())
        self.shape = (cfg["nbasis"], cfg["nQ"], cfg["nz"], cfg["ny"], cfg["nx"])

    def __del__(self):
        h = getattr(self, "h", None)
        L = getattr(self, "L", None)
        if h and L is not None:
            L.orc_destroy(h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise OracleError(f"{self.L.orc_last_error().decode()} (cell {self.L.orc_last_bad_cell()})")

    def zeros(self):
        return np.zeros(self.shape, dtype=np.float64)

    def limiter(self, q):
        self._check(self.L.orc_limiter(self.h, _p(q)))

    def glflux(self, q):
        out = self.zeros(); self._check(self.L.orc_glflux(self.h, _p(q), _p(out))); return out

    def source(self, q):
        out = self.zeros(); self._check(self.L.orc_source(self.h, _p(q), _p(out))); return out

    def rhs(self, q):
        out = self.zeros(); self._check(self.L.orc_rhs(self.h, _p(q), _p(out))); return out

    def euler_step(self, q_in, q_out, dt):
        self._check(self.L.orc_euler_step(self.h, _p(q_in), _p(q_out), float(dt)))

    def min_dt(self, q) -> float:
        return self.L.orc_min_dt(self.h, _p(q))

    def update(self, q_io, q1, q2, dt):
        self._check(self.L.orc_update(self.h, _p(q_io), _p(q1), _p(q2), float(dt)))

    def step(self, q_io, q1, q2, t, dt):
        """Reference step semantics: dt from get_min_dt, then RK update; t, dt
        are 0-d float64 arrays updated in place."""
        self._check(self.L.orc_step(self.h, _p(q_io), _p(q1), _p(q2), _p(t), _p(dt)))

    def set_step_index(self, n: int):
        self.L.orc_set_step_index(self.h, int(n))

    def set_noise_dt(self, dt: float):
        self.L.orc_set_noise_dt(self.h, float(dt))

    def set_noise_impulse(self, stream: int = 0, point_id: int = 0, comp: int = 0, value: float = 0.0,
                          on: bool = True):
        """Replace the drawn stochastic stress by one unit impulse (value in
        component comp of point point_id of a stream; every other point zero)
        until called with on=False.  For linear-response analysis."""
        self._check(self.L.orc_set_noise_impulse(self.h, int(on), int(stream), int(point_id), int(comp),
                                                 float(value)))

    def set_inflow(self, inflow_def):
        """x-low inflow exterior state, numpy (nQ, nface, nz, ny) C order
        (the reference's Qxlo_ext_def(ny,nz,nface,nQ))."""
        a = np.ascontiguousarray(inflow_def, dtype=np.float64)
        self._inflow = a
        self._check(self.L.orc_set_inflow(self.h, _p(a)))

    def set_mask(self, mask, rho_mask: float = 1.25):
        """Solid cells, numpy (nz, ny, nx) bool; None clears."""
        if mask is None:
            self._check(self.L.orc_set_mask(self.h, None, 0.0))
            return
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self._check(self.L.orc_set_mask(self.h, m.ctypes.data_as(C.c_void_p), float(rho_mask)))

    def apply_mask(self, q):
        self._check(self.L.orc_apply_mask(self.h, _p(q)))

    def sample(self, q, nvtk: int, fields: int, scales=None) -> np.ndarray:
        """output_vtk's fields: (nsel, nz*nvtk, ny*nvtk, nx*nvtk) float32."""
        nsel = bin(int(fields)).count("1")
        out = np.zeros((nsel, self.shape[2] * nvtk, self.shape[3] * nvtk, self.shape[4] * nvtk), dtype=np.float32)
        sc = None if scales is None else np.ascontiguousarray(scales, dtype=np.float64)
        self._check(self.L.orc_sample(self.h, _p(q), int(nvtk), int(fields),
                                     None if sc is None else sc.ctypes.data_as(C.c_void_p),
                                     out.ctypes.data_as(C.c_void_p)))
        return out

    def noise_stress(self, stream: int, ids, amp_eta: float, amp_zeta: float) -> np.ndarray:
        """(len(ids), 6) stresses xx,yy,zz,xy,xz,yz at the current step index."""
        ids = np.asarray(ids, dtype=np.uint64).ravel()
        out = np.zeros((ids.size, 6))
        row = np.zeros(6)
        for n, i in enumerate(ids):
            self.L.orc_noise_stress(self.h, int(stream), int(i), float(amp_eta), float(amp_zeta), _p(row))
            out[n] = row
        return out

    def flux_points(self, qp: np.ndarray, direction: int) -> np.ndarray:
        qp = np.ascontiguousarray(qp, dtype=np.float64)
        out = np.zeros_like(qp)
        self.L.orc_flux_points(self.h, _p(qp), _p(out), int(direction), qp.shape[0])
        return out

    def riemann_face(self, qlr: np.ndarray, direction: int, kind: int) -> np.ndarray:
        """qlr: (2*nface, nQ) left rows then right rows -> (nface, nQ)."""
        qlr = np.ascontiguousarray(qlr, dtype=np.float64)
        flr = self.flux_points(qlr, direction)
        nface = qlr.shape[0] // 2
        out = np.zeros((nface, qlr.shape[1]))
        self.L.orc_riemann_face(self.h, _p(qlr), _p(flr), _p(out), int(direction), int(kind))
        return out

    def tables(self) -> dict:
        q, nb = self.cfg["iquad"], self.cfg["nbasis"]
        nf, npg = q * q, q ** 3
        t = dict(xq=np.zeros(q), w1=np.zeros(q), w2=np.zeros(nf), w3=np.zeros(npg), cbasis=np.zeros(nb),
                 bint=np.zeros((npg, nb)), bxp=np.zeros((nf, nb)), bxm=np.zeros((nf, nb)),
                 byp=np.zeros((nf, nb)), bym=np.zeros((nf, nb)), bzp=np.zeros((nf, nb)),
                 bzm=np.zeros((nf, nb)), bfaces=np.zeros((6 * nf + npg, nb)))
        keys = ["xq", "w1", "w2", "w3", "cbasis", "bint", "bxp", "bxm", "byp", "bym", "bzp", "bzm", "bfaces"]
        self.L.orc_tables(self.h, *[_p(t[k]) for k in keys])
        return t


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def noise_normals(seed, realization, step, stream, index):
    out = np.zeros(4)
    lib().orc_noise_normals(seed, realization, step, stream, index, _p(out))
    return out


def set_threads(n: int) -> int:
    """Threads of the oracle's cell and face loops (OpenMP); returns the count in use."""
    return int(lib().orc_set_threads(int(n)))
