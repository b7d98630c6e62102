"""Device solver: one C-ABI context over torch FP64 tensors on a ROCm device.

A state is a contiguous float64 tensor of shape (nbasis, nQ, nz, ny, nx) on
the GPU, which has the memory layout of the reference's Fortran
Q(nx,ny,nz,nQ,nbasis) (src/params.f90:163).  PyTorch provides device memory
and the stream; all arithmetic runs in the HIP kernels of libhermeshd_amd.so.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .config import HDConfig


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Solver:
    def __init__(self, cfg: HDConfig, device: str | torch.device = "cuda"):
        if not torch.cuda.is_available():
            raise RuntimeError("hermeshd_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.cfg = cfg
        dev = torch.device(device)
        if dev.type != "cuda":
            raise ValueError("the solver runs on a ROCm device ('cuda' in torch)")
        self.device = dev if dev.index is not None else torch.device("cuda", torch.cuda.current_device())
        L = _lib.lib()
        self._cfg_c = _lib.config_struct(cfg.as_ints())
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(L.hd_create(C.byref(self._cfg_c), C.byref(h)))
        self._h = h
        self.dt = torch.zeros((), dtype=torch.float64, device=self.device)
        self.t = torch.zeros((), dtype=torch.float64, device=self.device)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().hd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------
    def zeros(self) -> torch.Tensor:
        return torch.zeros(self.cfg.shape, dtype=torch.float64, device=self.device)

    def _ptr(self, x: torch.Tensor, name: str = "q") -> C.c_void_p:
        if not isinstance(x, torch.Tensor):
            raise TypeError(f"{name} must be a torch tensor")
        if x.device != self.device or x.dtype != torch.float64 or not x.is_contiguous():
            raise ValueError(f"{name} must be a contiguous float64 tensor on {self.device}")
        if tuple(x.shape) != self.cfg.shape and x.dim() != 0:
            raise ValueError(f"{name} has shape {tuple(x.shape)}, expected {self.cfg.shape}")
        return C.c_void_p(x.data_ptr())

    def _scalar(self, v) -> torch.Tensor:
        if isinstance(v, torch.Tensor) and v.device == self.device and v.dtype == torch.float64 and v.numel() == 1:
            return v
        self.dt.fill_(float(v))
        return self.dt

    # ------------------------------------------------------------------
    def limit(self, q: torch.Tensor) -> None:
        """Positivity limiter in place (src/flux.f90:1077-1164)."""
        _lib.check(_lib.lib().hd_limit(self._h, self._ptr(q), _stream()))

    def rhs(self, q: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """-glflux + source of q as given (no limiter)."""
        out = self.zeros() if out is None else out
        _lib.check(_lib.lib().hd_rhs(self._h, self._ptr(q), self._ptr(out, "out"), _stream()))
        return out

    def euler_step(self, q_in: torch.Tensor, q_out: torch.Tensor, dt) -> None:
        d = self._scalar(dt)
        _lib.check(_lib.lib().hd_euler_step(self._h, self._ptr(q_in), self._ptr(q_out, "q_out"),
                                            C.c_void_p(d.data_ptr()), _stream()))

    def min_dt(self, q: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        out = torch.zeros((), dtype=torch.float64, device=self.device) if out is None else out
        _lib.check(_lib.lib().hd_min_dt(self._h, self._ptr(q), C.c_void_p(out.data_ptr()), _stream()))
        return out

    def update(self, q_io: torch.Tensor, q1: torch.Tensor, q2: torch.Tensor | None, dt) -> None:
        d = self._scalar(dt)
        p2 = self._ptr(q2, "q2") if q2 is not None else C.c_void_p(0)
        _lib.check(_lib.lib().hd_update(self._h, self._ptr(q_io), self._ptr(q1, "q1"), p2,
                                        C.c_void_p(d.data_ptr()), _stream()))

    def step(self, q_io: torch.Tensor, q1: torch.Tensor, q2: torch.Tensor | None = None,
             t: torch.Tensor | None = None, dt: torch.Tensor | None = None) -> None:
        """Reference step: dt = get_min_dt(q_io); RK update; t += dt.  t and dt
        are device 0-d tensors (default: the solver's own)."""
        t = self.t if t is None else t
        dt = self.dt if dt is None else dt
        p2 = self._ptr(q2, "q2") if q2 is not None else C.c_void_p(0)
        _lib.check(_lib.lib().hd_step(self._h, self._ptr(q_io), self._ptr(q1, "q1"), p2,
                                      C.c_void_p(t.data_ptr()), C.c_void_p(dt.data_ptr()), _stream()))

    def stage(self, rk_order: int, stage_index: int, q_in: torch.Tensor, q_n: torch.Tensor | None,
              q_out: torch.Tensor, dt) -> None:
        """One SSP-RK stage: limiter on q_in in place, q_out = a*q_n + b*E(q_in)."""
        d = self._scalar(dt)
        pn = self._ptr(q_n, "q_n") if q_n is not None else C.c_void_p(0)
        _lib.check(_lib.lib().hd_stage(self._h, int(rk_order), int(stage_index), self._ptr(q_in, "q_in"), pn,
                                       self._ptr(q_out, "q_out"), C.c_void_p(d.data_ptr()), _stream()))

    def rhs_vjp(self, q: torch.Tensor, lam: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """(dR/dq)^T lam at state q."""
        out = torch.empty_like(q) if out is None else out
        _lib.check(_lib.lib().hd_rhs_vjp(self._h, self._ptr(q), self._ptr(lam, "lam"), self._ptr(out, "out"),
                                         _stream()))
        return out

    def stage_vjp(self, q_in, lam_out, lam_in, dt, a, b, acc=(), q_raw=None, out=None, qn=None, dcoll=None):
        """Transpose of one stage at its limited input q_in applied to lam_out; writes lam_in.
        acc: up to two (weight, array) pairs added to lam_in; dcoll: 0-d device tensor (+=)."""
        d = self._scalar(dt)
        args = _lib.StageVjpArgs()
        args.a, args.b = float(a), float(b)
        acc = list(acc)
        if len(acc) > 2:
            raise ValueError("at most two accumulated arrays")
        for k, (w, arr) in enumerate(acc):
            setattr(args, f"acc{k}", self._ptr(arr, f"acc{k}").value)
            setattr(args, f"w{k}", float(w))
        if q_raw is not None:
            args.q_raw = self._ptr(q_raw, "q_raw").value
        if out is not None:
            args.out = self._ptr(out, "out").value
        if qn is not None:
            args.qn = self._ptr(qn, "qn").value
        if dcoll is not None:
            args.dcoll_dev = dcoll.data_ptr()
        _lib.check(_lib.lib().hd_stage_vjp(self._h, self._ptr(q_in, "q_in"), self._ptr(lam_out, "lam_out"),
                                           self._ptr(lam_in, "lam_in"), C.c_void_p(d.data_ptr()), C.byref(args),
                                           _stream()))

    def set_step(self, step: int) -> None:
        """Step index that keys the stochastic stress (advanced by every update)."""
        _lib.check(_lib.lib().hd_set_step(self._h, int(step), _stream()))

    def set_inflow(self, inflow_def) -> None:
        """Exterior state of the x-low inflow boundary: (nQ, nface, nz, ny)
        float64 (numpy or torch), the layout of the reference's
        Qxlo_ext_def(ny,nz,nface,nQ)."""
        nf = self.cfg.iquad ** 2
        want = (self.cfg.nQ, nf, self.cfg.nz, self.cfg.ny)
        t = torch.as_tensor(inflow_def, dtype=torch.float64).to(self.device).contiguous()
        if tuple(t.shape) != want:
            raise ValueError(f"inflow state must have shape {want}")
        self._inflow = t
        _lib.check(_lib.lib().hd_set_inflow(self._h, C.c_void_p(t.data_ptr()), _stream()))

    def set_mask(self, mask, rho_mask: float = 1.25) -> None:
        """Immersed-solid cells, (nz, ny, nx) bool (numpy or torch); None
        switches the mask off.  Applied after every step."""
        if mask is None:
            self._mask = None
            _lib.check(_lib.lib().hd_set_mask(self._h, None, 0.0, _stream()))
            return
        t = torch.as_tensor(mask).to(torch.uint8).to(self.device).contiguous()
        if tuple(t.shape) != (self.cfg.nz, self.cfg.ny, self.cfg.nx):
            raise ValueError("mask must have shape (nz, ny, nx)")
        self._mask = t
        _lib.check(_lib.lib().hd_set_mask(self._h, C.c_void_p(t.data_ptr()), float(rho_mask), _stream()))

    def apply_mask(self, q: torch.Tensor) -> None:
        _lib.check(_lib.lib().hd_apply_mask(self._h, self._ptr(q), _stream()))

    def check(self) -> None:
        """Synchronise and raise NonFiniteError if a stage produced NaN."""
        bad = C.c_int64(-1)
        rc = _lib.lib().hd_check(self._h, _stream(), C.byref(bad))
        if rc == _lib.HD_ERR_NONFINITE:
            raise _lib.NonFiniteError(rc, _lib.lib().hd_last_error().decode(), bad.value)
        _lib.check(rc)

    def timing(self, enable: bool) -> None:
        """Bracket every stage with HIP events on the launch stream."""
        _lib.check(_lib.lib().hd_timing_enable(self._h, int(bool(enable))))

    def timing_read(self):
        """(summed stage milliseconds, stages) since the last read."""
        ms, n = C.c_double(0.0), C.c_int64(0)
        _lib.check(_lib.lib().hd_timing_read(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    @property
    def device_bytes(self) -> int:
        return int(_lib.lib().hd_device_bytes(self._h))
