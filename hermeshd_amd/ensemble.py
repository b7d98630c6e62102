"""Ensembles of fluctuating-hydrodynamics realizations on one GPU.

Many independent realizations of one configuration, on periodic boxes of
different sizes, advanced together through the hd_ens_* entry points: one
kernel launch per RK stage for all of them, one noise launch per step, and a
HIP graph of the whole step replayed for long runs.  Realization r computes
exactly what a single ``Solver`` with the same configuration and realization
id computes (the reference runs one such realization per process; its time
loop is src/main.f90:22-49).

The states are packed back to back in one float64 tensor; ``views`` gives
each realization's (nbasis, nQ, nz, ny, nx) view into it.  ``structure_factor``
measures the static structure factors S(k) of density and velocity from the
cell averages (mode 1), for comparison with the fluctuation-dissipation
values.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .config import HDConfig


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


@dataclass(frozen=True)
class LaunchShape:
    block: int
    brick: tuple
    linear: bool
    wg_per_cu: int       # 0: one workgroup per tile
    xcd_chunks: bool


def launch_shapes() -> list[LaunchShape]:
    L = _lib.lib()
    out = []
    for c in range(L.hd_ens_launch_count()):
        info = (C.c_int32 * 6)()
        _lib.check(L.hd_ens_launch_info(c, info))
        out.append(LaunchShape(info[0], (info[1], info[2], info[3]), bool(info[4]), info[5] // 2, bool(info[5] % 2)))
    return out


class Ensemble:
    """``base`` gives physics, basis, boundaries and noise, and the cell size
    (lx/nx, ly/ny, lz/nz); ``dims`` lists each realization's (nx, ny, nz).
    Noise keys: ``realization_ids`` or base.realization + index."""

    def __init__(self, base: HDConfig, dims, realization_ids=None, device: str | torch.device = "cuda"):
        if not torch.cuda.is_available():
            raise RuntimeError("hermeshd_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        dev = torch.device(device)
        self.device = dev if dev.index is not None else torch.device("cuda", torch.cuda.current_device())
        self.base = base
        self.dims = [tuple(int(v) for v in d) for d in dims]
        self.nreal = len(self.dims)
        L = _lib.lib()
        self._cfg_c = _lib.config_struct(base.as_ints())
        flat = (C.c_int32 * (3 * self.nreal))(*[v for d in self.dims for v in d])
        ids = None
        if realization_ids is not None:
            if len(realization_ids) != self.nreal:
                raise ValueError("one realization id per box")
            ids = (C.c_uint32 * self.nreal)(*[int(v) for v in realization_ids])
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(L.hd_ens_create(C.byref(self._cfg_c), self.nreal, flat, ids, C.byref(h)))
        self._h = h
        self.size = int(L.hd_ens_state_size(h))
        offs = (C.c_int64 * self.nreal)()
        _lib.check(L.hd_ens_offsets(h, offs))
        self.offsets = [int(v) for v in offs]
        self.realization_ids = list(realization_ids) if realization_ids is not None else \
            [base.realization + r for r in range(self.nreal)]
        self.dt = torch.zeros((), dtype=torch.float64, device=self.device)
        self.cells = sum(a * b * c for a, b, c in self.dims)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().hd_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------
    def shape(self, r: int) -> tuple:
        nx, ny, nz = self.dims[r]
        return (self.base.nbasis, self.base.nQ, nz, ny, nx)

    def zeros(self) -> torch.Tensor:
        return torch.zeros(self.size, dtype=torch.float64, device=self.device)

    def views(self, q: torch.Tensor) -> list[torch.Tensor]:
        out = []
        for r in range(self.nreal):
            shp = self.shape(r)
            n = int(np.prod(shp))
            out.append(q[self.offsets[r]:self.offsets[r] + n].view(shp))
        return out

    def pack(self, states) -> torch.Tensor:
        q = self.zeros()
        for v, s in zip(self.views(q), states):
            v.copy_(torch.as_tensor(np.ascontiguousarray(s)) if not isinstance(s, torch.Tensor) else s)
        return q

    def _ptr(self, x: torch.Tensor, name: str) -> C.c_void_p:
        if not isinstance(x, torch.Tensor):
            raise TypeError(f"{name} must be a torch tensor")
        if x.device != self.device or x.dtype != torch.float64 or not x.is_contiguous():
            raise ValueError(f"{name} must be a contiguous float64 tensor on {self.device}")
        if x.numel() != self.size and x.dim() != 0:
            raise ValueError(f"{name} must hold {self.size} doubles (the packed ensemble)")
        return C.c_void_p(x.data_ptr())

    def _set_dt(self, dt):
        if isinstance(dt, torch.Tensor):
            self.dt.copy_(dt)
        else:
            self.dt.fill_(float(dt))

    # ------------------------------------------------------------------
    def limit(self, q: torch.Tensor) -> None:
        _lib.check(_lib.lib().hd_ens_limit(self._h, self._ptr(q, "q"), _stream()))

    def rhs(self, q: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        out = torch.empty_like(q) if out is None else out
        _lib.check(_lib.lib().hd_ens_rhs(self._h, self._ptr(q, "q"), self._ptr(out, "out"), _stream()))
        return out

    def update(self, q_io: torch.Tensor, q1: torch.Tensor, q2: torch.Tensor | None, dt) -> None:
        self._set_dt(dt)
        p2 = self._ptr(q2, "q2") if q2 is not None else C.c_void_p(0)
        _lib.check(_lib.lib().hd_ens_update(self._h, self._ptr(q_io, "q_io"), self._ptr(q1, "q1"), p2,
                                            C.c_void_p(self.dt.data_ptr()), _stream()))

    def run(self, q_io: torch.Tensor, q1: torch.Tensor, q2: torch.Tensor | None, dt, nsteps: int) -> None:
        """nsteps steps replayed from a captured HIP graph of one step."""
        self._set_dt(dt)
        p2 = self._ptr(q2, "q2") if q2 is not None else C.c_void_p(0)
        _lib.check(_lib.lib().hd_ens_run(self._h, self._ptr(q_io, "q_io"), self._ptr(q1, "q1"), p2,
                                         C.c_void_p(self.dt.data_ptr()), int(nsteps), _stream()))

    def set_step(self, step: int) -> None:
        _lib.check(_lib.lib().hd_ens_set_step(self._h, int(step), _stream()))

    def check(self) -> None:
        br, bc = C.c_int64(-1), C.c_int64(-1)
        rc = _lib.lib().hd_ens_check(self._h, _stream(), C.byref(br), C.byref(bc))
        if rc == _lib.HD_ERR_NONFINITE:
            raise _lib.NonFiniteError(rc, _lib.lib().hd_last_error().decode(), int(bc.value))
        _lib.check(rc)

    def tune(self, q: torch.Tensor, scratch: torch.Tensor, dt, reps: int = 5):
        """Times every launch shape on q (read only) and keeps the fastest.
        Returns (choice, microseconds per stage for every shape)."""
        self._set_dt(dt)
        n = _lib.lib().hd_ens_launch_count()
        us = (C.c_double * n)()
        choice = C.c_int32(-1)
        _lib.check(_lib.lib().hd_ens_tune(self._h, self._ptr(q, "q"), self._ptr(scratch, "scratch"),
                                          C.c_void_p(self.dt.data_ptr()), int(reps), _stream(),
                                          C.byref(choice), us))
        return int(choice.value), list(us)

    def set_launch(self, choice: int) -> None:
        _lib.check(_lib.lib().hd_ens_set_launch(self._h, int(choice)))

    def launch(self) -> int:
        return int(_lib.lib().hd_ens_get_launch(self._h))

    def device_bytes(self) -> int:
        return int(_lib.lib().hd_ens_device_bytes(self._h))


# ---------------------------------------------------------------------------
# static structure factors
# ---------------------------------------------------------------------------
def _kgrid(n, h, device):
    return 2 * np.pi * torch.fft.fftfreq(n, d=h, device=device, dtype=torch.float64)


class StructureFactor:
    """Accumulates S(k) of density and of the longitudinal and transverse
    velocity from mode-1 cell averages, over realizations and samples:
    S_rr(k) = <|drho_k|^2> dV / (N kB), S_uL = <|k.u_k/|k||^2> dV/(N kB),
    S_uT = per transverse component.  For an ideal gas at rest with only
    viscous (stress) noise the fluctuation-dissipation values at small k are
    S_uL = S_uT = T/rho, S_rr = rho/gamma (sound modes only: no heat noise)."""

    def __init__(self, ens: Ensemble, kB: float, cell=(1.0, 1.0, 1.0)):
        self.ens, self.kB, self.cell = ens, kB, cell
        self.acc = [None] * ens.nreal
        self.count = 0
        self._k = []
        for (nx, ny, nz) in ens.dims:
            kx = _kgrid(nx, cell[0], ens.device)
            ky = _kgrid(ny, cell[1], ens.device)
            kz = _kgrid(nz, cell[2], ens.device)
            KZ, KY, KX = torch.meshgrid(kz, ky, kx, indexing="ij")
            self._k.append(torch.stack([KX, KY, KZ]))

    def sample(self, q: torch.Tensor) -> None:
        dV = self.cell[0] * self.cell[1] * self.cell[2]
        for r, v in enumerate(self.ens.views(q)):
            rho = v[0, 0]
            u = v[0, 1:4] / rho
            N = rho.numel()
            fr = torch.fft.fftn(rho - rho.mean())
            fu = torch.fft.fftn(u, dim=(1, 2, 3))
            K = self._k[r]
            km = torch.sqrt((K ** 2).sum(0))
            kh = K / torch.where(km > 0, km, torch.ones_like(km))
            uL = (kh * fu).sum(0)
            s_rr = fr.abs() ** 2 * dV / (N * self.kB)
            s_ul = uL.abs() ** 2 * dV / (N * self.kB)
            s_ut = ((fu.abs() ** 2).sum(0) - uL.abs() ** 2) * dV / (2 * N * self.kB)
            st = torch.stack([s_rr, s_ul, s_ut])
            self.acc[r] = st if self.acc[r] is None else self.acc[r] + st
        self.count += 1

    def binned(self, edges) -> dict:
        """Averages over |k| bins (k = 0 excluded): dict of arrays k_mid, n, S_rr, S_uL, S_uT."""
        edges = np.asarray(edges, dtype=float)
        nb = len(edges) - 1
        tot = np.zeros((3, nb))
        cnt = np.zeros(nb)
        for r in range(self.ens.nreal):
            km = torch.sqrt((self._k[r] ** 2).sum(0)).cpu().numpy().ravel()
            S = (self.acc[r] / self.count).cpu().numpy().reshape(3, -1)
            idx = np.digitize(km, edges) - 1
            ok = (km > 0) & (idx >= 0) & (idx < nb)
            for b in range(nb):
                m = ok & (idx == b)
                cnt[b] += m.sum()
                tot[:, b] += S[:, m].sum(axis=1)
        with np.errstate(invalid="ignore", divide="ignore"):
            mean = tot / cnt
        return dict(k_mid=0.5 * (edges[1:] + edges[:-1]), n=cnt, S_rr=mean[0], S_uL=mean[1], S_uT=mean[2])
