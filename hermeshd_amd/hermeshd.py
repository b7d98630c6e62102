"""Drop-in for the reference's Python front end (src/hermeshd.py:5-120).

Same names, argument order and meaning as the f90wrap module ``hermeshd``:
``hermeshd.hermeshd.setup(q_io, t, dt, t1, t_start, dtout, nout, comm)``,
``step(q_io, q_1, q_2, t, dt)``, ``generate_output``, ``cleanup``, ``main``.

The reference takes its configuration from compile-time Fortran parameters
(src/input.f90); here ``configure(HDConfig(...), icid=..., tf=..., ntout=...)``
sets them before ``setup``.  As in the reference there is one simulation per
process (module state).

Arrays: ``q_io``, ``q_1``, ``q_2`` may be
  * torch float64 CUDA tensors of shape (nbasis, nQ, nz, ny, nx) -- updated
    in place on the device; ``t`` and ``dt`` may then be 0-d float64 CUDA
    tensors, in which case nothing synchronises unless ``check_every`` asks;
  * numpy float64 arrays of shape (nx, ny, nz, nQ, nbasis) in Fortran order,
    as the reference's driver allocates them (src/test_hermeshd.py:49-51) --
    copied to and from the device on every call.
``t``/``dt``/``t1``/``t_start``/``dtout`` are 0-d arrays updated in place and
``nout`` a 0-d integer array, like the reference's driver.

Errors: a NaN in a stage output raises NonFiniteError (the reference calls
exit(-1), src/integrator.f90:234-250).
"""
from __future__ import annotations

import time

import numpy as np
import torch

from ._lib import NonFiniteError  # noqa: F401  (re-exported)
from .checkpoint import read_checkpoint, write_checkpoint
from .config import HDConfig
from .ics import setup_problem
from .output import output_vtk
from .solver import Solver

_input = {"cfg": None, "icid": None, "tf": None, "ntout": 200, "clt": 2.0, "check_every": 1,
          "iread": 0, "iwrite": 0, "fpre": "Qout", "data_dir": "data", "resuming": True, "mpi": (1, 1, 1),
          "vis": None, "cylinder": False, "inflow": None, "vtk": (), "outdir": "output", "nvtk": 1}
_state = {"solver": None, "q1_dev": None, "steps": 0}


def configure(cfg: HDConfig, icid: int | None = None, tf: float | None = None, ntout: int = 200,
              clt: float = 2.0, check_every: int = 1, iread: int = 0, iwrite: int = 0, fpre: str = "Qout",
              data_dir: str = "data", resuming: bool = True, mpi=(1, 1, 1), vis: float | None = None,
              cylinder: bool = False, inflow=None, vtk=(), outdir: str = "output", nvtk: int = 1) -> None:
    """Set the run's switches (the reference's input.f90 module).  iread > 0
    starts from checkpoint dump iread (keeping its t, dt, nout when resuming);
    iwrite = 1 writes a checkpoint at every output time, alternating dumps 1
    and 2, split over `mpi` = (mpi_nx, mpi_ny, mpi_nz) rank files as the
    reference would (src/hermeshd.f90:206-266, src/initialize.f90:394-565).
    vis: the viscosity from which icid 0, 5, 6, 7 set coll and colvis (as
    set_ic does); cylinder: apply the cylinder mask of icid 6, 7 after every
    step; inflow: the x-low inflow state (nQ, nface, nz, ny) when cfg.bc[0] is
    "inflow" and the initial condition does not define one; vtk: the output
    switches of src/input.f90:87-96 (e.g. ("nstdout", "nstvout")) written as
    <outdir>/perseus_p0000_t<nout>.vtr at every output time, sampled at
    nvtk^3 points per cell."""
    _input.update(cfg=cfg, icid=icid, tf=tf, ntout=ntout, clt=clt, check_every=max(0, int(check_every)),
                  iread=int(iread), iwrite=int(iwrite), fpre=fpre, data_dir=data_dir, resuming=bool(resuming),
                  mpi=tuple(mpi), vis=vis, cylinder=bool(cylinder), inflow=inflow, vtk=tuple(vtk), outdir=outdir,
                  nvtk=int(nvtk))


def _cfg() -> HDConfig:
    if _input["cfg"] is None:
        raise RuntimeError("call hermeshd_amd.hermeshd.configure(HDConfig(...)) before setup")
    return _input["cfg"]


def _as_device(a, cfg: HDConfig):
    """(tensor on device, writeback function) for a torch or Fortran-ordered numpy state."""
    if isinstance(a, torch.Tensor):
        return a, None
    if isinstance(a, np.ndarray):
        if a.shape != (cfg.nx, cfg.ny, cfg.nz, cfg.nQ, cfg.nbasis):
            raise ValueError(f"numpy state must have shape (nx,ny,nz,nQ,nbasis), got {a.shape}")
        if a.dtype != np.float64:
            raise ValueError("numpy state must be float64 (the build computes in FP64)")
        view = a.T  # C-ordered (nbasis,nQ,nz,ny,nx) view of the same memory when a is F-ordered
        dev = torch.from_numpy(np.ascontiguousarray(view)).to("cuda")

        def back(d):
            view[...] = d.cpu().numpy()
        return dev, back
    raise TypeError("state must be a torch tensor or a numpy array")


def _get_scalar(x) -> float:
    if isinstance(x, torch.Tensor):
        return float(x.item())
    return float(np.asarray(x))


def _set_scalar(x, v) -> None:
    if isinstance(x, torch.Tensor):
        x.fill_(v)
    else:
        x[...] = v


class Hermeshd:
    """Module hermeshd (src/hermeshd.f90:2-267)."""

    @staticmethod
    def setup(q_io, t, dt, t1, t_start, dtout, nout, comm=None):
        """setup (src/hermeshd.f90:67-114): tables and context, initial
        condition (when icid is configured), integrator and boundaries.
        comm (the reference's MPI communicator) is accepted for the call's
        signature and not used: one device holds the whole grid."""
        cfg = _cfg()
        q_ic = inflow = mask = None
        if _input["icid"] is not None and not _input["iread"]:
            cfg, q_ic, inflow, mask = setup_problem(cfg, _input["icid"], _input["vis"])
            _input["cfg"] = cfg
        if _input["inflow"] is not None:
            inflow = _input["inflow"]
        if _state["solver"] is not None:
            _state["solver"].close()
        _state["solver"] = Solver(cfg)
        if cfg.bc[0] == "inflow":
            if inflow is None:
                raise ValueError("the x-low inflow boundary needs a state: configure(..., inflow=...) or icid 5-7")
            _state["solver"].set_inflow(inflow)
        if _input["cylinder"] and mask is not None:
            _state["solver"].set_mask(mask)
        _state["steps"] = 0
        _set_scalar(t, 0.0)
        # init_temporal_params (src/initialize.f90:332-345)
        _set_scalar(dt, cfg.cflm * (1.0 / cfg.dxi) / _input["clt"])
        _set_scalar(nout, 0)
        _set_scalar(t_start, time.time())
        _set_scalar(t1, time.time())
        if _input["tf"] is not None:
            _set_scalar(dtout, _input["tf"] / _input["ntout"])
        q0 = None
        if _input["iread"]:
            # set_ic_from_file (src/initialize.f90:394-439)
            Q, tp, dtp, noutp, _ = read_checkpoint(_input["data_dir"], _input["fpre"], _input["iread"],
                                                   (cfg.nx, cfg.ny, cfg.nz), cfg.nQ, cfg.nbasis)
            q0 = torch.from_numpy(np.ascontiguousarray(Q.T))
            if _input["resuming"]:
                _set_scalar(t, tp)
                _set_scalar(dt, dtp)
                _set_scalar(nout, noutp)
        elif q_ic is not None:
            q0 = torch.from_numpy(q_ic)
        if q0 is not None:
            if isinstance(q_io, torch.Tensor):
                q_io.copy_(q0.to(q_io.device))
            else:
                q_io.T[...] = q0.numpy()

    @staticmethod
    def step(q_io, q_1, q_2, t, dt):
        """step (src/hermeshd.f90:53-62): dt = get_min_dt(q_io); update; t += dt.

        Stage arrays: on return q_1 holds the first stage's output, limited
        (for SSP-RK2 that is the reference's Q_r1 after its step); q_2 holds
        unspecified values, because the SSP combination is fused into the
        stage kernels and the uncombined last stage is never stored.  A numpy
        q_1 receives the values from the device after the step."""
        s: Solver = _state["solver"]
        if s is None:
            raise RuntimeError("setup has not been called")
        cfg = s.cfg
        qd, back = _as_device(q_io, cfg)
        if isinstance(q_1, torch.Tensor):
            q1d = q_1
        else:
            if _state["q1_dev"] is None:
                _state["q1_dev"] = s.zeros()
            q1d = _state["q1_dev"]
        dev_scalars = isinstance(t, torch.Tensor) and isinstance(dt, torch.Tensor) and t.is_cuda and dt.is_cuda
        if dev_scalars:
            s.step(qd, q1d, q_2 if isinstance(q_2, torch.Tensor) else None, t=t, dt=dt)
        else:
            s.t.fill_(_get_scalar(t))
            s.step(qd, q1d, q_2 if isinstance(q_2, torch.Tensor) else None)
        _state["steps"] += 1
        ce = _input["check_every"]
        if back is not None or not dev_scalars or (ce and _state["steps"] % ce == 0):
            s.check()
        if not dev_scalars:
            _set_scalar(dt, float(s.dt.item()))
            _set_scalar(t, float(s.t.item()))
        if back is not None:
            back(qd)
        if isinstance(q_1, np.ndarray):
            if q_1.shape != (cfg.nx, cfg.ny, cfg.nz, cfg.nQ, cfg.nbasis) or q_1.dtype != np.float64:
                raise ValueError("numpy q_1 must be float64 with shape (nx,ny,nz,nQ,nbasis)")
            q_1.T[...] = q1d.cpu().numpy()

    @staticmethod
    def generate_output(q_r, t, dt, t1, dtout, nout):
        """Output scheduling of generate_output (src/hermeshd.f90:206-266):
        advances nout when t passes dtout*nout, writes the configured VTK
        fields (output_vtk) and, with iwrite = 1, a checkpoint in the
        reference's format."""
        tv, dto, n = _get_scalar(t), _get_scalar(dtout), int(np.asarray(nout)) if not isinstance(nout, torch.Tensor) else int(nout.item())
        if dto > 0 and tv > dto * n:
            _set_scalar(nout, n + 1)
            _set_scalar(t1, time.time())
            if _input["vtk"]:
                s: Solver = _state["solver"]
                qd, _ = _as_device(q_r, s.cfg)
                output_vtk(s, qd, n + 1, 0, _input["outdir"], _input["vtk"], _input["nvtk"])
            if _input["iwrite"] == 1:
                # writeQ with an odd/even dump id, so the last two sets are kept
                ioe = 2 - (n + 1) % 2
                write_checkpoint(_input["data_dir"], _input["fpre"], ioe, q_r, tv, _get_scalar(dt), n + 1,
                                 mpi=_input["mpi"])

    @staticmethod
    def cleanup(t_start):
        """cleanup (src/hermeshd.f90:119-138): release the device context."""
        if _state["solver"] is not None:
            _state["solver"].close()
        _state["solver"] = None
        _state["q1_dev"] = None

    @staticmethod
    def main(comm=None):
        """main (src/hermeshd.f90:26-48): setup, step until tf, cleanup."""
        cfg = _cfg()
        if _input["tf"] is None:
            raise RuntimeError("configure(tf=...) is required by main")
        q = torch.zeros(cfg.shape, dtype=torch.float64, device="cuda")
        q1 = torch.zeros_like(q)
        t, dt, t1, t_start, dtout = (np.array(0.0) for _ in range(5))
        nout = np.array(0)
        Hermeshd.setup(q, t, dt, t1, t_start, dtout, nout, comm)
        while float(t) < _input["tf"]:
            Hermeshd.step(q, q1, None, t, dt)
            Hermeshd.generate_output(q, t, dt, t1, dtout, nout)
        Hermeshd.cleanup(t_start)
        return q


hermeshd = Hermeshd()
