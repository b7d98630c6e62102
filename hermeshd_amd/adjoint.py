"""Gradients of a fixed-dt run through the hand-written adjoint kernels.

``run_with_grad(runner, q0, coll)`` advances ``q0`` by ``runner.nsteps``
SSP-RK steps of fixed ``runner.dt`` with relaxation rate ``coll`` and returns
the final state as a torch tensor that autograd can differentiate with
respect to ``q0`` and ``coll``.  The reference has no adjoint; the forward
steps are its RK2/RK3 (src/integrator.f90:62-86).

Memory is bounded by RK-stage checkpointing: the forward pass keeps the state
every ``checkpoint_every`` steps; the backward pass recomputes one segment at
a time, keeping that segment's stage inputs (before and after the limiter)
and stage outputs, and walks it backwards with ``hd_stage_vjp`` -- the
transpose of each stage, the limiter's factors frozen at their forward
values.  dt is a fixed input (as in a dt-frozen sensitivity); the gradient
with respect to coll collects d(out)/d(coll) of every stage (the implicit
relaxation 1/(1 + dt*coll) for ivis = 1, the explicit source -coll*P for
ivis = 0).
"""
from __future__ import annotations

import dataclasses

import torch

from .config import HDConfig
from .solver import Solver

# SSP weights (a, b) of out = a*Q^n + b*E(q_in), per stage
_WEIGHTS = {2: [(0.0, 1.0), (0.5, 0.5)], 3: [(0.0, 1.0), (0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0)]}


@dataclasses.dataclass
class AdjointRunner:
    cfg: HDConfig
    dt: float
    nsteps: int
    checkpoint_every: int = 20
    device: str = "cuda"

    def __post_init__(self):
        if self.cfg.noise:
            raise ValueError("the adjoint does not differentiate the stochastic stress")
        if self.cfg.ivis == 2 or self.cfg.ieos == 2:
            raise ValueError("no adjoint for the full 10-moment equations or the water EOS")
        self.order = 3 if self.cfg.integrator == "shu-osher" else 2
        self._solvers = {}
        self.last_stats = {}

    def _dt(self, s: Solver) -> torch.Tensor:
        """The time step as a device scalar (passed as a tensor, no fill per call)."""
        # keyed on the value too: dt may be changed on the runner between runs
        key = (s.device, float(self.dt))
        t = getattr(self, "_dt_dev", None)
        if t is None or getattr(self, "_dt_key", None) != key:
            t = torch.full((), float(self.dt), dtype=torch.float64, device=s.device)
            self._dt_dev = t
            self._dt_key = key
        return t

    def solver(self, coll: float) -> Solver:
        key = float(coll)
        s = self._solvers.get(key)
        if s is None:
            if len(self._solvers) > 4:
                self._solvers.clear()
            s = Solver(self.cfg.replace(coll=key), device=self.device)
            self._solvers[key] = s
        return s

    # ---------------------------------------------------------------- forward
    def forward(self, q0: torch.Tensor, coll: float):
        s = self.solver(coll)
        q = q0.detach().clone().contiguous()
        q1 = torch.zeros_like(q)
        q2 = torch.zeros_like(q) if self.order == 3 else None
        ck = [q.clone()]
        dt = self._dt(s)
        for n in range(self.nsteps):
            s.update(q, q1, q2, dt)
            if (n + 1) % self.checkpoint_every == 0 and n + 1 < self.nsteps:
                ck.append(q.clone())
        return q, ck

    # ---------------------------------------------------------------- records
    def _recompute(self, s: Solver, q: torch.Tensor, nsteps: int):
        """Stage records of nsteps steps from q: per step, per stage, (raw input,
        limited input, output)."""
        recs = []
        lim_on = self.cfg.limiter
        dt = self._dt(s)
        for _ in range(nsteps):
            raws, lims, outs = [], [], []
            raw = q
            for k in range(self.order):
                lim = raw.clone() if lim_on else raw   # the stage limits lim in place
                out = torch.empty_like(raw)
                s.stage(self.order, k + 1, lim, lims[0] if k > 0 else None, out, dt)
                raws.append(raw)
                lims.append(lim)
                outs.append(out)
                raw = out
            recs.append((raws, lims, outs))
            q = raw
        return recs, q

    def _step_vjp(self, s: Solver, rec, lam: torch.Tensor, dcoll: torch.Tensor) -> torch.Tensor:
        raws, lims, outs = rec
        W = _WEIGHTS[self.order]
        dt = self._dt(s)
        lam_out = lam
        qn_cot = []   # (weight, cotangent) pairs destined for the step input's limited state
        for k in reversed(range(self.order)):
            a, b = W[k]
            lam_in = torch.empty_like(lam)
            acc = []
            if k == 0:
                acc = qn_cot
            s.stage_vjp(lims[k], lam_out, lam_in, dt, a, b, acc=acc, q_raw=raws[k], out=outs[k],
                        qn=lims[0] if a != 0.0 else None, dcoll=dcoll)
            if a != 0.0:
                qn_cot.append((a, lam_out))
            lam_out = lam_in
        return lam_out

    def backward(self, coll: float, ck, grad_out: torch.Tensor):
        s = self.solver(coll)
        lam = grad_out.detach().clone().contiguous()
        dcoll = torch.zeros((), dtype=torch.float64, device=lam.device)
        K = self.checkpoint_every
        nseg = len(ck)
        for seg in reversed(range(nseg)):
            start = seg * K
            n = min(self.nsteps, start + K) - start
            recs, _ = self._recompute(s, ck[seg].clone(), n)
            for rec in reversed(recs):
                lam = self._step_vjp(s, rec, lam, dcoll)
            del recs
        return lam, dcoll


class _RunFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q0, coll, runner):
        c = float(coll.detach())
        q, ck = runner.forward(q0, c)
        ctx.runner, ctx.coll, ctx.ck = runner, c, ck
        return q

    @staticmethod
    def backward(ctx, grad_out):
        lam, dcoll = ctx.runner.backward(ctx.coll, ctx.ck, grad_out)
        ctx.ck = None
        return lam, dcoll.to(grad_out.dtype), None


def run_with_grad(runner: AdjointRunner, q0: torch.Tensor, coll: torch.Tensor) -> torch.Tensor:
    """Final state after runner.nsteps steps; differentiable in q0 and coll (0-d tensor)."""
    return _RunFn.apply(q0, coll, runner)


def wall_shear(q: torch.Tensor) -> torch.Tensor:
    """Mean wall shear stress of a channel with no-slip walls in y: the
    cell-average Pxy of the wall cells, low wall minus high wall, over x (and
    z), halved."""
    pxy = q[0, 8]
    return 0.5 * (pxy[:, 0, :].mean() - pxy[:, -1, :].mean())


def channel_state(cfg: HDConfig, u0: torch.Tensor, rho: float = 1.0, T: float = 1.0) -> torch.Tensor:
    """Cell-average state of a channel at uniform density and temperature with
    x velocity u0 (shape (nz, ny, nx)), zero stress: differentiable in u0."""
    z = torch.zeros_like(u0)
    P = rho * T
    mode0 = torch.stack([z + rho, rho * u0, z, z, P / (cfg.gamma - 1.0) + 0.5 * rho * u0 * u0] + [z] * (cfg.nQ - 5))
    rest = torch.zeros((cfg.nbasis - 1, cfg.nQ) + tuple(u0.shape), dtype=u0.dtype, device=u0.device)
    return torch.cat([mode0.unsqueeze(0), rest], dim=0).contiguous()


def kinetic_energy(q: torch.Tensor, cfg: HDConfig) -> torch.Tensor:
    """Kinetic energy of the cell averages, sum over cells of dV |m|^2 / (2 rho)."""
    dV = (cfg.lx / cfg.nx) * (cfg.ly / cfg.ny) * (cfg.lz / cfg.nz)
    m0 = q[0]
    return 0.5 * dV * ((m0[1] ** 2 + m0[2] ** 2 + m0[3] ** 2) / m0[0]).sum()


def velocity_state(cfg: HDConfig, vel: torch.Tensor, rho: float = 1.0, p: float | None = None) -> torch.Tensor:
    """Cell-average state at uniform density and pressure (default 100/gamma,
    the Taylor-Green vortex) with velocity vel (shape (3, nz, ny, nx)), zero
    stress and heat flux: differentiable in vel."""
    p = 100.0 / cfg.gamma if p is None else p
    z = torch.zeros_like(vel[0])
    ke = 0.5 * rho * (vel ** 2).sum(0)
    mode0 = torch.stack([z + rho, rho * vel[0], rho * vel[1], rho * vel[2], p / (cfg.gamma - 1.0) + ke]
                        + [z] * (cfg.nQ - 5))
    rest = torch.zeros((cfg.nbasis - 1, cfg.nQ) + tuple(z.shape), dtype=vel.dtype, device=vel.device)
    return torch.cat([mode0.unsqueeze(0), rest], dim=0).contiguous()
