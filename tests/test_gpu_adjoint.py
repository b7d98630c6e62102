"""Hand-written adjoint (HIP) against autograd through the PyTorch FP64
transcription, and the gradient of a whole run against central finite
differences.

Tolerances: vector-Jacobian products 1e-11 relative per physical quantity
(both sides are exact derivatives of the same arithmetic up to rounding);
run gradients 1e-6 relative against central differences.
"""
import numpy as np
import pytest
import torch

from hermeshd_amd import ics
from hermeshd_amd.config import poiseuille_2d

from tests.test_gpu_parity import field_rel_err, smooth_state
from tests.test_torch_transcription import CFGS, inflow_of
from tests.torch_dg import DG

pytestmark = pytest.mark.gpu
VJP_TOL = 1e-11


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch.device("cuda:0")


def _state(cfg, seed):
    q = smooth_state(cfg, seed=seed)
    if cfg.bc[0] == "outflow":
        q[0, 1] += 0.2 * np.sign(np.random.default_rng(seed).standard_normal(q[0, 1].shape))
    return q


def _limited_state(cfg, seed):
    """A state on which every limiter branch acts somewhere (with rh_floor 0.55,
    T_floor 0.5): density floor, density slope factor, pressure factor."""
    q = _state(cfg, seed)
    q[1:, :5] *= 3.0
    q[0, 0].flat[::4] = 0.62
    ke = 0.5 * (q[0, 1] ** 2 + q[0, 2] ** 2 + q[0, 3] ** 2) / q[0, 0]
    q[0, 4].flat[1::4] = (0.45 / (cfg.gamma - 1) + ke).flat[1::4]
    q[1:4, 4].reshape(3, -1)[:, 1::4] *= 4.0
    q[0, 0].flat[2::9] = 0.5
    return q


@pytest.mark.parametrize("name", list(CFGS))
def test_rhs_vjp_matches_autograd(dev, oracle_mod, name):
    from hermeshd_amd.solver import Solver
    cfg = CFGS[name]
    print(f"[rhs_vjp] {name}", flush=True)
    q = _state(cfg, 5)
    lam = np.random.default_rng(6).standard_normal(q.shape)
    dg = DG(cfg, oracle_mod.Oracle(cfg.as_ints()).tables(), inflow=inflow_of(cfg))
    qt = torch.from_numpy(q).requires_grad_(True)
    (ref,) = torch.autograd.grad(dg.rhs(qt), qt, torch.from_numpy(lam))
    s = Solver(cfg)
    if cfg.bc[0] == "inflow":
        s.set_inflow(inflow_of(cfg))
    got = s.rhs_vjp(torch.from_numpy(q).to(dev), torch.from_numpy(lam).to(dev)).cpu().numpy()
    assert field_rel_err(got, ref.numpy()) < VJP_TOL


@pytest.mark.parametrize("name,stage", [("nb4_nq11_hllc_noslip_ivis1", 2), ("nb4_nq11_hllc_walls_ivis0", 1),
                                        ("nb4_nq11_hllc_inflow_ivis1", 3),
                                        ("nb10_nq5_hllc_periodic", 3), ("nb8_nq5_central_outflow", 2),
                                        ("nb4_nq11_roe_walls_ivis1", 2), ("nb4_nq13_roe_periodic_ivis1", 3),
                                        ("nb10_nq13_roe_walls_ivis0", 2)])
def test_stage_vjp_matches_autograd(dev, oracle_mod, name, stage):
    """Transpose of one stage, limiter on with active cells (frozen factors),
    SSP weights and the accumulated q_n cotangent, and d/dcoll."""
    from hermeshd_amd.solver import Solver
    cfg = CFGS[name].replace(limiter=True, rh_floor=0.55, T_floor=0.5)
    print(f"[stage_vjp] {name} stage {stage}", flush=True)
    rng = np.random.default_rng(7)
    raw = _limited_state(cfg, 8)
    qn = _state(cfg, 9)
    lam = rng.standard_normal(raw.shape)
    acc = rng.standard_normal(raw.shape)
    a, b = {1: (0.0, 1.0), 2: (0.75, 0.25), 3: (1.0 / 3.0, 2.0 / 3.0)}[stage]
    dt = 2e-3
    dg = DG(cfg, oracle_mod.Oracle(cfg.as_ints()).tables(), inflow=inflow_of(cfg))
    f, _, th_r, th_p = dg.limit_factors(torch.from_numpy(raw))
    assert bool(f.any()) and bool((th_r < 1).any()) and bool((th_p < 1).any()), "every limiter branch should act"
    rt = torch.from_numpy(raw).requires_grad_(True)
    ct = torch.tensor(float(cfg.coll), dtype=torch.float64, requires_grad=True)
    out = a * torch.from_numpy(qn) + b * dg.euler(dg.limit(rt), dt, coll=ct)
    g_raw, g_coll = torch.autograd.grad(out, (rt, ct), torch.from_numpy(lam), allow_unused=True)
    g_coll = torch.zeros(()) if g_coll is None else g_coll
    # the same through the kernel: limit a copy, run the stage forward for out, then the transpose
    s = Solver(cfg)
    if cfg.bc[0] == "inflow":
        s.set_inflow(inflow_of(cfg))
    lim_d = torch.from_numpy(raw).to(dev)
    qn_d = torch.from_numpy(qn).to(dev)
    s.limit(lim_d)
    # the stage output (for d/dcoll): out = a*qn + b*E(lim)
    e_d = torch.empty_like(lim_d)
    s.euler_step(lim_d.clone(), e_d, dt)
    out_d = a * qn_d + b * e_d
    lam_in = torch.empty_like(lim_d)
    dcoll = torch.zeros((), dtype=torch.float64, device=dev)
    acc_d = torch.from_numpy(acc).to(dev)
    s.stage_vjp(lim_d, torch.from_numpy(lam).to(dev), lam_in, dt, a, b, acc=[(1.0, 0.5 * acc_d)],
                q_raw=torch.from_numpy(raw).to(dev), out=out_d, qn=qn_d, dcoll=dcoll)
    got = lam_in.cpu().numpy()
    # the accumulated term is added before the limiter transpose
    lt = torch.from_numpy(raw).requires_grad_(True)
    (g_acc,) = torch.autograd.grad(dg.limit(lt), lt, torch.from_numpy(0.5 * acc))
    ref = g_raw.numpy() + g_acc.numpy()
    assert field_rel_err(got, ref) < VJP_TOL
    if cfg.nQ >= 11 and cfg.coll != 0.0:
        assert abs(float(dcoll) - float(g_coll)) <= 1e-9 * max(abs(float(g_coll)), 1e-30)


def _c3_small(nx=12, ny=16):
    return poiseuille_2d(nx, ny, vis=0.05)


def _objective(q, w):
    from hermeshd_amd.adjoint import wall_shear
    return wall_shear(q) + (w * q).sum()


@pytest.mark.parametrize("riemann,h", [("central", 1e-6), ("hllc", 1e-8)])
def test_run_gradient_vs_finite_differences(dev, riemann, h):
    """d J / d(u0, coll) of a 24-step run with checkpoints every 5 steps,
    along random directions, against central differences.  With 11 fields
    the HLLC star states of the two sides differ at a zero contact speed
    (the stress terms of the physical flux are not in the star state), so
    the run is discontinuous where a contact speed changes sign; the HLLC
    case therefore uses a step small enough that no face crosses over.
    The coll direction takes a relative step of 1e-6 in both cases: at the
    u step of 1e-8 its difference quotient is rounding noise (J = 0.0165 and
    a coll step of 2e-7 scatter it by 1e-6 relative), while relative steps
    from 5e-8 to 1e-5 cross no contact and agree with the adjoint to
    3e-8 .. 1.5e-9."""
    from hermeshd_amd.adjoint import AdjointRunner, channel_state, run_with_grad
    cfg = _c3_small().replace(riemann=riemann)
    dt = 0.05
    runner = AdjointRunner(cfg, dt=dt, nsteps=24, checkpoint_every=5)
    rng = np.random.default_rng(11)
    u0 = torch.from_numpy(0.1 * ics.gaussian_field((1, cfg.ny, cfg.nx), 3.0, 5)).to(dev)
    w = torch.from_numpy(1e-3 * rng.standard_normal(cfg.shape)).to(dev)
    coll0 = float(cfg.coll)

    def J(u, c):
        q0 = channel_state(cfg, u)
        return _objective(run_with_grad(runner, q0, torch.tensor(c, dtype=torch.float64, device=dev)), w)

    u = u0.clone().requires_grad_(True)
    c = torch.tensor(coll0, dtype=torch.float64, device=dev, requires_grad=True)
    q0 = channel_state(cfg, u)
    val = _objective(run_with_grad(runner, q0, c), w)
    gu, gc = torch.autograd.grad(val, (u, c))
    du = torch.from_numpy(ics.gaussian_field((1, cfg.ny, cfg.nx), 2.0, 9)).to(dev)
    with torch.no_grad():
        fd_u = (J(u0 + h * du, coll0) - J(u0 - h * du, coll0)) / (2 * h)
        hc = 1e-6 * coll0
        fd_c = (J(u0, coll0 + hc) - J(u0, coll0 - hc)) / (2 * hc)
    adj_u = float((gu * du).sum())
    print(f"[run_grad] {riemann} u: adjoint {adj_u:.12e} fd {float(fd_u):.12e}; coll: adjoint {float(gc):.12e} "
          f"fd {float(fd_c):.12e}", flush=True)
    assert abs(adj_u - float(fd_u)) <= 1e-6 * abs(float(fd_u))
    assert abs(float(gc) - float(fd_c)) <= 1e-6 * abs(float(fd_c))


def test_checkpoint_interval_does_not_change_gradient(dev):
    from hermeshd_amd.adjoint import AdjointRunner, channel_state, run_with_grad
    cfg = _c3_small(8, 12)
    u0 = torch.from_numpy(0.1 * ics.gaussian_field((1, cfg.ny, cfg.nx), 3.0, 7)).to(dev)
    grads = []
    for K in (1, 4, 100):
        runner = AdjointRunner(cfg, dt=0.05, nsteps=10, checkpoint_every=K)
        u = u0.clone().requires_grad_(True)
        c = torch.tensor(float(cfg.coll), dtype=torch.float64, device=dev, requires_grad=True)
        J = _objective(run_with_grad(runner, channel_state(cfg, u), c), 0.0)
        grads.append(torch.autograd.grad(J, (u, c)))
    for g in grads[1:]:
        assert torch.equal(g[0], grads[0][0]) and torch.equal(g[1], grads[0][1])


def test_run_gradient_rk3_limiter_vs_autograd(dev, oracle_mod):
    """SSP-RK3 with the limiter on (frozen active set): the run gradient of
    3 steps against autograd through the transcription's steps."""
    from hermeshd_amd.adjoint import AdjointRunner, run_with_grad
    cfg = CFGS["nb10_nq5_hllc_periodic"].replace(integrator="shu-osher", limiter=True, rh_floor=0.55,
                                                  T_floor=0.5)
    q = _limited_state(cfg, 12)
    lam = np.random.default_rng(13).standard_normal(q.shape)
    dt = 1e-3
    dg = DG(cfg, oracle_mod.Oracle(cfg.as_ints()).tables())
    qt = torch.from_numpy(q).requires_grad_(True)
    x = qt
    for _ in range(3):
        x0 = dg.limit(x)
        x1 = dg.euler(x0, dt)
        x2 = 0.75 * x0 + 0.25 * dg.euler(dg.limit(x1), dt)
        x = x0 / 3.0 + 2.0 / 3.0 * dg.euler(dg.limit(x2), dt)
    (ref,) = torch.autograd.grad(x, qt, torch.from_numpy(lam))
    runner = AdjointRunner(cfg, dt=dt, nsteps=3, checkpoint_every=2)
    q0 = torch.from_numpy(q).to(dev).requires_grad_(True)
    out = run_with_grad(runner, q0, torch.tensor(0.0, dtype=torch.float64, device=dev))
    assert field_rel_err(out.detach().cpu().numpy(), x.detach().numpy()) < 1e-12
    (got,) = torch.autograd.grad(out, q0, torch.from_numpy(lam).to(dev))
    assert field_rel_err(got.cpu().numpy(), ref.numpy()) < VJP_TOL


def test_taylor_green_kinetic_energy_gradient_vs_finite_differences(dev):
    """The 13-moment Taylor-Green vortex (nbasis 10, Roe, limiter on,
    SSP-RK3) on 8^3 cells: d KE(t) / d u0 of a 12-step run with checkpoints
    every 5 steps, along a random direction of the initial velocity field,
    against central differences (1e-6).  The Roe dissipation has |u_n|, which
    has a kink at zero normal velocity, and the vortex has planes where the
    normal velocity is only the 1e-4 seeded noise: a difference step of 1e-6
    crosses that kink at a few tens of face points (an error of 2e-6 in the
    difference quotient), a step of 1e-8 at almost none."""
    from hermeshd_amd.adjoint import AdjointRunner, kinetic_energy, run_with_grad, velocity_state
    from hermeshd_amd.config import taylor_green_3d
    from hermeshd_amd.solver import Solver
    cfg = taylor_green_3d(8)
    q_ic = torch.from_numpy(ics.taylor_green(cfg)).to(dev)
    vel0 = (q_ic[0, 1:4] / q_ic[0, 0]).clone()
    dt = float(Solver(cfg).min_dt(q_ic))
    runner = AdjointRunner(cfg, dt=dt, nsteps=12, checkpoint_every=5)
    coll = torch.tensor(float(cfg.coll), dtype=torch.float64, device=dev)

    def J(v):
        return kinetic_energy(run_with_grad(runner, velocity_state(cfg, v), coll), cfg)

    v = vel0.clone().requires_grad_(True)
    (gv,) = torch.autograd.grad(J(v), v)
    dv = torch.from_numpy(np.random.default_rng(21).standard_normal(tuple(vel0.shape))).to(dev)
    h = 1e-8
    with torch.no_grad():
        fd = (J(vel0 + h * dv) - J(vel0 - h * dv)) / (2 * h)
    adj = float((gv * dv).sum())
    print(f"[tgv_grad] adjoint {adj:.12e} fd {float(fd):.12e}", flush=True)
    assert abs(adj - float(fd)) <= 1e-6 * abs(float(fd))


@pytest.mark.parametrize("limiter,a", [(False, 0.0), (True, 0.75)])
def test_thirteen_moment_split_transpose_matches_fused(dev, limiter, a):
    """On periodic 13-moment grids the stage transpose runs split (one wave per
    face point, then one field per wave); HD_VJP=fused selects the per-cell
    kernel.  Both give the same cotangent and d/dcoll to rounding.  With
    a = 0 (a first stage) no q_n is passed: d/dcoll then reads out only."""
    import os
    from hermeshd_amd.config import taylor_green_3d
    from hermeshd_amd.solver import Solver
    cfg = taylor_green_3d(8).replace(limiter=limiter, rh_floor=0.55, T_floor=0.5)
    print(f"[split_vjp] limiter={limiter} a={a}", flush=True)
    rng = np.random.default_rng(11)
    raw = _limited_state(cfg, 12) if limiter else ics.taylor_green(cfg)
    qn = ics.taylor_green(cfg)
    lam = rng.standard_normal(raw.shape)
    acc = rng.standard_normal(raw.shape)
    dt, b = 1e-4, 1.0 - a
    res = []
    for mode in ("split", "fused"):
        old = os.environ.get("HD_VJP")
        if mode == "fused":
            os.environ["HD_VJP"] = "fused"
        try:
            s = Solver(cfg)
        finally:
            if old is None:
                os.environ.pop("HD_VJP", None)
            else:
                os.environ["HD_VJP"] = old
        lim_d = torch.from_numpy(raw).to(dev)
        qn_d = torch.from_numpy(qn).to(dev)
        if limiter:
            s.limit(lim_d)
        e_d = torch.empty_like(lim_d)
        s.euler_step(lim_d.clone(), e_d, dt)
        out_d = a * qn_d + b * e_d
        lam_in = torch.empty_like(lim_d)
        dcoll = torch.zeros((), dtype=torch.float64, device=dev)
        s.stage_vjp(lim_d, torch.from_numpy(lam).to(dev), lam_in, dt, a, b,
                    acc=[(1.0, torch.from_numpy(acc).to(dev))],
                    q_raw=torch.from_numpy(raw).to(dev) if limiter else None, out=out_d,
                    qn=qn_d if a != 0.0 else None, dcoll=dcoll)
        rv = s.rhs_vjp(lim_d, torch.from_numpy(lam).to(dev))
        res.append((lam_in.cpu().numpy(), float(dcoll), rv.cpu().numpy()))
    assert field_rel_err(res[0][0], res[1][0]) < 1e-12
    assert field_rel_err(res[0][2], res[1][2]) < 1e-12
    assert abs(res[0][1] - res[1][1]) <= 1e-12 * max(abs(res[1][1]), 1e-30)
