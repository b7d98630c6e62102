"""Ensemble path (hd_ens_*) on the GPU.

Each realization of an ensemble must compute exactly what one context with
the same configuration and realization id computes (same kernels per cell,
same noise keys), whatever the launch shape; the graph-replayed steps must
equal eagerly launched steps; and every realization matches the oracle to
the parity tolerance of the single-context tests (1e-12 relative per
physical quantity).
"""
import numpy as np
import pytest

from hermeshd_amd import ics
from hermeshd_amd.config import HDConfig, fluct_ensemble

from tests.test_gpu_parity import TOL, field_rel_err, smooth_state

pytestmark = pytest.mark.gpu

DIMS = [(4, 4, 4), (5, 3, 6), (6, 6, 2), (3, 5, 4)]


def noisy_base(**kw):
    base = dict(nx=4, ny=4, nz=4, nbasis=4, iquad=2, nQ=11, ivis=0, coll=3.0, colvis=0.9, riemann="hllc",
                noise=True, eta_sd=0.02, zeta_sd=0.004, lx=4.0, ly=4.0, lz=4.0, T_floor=1e-4, seed=77,
                realization=10, integrator="heun", limiter=False)
    base.update(kw)
    return HDConfig(**base)


CASES = {
    "noise_nb4_hllc_rk2": noisy_base(),
    "noise_nb4_roe_walls_ivis1": noisy_base(riemann="roe", ivis=1, limiter=True,
                                           bc=("periodic", "periodic", "noslip", "noslip", "wall", "wall")),
    "nb10_rk3_limiter": HDConfig(nx=4, ny=4, nz=4, nbasis=10, iquad=3, nQ=5, riemann="hllc", limiter=True,
                                 integrator="shu-osher", lx=4.0, ly=4.0, lz=4.0, T_floor=1e-4),
}


def realization_cfg(base, dims, r):
    nx, ny, nz = dims[r]
    h = (base.lx / base.nx, base.ly / base.ny, base.lz / base.nz)
    return base.replace(nx=nx, ny=ny, nz=nz, lx=h[0] * nx, ly=h[1] * ny, lz=h[2] * nz,
                        realization=base.realization + r)


def states(base, dims, seed):
    return [smooth_state(realization_cfg(base, dims, r), seed=seed + r) for r in range(len(dims))]


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch


@pytest.mark.parametrize("name", list(CASES))
def test_ensemble_equals_single_contexts(torch_mod, name):
    """rhs and three steps of every realization are bitwise those of a
    single context with the same configuration and realization id."""
    from hermeshd_amd.ensemble import Ensemble
    from hermeshd_amd.solver import Solver
    base = CASES[name]
    print(f"[ens_vs_single] {name}", flush=True)
    ens = Ensemble(base, DIMS)
    qs = states(base, DIMS, seed=11)
    q = ens.pack(qs)
    r_ens = ens.rhs(q)
    q1 = ens.zeros()
    q2 = ens.zeros() if base.integrator == "shu-osher" else None
    for _ in range(3):
        ens.update(q, q1, q2, 2e-3)
    ens.check()
    for r in range(len(DIMS)):
        cfg = realization_cfg(base, DIMS, r)
        s = Solver(cfg)
        qd = torch_mod.from_numpy(qs[r]).cuda()
        assert torch_mod.equal(s.rhs(qd), ens.views(r_ens)[r]), f"rhs of realization {r}"
        a1 = torch_mod.zeros_like(qd)
        a2 = torch_mod.zeros_like(qd) if q2 is not None else None
        for _ in range(3):
            s.update(qd, a1, a2, 2e-3)
        assert torch_mod.equal(qd, ens.views(q)[r]), f"state of realization {r} after 3 steps"


@pytest.mark.parametrize("name", ["noise_nb4_hllc_rk2", "noise_nb4_roe_walls_ivis1"])
def test_ensemble_vs_oracle(torch_mod, oracle_mod, name):
    from hermeshd_amd.ensemble import Ensemble
    base = CASES[name]
    print(f"[ens_vs_oracle] {name}", flush=True)
    ens = Ensemble(base, DIMS)
    qs = states(base, DIMS, seed=21)
    q = ens.pack(qs)
    r_ens = [v.cpu().numpy() for v in ens.views(ens.rhs(q))]
    q1 = ens.zeros()
    for _ in range(2):
        ens.update(q, q1, None, 2e-3)
    out = [v.cpu().numpy() for v in ens.views(q)]
    for r in range(len(DIMS)):
        o = oracle_mod.Oracle(realization_cfg(base, DIMS, r).as_ints())
        qo = qs[r].copy()
        assert field_rel_err(r_ens[r], o.rhs(qo)) < TOL
        a1, a2 = np.zeros_like(qo), np.zeros_like(qo)
        for _ in range(2):
            o.update(qo, a1, a2, 2e-3)
        assert field_rel_err(out[r], qo) < TOL


def test_ensemble_every_launch_shape_agrees(torch_mod):
    """The result does not depend on the schedule (tile shape, persistence,
    XCD chunking)."""
    from hermeshd_amd.ensemble import Ensemble, launch_shapes
    base = CASES["noise_nb4_hllc_rk2"]
    ens = Ensemble(base, DIMS)
    q0 = ens.pack(states(base, DIMS, seed=31))
    ref = None
    for c, shape in enumerate(launch_shapes()):
        print(f"[ens_launch] {c} {shape}", flush=True)
        ens.set_launch(c)
        ens.set_step(5)
        q = q0.clone()
        q1 = ens.zeros()
        ens.update(q, q1, None, 2e-3)
        if ref is None:
            ref = q
        else:
            assert torch_mod.equal(q, ref), f"launch shape {c}"
    ens.check()


def test_ensemble_graph_replay_equals_eager(torch_mod):
    """hd_ens_run replays a captured step: 3 + 2 replays (eager work and a
    synchronisation between) equal 5 eager steps, noise keys included."""
    from hermeshd_amd.ensemble import Ensemble
    base = CASES["noise_nb4_hllc_rk2"]
    ens = Ensemble(base, DIMS)
    q0 = ens.pack(states(base, DIMS, seed=41))
    qa, qa1 = q0.clone(), ens.zeros()
    for _ in range(5):
        ens.update(qa, qa1, None, 1e-3)
    ens.set_step(0)
    qb, qb1 = q0.clone(), ens.zeros()
    ens.run(qb, qb1, None, 1e-3, 3)
    torch_mod.cuda.synchronize()
    _ = (qb * 1.0).sum().item()
    ens.run(qb, qb1, None, 1e-3, 2)
    ens.check()
    assert torch_mod.equal(qa, qb)


def test_ensemble_graph_replay_new_dt(torch_mod):
    """A replayed step draws the next step's noise ahead of time; a later call
    with another dt redraws it, so replays still equal eager steps when dt
    changes between calls."""
    from hermeshd_amd.ensemble import Ensemble
    base = CASES["noise_nb4_hllc_rk2"]
    ens = Ensemble(base, DIMS)
    q0 = ens.pack(states(base, DIMS, seed=43))
    qa, qa1 = q0.clone(), ens.zeros()
    for dt in (1e-3, 1e-3, 2e-3, 2e-3):
        ens.update(qa, qa1, None, dt)
    ens.set_step(0)
    qb, qb1 = q0.clone(), ens.zeros()
    ens.run(qb, qb1, None, 1e-3, 2)
    ens.run(qb, qb1, None, 2e-3, 2)
    ens.check()
    assert torch_mod.equal(qa, qb)


def test_ensemble_tuner(torch_mod):
    from hermeshd_amd.ensemble import Ensemble, launch_shapes
    base = CASES["noise_nb4_hllc_rk2"]
    ens = Ensemble(base, DIMS)
    q = ens.pack(states(base, DIMS, seed=51))
    choice, us = ens.tune(q, ens.zeros(), 1e-3, reps=2)
    assert 0 <= choice < len(launch_shapes()) and ens.launch() == choice
    assert all(t > 0 for t in us) and us[choice] == min(us)
    ens.check()


def test_ensemble_rejects_bad_input(torch_mod):
    from hermeshd_amd._lib import HDError
    from hermeshd_amd.ensemble import Ensemble
    base = CASES["noise_nb4_hllc_rk2"]
    with pytest.raises(HDError):
        Ensemble(base, [(4, 4, 0)])
    with pytest.raises(HDError):
        Ensemble(base.replace(nQ=5, noise=False, ivis=0), [(4, 4, 4)]).set_launch(999)


def test_structure_factor_linear_response(torch_mod, oracle_mod):
    """Static structure factors of 256 realizations over 2,000 steps against
    the exact linear-response prediction of the discrete scheme
    (tests/fdt_linear.py).  The prediction is built from the oracle's one-step
    stencils and includes the approach to equilibrium from the quiescent
    start.  The Roe flux is used because it is smooth at rest, so the
    linearisation is exact for the velocity channels.  The HLLC flux is not
    smooth at rest: the stress part of its momentum flux switches sides with
    the sign of the contact speed.  S_rr is not gated: the density slopes and
    the mean density (entropy modes) are neutral in the linearisation and are
    damped only by the |u|-weighted contact term, which is nonlinear.  The
    low-k bins must agree to 5%; sampling error is about 1%."""
    from hermeshd_amd.ensemble import Ensemble, StructureFactor
    from tests import fdt_linear
    kB, dt, equil, nsamp, every = 1e-4, 0.1, 1000, 100, 10
    base, dims = fluct_ensemble(sizes=((256, 16),), kB=kB)
    base = base.replace(riemann="roe")
    ens = Ensemble(base, dims)
    q = ens.pack([ics.uniform(realization_cfg(base, dims, r), rho=1.0, T=1.0) for r in range(len(dims))])
    q1 = ens.zeros()
    sf = StructureFactor(ens, kB=kB)
    print(f"[sk] {len(dims)} realizations of 16^3, {equil} + {nsamp}x{every} steps", flush=True)
    ens.run(q, q1, None, dt, equil)
    for n in range(nsamp):
        ens.run(q, q1, None, dt, every)
        sf.sample(q)
    ens.check()
    edges = [0.2, 0.6, 1.0, 1.4]
    b = sf.binned(edges)
    model = fdt_linear.LinearModel(realization_cfg(base, dims, 0), dt)
    pred = fdt_linear.binned_prediction(model, dims, edges, kB, equil, nsamp, every)
    for key in ("S_uL", "S_uT", "S_rr"):
        print(f"[sk] {key} measured {np.round(b[key], 4).tolist()} predicted {np.round(pred[key], 4).tolist()}",
              flush=True)
    assert np.array_equal(b["n"], pred["n"])
    for key in ("S_uL", "S_uT"):
        assert np.all(np.abs(b[key][:2] / pred[key][:2] - 1.0) < 0.05), key
