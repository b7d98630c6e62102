"""The PyTorch transcription (tests/torch_dg.py) against the C oracle (CPU):
its forward values must match before autograd through it can check the HIP
adjoint."""
import numpy as np
import pytest
import torch

from hermeshd_amd import ics
from hermeshd_amd.config import HDConfig, poiseuille_2d

from tests.test_gpu_parity import field_rel_err, smooth_state
from tests.torch_dg import DG

CFGS = {
    "nb4_nq11_hllc_noslip_ivis1": poiseuille_2d(6, 8).replace(body_force=(1e-2, 3e-3, 0.0)),
    "nb4_nq11_hllc_walls_ivis0": HDConfig(nx=5, ny=4, nz=3, nbasis=4, iquad=2, nQ=11, ivis=0, coll=20.0, colvis=0.7,
                                          riemann="hllc", bc=("wall", "wall", "noslip", "noslip", "periodic",
                                                              "periodic"), lx=1.0, ly=0.8, lz=0.6, T_floor=1e-4),
    "nb8_nq5_central_outflow": HDConfig(nx=4, ny=5, nz=3, nbasis=8, iquad=2, nQ=5, riemann="central",
                                        bc=("outflow",) * 6, lx=0.8, ly=1.0, lz=0.6, T_floor=1e-4),
    "nb10_nq5_hllc_periodic": HDConfig(nx=4, ny=3, nz=3, nbasis=10, iquad=3, nQ=5, riemann="hllc",
                                       lx=1.2, ly=0.9, lz=0.9, T_floor=1e-4),
    "nb10_nq11_hllc_mixed": HDConfig(nx=3, ny=4, nz=3, nbasis=10, iquad=3, nQ=11, ivis=1, coll=30.0, colvis=0.5,
                                     riemann="hllc", bc=("outflow", "outflow", "wall", "wall", "noslip", "noslip"),
                                     lx=0.9, ly=1.2, lz=0.9, T_floor=1e-4),
    # the pipe of icid 5: parabolic inflow, outflow, no-slip walls
    "nb4_nq11_hllc_inflow_ivis1": HDConfig(nx=5, ny=6, nz=1, nbasis=4, iquad=2, nQ=11, ivis=1, coll=10.0,
                                           colvis=0.5, riemann="hllc",
                                           bc=("inflow", "outflow", "noslip", "noslip", "periodic", "periodic"),
                                           lx=1.0, ly=1.2, lz=0.2, T_floor=1e-4),
    # Roe (legacy solver) and the 13-moment model
    "nb4_nq11_roe_walls_ivis1": HDConfig(nx=5, ny=4, nz=3, nbasis=4, iquad=2, nQ=11, ivis=1, coll=15.0, colvis=0.4,
                                         riemann="roe", bc=("noslip", "noslip", "periodic", "periodic", "wall",
                                                            "wall"), lx=1.0, ly=0.8, lz=0.6, T_floor=1e-4),
    "nb4_nq13_roe_periodic_ivis1": HDConfig(nx=4, ny=5, nz=3, nbasis=4, iquad=2, nQ=13, ivis=1, coll=20.0,
                                            colvis=0.5, coll_q=10.0, col_q=0.8, riemann="roe", lx=0.8, ly=1.0,
                                            lz=0.6, T_floor=1e-4),
    "nb10_nq13_roe_walls_ivis0": HDConfig(nx=3, ny=4, nz=3, nbasis=10, iquad=3, nQ=13, ivis=0, coll=3.0,
                                          colvis=0.2, coll_q=2.0, col_q=0.3, riemann="roe",
                                          bc=("periodic", "periodic", "noslip", "noslip", "wall", "wall"),
                                          lx=0.9, ly=1.2, lz=0.9, T_floor=1e-4),
}


def inflow_of(cfg):
    """The x-low inflow state used with an inflow configuration (peak 0.3,
    rho = 1, p = 1), or None."""
    return ics.inflow_profile(cfg, 0.3, 1.0, 1.0) if cfg.bc[0] == "inflow" else None


def oracle_for(oracle_mod, cfg):
    o = oracle_mod.Oracle(cfg.as_ints())
    if cfg.bc[0] == "inflow":
        o.set_inflow(inflow_of(cfg))
    return o


@pytest.mark.parametrize("name", list(CFGS))
def test_transcription_rhs_matches_oracle(oracle_mod, name):
    cfg = CFGS[name]
    o = oracle_for(oracle_mod, cfg)
    q = smooth_state(cfg, seed=3)
    if "outflow" in name:
        q[0, 1] += 0.2 * np.sign(np.random.default_rng(1).standard_normal(q[0, 1].shape))
    dg = DG(cfg, o.tables(), inflow=inflow_of(cfg))
    r = dg.rhs(torch.from_numpy(q)).numpy()
    assert field_rel_err(r, o.rhs(q)) < 1e-12


@pytest.mark.parametrize("name", ["nb4_nq11_hllc_walls_ivis0", "nb10_nq5_hllc_periodic"])
def test_transcription_limiter_matches_oracle(oracle_mod, name):
    cfg = CFGS[name].replace(rh_floor=0.6, T_floor=0.9)
    o = oracle_mod.Oracle(cfg.as_ints())
    rng = np.random.default_rng(4)
    q = ics.uniform(cfg, rho=0.8, T=1.0)
    q[1:, :5] = 0.15 * rng.standard_normal(q[1:, :5].shape)
    q[0, 0].flat[::5] = 0.5
    dg = DG(cfg, o.tables())
    lim = dg.limit(torch.from_numpy(q.copy())).numpy()
    o.limiter(q)
    assert field_rel_err(lim, q) < 1e-13


@pytest.mark.parametrize("name", ["nb4_nq11_hllc_noslip_ivis1", "nb4_nq13_roe_periodic_ivis1"])
def test_transcription_euler_step_matches_oracle(oracle_mod, name):
    cfg = CFGS[name]
    o = oracle_mod.Oracle(cfg.as_ints())
    q = smooth_state(cfg, seed=9)
    out = np.zeros_like(q)
    dg = DG(cfg, o.tables())
    v = dg.euler(dg.limit(torch.from_numpy(q.copy())), 0.01).numpy()
    o.euler_step(q, out, 0.01)
    assert field_rel_err(v, out) < 1e-12
