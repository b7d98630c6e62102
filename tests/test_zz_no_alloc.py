"""hd_min_dt_static (no context, no allocation) against hd_min_dt (GPU).

Two calls in a row on the same GPU, the second on a smaller grid: the
static scratch keeps the first call's workgroup maxima, and the second call
must read only its own.  The time step is a maximum over cells followed by
one division, so both entry points give the same bits."""
import ctypes as C

import numpy as np
import pytest

from hermeshd_amd import _lib, ics
from hermeshd_amd.config import HDConfig, vortex_3d

pytestmark = pytest.mark.gpu


def _static_dt(torch, cfg, q):
    L = _lib.lib()
    dt = torch.full((), -1.0, dtype=torch.float64, device=q.device)
    cs = _lib.config_struct(cfg.as_ints())
    _lib.check(L.hd_min_dt_static(C.byref(cs), C.c_void_p(q.data_ptr()), C.c_void_p(dt.data_ptr()),
                                  C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return dt


def test_min_dt_static_matches_context_min_dt():
    import torch
    from hermeshd_amd.solver import Solver
    assert torch.cuda.is_available()
    # first call: 2048 workgroup maxima; second: 3 (a smaller grid, faster gas
    # in its first cells, so stale maxima of the first call would change it)
    big = vortex_3d()
    small = HDConfig(nx=9, ny=7, nz=11, nbasis=4, iquad=2, nQ=5, riemann="hllc", lx=0.9, ly=0.7, lz=1.1,
                     T_floor=1e-4)
    qs = {}
    qs["big"] = torch.from_numpy(ics.isentropic_vortex(big, jitter=1e-6)).cuda()
    rng = np.random.default_rng(5)
    qsm = ics.uniform(small, rho=1.0, T=0.5, v=(0.1, -0.2, 0.05))
    qsm[0, 1] *= 1.0 + 0.5 * rng.random(qsm[0, 1].shape)
    qs["small"] = torch.from_numpy(qsm).cuda()
    for name, cfg in (("big", big), ("small", small)):
        print(f"[no_alloc] {name} {cfg.shape}", flush=True)
        q = qs[name]
        want = Solver(cfg).min_dt(q)
        got = _static_dt(torch, cfg, q)
        torch.cuda.synchronize()
        assert got.item() == want.item() and got.item() > 0, (name, got.item(), want.item())
    # the big grid again after the small one: all of its maxima are rewritten
    got = _static_dt(torch, big, qs["big"])
    assert got.item() == Solver(big).min_dt(qs["big"]).item()


def test_min_dt_static_refuses_grids_beyond_its_scratch():
    import torch
    cfg = HDConfig(nx=256, ny=256, nz=257, nbasis=4, iquad=2, nQ=5)
    L = _lib.lib()
    cs = _lib.config_struct(cfg.as_ints())
    dt = torch.zeros((), dtype=torch.float64, device="cuda")
    rc = L.hd_min_dt_static(C.byref(cs), C.c_void_p(dt.data_ptr()), C.c_void_p(dt.data_ptr()), None)
    assert rc == _lib.HD_ERR_INVALID and b"16777216" in L.hd_last_error()
