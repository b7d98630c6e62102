"""Benchmark of the DG RHS + SSP-RK stage on one MI355X.

Workload (default): the 3D isentropic vortex extruded in z, 128x128x32 cells,
complete-quadratic basis (10 modes), 3x3x3 volume / 3x3 face Gauss points,
5 fields, HLLC, limiter on, SSP-RK3, FP64 -- the configuration the metric is
quoted on.  One "step" is one reference step: get_min_dt + three RK stages.

Prints one JSON line (rank 0, N = 1):
  value      DOF-updates/s per RK stage = cells*nQ*nbasis*stages/step_time
  roofline   algorithmic bytes of a stage (S*(2+c_s) + 16*cells, S = state
             bytes, c_s = 1 on stages that read Q^n) over the measured mean
             duration of the stage kernel (HIP events on the launch stream),
             against 8 TB/s; traffic = HBM bytes per stage-kernel dispatch
             from rocprofv3 FETCH_SIZE + WRITE_SIZE (two child runs, each
             scaled by the counter on a copy of known size)
  cpu_baseline  the C oracle on the host cores (OpenMP, at most 16
             threads), one step of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

T_START = time.perf_counter()
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "DG DOF-updates/s per RK stage, % HBM roofline; adjoint/forward cost ratio"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(cfg, stages):
    """Mean algorithmic bytes per stage over one step."""
    S = cfg.ncell * cfg.nQ * cfg.nbasis * 8
    per = [S * 2 + 16 * cfg.ncell] + [S * 3 + 16 * cfg.ncell] * (stages - 1)
    return sum(per) / len(per)


def cpu_baseline(cfg, make_ic, stages):
    """Times the C oracle on the host cores: one reference step (get_min_dt +
    all RK stages) of the whole workload, with the oracle's cell and face
    loops spread over min(16, host cores) OpenMP threads (16 is this GPU's
    share of the host).  Results are bitwise those of one thread."""
    import numpy as np
    from oracle import pyoracle
    pyoracle.build()
    cores = pyoracle.set_threads(min(16, os.cpu_count() or 1))
    o = pyoracle.Oracle(cfg.as_ints())
    q = make_ic(cfg)
    q1, q2 = np.zeros_like(q), np.zeros_like(q)
    t, dt = np.array(0.0), np.array(0.0)
    t0 = time.perf_counter()
    o.step(q, q1, q2, t, dt)
    el = time.perf_counter() - t0
    pyoracle.set_threads(1)
    dofs = cfg.ncell * cfg.nQ * cfg.nbasis
    return {"value": dofs * stages / el, "unit": "DOF-updates/s", "cores": cores, "kind": "port",
            "sample": f"one reference step ({stages} RK stages + get_min_dt) of the whole "
                      f"{cfg.nx}x{cfg.ny}x{cfg.nz} workload by the C oracle, {cores} OpenMP threads, "
                      f"{el:.1f} s"}


# ---------------------------------------------------------------------------
# HBM traffic of the stage kernel from hardware counters: two child runs of
# this script under rocprofv3 --pmc (FETCH_SIZE, then WRITE_SIZE; KiB per
# dispatch), each scaled by the same counter on a device-to-device copy of
# known size in the same run.
# ---------------------------------------------------------------------------
COPY_BYTES = 2**30  # larger than the 256 MB last-level cache, so the copy streams from HBM


def pmc_child(cfg_fn, make_ic, steps):
    import torch
    from hermeshd_amd.solver import Solver
    cfg = cfg_fn()
    dev = torch.device("cuda:0")
    s = Solver(cfg, device=dev)
    q = torch.from_numpy(make_ic(cfg)).to(dev)
    q1 = torch.zeros_like(q)
    buf = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device=dev)
    dst = torch.empty_like(buf)
    for _ in range(3):
        dst.copy_(buf)
    for _ in range(steps):
        s.step(q, q1)
    torch.cuda.synchronize()
    s.check()
    print("pmc child done", flush=True)


def _pmc_rows(csv_path):
    import csv
    rows = list(csv.DictReader(open(csv_path)))
    if not rows:
        return []
    keys = rows[0].keys()
    kn = next(k for k in keys if k.lower() in ("kernel_name", "kernel-name"))
    cn = next(k for k in keys if k.lower() in ("counter_name", "counter-name"))
    cv = next(k for k in keys if k.lower() in ("counter_value", "counter-value"))
    return [(r[kn], r[cn], float(r[cv])) for r in rows]


def measure_traffic(workload, kernel_key):
    """Bytes per dispatch of the stage kernel, or None if the counters are unavailable."""
    import glob
    import shutil
    import subprocess
    import tempfile
    if not shutil.which("rocprofv3"):
        log("bench: rocprofv3 not found; traffic = null")
        return None, None
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="hd_pmc_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", "150", "rocprofv3", "--pmc", counter, "-d", d, "-o", "pmc",
               "--output-format", "csv", "--", sys.executable, os.path.abspath(__file__), "--pmc-child",
               "--workload", workload]
        env = dict(os.environ, TMPDIR="/tmp")
        log(f"bench: counter pass {counter}")
        r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if r.returncode != 0 or not files:
            log(f"bench: counter pass {counter} failed (rc {r.returncode}); traffic = null")
            log(r.stdout[-2000:])
            return None, None
        rows = _pmc_rows(files[0])
        kern = [v for (k, c, v) in rows if kernel_key in k and c == counter]
        copy = [v for (k, c, v) in rows if "copy" in k.lower() and "hd::" not in k and c == counter]
        if not kern or not copy:
            log(f"bench: counter pass {counter}: no rows for the stage or copy kernel; traffic = null")
            return None, None
        # the copy moves COPY_BYTES each way; its largest dispatch is that copy
        factor = max(copy) * 1024.0 / COPY_BYTES
        out[counter] = (sum(kern) / len(kern) * 1024.0 / factor, factor)
    fetch, ff = out["FETCH_SIZE"]
    write, fw = out["WRITE_SIZE"]
    return fetch + write, {"fetch_bytes": int(fetch), "write_bytes": int(write),
                           "copy_factor_fetch": round(ff, 4), "copy_factor_write": round(fw, 4)}


def adjoint_leg(nsteps=400, checkpoint_every=20):
    """Cost of the gradient of the relaxation channel (64x128, nbasis 4,
    nQ 11, ivis 1, HLLC, no-slip walls, 400 steps, dt 0.05, checkpoints
    every 20 steps) with respect to the initial velocity and the relaxation
    rate: forward run vs backward (recompute + stage transposes)."""
    import torch
    from hermeshd_amd import ics
    from hermeshd_amd.adjoint import AdjointRunner, channel_state, run_with_grad, wall_shear
    from hermeshd_amd.config import poiseuille_2d
    cfg = poiseuille_2d()
    dev = torch.device("cuda:0")
    u0 = torch.from_numpy(0.1 * ics.gaussian_field((1, cfg.ny, cfg.nx), 8.0, 1234)).to(dev)
    runner = AdjointRunner(cfg, dt=0.05, nsteps=nsteps, checkpoint_every=checkpoint_every)

    vis0 = 0.01  # coll = rho*T/vis; colvis = coll*vis = rho*T does not depend on vis

    def once():
        u = u0.clone().requires_grad_(True)
        vis = torch.tensor(vis0, dtype=torch.float64, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = wall_shear(run_with_grad(runner, channel_state(cfg, u), 1.0 / vis))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        gu, gv = torch.autograd.grad(J, (u, vis))
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, float(J), float(gv), float(gu.norm())
    once()
    tf, tb, J, gv, gun = once()
    return {"workload": f"relaxation channel 64x128, nbasis 4, nQ 11, ivis 1, HLLC, no-slip walls, "
                        f"{nsteps} SSP-RK2 steps, dt 0.05, checkpoints every {checkpoint_every}",
            "forward_ms": round(tf * 1e3, 2), "backward_ms": round(tb * 1e3, 2), "ratio": round(tb / tf, 3),
            "J_wall_shear": J, "dJ_dvis": gv, "norm_dJ_du0": gun}


def taylor_green_leg(n=192, steps=4, adj_n=48, adj_steps=20, checkpoint_every=5):
    """13-moment Taylor-Green vortex (nbasis 10, nQ 13, Roe, limiter on,
    SSP-RK3, Re 1600): the forward step on the n^3 grid (DOF-updates/s per
    RK stage at the CFL time step), and the cost of the gradient of the
    kinetic energy after adj_steps steps with respect to the initial velocity
    field on adj_n^3 cells (forward run vs backward: recompute from
    checkpoints + stage transposes)."""
    import torch
    from hermeshd_amd import ics
    from hermeshd_amd.adjoint import AdjointRunner, kinetic_energy, run_with_grad, velocity_state
    from hermeshd_amd.config import taylor_green_3d
    from hermeshd_amd.solver import Solver
    dev = torch.device("cuda:0")
    cfg = taylor_green_3d(n)
    s = Solver(cfg, device=dev)
    q = torch.from_numpy(ics.taylor_green(cfg)).to(dev)
    q1 = torch.zeros_like(q)
    s.step(q, q1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(q, q1)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) / steps * 1e3
    s.check()
    dofs = cfg.ncell * cfg.nQ * cfg.nbasis
    dt = float(s.dt)
    del q, q1, s
    torch.cuda.empty_cache()
    # adjoint cost on a smaller grid (the stage records of a checkpoint segment
    # hold nine states per step)
    acfg = taylor_green_3d(adj_n)
    qa = torch.from_numpy(ics.taylor_green(acfg)).to(dev)
    vel0 = (qa[0, 1:4] / qa[0, 0]).clone()
    adt = float(Solver(acfg, device=dev).min_dt(qa))
    runner = AdjointRunner(acfg, dt=adt, nsteps=adj_steps, checkpoint_every=checkpoint_every)
    coll = torch.tensor(float(acfg.coll), dtype=torch.float64, device=dev)

    def once():
        v = vel0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        a0 = time.perf_counter()
        J = kinetic_energy(run_with_grad(runner, velocity_state(acfg, v), coll), acfg)
        torch.cuda.synchronize()
        a1 = time.perf_counter()
        (g,) = torch.autograd.grad(J, v)
        torch.cuda.synchronize()
        a2 = time.perf_counter()
        return a1 - a0, a2 - a1, float(J), float(g.norm())
    once()
    tf, tb, J, gn = once()
    return {"workload": f"13-moment Taylor-Green vortex {n}^3 on [0,2pi]^3, nbasis 10, nQ 13, Roe, limiter, SSP-RK3, "
                        f"Re 1600, Pr 0.71",
            "ms_per_step": round(ms_step, 2), "dt": dt, "dof_updates_per_s": dofs * 3 / (ms_step * 1e-3),
            "adjoint": {"grid": f"{adj_n}^3", "steps": adj_steps, "checkpoint_every": checkpoint_every,
                        "forward_ms": round(tf * 1e3, 2), "backward_ms": round(tb * 1e3, 2),
                        "ratio": round(tb / tf, 3), "KE": J, "norm_dKE_du0": gn}}


def kinetic_energy_quadrature(q, nbasis):
    """Mean over the cells of 0.5 |m|^2 / rho of the whole polynomial state,
    by the 27-point Gauss rule (exact for the squares of complete-quadratic
    fields at constant density).  q: (nbasis, nQ, nz, ny, nx) tensor, modes
    1, x, y, z, yz, zx, xy, then xyz (nbasis 8) or P2(x), P2(y), P2(z)."""
    gx, gw = (-math.sqrt(0.6), 0.0, math.sqrt(0.6)), (5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0)
    m = q[:, :5]
    acc = 0.0
    for a, wa in zip(gx, gw):
        for b, wb in zip(gx, gw):
            for c, wc in zip(gx, gw):
                phi = [1.0, a, b, c, b * c, c * a, a * b]
                phi += [a * b * c] if nbasis == 8 else [1.5 * a * a - 0.5, 1.5 * b * b - 0.5, 1.5 * c * c - 0.5]
                v = m[0] * phi[0]
                for r in range(1, nbasis):
                    v = v + phi[r] * m[r]
                acc += wa * wb * wc / 8.0 * float((0.5 * (v[1] ** 2 + v[2] ** 2 + v[3] ** 2) / v[0]).mean())
    return acc


def taylor_green_run(n, t_end, check_every=50, report_s=30.0, report_t=0.0, wall_limit_s=0.0):
    """The 13-moment Taylor-Green vortex (taylor_green_3d(n)) stepped with the
    reference's step semantics until t >= t_end.  One JSON line every
    report_s seconds of wall time (or, with report_t > 0, every report_t of
    simulated time) and at the end: time, steps, the mean kinetic energy of
    the cell averages (KE) and of the whole polynomial state by the 27-point
    Gauss rule (KE_q), their decay rates since the last report, and the drift
    of the total mass, momentum and energy (relative to the totals of |.|).
    The decay rate at early times is compared with nu <|grad u|^2> = 3/(4 Re)
    of the initial field (the incompressible Navier-Stokes dissipation
    rate).  With wall_limit_s > 0 the run also stops once that many seconds
    have passed since the process started, and says so in its last line."""
    import numpy as np
    import torch
    from hermeshd_amd import ics
    from hermeshd_amd.config import taylor_green_3d
    from hermeshd_amd.solver import Solver
    dev = torch.device("cuda:0")
    cfg = taylor_green_3d(n)
    s = Solver(cfg, device=dev)
    q = torch.from_numpy(ics.taylor_green(cfg)).to(dev)
    q1 = torch.zeros_like(q)
    dV = (cfg.lx / cfg.nx) * (cfg.ly / cfg.ny) * (cfg.lz / cfg.nz)

    def diag():
        m = q[0]
        tot = (m[:5].reshape(5, -1).sum(1) * dV).cpu().numpy()
        ke = float((0.5 * (m[1] ** 2 + m[2] ** 2 + m[3] ** 2) / m[0]).mean())
        return tot, ke, kinetic_energy_quadrature(q, cfg.nbasis)
    tot0, ke0, keq0 = diag()
    scale = (q[0, :5].abs().reshape(5, -1).sum(1) * dV).cpu().numpy()
    eps0 = 0.75 / 1600.0
    log(f"tgv-run: {n}^3, t_end {t_end}, KE0 {ke0:.6f}, nu<|grad u|^2> {eps0:.4e}")
    steps, t, t_last, ke_last, keq_last = 0, 0.0, 0.0, ke0, keq0
    w0 = w_last = time.perf_counter()
    stopped = False
    while t < t_end:
        if wall_limit_s > 0 and time.perf_counter() - T_START > wall_limit_s:
            stopped = True
            break
        for _ in range(check_every):
            s.step(q, q1)
        steps += check_every
        t = float(s.t)
        now = time.perf_counter()
        due = (t - t_last >= report_t) if report_t > 0 else (now - w_last >= report_s)
        if due or t >= t_end:
            s.check()
            tot, ke, keq = diag()
            rec = {"n": n, "t": round(t, 6), "steps": steps, "KE": ke, "dKE_dt": (ke - ke_last) / (t - t_last),
                   "KE_q": keq, "dKEq_dt": (keq - keq_last) / (t - t_last),
                   "rel_drift": ((tot - tot0) / scale).tolist(), "wall_s": round(now - w0, 1)}
            print(json.dumps(rec), flush=True)
            w_last, t_last, ke_last, keq_last = now, t, ke, keq
    s.check()
    tot, ke, keq = diag()
    wall = time.perf_counter() - w0
    return {"workload": f"13-moment Taylor-Green vortex {n}^3, Re 1600, Pr 0.71, nbasis 10, Roe, limiter, SSP-RK3",
            "t": t, "steps": steps, "KE0": ke0, "KE": ke, "KE_q0": keq0, "KE_q": keq,
            "mean_dKE_dt": (ke - ke0) / t, "nu_grad_u2_t0": eps0,
            "max_rel_drift": float(np.max(np.abs((tot - tot0) / scale))), "ms_per_step": wall / steps * 1e3,
            "wall_s": round(wall, 1), "setup_s": round(w0 - T_START, 1), "stopped_at_wall_limit": stopped}


def ensemble_leg(steps=20):
    """Batched ensemble (96x16^3 + 96x24^3 + 64x32^3 periodic boxes, nbasis 4,
    nQ 11, stochastic stress, HLLC, SSP-RK2, dt 0.1, graph-replayed steps,
    tuned launch shape) against the same realizations stepped one at a time
    through single contexts."""
    import torch
    from hermeshd_amd import ics
    from hermeshd_amd.config import fluct_ensemble
    from hermeshd_amd.ensemble import Ensemble
    from hermeshd_amd.solver import Solver
    base, dims = fluct_ensemble()
    ens = Ensemble(base, dims)

    def unif(r):
        nx, ny, nz = dims[r]
        return ics.uniform(base.replace(nx=nx, ny=ny, nz=nz, lx=float(nx), ly=float(ny), lz=float(nz)),
                           rho=1.0, T=1.0)
    states = [unif(r) for r in range(ens.nreal)]
    q = ens.pack(states)
    q1 = ens.zeros()
    choice, _ = ens.tune(q, ens.zeros(), 0.1, reps=3)
    ens.run(q, q1, None, 0.1, 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run(q, q1, None, 0.1, steps)
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / steps
    ens.check()
    sols, qs, q1s = [], [], []
    for r in range(ens.nreal):
        nx, ny, nz = dims[r]
        cfg = base.replace(nx=nx, ny=ny, nz=nz, lx=float(nx), ly=float(ny), lz=float(nz), realization=r)
        sols.append(Solver(cfg))
        qs.append(torch.from_numpy(states[r]).cuda())
        q1s.append(torch.zeros_like(qs[-1]))
    # one device dt per context, filled once: each realization-step then costs
    # what a batched step costs per realization (no per-call dt fill)
    dts = [torch.full((), 0.1, dtype=torch.float64, device=qs[0].device) for _ in sols]
    for s_, a, b, d in zip(sols, qs, q1s, dts):
        s_.update(a, b, None, d)
    torch.cuda.synchronize()
    n1 = max(2, steps // 4)
    t0 = time.perf_counter()
    for _ in range(n1):
        for s_, a, b, d in zip(sols, qs, q1s, dts):
            s_.update(a, b, None, d)
    torch.cuda.synchronize()
    t1 = (time.perf_counter() - t0) / n1
    dofs = ens.cells * base.nQ * base.nbasis
    return {"workload": "256 periodic realizations (96x16^3, 96x24^3, 64x32^3), nbasis 4, nQ 11, stochastic "
                        "stress, HLLC, SSP-RK2, dt 0.1",
            "batched_ms_per_step": round(tb * 1e3, 3), "one_at_a_time_ms_per_step": round(t1 * 1e3, 3),
            "speedup": round(t1 / tb, 2), "batched_dof_updates_per_s": dofs * 2 / tb, "launch_shape": choice}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="vortex", choices=["vortex", "sod", "channel"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter passes")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--no-tgv", action="store_true", help="skip the 13-moment Taylor-Green leg")
    ap.add_argument("--tgv-run", type=int, default=0, metavar="N",
                    help="only run the 13-moment Taylor-Green vortex on N^3 cells to --tgv-t and print its diagnostics")
    ap.add_argument("--tgv-t", type=float, default=1.0)
    ap.add_argument("--tgv-report-dt", type=float, default=0.0,
                    help="report every this much simulated time (default: every 30 s of wall time)")
    ap.add_argument("--tgv-wall", type=float, default=0.0, metavar="S",
                    help="stop the --tgv-run loop S seconds after the process started (0: no limit)")
    ap.add_argument("--tgv-adj", type=int, nargs=3, default=(48, 20, 5), metavar=("N", "STEPS", "EVERY"),
                    help="Taylor-Green leg's adjoint: N^3 cells, STEPS steps, a checkpoint every EVERY steps")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the grid would shard "
              "by slabs with one face-trace exchange per stage); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.tgv_run:
        print(json.dumps({"tgv_run": taylor_green_run(args.tgv_run, args.tgv_t, report_t=args.tgv_report_dt,
                                                      wall_limit_s=args.tgv_wall)}), flush=True)
        return

    from hermeshd_amd import ics
    from hermeshd_amd.config import poiseuille_2d, sod_1d, vortex_3d

    if args.workload == "vortex":
        cfg_fn = vortex_3d
        make_ic = lambda c: ics.isentropic_vortex(c, jitter=1e-6, seed=1234)  # noqa: E731
        desc = ("isentropic vortex 128x128x32, periodic [0,10]^2x[0,2.5], nbasis 10 (complete quadratic), "
                "27 volume / 9 face Gauss points, nQ 5, HLLC, limiter on, SSP-RK3, beta 5, gamma 1.4, "
                "rho jitter 1e-6 (seed 1234)")
    elif args.workload == "sod":
        cfg_fn = sod_1d
        make_ic = ics.sod
        desc = "Sod tube 256 cells, nbasis 4, HLLC, limiter, SSP-RK2, CFL 0.3"
    else:
        cfg_fn = poiseuille_2d
        make_ic = ics.channel_flow
        desc = "relaxation channel 64x128, nbasis 4, nQ 11, ivis 1, no-slip y walls, body force 1e-3"
    if args.pmc_child:
        pmc_child(cfg_fn, make_ic, steps=2)
        return
    # counter passes run in child processes before this process touches the GPU
    traffic, traffic_detail = (None, None) if args.no_traffic else measure_traffic(args.workload, "k_stage")

    import numpy as np  # noqa: F401
    import torch

    from hermeshd_amd.solver import Solver

    cfg = cfg_fn()
    q0 = make_ic(cfg)
    stages = 3 if cfg.integrator == "shu-osher" else 2

    dev = torch.device("cuda:0")
    s = Solver(cfg, device=dev)
    q = torch.from_numpy(q0).to(dev)
    q1 = torch.zeros_like(q)
    log(f"bench: {desc}; {cfg.ncell} cells, {cfg.ncell * cfg.nQ * cfg.nbasis} DOFs, "
        f"context {s.device_bytes / 2**20:.0f} MiB")
    for _ in range(args.warmup):
        s.step(q, q1)
    s.check()
    torch.cuda.synchronize()

    # pass 1: whole steps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.steps):
        s.step(q, q1)
    e1.record()
    torch.cuda.synchronize()
    ms_per_step = e0.elapsed_time(e1) / args.steps
    s.check()
    log(f"bench: {args.steps} steps, {ms_per_step:.3f} ms/step")

    # pass 2: the same steps with every stage bracketed by HIP events on the launch stream
    s.timing(True)
    for _ in range(args.steps):
        s.step(q, q1)
    stage_ms, nstage = s.timing_read()
    s.timing(False)
    s.check()
    mean_stage_ms = stage_ms / max(nstage, 1)
    log(f"bench: {nstage} stages, mean {mean_stage_ms * 1e3:.1f} us/stage")

    # copy probe (context for the spec peak)
    buf = torch.empty(256 * 2**20 // 8, dtype=torch.float64, device=dev)
    dst = torch.empty_like(buf)
    for _ in range(3):
        dst.copy_(buf)
    torch.cuda.synchronize()
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record()
    for _ in range(20):
        dst.copy_(buf)
    c1.record()
    torch.cuda.synchronize()
    copy_gbps = 2 * buf.numel() * 8 * 20 / (c0.elapsed_time(c1) * 1e-3) / 1e9
    del buf, dst

    dofs = cfg.ncell * cfg.nQ * cfg.nbasis
    value = dofs * stages / (ms_per_step * 1e-3)
    bytes_stage = stage_bytes(cfg, stages)
    achieved = bytes_stage / (mean_stage_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4),
                "traffic": int(traffic) if traffic is not None else None,
                "kernel": "k_stage_xs (fused RK stage, x faces shared along each periodic row: volume, "
                          "faces, sources, update, SSP combination, next stage's limiter)",
                "algorithmic_bytes_per_launch": int(bytes_stage), "mean_launch_us": round(mean_stage_ms * 1e3, 2),
                "copy_probe_GBps": round(copy_gbps, 1)}
    if traffic_detail:
        roofline["traffic_detail"] = traffic_detail
    out = {"metric": METRIC, "value": value, "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": desc, "cells": cfg.ncell, "nbasis": cfg.nbasis, "nQ": cfg.nQ,
                      "stages_per_step": stages, "parallelism": "single GPU"},
           "roofline": roofline}
    if not args.no_cpu_baseline:
        log("bench: timing the CPU baseline (oracle, OpenMP threads)...")
        out["cpu_baseline"] = cpu_baseline(cfg, make_ic, stages)
    if not args.no_adjoint:
        log("bench: adjoint leg (relaxation channel, 400 steps)...")
        out["adjoint"] = adjoint_leg()
        log(f"bench: adjoint {out['adjoint']}")
    if not args.no_tgv:
        log("bench: 13-moment Taylor-Green leg...")
        an, asteps, aevery = args.tgv_adj
        out["taylor_green"] = taylor_green_leg(adj_n=an, adj_steps=asteps, checkpoint_every=aevery)
        log(f"bench: taylor_green {out['taylor_green']}")
    if not args.no_ensemble:
        log("bench: ensemble leg (256 realizations)...")
        out["ensemble"] = ensemble_leg()
        log(f"bench: ensemble {out['ensemble']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
