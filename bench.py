"""Benchmark of the DG RHS + SSP-RK stage on one MI355X.

Workload (default): the 3D isentropic vortex extruded in z, 128x128x32 cells,
complete-quadratic basis (10 modes), 3x3x3 volume / 3x3 face Gauss points,
5 fields, HLLC, limiter on, SSP-RK3, FP64 -- the configuration the metric is
quoted on.  One "step" is one reference step: get_min_dt + three RK stages.

Prints one JSON line (rank 0, N = 1):
  value      DOF-updates/s per RK stage = cells*nQ*nbasis*stages/step_time
  roofline   algorithmic bytes of a stage (S*(2+c_s) + 16*cells, S = state
             bytes, c_s = 1 on stages that read Q^n) over the measured mean
             stage time (HIP events on the launch stream), against 8 TB/s
  cpu_baseline  the C oracle (one thread) on a bounded sample of the same
             workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

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


def cpu_baseline(cfg, q0, stages):
    """Time the oracle on one step of the same grid (bounded: ~10-30 s)."""
    import numpy as np
    from oracle import pyoracle
    pyoracle.build()
    o = pyoracle.Oracle(cfg.as_ints())
    q = q0.copy()
    q1, q2 = np.zeros_like(q), np.zeros_like(q)
    t, dt = np.array(0.0), np.array(0.0)
    t0 = time.perf_counter()
    o.step(q, q1, q2, t, dt)
    el = time.perf_counter() - t0
    dofs = cfg.ncell * cfg.nQ * cfg.nbasis
    return {"value": dofs * stages / el, "unit": "DOF-updates/s", "cores": 1, "kind": "port",
            "sample": f"one reference step ({stages} RK stages incl. get_min_dt) of the same "
                      f"{cfg.nx}x{cfg.ny}x{cfg.nz} workload by the C oracle, 1 thread, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="vortex", choices=["vortex", "sod", "channel"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md for how "
              "the grid would shard); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from hermeshd_amd import ics
    from hermeshd_amd.config import poiseuille_2d, sod_1d, vortex_3d
    from hermeshd_amd.solver import Solver

    if args.workload == "vortex":
        cfg = vortex_3d()
        q0 = ics.isentropic_vortex(cfg, jitter=1e-6, seed=1234)
        desc = ("isentropic vortex 128x128x32, periodic [0,10]^2x[0,2.5], nbasis 10 (complete quadratic), "
                "27 volume / 9 face Gauss points, nQ 5, HLLC, limiter on, SSP-RK3, beta 5, gamma 1.4, "
                "rho jitter 1e-6 (seed 1234)")
    elif args.workload == "sod":
        cfg = sod_1d()
        q0 = ics.sod(cfg)
        desc = "Sod tube 256 cells, nbasis 4, HLLC, limiter, SSP-RK2, CFL 0.3"
    else:
        cfg = poiseuille_2d()
        q0 = ics.channel_flow(cfg)
        desc = "relaxation channel 64x128, nbasis 4, nQ 11, ivis 1, no-slip y walls, body force 1e-3"
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
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                "kernel": "RK stage (limiter + face fluxes + cell update)",
                "algorithmic_bytes_per_launch": int(bytes_stage), "mean_launch_us": round(mean_stage_ms * 1e3, 2),
                "copy_probe_GBps": round(copy_gbps, 1)}
    out = {"metric": METRIC, "value": value, "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": desc, "cells": cfg.ncell, "nbasis": cfg.nbasis, "nQ": cfg.nQ,
                      "stages_per_step": stages, "parallelism": "single GPU"},
           "roofline": roofline}
    if not args.no_cpu_baseline:
        log("bench: timing the CPU baseline (oracle, 1 thread)...")
        out["cpu_baseline"] = cpu_baseline(cfg, q0, stages)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
