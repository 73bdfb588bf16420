#!/usr/bin/env python3
"""Benchmark: cell-updates/s of the MHD-CT time step on one MI355X.

A step is one SolverGodunovMHD iteration on the workload (dt reduction, outside fill, the
U2 <- U copy and do_time_step), inputs resident in HBM.  Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--impl plain|fused] [--workload bl3d|kh3d|ot2d]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
ALGO_BYTES_PER_CELL = 128  # FP64 state read + written once per cell-update (5 + 3 doubles each way)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--impl", default="plain", choices=["plain", "fused"])
    ap.add_argument("--workload", default="bl3d", choices=["bl3d", "kh3d", "ot2d"])
    ap.add_argument("--cpu-steps", type=int, default=1, help="oracle steps for the CPU baseline (0: skip)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="before the steps, copy a buffer of this many MiB device to device once (the "
                         "known-size reference for HBM byte counters)")
    return ap.parse_args()


def cpu_baseline(mesh, U, B, st, steps):
    """The CPU restatement (oracle) timed on the host cores on `steps` steps of the same
    workload; reported beside the GPU figure, not a target."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_ref import OracleSolver, settings_array  # test infrastructure: the CPU baseline leg

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ora = OracleSolver(mesh, settings_array(gamma0=st.gamma0, cfl=st.cfl, slope_type=st.slope_type),
                       nthreads=threads)
    nt = mesh.num_total
    Uo = U.reshape(nt, 5, -1).copy()
    Bo = B.reshape(nt, -1).copy()
    t0 = time.perf_counter()
    for _ in range(steps):
        Uo, Bo, _ = ora.step(Uo, Bo)
    el = time.perf_counter() - t0
    cells = mesh.owned_cells() * steps
    return {"value": cells / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} full step(s) of the same workload ({mesh.owned_cells()} cells), "
                      f"C++ FP64 restatement with OpenMP, {el:.2f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from godunov_mhd_ct_amd import workloads
    from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_PLAIN
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD

    mesh, U, B, st = workloads.WORKLOADS[args.workload]()
    impl = IMPL_PLAIN if args.impl == "plain" else IMPL_FUSED
    sol = SolverGodunovMHD(mesh, U, B, st, impl_version=impl)
    cells = mesh.owned_cells()
    print(f"[bench] workload={args.workload} impl={args.impl} leaves={mesh.num_leaves} cells={cells}",
          file=sys.stderr, flush=True)
    if args.copy_mb > 0:
        src = torch.ones(args.copy_mb << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copied {args.copy_mb} MiB device to device", file=sys.stderr, flush=True)
        del src, dst
    for _ in range(args.warmup):
        sol.step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    step_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev0.record()
        sol.step()
        ev1.record()
        ev1.synchronize()
        step_ms.append(ev0.elapsed_time(ev1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = wall * 1e3 / args.steps
    value = cells * args.steps / wall
    step_ms.sort()
    p50 = step_ms[len(step_ms) // 2]
    p99 = step_ms[min(len(step_ms) - 1, int(0.99 * len(step_ms)))]
    achieved = cells * ALGO_BYTES_PER_CELL / (p50 * 1e-3)
    out = {
        "metric": "cell-updates/s FP64 on 1 GPU",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "impl": args.impl, "leaves": mesh.num_leaves, "cells": cells,
                   "block": mesh.b, "dim": mesh.dim},
        "step_ms_p50": p50,
        "step_ms_p99": p99,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "scope": "whole step (128 B algorithmic per cell-update / median step time)"},
    }
    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(mesh, U, B, st, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
