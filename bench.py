#!/usr/bin/env python3
"""Benchmark: cell-updates/s of the MHD-CT time step on one MI355X.  Prints one JSON line.

  python bench.py [--workload mix|bl3d|kh3d|ot2d|srv] [--steps K] [--warmup W] [--impl fused|fp32|plain]
                  [--integrator hancock|rk2]

A step advances every owned cell of the workload once: the outside-block fill, the CFL dt
reduction (dt = cfl / inv_dt, on the device) and the time step, inputs resident in HBM.

  mix (default)   the mixed-block 3D AMR mesh: 4096 blocks, 2048 x 8^3 + 1434 x 16^3 + 614 x 32^3
                  (27,041,792 cells, levels 2-6) as three periodic meshes in one batch, FP64
  bl3d            3D MHD blast, 16^3 blocks, levels 3-6 (2808 leaves, 11.5 M cells), FP64
  kh3d            3D Kelvin-Helmholtz, 256^3 as 16^3 blocks
  ot2d            2D Orszag-Tang, 128^2 as 8^2 blocks
  srv             the serving loop: a seeded Poisson request stream, p50/p99 latency

The fused forms run through the batch engine (one kernel per stage over the work list, device
dt).  The timed steps record HIP events between the stage launches on the launch stream, so
the dominant kernel's duration comes from the timed region itself.  The plain form (one kernel
per stage, the GodunovImplemV0 drop-in) is timed in the same run on the same workload, and the
CPU restatement of the step (oracle/, OpenMP) on a bounded sample of it.

--integrator rk2 times the RK2_SSP integrator (two forward-Euler stages and the stage
combinations per step, SolverGodunovMHD.cpp:551-647) through the per-mesh solver driver, with the
CFL dt read back on the host every step as the reference does; a cell-update is still one owned
cell advanced one full step.
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
METRIC = "cell-updates/s FP64 on 1 GPU"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--impl", default="fused", choices=["fused", "fp32", "plain"])
    ap.add_argument("--workload", default="mix", choices=["bl3d", "mix", "kh3d", "ot2d", "srv"])
    ap.add_argument("--integrator", default="hancock", choices=["hancock", "rk2"])
    ap.add_argument("--amr-cycle", action="store_true",
                    help="time one AMR cycle (refinement flags, host adaptation, remap) on the workload after "
                         "--steps steps, [amr] keys of app/test_mhd_blast_3d.ini")
    ap.add_argument("--gravity", type=float, default=None,
                    help="add a uniform gravity field of this magnitude along -z (-y in 2D) to the workload")
    ap.add_argument("--viscosity", type=float, default=None,
                    help="the FP64 forms (plain and fused) with viscous fluxes of this mu (and the viscous "
                         "predictor term), next to the same steps without viscosity")
    ap.add_argument("--checkpoint", action="store_true",
                    help="time writing a checkpoint and the XDMF output of the workload's first mesh after "
                         "--steps steps, and a restart from that checkpoint")
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps for the CPU baseline (0: skip)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the serving and BL3D sub-lines of the default (mix) run")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch autotuner")
    ap.add_argument("--front", default="off", choices=["on", "off"],
                    help="primitive, electric-field and cell stages as one LDS-tiled kernel (on) or three (off)")
    ap.add_argument("--stage-threads", default="",
                    help="CELL,EDGE: workgroup sizes of k_cell (256 or 512) and k_edge (256 or 512; default: the "
                         "library's choice)")
    ap.add_argument("--precision-steps", type=int, default=0,
                    help="with --impl fp32: run the FP32-bulk and the FP64 form this many steps from the same "
                         "state and report the per-field relative L1 error (the precision config: 500 on kh3d)")
    ap.add_argument("--requests", type=int, default=40, help="srv: number of requests")
    ap.add_argument("--rate", type=float, default=50.0, help="srv: arrivals per second")
    ap.add_argument("--graph", action="store_true", help="srv: replay each batch step from a HIP graph")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="before the steps, copy a buffer of this many MiB device to device once (the "
                         "known-size reference for HBM byte counters)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(cases, steps):
    """The CPU restatement (oracle/, its -O3 OpenMP build) timed on the host cores over `steps`
    steps of the workload's meshes; reported beside the GPU figure, not a target."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_ref  # test infrastructure: the CPU baseline leg
    from oracle_ref import OracleSolver, settings_array

    oracle_ref.use_library(oracle_ref.ORACLE_LIB_O3)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cells = 0
    el = 0.0
    for mesh, U, B, st in cases:
        grav = tuple(st.gravity) if st.gravity_enabled else None
        mu = st.viscosity_mu if st.viscosity_enabled else None
        ora = OracleSolver(mesh, settings_array(gamma0=st.gamma0, cfl=st.cfl, slope_type=st.slope_type,
                                                time_integrator=st.time_integrator, gravity=grav, mu=mu,
                                                viscous_predictor=bool(st.viscosity_hancock_predictor),
                                                prolongation=int(st.prolongation)),
                           nthreads=threads)
        nt = mesh.num_total
        Uo = U.reshape(nt, 5, -1).copy()
        Bo = B.reshape(nt, -1).copy()
        t0 = time.perf_counter()
        for _ in range(steps):
            Uo, Bo, _ = ora.step(Uo, Bo)
        el += time.perf_counter() - t0
        cells += mesh.owned_cells() * steps
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"value": cells / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "build": "g++ -O3 -fopenmp (oracle/Makefile, libmhd_oracle_o3.so), OpenMP schedule(static)",
            "sample": f"{steps} full step(s) of every mesh of the workload ({cells // steps} cells), "
                      f"C++ FP64 restatement with OpenMP over {threads} threads, {el:.2f} s"}


# the kernel sources whose counters profiles/hbm_traffic.json records
COUNTED_SOURCES = ["fused.hip", "fused.hpp", "phys.hpp", "phys_fast.hpp", "common.hpp", "batch.hip"]


def sources_digest(read=None):
    """SHA-256 of the code of the counted kernel sources: `//` comments and whitespace are left
    out, so that an edit of a comment keeps the counters of unchanged code.  `read(name)` gives a
    source's text (default: the file in this tree)."""
    import hashlib
    import re

    read = read or (lambda name: (ROOT / "godunov_mhd_ct_amd" / "csrc" / name).read_text())
    h = hashlib.sha256()
    for name in COUNTED_SOURCES:
        code = re.sub(r"//[^\n]*", "", read(name))
        h.update(" ".join(code.split()).encode())
    return h.hexdigest()


def _counters(workload, kernel):
    """The committed counter record of `kernel` on `workload` and its provenance, or (None, why):
    the record counts only if it was taken from the kernel sources of this tree (their SHA-256 is
    stored beside the counters)."""
    p = ROOT / "profiles" / "hbm_traffic.json"
    try:
        d = json.loads(p.read_text())
    except Exception:
        return None, "no counter summary"
    if d.get("sources_sha256") != sources_digest():
        return None, f"counters of other kernel sources (taken at {d.get('commit', '?')})"
    rec = d.get(workload, {}).get(kernel)
    if rec is None:
        return None, f"no counters for {kernel} on {workload}"
    return rec, f"profiles/hbm_traffic.json, taken at {d.get('commit', '?')} from these kernel sources"


def load_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs, scaled by a copy of known size), or None."""
    rec, _ = _counters(workload, kernel)
    return rec["bytes_per_launch"] if rec else None


# wave64 FP64 VALU instructions per second: 1024 SIMDs x 2.4 GHz / 4 cycles per instruction
VALU_PEAK = 1024 * 2.4e9 / 4


def valu_utilisation(kernel, workload, kernel_ms):
    """SQ_INSTS_VALU per launch of `kernel` (committed counter summary) over this run's kernel
    time, as a fraction of the full-rate VALU issue peak, or None."""
    rec, _ = _counters(workload, kernel)
    return rec["valu_insts"] / (kernel_ms * 1e-3) / VALU_PEAK if rec else None


def run_plain(cases, steps, warmup, impl=None):
    """Steps through the per-mesh solver driver (SolverGodunovMHD: host dt read-back each step),
    the plain form unless `impl` names another."""
    import torch
    from godunov_mhd_ct_amd.implem import IMPL_PLAIN
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD

    impl = IMPL_PLAIN if impl is None else impl
    sols = [SolverGodunovMHD(m, U, B, st, impl_version=impl) for (m, U, B, st) in cases]
    for _ in range(warmup):
        for s in sols:
            s.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for s in sols:
            s.step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = sum(m.owned_cells() for (m, _, _, _) in cases)
    del sols
    torch.cuda.empty_cache()
    return {"value": cells * steps / wall, "ms_per_step": wall * 1e3 / steps}


def cases_for(name):
    from godunov_mhd_ct_amd import workloads

    if name == "mix":
        return workloads.mix3d(seed=0)
    return [workloads.WORKLOADS[name]()]


def bench_batch(args, cases, impl):
    import torch
    from godunov_mhd_ct_amd.batch import BatchGodunovMHD

    dim = cases[0][0].dim
    bt = BatchGodunovMHD(dim, impl)
    bt.set_front(args.front == "on")
    if args.stage_threads:
        bt.set_stage_threads(*(int(x) for x in args.stage_threads.split(",")))
    for (m, U, B, st) in cases:
        bt.add(m, U, B, st)
    cells = bt.cells()
    tune = None
    if not args.no_tune:
        tune = bt.autotune(reps=2)
        log(f"[bench] autotuned grid_cap={tune[0]} tile_cap={tune[1]} tile_z={tune[2]}")
    bt.step(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bt.step(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    bt.check()
    # the per-stage HIP-event times from as many steps again, outside the timed region (the events
    # between the stages cost about 0.1 ms per step)
    stage_ms = bt.step_timed(args.steps)       # synchronises at its end
    bt.check()
    # the same steps replayed from HIP graphs
    bt.set_graph(True)
    bt.step(1)
    torch.cuda.synchronize()
    g0 = time.perf_counter()
    bt.step(args.steps)
    torch.cuda.synchronize()
    gwall = time.perf_counter() - g0
    bt.check()
    return cells, wall, stage_ms, gwall, tune


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_FUSED_FP32

    if args.copy_mb > 0:
        src = torch.ones(args.copy_mb << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.copy_mb} MiB device to device")
        del src, dst

    if args.workload == "srv":
        return bench_serving(args)

    cases = cases_for(args.workload)
    if args.gravity is not None:
        # [gravity] enabled with a uniform field along the last axis (AddGravitySourceTerm.cpp)
        for (m, _, _, st) in cases:
            st.gravity_enabled = 1
            st.gravity[m.dim - 1] = -abs(args.gravity)
    if args.amr_cycle:
        return bench_amr(args, cases)
    if args.checkpoint:
        return bench_checkpoint(args, cases)
    if args.viscosity is not None:
        return bench_viscosity(args, cases)
    if args.integrator == "rk2":
        return bench_rk2(args, cases)
    cells = sum(m.owned_cells() for (m, _, _, _) in cases)
    log(f"[bench] workload={args.workload} impl={args.impl} meshes={len(cases)} "
        f"leaves={sum(m.num_leaves for (m, _, _, _) in cases)} cells={cells}")
    out = {"metric": METRIC, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if args.impl != "fp32" else "f32 bulk / f64 accumulation", "data": "synthetic"}
    config = {"workload": args.workload + ("+gravity" if args.gravity is not None else ""), "impl": args.impl,
              "meshes": len(cases),
              "leaves": sum(m.num_leaves for (m, _, _, _) in cases), "cells": cells,
              "blocks": sorted({m.b for (m, _, _, _) in cases}), "dim": cases[0][0].dim}
    if args.impl != "plain":
        config["front"] = args.front == "on"
    if args.impl == "plain":
        p = run_plain(cases, args.steps, args.warmup)
        out.update({"value": p["value"], "ms_per_step": p["ms_per_step"]})
        out["roofline"] = {"bound": "hbm", "achieved": p["value"] * ALGO_BYTES_PER_CELL / 1e9,
                           "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                           "frac": p["value"] * ALGO_BYTES_PER_CELL / HBM_PEAK, "traffic": None,
                           "kernel": "whole step (plain form: no per-kernel events)"}
    else:
        impl = IMPL_FUSED if args.impl == "fused" else IMPL_FUSED_FP32
        ncells, wall, stage_ms, gwall, tune = bench_batch(args, cases, impl)
        value = ncells * args.steps / wall
        kern = max((k for k in stage_ms if k.startswith("k_")), key=lambda k: stage_ms[k])
        kms = stage_ms[kern]
        achieved = ALGO_BYTES_PER_CELL * ncells / (kms * 1e-3)
        out.update({"value": value, "ms_per_step": wall * 1e3 / args.steps,
                    "ms_per_step_graph": gwall * 1e3 / args.steps})
        out["roofline"] = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK, "traffic": load_traffic(kern, args.workload),
                           "kernel": kern, "kernel_ms": kms,
                           "valu_utilisation": valu_utilisation(kern, args.workload, kms),
                           "counters": _counters(args.workload, kern)[1],
                           "algorithmic_bytes": f"{ALGO_BYTES_PER_CELL} B x {ncells} cell-updates per launch",
                           "end_to_end": {"achieved": value * ALGO_BYTES_PER_CELL / 1e9,
                                          "frac": value * ALGO_BYTES_PER_CELL / HBM_PEAK}}
        out["stage_ms"] = {k: round(v, 4) for k, v in stage_ms.items()}
        if tune:
            config["tune"] = {"grid_cap": tune[0], "tile_cap": tune[1], "tile_z": tune[2]}
        if not args.no_plain:
            p = run_plain(cases, min(args.steps, 10), min(args.warmup, 2))
            out["plain"] = {"value": p["value"], "ms_per_step": p["ms_per_step"],
                            "speedup": value / p["value"]}
    if args.precision_steps > 0 and args.impl == "fp32":
        out["l1_error_vs_fp64"] = precision_errors(cases, args.precision_steps)
    out["config"] = config
    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(cases, args.cpu_steps)
    if args.workload == "mix" and args.impl == "fused" and not args.no_extras:
        del cases
        out["serving"] = serving_lines(args)
        out["bl3d"] = bl3d_line(args)
    print(json.dumps(out))


def serving_lines(args):
    """The serving workload at the stated rate (50 requests/s) and at a rate the GPU keeps up
    with (8 requests/s): p50/p99 latency from arrival to the result on the host."""
    import torch
    from godunov_mhd_ct_amd.serving import Server, request_stream

    free, _ = torch.cuda.mem_get_info()
    budget = int(0.5 * free)
    Server(hbm_budget=budget).run(request_stream(2, rate=1000.0, seed=99, nsteps=5, sizes2d=(32,), sizes3d=(32,)))
    res = {}
    for rate in (50.0, 8.0):
        log(f"[bench] serving {args.requests} requests at {rate}/s")
        r = Server(hbm_budget=budget, max_tick=20).run(request_stream(args.requests, rate=rate, seed=0, nsteps=200))
        res[f"rate_{int(rate)}"] = {"rate_per_s": rate, "requests": r["requests"],
                                    "latency_ms_p50": r["latency_ms_p50"], "latency_ms_p99": r["latency_ms_p99"],
                                    "requests_per_s": r["requests"] / r["wall_s"],
                                    "cell_updates_per_s": r["cell_updates_per_s"]}
    res["workload"] = ("Poisson arrivals (seed 0); 2D field loop or 2D/3D blast (beta log-uniform in [0.1, 10]); "
                       "2D 64^2-512^2, 3D 32^3-128^3; 8^3 or 16^3 blocks; 200 steps each; FP64")
    torch.cuda.empty_cache()
    return res


def amr_end_to_end(case, amr, steps, impl=None):
    """The per-mesh solver (host dt each step, as the reference) with an AMR cycle when due
    (next_iteration_impl, SolverGodunovMHD.cpp:409-484), counted as the reference counts: the
    owned cells of each step's mesh (:482), over the whole run including the cycles."""
    import torch
    from godunov_mhd_ct_amd.implem import IMPL_FUSED
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD

    m, U, B, st = case
    sol = SolverGodunovMHD(m, U, B, st, impl_version=IMPL_FUSED if impl is None else impl)
    leaves0 = m.num_leaves
    sol.step()  # warm-up (the first cycle is disabled, as in the blast configuration)
    torch.cuda.synchronize()
    sol.iteration = 1
    cells = cycles = 0
    w0 = time.perf_counter()
    for _ in range(steps):
        if sol.should_do_amr_cycle(amr, disable_first=True):
            sol.do_amr_cycle(amr)
            cycles += 1
        cells += sol.mesh.owned_cells()
        sol.step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    out = {"steps": steps, "amr_cycles": cycles, "cycle_frequency": amr.cycle_frequency,
           "criterion": f"{amr.criterion} on {', '.join(amr.names())} > {amr.epsilon_refine}",
           "levels": [amr.level_min, amr.level_max], "leaves_at_start": leaves0,
           "leaves_at_end": sol.mesh.num_leaves, "cell_updates_per_s": cells / wall,
           "ms_per_step": wall * 1e3 / steps, "driver": "per-mesh solver, host dt each step, AMR cycles included"}
    del sol
    torch.cuda.empty_cache()
    return out


def bl3d_line(args):
    """The 3D blast (16^3 blocks, levels 3-6, 11.5 M cells) through the same batch engine, with
    the plain form beside it."""
    import torch
    from godunov_mhd_ct_amd.implem import IMPL_FUSED

    log("[bench] bl3d sub-line")
    cases = cases_for("bl3d")
    ncells, wall, stage_ms, gwall, tune = bench_batch(args, cases, IMPL_FUSED)
    value = ncells * args.steps / wall
    kern = max((k for k in stage_ms if k.startswith("k_")), key=lambda k: stage_ms[k])
    out = {"value": value, "ms_per_step": wall * 1e3 / args.steps, "cells": ncells,
           "leaves": sum(m.num_leaves for (m, _, _, _) in cases),
           "roofline_frac": value * ALGO_BYTES_PER_CELL / HBM_PEAK,
           "dominant_kernel": kern, "dominant_kernel_ms": stage_ms[kern],
           "dominant_kernel_frac": ALGO_BYTES_PER_CELL * ncells / (stage_ms[kern] * 1e-3) / HBM_PEAK}
    if not args.no_plain:
        p = run_plain(cases, min(args.steps, 10), min(args.warmup, 2))
        out["plain"] = {"value": p["value"], "speedup": value / p["value"]}
    torch.cuda.empty_cache()
    from godunov_mhd_ct_amd.workloads import bl3d_amr_settings

    log("[bench] bl3d end to end with AMR cycles (200 steps)")
    out["end_to_end_amr"] = amr_end_to_end(cases[0], bl3d_amr_settings(), 200)
    return out


def bench_rk2(args, cases):
    from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_FUSED_FP32, IMPL_PLAIN, TIME_INTEGRATOR

    for (_, _, _, st) in cases:
        st.time_integrator = TIME_INTEGRATOR["RK2_SSP"]
    impl = {"fused": IMPL_FUSED, "fp32": IMPL_FUSED_FP32, "plain": IMPL_PLAIN}[args.impl]
    cells = sum(m.owned_cells() for (m, _, _, _) in cases)
    log(f"[bench] RK2_SSP workload={args.workload} impl={args.impl} meshes={len(cases)} cells={cells}")
    batch = None
    if args.impl != "plain":
        batch = bench_rk2_batch(args, cases, impl)
    r = run_plain(cases, args.steps, args.warmup, impl)
    out = {"metric": METRIC, "value": r["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64" if args.impl != "fp32" else "f32 bulk / f64 accumulation",
           "data": "synthetic",
           "config": {"workload": args.workload, "impl": args.impl, "integrator": "RK2_SSP", "meshes": len(cases),
                      "leaves": sum(m.num_leaves for (m, _, _, _) in cases), "cells": cells,
                      "driver": "per-mesh solver, host dt each step"}}
    # two stages per step: the state is read and written twice
    algo = 2 * ALGO_BYTES_PER_CELL
    out["roofline"] = {"bound": "hbm", "achieved": r["value"] * algo / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                       "frac": r["value"] * algo / HBM_PEAK, "traffic": None,
                       "kernel": "whole step (two stages, per-mesh driver: no per-kernel events)",
                       "algorithmic_bytes": f"{algo} B per cell-update"}
    if batch is not None:
        # the batch engine (dt on the device, HIP-graph replay) is the headline form; the
        # per-mesh driver's figure stays beside it
        out["per_mesh_driver"] = {"value": r["value"], "ms_per_step": r["ms_per_step"]}
        out.update({"value": batch["value"], "ms_per_step": batch["ms_per_step"]})
        out["config"]["driver"] = "batch engine, device dt, HIP-graph replay"
        out["roofline"]["achieved"] = batch["value"] * algo / 1e9
        out["roofline"]["frac"] = batch["value"] * algo / HBM_PEAK
        out["roofline"]["kernel"] = "whole step (two stages; the batch times no single kernel under RK2_SSP)"
    if args.impl != "plain" and not args.no_plain:
        p = run_plain(cases, min(args.steps, 10), min(args.warmup, 2))
        out["plain"] = {"value": p["value"], "ms_per_step": p["ms_per_step"], "speedup": out["value"] / p["value"]}
    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(cases, args.cpu_steps)
    print(json.dumps(out))


def bench_rk2_batch(args, cases, impl):
    """RK2_SSP steps through the batch engine with graph replay: the owned cells of every mesh
    advanced per second."""
    import torch
    from godunov_mhd_ct_amd.batch import BatchGodunovMHD

    bt = BatchGodunovMHD(cases[0][0].dim, impl)
    for (m, U, B, st) in cases:
        bt.add(m, U, B, st)
    cells = bt.cells()
    bt.set_graph(True)
    bt.step(max(args.warmup, 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bt.step(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    bt.check()
    del bt
    torch.cuda.empty_cache()
    return {"value": cells * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps}


def bench_viscosity(args, cases):
    """The FP64 forms (plain and fused, the per-mesh solver) with [viscosity] enabled, mu, and the
    predictor term, against the same steps without viscosity.  The line's value is the fused
    form's."""
    from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_PLAIN

    cells = sum(m.owned_cells() for (m, _, _, _) in cases)
    log(f"[bench] viscosity mu={args.viscosity} workload={args.workload} plain and fused forms, cells={cells}")
    base = {f: run_plain(cases, args.steps, args.warmup, impl) for f, impl in (("plain", IMPL_PLAIN),
                                                                               ("fused", IMPL_FUSED))}
    for (_, _, _, st) in cases:
        st.viscosity_enabled = 1
        st.viscosity_mu = args.viscosity
        st.viscosity_hancock_predictor = 1
    r = {f: run_plain(cases, args.steps, args.warmup, impl) for f, impl in (("plain", IMPL_PLAIN),
                                                                            ("fused", IMPL_FUSED))}
    v = r["fused"]["value"]
    out = {"metric": METRIC, "value": v, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": r["fused"]["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": args.workload, "impl": "fused", "viscosity_mu": args.viscosity,
                      "viscous_predictor": True, "meshes": len(cases), "cells": cells,
                      "driver": "per-mesh solver, host dt each step"},
           "roofline": {"bound": "hbm", "achieved": v * ALGO_BYTES_PER_CELL / 1e9, "peak": HBM_PEAK / 1e9,
                        "unit": "GB/s", "frac": v * ALGO_BYTES_PER_CELL / HBM_PEAK, "traffic": None,
                        "kernel": "whole step (fused form)",
                        "algorithmic_bytes": f"{ALGO_BYTES_PER_CELL} B per cell-update"},
           "forms": {f: {"value": r[f]["value"], "ms_per_step": r[f]["ms_per_step"],
                         "without_viscosity": base[f], "viscous_cost_ms": r[f]["ms_per_step"] - base[f]["ms_per_step"]}
                     for f in r}}
    print(json.dumps(out))


def bench_amr(args, cases):
    """One AMR cycle on the first mesh of the workload after args.steps steps: device
    refinement flags, host adaptation + 2:1 balance + remap plan, device remap, new mesh tables.
    The remap's algorithmic bytes are the owned state read once and the new owned state written
    once."""
    import torch
    from godunov_mhd_ct_amd.amr import AMRSettings, Remap, adapted, refine_flags
    from godunov_mhd_ct_amd.implem import IMPL_FUSED
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD

    m, U, B, st = cases[0]
    lv, _ = m.leaves()
    if args.workload == "bl3d":
        from godunov_mhd_ct_amd.workloads import bl3d_amr_settings

        amr = bl3d_amr_settings(int(lv.min()), int(lv.max()))
    else:
        amr = AMRSettings(level_min=int(lv.min()), level_max=int(lv.max()), epsilon_refine=0.01,
                          epsilon_coarsen=0.00125, epsilon_lohner=0.02)
    sol = SolverGodunovMHD(m, U, B, st, impl_version=IMPL_FUSED)
    for _ in range(args.steps):
        sol.step()
    torch.cuda.synchronize()
    log(f"[bench] AMR cycle after {args.steps} steps: {m.num_leaves} leaves, levels {amr.level_min}-{amr.level_max}")
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    reps = 5
    e0, e1 = ev(), ev()
    e0.record()
    for _ in range(reps):
        flags, ind = refine_flags(sol.implem, sol.U, amr, B=sol.B)
    e1.record()
    torch.cuda.synchronize()
    flags_ms = e0.elapsed_time(e1) / reps
    t0 = time.perf_counter()
    new = adapted(sol.mesh, flags)
    t1 = time.perf_counter()
    rm = Remap(sol.mesh, new)
    t2 = time.perf_counter()
    nt = new.num_total
    Un = torch.zeros(nt * 5 * new.cells_per_block, dtype=torch.float64, device="cuda")
    Bn = torch.zeros(nt * new.faces_per_block, dtype=torch.float64, device="cuda")
    rm.apply(sol.U, sol.B, Un, Bn)  # first call uploads the plan
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        rm.apply(sol.U, sol.B, Un, Bn)
    e1.record()
    torch.cuda.synchronize()
    remap_ms = e0.elapsed_time(e1) / reps
    t3 = time.perf_counter()
    sol.implem.resize_auxiliary_data(new)
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    per_block = 8 * (5 * m.cells_per_block + m.faces_per_block)
    algo = per_block * (m.num_leaves + new.num_leaves)
    counts = rm.counts()
    step_ms = None
    sol.mesh, sol.U, sol.B = new, Un, Bn
    sol.U2, sol.B2 = Un.clone(), Bn.clone()
    sol.step()
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    for _ in range(3):
        sol.step()
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - s0) * 1e3 / 3
    # end to end: the per-mesh driver with an AMR cycle every 10 steps (next_iteration_impl),
    # counted as the reference counts (owned cells of the step's mesh per step)
    e2e_steps, cycles_done = 40, 0
    amr.cycle_frequency = 10
    sol.iteration = 0
    torch.cuda.synchronize()
    cells_done = 0
    w0 = time.perf_counter()
    for _ in range(e2e_steps):
        if sol.should_do_amr_cycle(amr):
            sol.do_amr_cycle(amr)
            cycles_done += 1
        cells_done += sol.mesh.owned_cells()
        sol.step()
    torch.cuda.synchronize()
    e2e_wall = time.perf_counter() - w0
    out = {"metric": "AMR-cycle remap throughput", "value": algo / (remap_ms * 1e-3) / 1e9, "unit": "GB/s",
           "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": args.workload, "steps_before": args.steps, "leaves_before": m.num_leaves,
                      "leaves_after": new.num_leaves, "copied": counts[0], "prolonged": counts[1],
                      "restricted": counts[2], "refine": int((flags > 0).sum()), "coarsen": int((flags < 0).sum())},
           "roofline": {"bound": "hbm", "achieved": algo / (remap_ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                        "unit": "GB/s", "frac": algo / (remap_ms * 1e-3) / HBM_PEAK, "traffic": None,
                        "kernel": "k_remap_cells + k_remap_faces + k_prolong_faces",
                        "algorithmic_bytes": f"{algo} B: old and new owned state, once each"},
           "ms": {"refine_flags_device_and_readback": flags_ms, "host_adapt_balance": (t1 - t0) * 1e3,
                  "host_remap_plan": (t2 - t1) * 1e3, "remap_device": remap_ms,
                  "new_mesh_tables": (t4 - t3) * 1e3, "step_after": step_ms},
           "end_to_end": {"steps": e2e_steps, "amr_cycles": cycles_done, "cycle_frequency": 10,
                          "leaves_at_end": sol.mesh.num_leaves, "cell_updates_per_s": cells_done / e2e_wall,
                          "ms_per_step": e2e_wall * 1e3 / e2e_steps,
                          "driver": "per-mesh solver (host dt each step), AMR cycles included"}}
    print(json.dumps(out))


def bench_checkpoint(args, cases):
    """Checkpoint write, XDMF write and restart of the workload's first mesh (files under a
    temporary directory). The checkpoint's bytes are the owned state (U and face B) plus the leaf
    list; the restart reads them, rebuilds the mesh and its device tables, and uploads the state."""
    import shutil
    import tempfile

    import numpy as np
    import torch
    from godunov_mhd_ct_amd.implem import IMPL_FUSED
    from godunov_mhd_ct_amd.io import restart_solver, save_checkpoint, write_xdmf
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD

    m, U, B, st = cases[0]
    sol = SolverGodunovMHD(m, U, B, st, impl_version=IMPL_FUSED)
    for _ in range(args.steps):
        sol.step()
    torch.cuda.synchronize()
    d = Path(tempfile.mkdtemp(prefix="mhdct_chk_"))
    try:
        log(f"[bench] checkpoint of {m.num_leaves} leaves after {args.steps} steps into {d}")
        t0 = time.perf_counter()
        Uh, Bh = sol.state()
        t1 = time.perf_counter()
        save_checkpoint(d / "chk", sol.mesh, Uh, Bh, sol.t, sol.iteration, settings=st)
        t2 = time.perf_counter()
        write_xdmf(d / "chk", sol.mesh, Uh, Bh, sol.t, st.gamma0, scaling=st.scaling_factor)
        t3 = time.perf_counter()
        log("[bench] checkpoint and XDMF written")
        nbytes = (d / "chk.bin").stat().st_size
        vbytes = (d / "chk_cells.bin").stat().st_size
        t4 = time.perf_counter()
        rst = restart_solver(d / "chk", impl_version=IMPL_FUSED)
        torch.cuda.synchronize()
        t5 = time.perf_counter()
        n = m.num_leaves
        Ur, Br = rst.state()
        same = bool(np.array_equal(Ur[:n], Uh[:n]) and np.array_equal(Br[:n], Bh[:n]))
        rst.step()
        sol.step()
        torch.cuda.synchronize()
        Ua, Ba = sol.state()
        Ub, Bb = rst.state()
        cont = bool(np.array_equal(Ua[:n], Ub[:n]) and np.array_equal(Ba[:n], Bb[:n]))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    write_s = t2 - t1
    out = {"metric": "checkpoint write throughput", "value": nbytes / write_s / 1e9, "unit": "GB/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": args.workload, "leaves": n, "steps_before": args.steps,
                      "checkpoint_bytes": nbytes, "xdmf_data_bytes": vbytes},
           "roofline": None,
           "ms": {"device_to_host": (t1 - t0) * 1e3, "checkpoint_write": write_s * 1e3,
                  "xdmf_write": (t3 - t2) * 1e3, "restart_read_rebuild_upload": (t5 - t4) * 1e3},
           "GB_per_s": {"device_to_host": nbytes / (t1 - t0) / 1e9, "xdmf_write": vbytes / (t3 - t2) / 1e9,
                        "restart": nbytes / (t5 - t4) / 1e9},
           "restart_state_bitwise": same, "next_step_bitwise": cont,
           "note": "host file-system work: bounded by the box's page cache and PCIe, not by HBM"}
    print(json.dumps(out))


def precision_errors(cases, steps):
    """Relative L1 error per field of the FP32-bulk form against the FP64 form after `steps`
    steps from the same state, each with its own CFL dt (the FP64 form matches the oracle at the
    1e-12 parity bar).  Momentum and B are normalised as vectors."""
    import numpy as np
    import torch
    from godunov_mhd_ct_amd.batch import BatchGodunovMHD
    from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_FUSED_FP32

    res = {}
    states = {}
    for impl in (IMPL_FUSED, IMPL_FUSED_FP32):
        bt = BatchGodunovMHD(cases[0][0].dim, impl)
        ids = [bt.add(m, U, B, st) for (m, U, B, st) in cases]
        bt.set_graph(True)
        done = 0
        while done < steps:
            n = min(50, steps - done)
            bt.step(n)
            done += n
            torch.cuda.synchronize()
            log(f"[bench] precision run impl={impl}: {done}/{steps} steps")
        bt.check()
        states[impl] = [(bt.state(i), bt.slot_state(i)[1]) for i in ids]
        del bt
    l1 = lambda a, b: float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))  # noqa: E731
    for k, (m, _, _, _) in enumerate(cases):
        (U64, B64), t64 = states[IMPL_FUSED][k]
        (U32, B32), t32 = states[IMPL_FUSED_FP32][k]
        no = m.num_leaves
        res = {"steps": steps, "time_fp64": t64, "time_fp32": t32,
               "rho": l1(U32[:no, 0], U64[:no, 0]), "E": l1(U32[:no, 1], U64[:no, 1]),
               "momentum": l1(U32[:no, 2:5], U64[:no, 2:5]),
               "mx": l1(U32[:no, 2], U64[:no, 2]), "my": l1(U32[:no, 3], U64[:no, 3]),
               "mz_abs_over_momentum": float(np.abs(U32[:no, 4] - U64[:no, 4]).sum()
                                             / max(np.abs(U64[:no, 2:5]).sum(), 1e-300)),
               "B": l1(B32[:no], B64[:no])}
    return res


def bench_serving(args):
    import torch
    from godunov_mhd_ct_amd.serving import Server, request_stream
    from godunov_mhd_ct_amd.implem import IMPL_FUSED

    reqs = request_stream(args.requests, rate=args.rate, seed=0, nsteps=200)
    free, total = torch.cuda.mem_get_info()
    budget = int(0.5 * free)
    log(f"[bench] serving {len(reqs)} requests at {args.rate}/s, HBM budget {budget / 2**30:.1f} GiB")
    # warm-up: compile-free, but the first batch steps page code in
    Server(hbm_budget=budget).run(request_stream(2, rate=1000.0, seed=99, nsteps=5, sizes2d=(32,), sizes3d=(32,)))
    srv = Server(hbm_budget=budget, impl=IMPL_FUSED, max_tick=20, use_graph=args.graph)
    res = srv.run(reqs)
    out = {"metric": METRIC, "value": res["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1,
           "steps": 200, "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "latency_ms_p50": res["latency_ms_p50"], "latency_ms_p99": res["latency_ms_p99"],
           "requests_per_s": res["requests"] / res["wall_s"],
           "config": {"workload": "srv", "requests": res["requests"], "rate_per_s": args.rate,
                      "steps_per_request": 200, "hbm_budget_gib": budget / 2 ** 30, "graph": args.graph},
           "roofline": {"bound": "hbm", "achieved": res["cell_updates_per_s"] * ALGO_BYTES_PER_CELL / 1e9,
                        "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": res["cell_updates_per_s"] * ALGO_BYTES_PER_CELL / HBM_PEAK, "traffic": None,
                        "kernel": "whole serving run"}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
