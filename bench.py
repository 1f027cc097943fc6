#!/usr/bin/env python3
"""Hydro-step throughput on one MI355X: cell-updates/s of the FP64 path.

A step is one iteration of the reference's sim::run (simulation.cpp:515-556):
CFL time step with the CFLmaxVar limiter and the monitor clip, then the gas
part of step_Euler (potential, sources, artificial viscosity, viscosity,
SubStep3, boundaries, van Leer / FARGO transport, damping, derived fields),
all resident in HBM and replayed as one captured HIP graph per step.

Default workload: the adiabatic 512 x 1536 disk (gamma 1.4, von
Neumann-Richtmyer viscosity, alpha 1e-3, planet q 1e-3), seeded initial
condition, planets on fixed positions.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s per hydro step: FP64 vs split-FP32 at stated err; % HBM peak"
HBM_PEAK_GBS = 8000.0


def cpu_baseline(name: str, seconds: float) -> dict:
    """The oracle (C restatement of the reference loops, OpenMP) on this host,
    from the same initial state, for a bounded number of steps."""
    from oracle import OracleDisk
    from fargocpt_amd import workloads as W
    d = OracleDisk(W.disk_config(name))
    W.start(d, name)
    d.run_step()  # first-touch and thread start-up outside the timed sample
    n = 0
    t0 = time.perf_counter()
    while True:
        d.run_step()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 500:
            break
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": W.cells(name) * n / el, "unit": "cell-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"{n} steps of {name} on the host ({el:.1f} s), oracle/liboracle.so"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="adi_512x1536")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the hydro step of one disk runs on a single "
              "MI355X here; multi-GPU runs are not part of this build", file=sys.stderr)
        return 2

    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import KERNELS, Disk

    name = args.workload
    use_graph = not args.no_graph
    disk = Disk(W.disk_config(name))
    W.start(disk, name)
    disk.run_steps(args.warmup, use_graph)
    disk.synchronize()
    t0 = time.perf_counter()
    ms_events = disk.time_steps(args.steps, use_graph)
    disk.synchronize()
    wall = time.perf_counter() - t0
    cells = W.cells(name)
    value = cells * args.steps / wall
    t_sim, _, _ = disk.get_time()

    # roofline of the dominant kernel, timed live with HIP events on the stream
    adi = W.is_adiabatic(name)
    kt = {k: disk.time_kernel(v, 20) for k, v in KERNELS.items()}
    dom = max(kt, key=kt.get)
    achieved = W.algorithmic_bytes(dom, adi) * cells / (kt[dom] * 1e-3) / 1e9
    copy_ms, copy_bytes = disk.time_copy(20)
    traffic = None
    prof = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(prof):
        traffic = json.load(open(prof)).get(dom)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": name, "description": W.DESCRIPTION[name],
                   "cells": cells, "graph": use_graph, "sim_time_reached": t_sim},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_cell": W.algorithmic_bytes(dom, adi),
                     "ms_per_launch": kt[dom]},
        "step": {"ms_per_step_events": ms_events / args.steps,
                 "algorithmic_GBs": W.STEP_BYTES[adi] * value / 1e9,
                 "frac_of_hbm_peak": W.STEP_BYTES[adi] * value / 1e9 / HBM_PEAK_GBS,
                 "kernel_ms": kt,
                 "copy_probe_GBs": copy_bytes / (copy_ms * 1e-3) / 1e9},
    }
    if not args.skip_cpu:
        out["cpu_baseline"] = cpu_baseline(name, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
