#!/usr/bin/env python3
"""Hydro-step throughput on one MI355X: cell-updates/s of the FP64 path and of
the split-storage path, with the split path's error against FP64.

A step is one iteration of the reference's sim::run (simulation.cpp:515-556):
CFL time step with the CFLmaxVar limiter and the monitor clip, then the gas
part of step_Euler (potential, sources, artificial viscosity, viscosity,
SubStep3, boundaries, van Leer / FARGO transport, damping, derived fields),
all resident in HBM and replayed as one captured HIP graph per step.

Default workload: the large single disk, 4096 x 12288 rings x sectors in
[0.4, 4.0], adiabatic (gamma 1.4, von Neumann-Richtmyer viscosity, viscous
heating) with beta cooling (beta 10) and three planets (q 1e-3, 3e-4, 1e-4 at
r 1.0, 1.6, 2.5), seeded initial condition, planets on fixed positions.
`value` is the throughput of the storage mode chosen with --precision (FP64
by default); the other mode runs the same steps from the same state and is
reported under "compare", with its per-field error after --error-steps steps.
Prints one JSON line on stdout; progress lines go to stderr.
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
PRECISIONS = {"fp64": 0, "split": 1}
ELEMENT = {0: 8, 1: 4}


def log(msg: str) -> None:
    """Progress line on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu() -> dict:
    """Model name, sockets and physical cores of this host (lscpu fields)."""
    info = {"model": None, "sockets": None, "physical_cores": None, "logical_cpus": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for ln in out.splitlines():
            if ":" in ln:
                k, v = ln.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        sockets = int(kv.get("Socket(s)", "0") or 0)
        per = int(kv.get("Core(s) per socket", "0") or 0)
        info["sockets"] = sockets or None
        info["physical_cores"] = sockets * per or None
    except Exception:  # lscpu absent: /proc/cpuinfo's model name only
        try:
            for ln in open("/proc/cpuinfo"):
                if ln.startswith("model name"):
                    info["model"] = ln.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return info


def cpu_baseline(name: str, seconds: float, one_thread_steps: int) -> dict:
    """The oracle (C restatement of the reference loops, OpenMP) on this host,
    from the same initial state, for a bounded number of steps; then a few
    steps on one thread."""
    from oracle import OracleDisk
    from fargocpt_amd import workloads as W
    d = OracleDisk(W.disk_config(name))
    W.start(d, name)
    cores = d.threads()
    d.run_step()  # first-touch and thread start-up outside the timed sample
    n = 0
    t0 = time.perf_counter()
    while True:
        d.run_step()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 500:
            break
    log(f"cpu baseline: {n} steps on {cores} threads in {el:.1f} s")
    out = {"value": W.cells(name) * n / el, "unit": "cell-updates/s", "cores": cores,
           "kind": "port",
           "sample": f"{n} steps of {name} on the host ({el:.1f} s), oracle/liboracle.so, "
                     f"OpenMP {cores} threads",
           "host": host_cpu()}
    if one_thread_steps > 0:
        d.set_threads(1)
        t0 = time.perf_counter()
        for _ in range(one_thread_steps):
            d.run_step()
        el1 = time.perf_counter() - t0
        d.set_threads(cores)
        out["one_thread"] = {"value": W.cells(name) * one_thread_steps / el1,
                             "sample": f"{one_thread_steps} steps ({el1:.1f} s)"}
        log(f"cpu baseline: {one_thread_steps} steps on 1 thread in {el1:.1f} s")
    d.close()
    return out


def timed_run(name: str, prec: int, steps: int, warmup: int, use_graph: bool, tune: bool):
    """Start the workload, run the warm-up steps, pick launch shapes with the
    library's autotuner (untimed), capture the step graph of those shapes with
    one more untimed step, then time `steps` steps.  Returns the disk, the
    wall time, the event time and the number of steps taken."""
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import Disk
    disk = Disk(W.disk_config(name, prec))
    W.start(disk, name)
    disk.run_steps(warmup, use_graph)
    disk.synchronize()
    disk.tuned = disk.autotune(10) if tune else {}
    if tune:
        disk.tuned.update(disk.autotune_walks(3))
    from fargocpt_amd.capi import PARAM_KICK_TILE, PARAM_VISC_TILE
    disk.tuned.update({"kick_tile": disk.launch_param(PARAM_KICK_TILE),
                       "visc_tile": disk.launch_param(PARAM_VISC_TILE)})
    # the autotuner drops the captured graph: capture it again outside the timing
    disk.run_steps(1, use_graph)
    disk.synchronize()
    t0 = time.perf_counter()
    ms_events = disk.time_steps(steps, use_graph)
    disk.synchronize()
    wall = time.perf_counter() - t0
    return disk, wall, ms_events, warmup + 1 + steps


def batch_run(name: str, prec: int, steps: int, warmup: int, use_graph: bool, tune: bool = False):
    """The batched sweep: every member disk started, then `steps` batched
    steps timed after `warmup` untimed ones (and, with `tune`, the batch's
    radial-sweep autotuner)."""
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import Batch, Disk
    members = W.batch_members(name)
    disks = []
    for m in members:
        d = Disk(W.member_config(m, prec))
        W.member_start(d, m)
        disks.append(d)
    batch = Batch(disks)
    batch.run_steps(warmup, use_graph)
    batch.synchronize()
    batch.shapes = batch.autotune(5) if tune else None
    batch.run_steps(1, use_graph)  # captures the graph of the tuned plan
    batch.synchronize()
    t0 = time.perf_counter()
    ms_events = batch.time_steps(steps, use_graph)
    wall = time.perf_counter() - t0
    cells = sum(m["nr"] * m["nphi"] for m in members)
    return batch, disks, cells, wall, ms_events


def bench_batch(args) -> int:
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import BATCH_KERNELS
    name = args.workload
    prec = PRECISIONS[args.precision]
    use_graph = not args.no_graph
    tune = not args.no_autotune
    batch, disks, cells, wall, ms_events = batch_run(name, prec, args.steps, args.warmup, use_graph, tune)
    value = cells * args.steps / wall
    el = ELEMENT[prec]
    kt = {k: batch.time_kernel(v, 10) for k, v in BATCH_KERNELS.items()}
    dom = max(kt, key=kt.get)
    # the step's algorithmic bytes per cell (evolved fields read + written
    # once), as for the single-disk line; the kernel's own field count apart
    achieved = W.step_bytes(False, el) * cells / (kt[dom] * 1e-3) / 1e9
    kernel_bytes = W.algorithmic_bytes(dom, False, el)
    n = len(disks)
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": name, "disks": n, "cells": cells,
                   "description": f"{n} locally isothermal disks, Nr ~ U[128,1024], Nphi = 3 Nr, "
                                  "q ~ logU[1e-5,1e-3], alpha ~ logU[1e-4,1e-2], h ~ U[0.03,0.07], "
                                  "one launch per stage over a work table of all disks' blocks",
                   "storage": "f64" if prec == 0 else "f32 deviations on f64 ring backgrounds",
                   "graph": use_graph, "launch_shapes": batch.shapes},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_cell": W.step_bytes(False, el), "ms_per_launch": kt[dom],
                     "bytes_basis": "evolved fields read + written once per cell-update",
                     "kernel_bytes_per_cell": kernel_bytes,
                     "kernel_achieved_GBs": kernel_bytes * cells / (kt[dom] * 1e-3) / 1e9},
        "step": {"ms_per_step_events": ms_events / args.steps,
                 "algorithmic_GBs": W.step_bytes(False, el) * value / 1e9,
                 "frac_of_hbm_peak": W.step_bytes(False, el) * value / 1e9 / HBM_PEAK_GBS,
                 "kernel_ms": kt},
    }
    if not args.no_compare:
        other = 1 - prec
        b2, d2, _, wall2, _ = batch_run(name, other, args.steps, args.warmup, use_graph, tune)
        value2 = cells * args.steps / wall2
        fp64, split = (disks, d2) if prec == 0 else (d2, disks)
        worst = {f: max(rel_err(s.download_field(f), d.download_field(f)) for s, d in zip(split, fp64))
                 for f in ("SIGMA", "VRAD", "VAZI")}
        out["compare"] = {"precision": "split" if other == 1 else "fp64", "value": value2,
                          "ms_per_step": 1e3 * wall2 / args.steps,
                          "split_speedup_vs_fp64": (value2 / value) if other == 1 else (value / value2),
                          "split_rel_err_vs_fp64_worst_disk": worst,
                          "after_steps": args.warmup + 1 + args.steps}
        b2.close()
    print(json.dumps(out), flush=True)
    return 0


def kernel_ms(disk, kernel: int, reps: int = 20, rounds: int = 3) -> float:
    """Median over `rounds` of the mean launch time of `reps` back-to-back
    launches (HIP events on the disk's stream)."""
    t = sorted(disk.time_kernel(kernel, reps) for _ in range(rounds))
    return t[len(t) // 2]


def rel_err(a, b) -> float:
    import numpy as np
    scale = float(np.max(np.abs(b)))
    diff = float(np.max(np.abs(a - b)))
    return diff / scale if scale else diff


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="adi_beta_4096x12288")
    ap.add_argument("--precision", choices=sorted(PRECISIONS), default="fp64")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the run of the other storage mode")
    ap.add_argument("--error-steps", type=int, default=1000,
                    help="steps after which the split-vs-FP64 error is stated (untimed "
                         "steps are added to both runs to reach it)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-one-thread-steps", type=int, default=1)
    ap.add_argument("--skip-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the whole disk lives on one MI355X, and runs on "
              "more than one GPU are not implemented (the sharding design is written up, "
              "not built)", file=sys.stderr)
        return 2

    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import KERNEL_DISK_FORCE, KERNELS

    if args.workload in W.BATCHES:
        return bench_batch(args)
    name = args.workload
    prec = PRECISIONS[args.precision]
    use_graph = not args.no_graph
    cells = W.cells(name)
    adi = W.is_adiabatic(name)
    tune = not args.no_autotune
    log(f"{name} ({cells} cells), {args.precision}: warm-up {args.warmup}, timed {args.steps}")
    disk, wall, ms_events, taken = timed_run(name, prec, args.steps, args.warmup, use_graph, tune)
    value = cells * args.steps / wall
    t_sim, _, _ = disk.get_time()
    log(f"{args.precision}: {1e3 * wall / args.steps:.4f} ms/step, {value:.4g} cell-updates/s")

    # roofline of the dominant kernel, timed live with HIP events on the
    # library's stream: the step's algorithmic bytes per cell (each evolved
    # field read and written once at storage precision) x cells / its launch time
    el = ELEMENT[prec]
    kt = {k: kernel_ms(disk, v) for k, v in KERNELS.items()}
    dom = max(kt, key=kt.get)
    b_alg = W.step_bytes(adi, el)
    achieved = b_alg * cells / (kt[dom] * 1e-3) / 1e9
    b_kernel = W.algorithmic_bytes(dom, adi, el)
    copy_ms, copy_bytes = disk.time_copy(20)
    # the disk's gravity on star + planet (ComputeDiskOnPlanetAccel), once per step
    # in the reference's N-body coupling
    force_ms = disk.time_kernel(KERNEL_DISK_FORCE, 20)
    force_bytes = W.algorithmic_bytes("disk_force", adi, el) * cells
    log(f"kernel timings done (dominant {dom}: {kt[dom]:.4f} ms)")
    traffic = None
    measured = {}
    prof = os.path.join(ROOT, "profiles", f"traffic_{name}_{args.precision}.json")
    if os.path.exists(prof):
        tj = json.load(open(prof))
        traffic = tj.get(dom)
        if tj.get("step"):
            # rocprof-measured HBM bytes of one whole step over this run's step time
            gbs = tj["step"] / (wall / args.steps) / 1e9
            measured = {"hbm_bytes_per_step_rocprof": tj["step"],
                        "measured_GBs": gbs, "measured_frac_of_hbm_peak": gbs / HBM_PEAK_GBS,
                        "passes_over_algorithmic": tj["step"] / (b_alg * cells)}

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
        "config": {"workload": name, "description": W.DESCRIPTION[name], "cells": cells,
                   "storage": "f64" if prec == 0 else "f32 deviations on f64 ring backgrounds",
                   "graph": use_graph, "launch_shapes": disk.tuned, "sim_time_reached": t_sim},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_cell": b_alg, "ms_per_launch": kt[dom],
                     "bytes_basis": "evolved fields read + written once per cell-update",
                     "kernel_bytes_per_cell": b_kernel,
                     "kernel_achieved_GBs": b_kernel * cells / (kt[dom] * 1e-3) / 1e9},
        "step": {"ms_per_step_events": ms_events / args.steps,
                 "algorithmic_GBs": b_alg * value / 1e9,
                 "frac_of_hbm_peak": b_alg * value / 1e9 / HBM_PEAK_GBS,
                 "kernel_ms": kt,
                 "copy_probe_GBs": copy_bytes / (copy_ms * 1e-3) / 1e9, **measured},
        "disk_force": {"bodies": 2, "ms_per_call": force_ms,
                       "bytes_per_cell": W.algorithmic_bytes("disk_force", adi, el),
                       "achieved_GBs": force_bytes / (force_ms * 1e-3) / 1e9,
                       "frac_of_hbm_peak": force_bytes / (force_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    if not args.no_compare:
        other = 1 - prec
        log(f"{'split' if other else 'fp64'}: same start, same steps")
        disk2, wall2, _, taken2 = timed_run(name, other, args.steps, args.warmup, use_graph, tune)
        value2 = cells * args.steps / wall2
        log(f"{'split' if other else 'fp64'}: {1e3 * wall2 / args.steps:.4f} ms/step")
        # both runs continue, untimed, to the step count the error is stated at
        target = max(args.error_steps, taken, taken2)
        disk.run_steps(target - taken, use_graph)
        disk2.run_steps(target - taken2, use_graph)
        disk.synchronize()
        disk2.synchronize()
        fp64, split = (disk, disk2) if prec == 0 else (disk2, disk)
        fields = ["SIGMA", "VRAD", "VAZI"] + (["ENERGY"] if adi else [])
        errs = {f: rel_err(split.download_field(f), fp64.download_field(f)) for f in fields}
        m64, l64 = fp64.totals(1e300)
        m32, l32 = split.totals(1e300)
        t64, t32 = fp64.get_time()[0], split.get_time()[0]
        log(f"errors after {target} steps: {errs}")
        out["compare"] = {
            "precision": "split" if other == 1 else "fp64",
            "value": value2,
            "ms_per_step": 1e3 * wall2 / args.steps,
            "split_speedup_vs_fp64": (value2 / value) if other == 1 else (value / value2),
            "split_rel_err_vs_fp64": errs,
            "split_total_mass_rel_err": abs(m32 - m64) / abs(m64),
            "split_total_angular_momentum_rel_err": abs(l32 - l64) / abs(l64),
            "sim_time_rel_diff": abs(t32 - t64) / abs(t64),
            "after_steps": target,
            "error_definition": "max |split - fp64| / max |fp64| per field, same start, same steps",
        }
        disk2.close()
    # host arrays across the boundary (fcpt_download_state / fcpt_upload_state,
    # FP64 reference layout); the step itself never leaves HBM
    t0 = time.perf_counter()
    state = disk.download_state()
    dl = time.perf_counter() - t0
    t0 = time.perf_counter()
    disk.upload_state(*state)
    disk.synchronize()
    ul = time.perf_counter() - t0
    nbytes = sum(a.nbytes for a in state if a is not None)
    out["host_transfer"] = {
        "download_ms": 1e3 * dl, "upload_ms": 1e3 * ul, "bytes_each_way": nbytes,
        "value_if_state_crossed_pcie_every_step": cells / (wall / args.steps + dl + ul),
    }
    disk.close()
    del state
    if not args.skip_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(name, args.cpu_seconds, args.cpu_one_thread_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
