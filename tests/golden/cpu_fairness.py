#!/usr/bin/env python3
"""Fairness check of the CPU baseline: the oracle (oracle/liboracle.so, the
C restatement of the reference's loops that bench.py times) beside the
reference itself, built from its sources by oracle/ref.mk (-O2, no fast-math)
and driven through the same sim::run iterations by tests/golden/refdump.cpp,
on the same configuration, initial condition, step count and threads.

Runs only where the reference sources are (/root/reference/src).  Prints one
JSON line per case: both cell-update rates and their ratio.

Usage:  python tests/golden/cpu_fairness.py [--threads 8] [case:steps ...]
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import make_golden as G  # noqa: E402
from fixtures import disk_config, load_case, start_run, step_inputs  # noqa: E402

OMP_ENV = {"OMP_PROC_BIND": "close", "OMP_PLACES": "cores", "OMP_WAIT_POLICY": "active"}


def lscpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_reference(exe, name, steps, threads, reps):
    work = tempfile.mkdtemp(prefix=f"fair_{name}_")
    try:
        cfg, icdir = G.case_inputs(name, exe, 0, work)
        best = []
        for r in range(reps):
            env = dict(OMP_ENV, OMP_NUM_THREADS=str(threads), REFDUMP_TIME="1")
            dump = G.run_driver(exe, 0, cfg, os.path.join(work, f"run{r}"), icdir, steps, 0, [],
                                env_extra=env)
            best.append(G.read_meta(os.path.join(dump, "meta.txt"))["loop_seconds"])
        return sorted(best)[len(best) // 2]
    finally:
        shutil.rmtree(work, ignore_errors=True)


def time_oracle(name, steps, threads, reps):
    from oracle import OracleDisk
    meta, data = load_case(name)
    out = []
    for _ in range(reps):
        d = OracleDisk(disk_config(meta))
        d.set_threads(threads)
        start_run(d, meta)
        planets, ind = step_inputs(data, 1)
        d.set_step_params(planets, ind)
        t0 = time.perf_counter()
        for _ in range(steps):
            d.run_step()
        out.append(time.perf_counter() - t0)
        d.close()
    return sorted(out)[len(out) // 2]


def main(argv):
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("cases", nargs="*", default=["c1_iso:100", "c2_adi:100"])
    args = ap.parse_args(argv)
    exe, _ = G.build()
    for spec in args.cases:
        name, steps = spec.split(":")
        steps = int(steps)
        meta, _ = load_case(name)
        cells = meta["Nr"] * meta["Nphi"]
        t_ref = time_reference(exe, name, steps, args.threads, args.reps)
        t_orc = time_oracle(name, steps, args.threads, args.reps)
        print(json.dumps({
            "case": name, "cells": cells, "steps": steps, "threads": args.threads,
            "cpu": lscpu_model(),
            "reference_cell_updates_per_s": cells * steps / t_ref,
            "oracle_cell_updates_per_s": cells * steps / t_orc,
            "oracle_over_reference": t_ref / t_orc,
        }), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
