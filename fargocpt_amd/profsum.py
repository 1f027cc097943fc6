"""Summarise the rocprofv3 runs of one bench.py configuration into a short
Markdown table and a JSON of HBM bytes (kernel -> bytes per launch, plus the
bytes of one whole step), the files bench.py reads back as `traffic`.

Three runs of `bench.py --workload W --precision P --skip-cpu --no-compare
--no-autotune` are expected under DIR: a kernel trace with statistics
(`DIR/prof_W_P_kt`, 50 timed + 5 warm-up steps), and two counter runs
(`DIR/prof_W_P_fetch` with FETCH_SIZE, `DIR/prof_W_P_write` with WRITE_SIZE,
10 + 2 steps).  Byte counters are calibrated on `k_copy4`, which moves
4 x element x cells bytes each way per launch.

Usage: python -m fargocpt_amd.profsum W P [DIR] [--out profiles]
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re

# bench.py times these kernels with fcpt_time_kernel: 3 rounds of 1 warm-up + 20
# launches (the viscous pass timed is the walk without SubStep3)
TIMED = (r"k_radial_sweep", r"k_azimuthal_sweep", r"k_cfl_rings", r"k_visc_walk<(true|false), false>",
         r"k_potential_fold<")
TIMING_LAUNCHES = 63
KT_STEPS = 55
JSON_NAMES = {r"k_radial_sweep": "radial_sweep", r"k_azimuthal_sweep": "azimuthal_sweep",
              r"k_cfl_rings": "cfl", r"k_visc_walk<(true|false), false>": "visc_fused",
              r"k_potential_fold<": "potential", r"k_disk_force<": "disk_force"}


def _short(name: str) -> str:
    return name.replace("fcpt::", "").split("(")[0]


def _per_kernel(files) -> dict:
    agg = collections.defaultdict(list)
    for f in files:
        for r in csv.DictReader(open(f)):
            agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in agg.items()}


def summarise(workload: str, prec: str, base: str, outdir: str) -> dict:
    from fargocpt_amd import workloads as W
    cells = W.cells(workload)
    esize = 8 if prec == "fp64" else 4
    tag = f"{workload}_{prec}"
    kt = glob.glob(f"{base}/prof_{tag}_kt/**/*kernel_stats.csv", recursive=True)
    fe = glob.glob(f"{base}/prof_{tag}_fetch/**/*counter_collection.csv", recursive=True)
    wr = glob.glob(f"{base}/prof_{tag}_write/**/*counter_collection.csv", recursive=True)
    cmd = f"python3 bench.py --workload {workload} --precision {prec}"
    opts = "--skip-cpu --no-compare --no-autotune"
    L = [f"# rocprofv3 summary: {workload} ({cells} cells), "
         f"{'FP64' if prec == 'fp64' else 'split'} storage, graph-replayed step\n",
         "Commands (repository root, one MI355X; default launch shapes, no autotuning):\n",
         "```",
         f"rocprofv3 --kernel-trace --stats --output-format csv -d {base}/prof_{tag}_kt -o kt -- "
         f"{cmd} --steps 50 --warmup 5 {opts}",
         f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d {base}/prof_{tag}_fetch -o fetch -- "
         f"{cmd} --steps 10 --warmup 2 {opts}",
         f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d {base}/prof_{tag}_write -o write -- "
         f"{cmd} --steps 10 --warmup 2 {opts}",
         f"python3 -m fargocpt_amd.profsum {workload} {prec} {base}",
         "```\n"]
    rows = list(csv.DictReader(open(kt[0])))
    L += ["## Kernel time (kernel_stats.csv)\n",
          f"Calls include the {TIMING_LAUNCHES} launches with which bench.py times each of the "
          "radial sweep, azimuthal sweep, CFL rings, potential and the viscous walk without "
          "SubStep3 (an adiabatic step runs the walk with SubStep3, so there this one shows 0 "
          "launches per step); launches per "
          f"step excludes them and divides by the {KT_STEPS} traced steps.\n",
          "| kernel | calls | launches per step | average µs | % of GPU time |",
          "|---|---|---|---|---|"]
    per_step = {}
    avg_us = {}
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        name = r["Name"]
        calls = int(r["Calls"])
        extra = TIMING_LAUNCHES if any(re.search(t, name) for t in TIMED) else 0
        n = (calls - extra) // KT_STEPS
        per_step[_short(name)] = n
        avg_us[_short(name)] = float(r["AverageNs"]) / 1e3
        L.append(f"| `{_short(name)}` | {calls} | {n} | {avg_us[_short(name)]:.1f} | "
                 f"{float(r['Percentage']):.1f} |")
    F = _per_kernel(fe)
    Wr = _per_kernel(wr)
    copy = [k for k in F if "k_copy4" in k][0]
    exp_kib = 4 * esize * cells / 1024.0
    ff, fw = exp_kib / F[copy], exp_kib / Wr[copy]
    L += ["\n## HBM bytes per launch (counter_collection.csv of the --pmc runs, mean per kernel)\n",
          f"Calibration: `k_copy4` moves {exp_kib:.0f} KiB each way per launch; FETCH_SIZE reads "
          f"{F[copy]:.0f} KiB and WRITE_SIZE {Wr[copy]:.0f} KiB for it, so fetch is scaled by "
          f"{ff:.3f} and write by {fw:.3f}.\n",
          "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | scaled HBM bytes / cell | achieved GB/s |",
          "|---|---|---|---|---|"]
    traffic = {}
    step_bytes = 0.0
    step_us = 0.0
    for k in sorted(F, key=lambda k: -(F[k] * ff + Wr.get(k, 0) * fw)):
        b = (F[k] * ff + Wr.get(k, 0) * fw) * 1024
        s = _short(k)
        us = avg_us.get(s, 0.0)
        gbs = b / (us * 1e-6) / 1e9 if us else 0.0
        L.append(f"| `{s}` | {F[k]:.0f} | {Wr.get(k, 0):.0f} | {b / cells:.1f} | {gbs:.0f} |")
        traffic[s] = b
        step_bytes += per_step.get(s, 0) * b
        step_us += per_step.get(s, 0) * us
    L += ["\n## One step\n",
          f"HBM bytes per step (scaled counters x launches per step): {step_bytes / 1e6:.1f} MB "
          f"= {step_bytes / cells:.0f} B/cell; kernel time per step {step_us:.1f} µs; "
          f"{step_bytes / (step_us * 1e-6) / 1e9:.0f} GB/s over the kernels "
          f"({step_bytes / (step_us * 1e-6) / 8e12 * 100:.0f} % of 8 TB/s). "
          f"Algorithmic bytes (evolved fields read + written once): "
          f"{W.step_bytes(W.is_adiabatic(workload), esize)} B/cell."]
    os.makedirs(outdir, exist_ok=True)
    with open(os.path.join(outdir, f"{tag}_step.md"), "w") as fh:
        fh.write("\n".join(L) + "\n")
    tj = {}
    for s, b in traffic.items():
        for kk, v in JSON_NAMES.items():
            if re.search(kk, s):
                tj[v] = b
    tj["step"] = step_bytes
    tj["step_kernel_us"] = step_us
    with open(os.path.join(outdir, f"traffic_{tag}.json"), "w") as fh:
        json.dump(tj, fh, indent=1)
    return tj


# instruction counters of the issue summary: two rocprofv3 passes (at most 8
# SQ counters each)
ISSUE_PASSES = {
    "issue": ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR",
              "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES"),
    "mix": ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64",
            "SQ_INSTS_VALU_TRANS_F64", "SQ_INSTS_VALU_INT32", "SQ_INSTS_VALU_CVT", "SQ_INSTS_VALU_INT64",
            "SQ_INSTS_SMEM"),
}
STEP_KERNELS = ("k_azimuthal_sweep", "k_radial_sweep3", "k_visc_walk<(true|false), true>", "k_velocities_final",
                "k_kick_walk", "k_cfl_rings", "k_potential_fold", "k_ghost_ring_mean")
SIMDS, CLOCK_HZ = 1024, 2.4e9  # MI355X: 256 CUs x 4 SIMDs, 2.4 GHz peak engine clock


def _counters(files) -> dict:
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for f in files:
        for r in csv.DictReader(open(f)):
            agg[_short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in agg.items()}


def summarise_issue(workload: str, base: str, outdir: str) -> None:
    """VALU issue of the step kernels against their measured time, both storages.

    Expects DIR/prof_W_P_issue and DIR/prof_W_P_mix (--pmc runs of
    ISSUE_PASSES, 3 + 1 steps) and DIR/prof_W_P_kt (the kernel trace of
    summarise) for P in fp64, split.  The issue time is the VALU instructions
    at one wavefront instruction per 4 cycles per SIMD over all SIMDs at the
    peak clock: the time the kernel would take if VALU issue were all it did.
    """
    cmd = f"python3 bench.py --workload {workload} --precision P --steps 3 --warmup 1 " \
          "--skip-cpu --no-compare --no-autotune"
    L = [f"# Instruction issue of the step kernels: {workload}\n",
         "Commands (repository root, one MI355X; P = fp64, then split):\n", "```"]
    for name, ctrs in ISSUE_PASSES.items():
        L.append(f"rocprofv3 --pmc {' '.join(ctrs)} --output-format csv -d {base}/prof_{workload}_P_{name} "
                 f"-o p -- {cmd}")
    L += [f"python3 -m fargocpt_amd.profsum {workload} issue {base}", "```\n",
          "Per wavefront (counter ÷ SQ_WAVES). VALU issue µs = SQ_INSTS_VALU x 4 cycles ÷ "
          f"({SIMDS} SIMDs x {CLOCK_HZ / 1e9:.1f} GHz); measured µs from the kernel trace "
          f"(`prof_{workload}_P_kt`). A kernel whose issue time is near its measured time is bound by "
          "its VALU instructions; well below it, by its bytes or its latency.\n",
          "| kernel | storage | waves | VALU/wave | F64 add+mul+fma | F64 trans | int32 | cvt | SALU/wave "
          "| SMEM/wave | VALU issue µs | measured µs | issue ÷ measured |",
          "|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for prec in ("fp64", "split"):
        C = {}
        for name in ISSUE_PASSES:
            files = glob.glob(f"{base}/prof_{workload}_{prec}_{name}/**/*counter_collection.csv", recursive=True)
            for k, d in _counters(files).items():
                C.setdefault(k, {}).update(d)
        kt = glob.glob(f"{base}/prof_{workload}_{prec}_kt/**/*kernel_stats.csv", recursive=True)
        avg_us = {_short(r["Name"]): float(r["AverageNs"]) / 1e3 for r in csv.DictReader(open(kt[0]))}
        for pat in STEP_KERNELS:
            for k in sorted(C):
                if not re.search(pat, k):
                    continue
                c = C[k]
                w = c["SQ_WAVES"]
                valu = c["SQ_INSTS_VALU"]  # over the launch
                issue_us = valu * 4 / SIMDS / CLOCK_HZ * 1e6
                arith = c["SQ_INSTS_VALU_ADD_F64"] + c["SQ_INSTS_VALU_MUL_F64"] + c["SQ_INSTS_VALU_FMA_F64"]
                us = avg_us.get(k, float("nan"))
                L.append(f"| `{k}` | {prec} | {w:.0f} | {valu / w:.0f} | {arith / w:.0f} | "
                         f"{c['SQ_INSTS_VALU_TRANS_F64'] / w:.0f} | {c['SQ_INSTS_VALU_INT32'] / w:.0f} | "
                         f"{c['SQ_INSTS_VALU_CVT'] / w:.0f} | {c['SQ_INSTS_SALU'] / w:.0f} | "
                         f"{c['SQ_INSTS_SMEM'] / w:.0f} | {issue_us:.0f} | {us:.0f} | {issue_us / us:.2f} |")
    os.makedirs(outdir, exist_ok=True)
    with open(os.path.join(outdir, f"{workload}_issue.md"), "w") as fh:
        fh.write("\n".join(L) + "\n")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("workload")
    ap.add_argument("precision", choices=("fp64", "split", "issue"),
                    help="storage of the traffic summary, or issue: the instruction summary of both")
    ap.add_argument("base", help="folder holding the rocprofv3 output folders")
    ap.add_argument("--out", default="profiles")
    a = ap.parse_args()
    if a.precision == "issue":
        summarise_issue(a.workload, a.base, a.out)
        return 0
    print(json.dumps(summarise(a.workload, a.precision, a.base, a.out)))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
