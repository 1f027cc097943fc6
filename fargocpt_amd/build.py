"""Build libfcpt_hip.so in-tree for gfx950 with hipcc.

Flags:
  --offload-arch=gfx950            MI355X (CDNA4) only
  -ffp-contract=off                no FMA contraction: FP64 add/mul round like the
                                   reference's host build
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                   never emit packed FP32 VALU instructions
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfcpt_hip.so")
SOURCES = ["fcpt_kernels.hip", "fcpt_capi.hip", "fcpt_host.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libfcpt_hip.so failed")
    os.replace(tmp, OUT)
    return OUT


def build_host_check() -> str:
    """The C++ host-interface check program used by the GPU tests."""
    src = os.path.join(os.path.dirname(HERE), "tests", "cpp", "host_api_check.cpp")
    exe = os.path.join(os.path.dirname(HERE), "tests", "cpp", "host_api_check.bin")
    cmd = [HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", src, "-o", exe,
           f"-L{HERE}", f"-Wl,-rpath,{HERE}", "-lfcpt_hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build of host_api_check failed")
    return exe


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_host_check())
