"""Build recipe for libholo.so (gfx950) and the `holo` CLI.

Everything is compiled in-tree with hipcc so that the shared library travels
with the source tree:

    python holographic-projector-simulations_amd/build.py [--jobs N]

Flags: -O3 for gfx950 only; packed FP32 instructions are disabled
(-packed-fp32-ops) because some of their forms have given wrong results on
MI355X hosts; no fast-math, so FP64 operations keep IEEE rounding.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libholo.so")
BIN_DIR = os.path.join(HERE, "bin")
CLI = os.path.join(BIN_DIR, "holo")

ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA results in ordinary VGPRs: no accumulation-register copies
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall", "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
]
LIB_SOURCES = ["holo_kernels.hip", "holo_grid.hip", "holo_capi.hip"]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in (r.stderr or "").splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")
    return err


def build(jobs: int = 4, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    cc = hipcc()
    objs = []
    cmds = []
    for src in LIB_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        objs.append(obj)
        cmds.append([cc, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for msg in ex.map(_run, cmds):
            if verbose and msg:
                print(msg, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "holo_cli.cpp")
    if os.path.exists(cli_src):
        os.makedirs(BIN_DIR, exist_ok=True)
        _run([cc, "-O2", "-std=c++17", f"-I{INCLUDE}", cli_src, "-o", CLI,
              f"-L{HERE}", "-lholo", "-Wl,-rpath,$ORIGIN/.."])
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    print(build(a.jobs, verbose=True))
