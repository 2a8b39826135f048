"""Build libclbm_amd.so for gfx950 with hipcc (in-tree, next to this file).

Usage: python build.py [--arch gfx950] [--verbose]
"""

from __future__ import annotations

import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["csrc/clbm_lbm.hip", "csrc/clbm_ibm.hip", "csrc/clbm_closure.hip",
           "csrc/clbm_coupled.hip"]
LIB_NAME = "libclbm_amd.so"

# Packed-FP32 instructions are kept out of the device code on purpose, and
# float32 denormals are kept (the f16c decode scales by 2^-128, a denormal).
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-fno-gpu-flush-denormals-to-zero"]


def hipcc() -> str:
    root = os.environ.get("ROCM_PATH", "/opt/rocm")
    path = os.path.join(root, "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


def build(arch: str = "gfx950", verbose: bool = False, jobs: int = 0) -> str:
    """Compile every source to an object (in parallel, ``jobs`` at a time; 0 =
    one per source up to the usable cores), then link the shared library."""
    out = os.path.join(HERE, LIB_NAME)
    tmp = out + ".tmp"
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    common = [hipcc(), f"--offload-arch={arch}", "-O3", "-std=c++17", "-fPIC", "-Wall",
              "-Wno-unused-function", *DEVICE_FLAGS]
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in SOURCES]
    cmds = [common + ["-c", os.path.join(HERE, s), "-o", o] for s, o in zip(SOURCES, objs)]
    if verbose:
        for c in cmds:
            print(" ".join(c), flush=True)
    n = jobs or max(1, min(len(cmds), len(os.sched_getaffinity(0))))
    procs, failed = [], []
    pending = list(cmds)
    while pending or procs:
        while pending and len(procs) < n:
            c = pending.pop(0)
            procs.append((c, subprocess.Popen(c, cwd=HERE)))
        c, p = procs.pop(0)
        if p.wait() != 0:
            failed.append(c)
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    link = [hipcc(), f"--offload-arch={arch}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True, cwd=HERE)
    os.replace(tmp, out)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    print(build(a.arch, a.verbose))
    return 0


if __name__ == "__main__":
    sys.exit(main())
