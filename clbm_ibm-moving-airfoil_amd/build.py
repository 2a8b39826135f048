"""Build libclbm_amd.so for gfx950 with hipcc (in-tree, next to this file).

Usage: python build.py [--arch gfx950] [--verbose]
"""

from __future__ import annotations

import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["csrc/clbm_lbm.hip", "csrc/clbm_ibm.hip", "csrc/clbm_closure.hip"]
LIB_NAME = "libclbm_amd.so"

# Packed-FP32 instructions are kept out of the device code on purpose, and
# float32 denormals are kept (the f16c decode scales by 2^-128, a denormal).
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-fno-gpu-flush-denormals-to-zero"]


def hipcc() -> str:
    root = os.environ.get("ROCM_PATH", "/opt/rocm")
    path = os.path.join(root, "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


def build(arch: str = "gfx950", verbose: bool = False) -> str:
    out = os.path.join(HERE, LIB_NAME)
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={arch}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Wall", "-Wno-unused-function", *DEVICE_FLAGS,
           *[os.path.join(HERE, s) for s in SOURCES], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=HERE)
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
