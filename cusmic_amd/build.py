"""Build libcusmic.so (HIP, gfx950) in-tree.

The library is a single translation unit, csrc/api.hip, which includes the
kernels. Flags:
  -ffp-contract=off   no fused multiply-add contraction, as the reference build
                      (Makefile:92,102-104), so every product is rounded first
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keep the compiler from emitting packed FP32 instructions
Run: python -m cusmic_amd.build
"""

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "libcusmic.so"
ARCH = os.environ.get("CUSMIC_ARCH", "gfx950")


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return str(Path(rocm) / "bin" / "hipcc")


def command(out=LIB, extra=()):
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off", "-fno-fast-math",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wextra", "-Wno-unused-parameter",
        "-o", str(out), str(HERE / "csrc" / "api.hip"), *extra,
    ]


def build(verbose=False):
    tmp = LIB.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
