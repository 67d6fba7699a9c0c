"""Build recipe for the HIP library (gfx950).

Compiles csrc/*.hip into libocfd_hip.so next to this file.  FP64 kernels are
built with -ffp-contract=off so every operation rounds as the reference
Fortran does, and with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libocfd_hip.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("ocfd_flow.hip", "ocfd_chem.hip", "ocfd_mlp.hip")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def build(verbose: bool = True) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + HIPCC_FLAGS + SOURCES + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
