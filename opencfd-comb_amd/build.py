"""Build recipe for the HIP library (gfx950).

Compiles csrc/*.hip (one object each, in parallel, under _build/) and links
libocfd_hip.so next to this file.  FP64 kernels are
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
    """Compile each source to an object in parallel, then link the library."""
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    jobs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        jobs.append(([hipcc] + compile_flags + ["-c", src, "-o", obj], obj))

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        for f in [ex.submit(run, c) for c, _ in jobs]:
            f.result()
    run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"] + [o for _, o in jobs] + ["-o", LIB + ".tmp"])
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
