"""Build the HIP library ``neuwon_amd/libneuwon_amd.so`` for gfx950 (in-tree, no JIT cache).

Usage: ``python -m neuwon_amd.build`` (also called by ``__graft_entry__.build()``).
hipcc cross-compiles without a GPU present.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libneuwon_amd.so")
SOURCES = ["nw_propagator.hip", "nw_kernels.hip", "nw_clock.hip", "nw_kinetic.hip"]
ARCH = os.environ.get("NEUWON_AMD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay disabled in every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-Wno-unused-command-line-argument",
]
LIBS = ["-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def build(verbose=True):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp] + LIBS
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libneuwon_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
