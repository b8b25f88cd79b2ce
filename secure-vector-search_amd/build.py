"""Builds the gfx950 backend library in-tree: svs/libsvs_gpu.so.

hipcc cross-compiles for gfx950 without a GPU. Packed-FP32 instruction
selection is disabled for all device code (no kernel needs it), and
floating-point contraction is off, so the host and device routing sum
squared differences in the reference's order.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "svs", "libsvs_gpu.so")
SOURCES = ["svs_api.hip", "svs_schemes.hip", "svs_bfv.hip", "svs_host.cpp", "svs_server.cpp",
           "svs_cache.cpp"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, *HIPCC_FLAGS, *[os.path.join(CSRC, s) for s in SOURCES], "-o", OUT + ".tmp",
           "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
