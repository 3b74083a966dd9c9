"""Build recipe for librtg.so (HIP, gfx950 only).

The library is built in-tree next to this file so it travels with the repository.
Every HIP source is compiled with FMA contraction off (FP32 results follow the
reference's operation order) and with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librtg.so")

SOURCES = ["rtg_kernels.hip", "rtg_guide.hip", "rtg_bvh.cpp", "rtg_host.cpp", "rtg_guide_host.cpp", "rtg_serve.cpp", "rtg_xml.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-Wl,--no-undefined",
]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES]


def build(verbose=False):
    """Compile librtg.so for gfx950 (cross-compiles without a GPU)."""
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + sources() + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("librtg build failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
