"""Build recipe for the in-tree HIP library ``treetime_amd/libttgpu.so``.

``python -m treetime_amd.build`` compiles every HIP source under ``csrc/``
for gfx950 with hipcc into one shared library next to this file.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libttgpu.so")
SOURCES = ["tt_kernels.hip", "tt_api.hip", "tt_batch.hip", "tt_mfma.hip", "tt_branch.hip", "tt_compress.hip", "tt_site.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TT_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose=False, extra_flags=()):
    """Compile the library; raises CalledProcessError on failure."""
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, *extra_flags, "-I", os.path.join(ROOT, "include"), *sources(), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra_flags=sys.argv[1:])
    print("built", LIB)
