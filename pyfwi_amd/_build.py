"""Build the native libraries in-tree.

``build_native()`` compiles the HIP kernels and the C-ABI layer for gfx950
into ``pyfwi_amd/libpyfwi_hip.so`` with hipcc (cross-compiles without a GPU).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libpyfwi_hip.so")
SOURCES = [os.path.join(HERE, "csrc", n) for n in ("pyfwi_kernels.hip", "pyfwi_plan.hip")]
ARCH = os.environ.get("PYFWI_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # packed-FP32 instructions stay off: the kernels are written for scalar FP32
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build_native(verbose=False):
    """Compile libpyfwi_hip.so; raises CalledProcessError on failure."""
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), *SOURCES,
           "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build_native(verbose=True))
