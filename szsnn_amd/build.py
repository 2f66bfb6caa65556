"""Build recipe for the gfx950 HIP library (szsnn_amd/libszsnn_hip.so).

Run as `python -m szsnn_amd.build` or through __graft_entry__.build().  The
library is built in-tree so it travels with the source tree.

Compiler settings that the numerics depend on:
  -ffp-contract=off                        no FMA contraction: the written
                                           operation order is the evaluated one
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE division
  -fno-gpu-flush-denormals-to-zero         IEEE denormals, like the host
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                           no packed-FP32 instructions
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libszsnn_hip.so")
SOURCES = ["szsnn_kernels.hip", "szsnn_runtime.hip", "szsnn_adjoint.hip"]
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build szsnn_amd)")


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIP_FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"hipcc failed with code {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
