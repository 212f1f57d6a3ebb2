"""Builds libolfhip.so (gfx950) in-tree with hipcc: `python olf-gpu_amd/build.py`."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libolfhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["api.hip", "plan.cpp", "mlp.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # no packed FP32 instructions in device code (some of their forms gave wrong results on these boxes)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def build(verbose=True):
    cmd = [HIPCC] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    # the HBM-counter calibration program (a plain copy of known size)
    probe = os.path.join(HERE, "hbm_copy")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", os.path.join(CSRC, "hbm_copy.hip"), "-o", probe + ".tmp"], check=True)
    os.replace(probe + ".tmp", probe)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
