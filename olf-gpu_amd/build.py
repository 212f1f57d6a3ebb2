"""Builds libolfhip.so (gfx950) in-tree with hipcc: `python olf-gpu_amd/build.py`; and, where <jni.h> is found
($JAVA_HOME/include or /usr/lib/jvm/*/include), the JNI shim libolfgpu.so over it."""
import glob
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
         # the machine scheduler's instruction-level-parallelism strategy: the Newton-step launch of the 2,000-
         # contingency bench ran 28.92 / 28.93 ms against 29.27 / 29.23 ms with the default (the same spills)
         "-mllvm", "-amdgpu-sched-strategy=max-ilp",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def check_generated():
    """The generated term headers must be the java2hip rendering of the committed manifests (no hand edits, no
    stale rendering)."""
    gen = os.path.join(HERE, "codegen", "java2hip.py")
    if subprocess.run([sys.executable, gen, "--check"]).returncode != 0:
        raise RuntimeError("csrc/generated is out of date: run python olf-gpu_amd/codegen/java2hip.py")


def build(verbose=True):
    check_generated()
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
    build_jni(verbose)
    return LIB


def jni_include():
    """The directory holding jni.h, or None."""
    cands = []
    if os.environ.get("JAVA_HOME"):
        cands.append(os.path.join(os.environ["JAVA_HOME"], "include"))
    cands += sorted(glob.glob("/usr/lib/jvm/*/include"))
    for c in cands:
        if os.path.exists(os.path.join(c, "jni.h")):
            return c
    return None


def build_jni(verbose=True):
    """libolfgpu.so: the JNI binding of StructureNative / BatchSolveNative over libolfhip.so (jni/olfhip_jni.c),
    found beside it through its rpath. Skipped (None) without jni.h."""
    inc = jni_include()
    if inc is None:
        if verbose:
            print("jni.h not found: the JNI shim libolfgpu.so is not built", flush=True)
        return None
    out = os.path.join(HERE, "libolfgpu.so")
    cmd = ["gcc", "-std=c99", "-O2", "-fPIC", "-shared", "-Wall", "-I" + inc, "-I" + os.path.join(inc, "linux"),
           "-I" + os.path.join(os.path.dirname(HERE), "include"), os.path.join(HERE, "jni", "olfhip_jni.c"),
           "-L" + HERE, "-lolfhip", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
