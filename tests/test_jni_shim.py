"""The JNI shim (olf-gpu_amd/jni/olfhip_jni.c): it binds exactly the reference's hot-path natives under their JNI
names (packages and classes of StructureNative.java and BatchSolveNative.java), calls only entry points that
include/olfhip.h declares, and builds against <jni.h> when a JDK is present (this image has none: then the build
check is skipped)."""
import importlib.util
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "olf-gpu_amd", "jni", "olfhip_jni.c")

EXPECTED = {
    "Java_com_powsybl_openloadflow_ac_equations_vector_gpu_StructureNative_createStructure",
    "Java_com_powsybl_openloadflow_ac_equations_vector_gpu_StructureNative_destroyStructure",
    "Java_com_powsybl_openloadflow_ac_equations_vector_gpu_StructureNative_recommendBatchSize",
    "Java_com_powsybl_openloadflow_ac_solver_gpu_BatchSolveNative_createBatchContext",
    "Java_com_powsybl_openloadflow_ac_solver_gpu_BatchSolveNative_destroyBatchContext",
    "Java_com_powsybl_openloadflow_ac_solver_gpu_BatchSolveNative_solveBatchContext",
    "Java_com_powsybl_openloadflow_ac_solver_gpu_BatchSolveNative_beginBatchInstall",
    "Java_com_powsybl_openloadflow_ac_solver_gpu_BatchSolveNative_readBatchHandedBackSlots",
}


def test_shim_binds_the_hot_path_natives():
    src = open(SRC).read()
    assert set(re.findall(r"JNICALL\s+(Java_\w+)\s*\(", src)) == EXPECTED


def test_shim_calls_only_declared_entry_points():
    src = open(SRC).read()
    header = open(os.path.join(REPO, "include", "olfhip.h")).read()
    declared = set(re.findall(r"\b(olfhip_[a-z_0-9]+)\s*\(", header))
    called = set(re.findall(r"\b(olfhip_[a-z_0-9]+)\s*\(", src))
    assert called and called <= declared, called - declared


def test_shim_builds_where_jni_is_available():
    spec = importlib.util.spec_from_file_location("olfhip_build", os.path.join(REPO, "olf-gpu_amd", "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    if b.jni_include() is None:
        pytest.skip("no jni.h in this image (no JDK): the shim is built where one is installed")
    out = b.build_jni(verbose=False)
    assert out and os.path.exists(out)
