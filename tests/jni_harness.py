"""The JNI shim (olf-gpu_amd/jni/olfhip_jni.c) driven without a JVM — TEST INFRASTRUCTURE ONLY: the shim is compiled
with a test double of <jni.h> (tests/native/jni_test/jni.h) and a fake JVM (fake_jvm.c) into libfakejni.so, linked
against the in-tree libolfhip.so; Python builds the Java arguments (arrays as the fake JVM's objects) and calls the
natives with the argument lists of StructureNative.java and BatchSolveNative.java."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SRC = os.path.join(HERE, "native", "jni_test")
LIB = os.path.join(SRC, "libfakejni.so")
SHIM = os.path.join(REPO, "olf-gpu_amd", "jni", "olfhip_jni.c")
OLFHIP = os.path.join(REPO, "olf-gpu_amd", "libolfhip.so")

STRUCT = "Java_com_powsybl_openloadflow_ac_equations_vector_gpu_StructureNative_"
BATCH = "Java_com_powsybl_openloadflow_ac_solver_gpu_BatchSolveNative_"

# the reference's GpuLinearSolveMode codes the shim maps
MODE_UBATCH, MODE_ROBUST, MODE_BATCHED_LU, MODE_BATCHED_LU_FP32_REFINED = 3, 6, 7, 8


def build(force=False):
    """Compiles libfakejni.so (gcc) when it is missing or older than its sources or libolfhip.so."""
    srcs = [os.path.join(SRC, "fake_jvm.c"), os.path.join(SRC, "jni.h"), SHIM, OLFHIP]
    if not force and os.path.exists(LIB) and all(os.path.getmtime(LIB) >= os.path.getmtime(s) for s in srcs):
        return LIB
    cmd = ["gcc", "-std=c99", "-O1", "-fPIC", "-shared", "-Wall", "-Werror", "-I" + SRC,
           "-I" + os.path.join(REPO, "include"), os.path.join(SRC, "fake_jvm.c"), SHIM,
           "-L" + os.path.dirname(OLFHIP), "-lolfhip", "-Wl,-rpath,$ORIGIN/" + os.path.relpath(os.path.dirname(OLFHIP), SRC),
           "-o", LIB + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


class Jvm:
    """The fake JVM and the shim's natives."""

    def __init__(self):
        self.lib = C.CDLL(build())
        L = self.lib
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        L.fj_env.restype = vp
        for f in ("fj_doubles", "fj_ints", "fj_booleans", "fj_bytes"):
            getattr(L, f).restype = vp
            getattr(L, f).argtypes = [i, vp]
        L.fj_objects.restype = vp
        L.fj_objects.argtypes = [i]
        L.fj_set.argtypes = [vp, i, vp]
        L.fj_get.restype = vp
        L.fj_get.argtypes = [vp, i]
        L.fj_length.argtypes = [vp]
        L.fj_kind.argtypes = [vp]
        L.fj_read_doubles.argtypes = [vp, vp]
        L.fj_read_ints.argtypes = [vp, vp]
        L.fj_pinned.argtypes = [vp]
        L.fj_exception.argtypes = [C.c_char_p, i, C.c_char_p, i]
        self.env = L.fj_env()
        j = C.c_int64
        getattr(L, STRUCT + "createStructure").restype = j
        getattr(L, STRUCT + "createStructure").argtypes = [vp, vp, i] + [vp] * 32
        getattr(L, STRUCT + "destroyStructure").argtypes = [vp, vp, j]
        getattr(L, STRUCT + "recommendBatchSize").restype = i
        getattr(L, STRUCT + "recommendBatchSize").argtypes = [vp, vp, j, i, i]
        getattr(L, BATCH + "createBatchContext").restype = j
        getattr(L, BATCH + "createBatchContext").argtypes = ([vp, vp, j, i, i] + [vp] * 9 + [i, d] + [vp] * 3 + [i] +
                                                             [vp] * 2 + [d, d])
        getattr(L, BATCH + "destroyBatchContext").argtypes = [vp, vp, j]
        getattr(L, BATCH + "solveBatchContext").restype = vp
        getattr(L, BATCH + "solveBatchContext").argtypes = ([vp, vp, j] + [vp] * 20 + [i, d, d, d, d, i, d, d] +
                                                            [vp] * 7 + [i, i, i, i, i, vp, vp, i])
        getattr(L, BATCH + "beginBatchInstall").argtypes = [vp, vp, j, vp]
        getattr(L, BATCH + "readBatchHandedBackSlots").restype = vp
        getattr(L, BATCH + "readBatchHandedBackSlots").argtypes = [vp, vp, j]

    # ---- Java values
    def doubles(self, a):
        a = np.ascontiguousarray(a, np.float64).reshape(-1)
        return self.lib.fj_doubles(a.size, a.ctypes.data)

    def ints(self, a):
        a = np.ascontiguousarray(a, np.int32).reshape(-1)
        return self.lib.fj_ints(a.size, a.ctypes.data)

    def bytes(self, a):
        a = np.ascontiguousarray(a, np.int8).reshape(-1)
        return self.lib.fj_bytes(a.size, a.ctypes.data)

    def boolean_rows(self, mask):
        """boolean[][] from a [rows][cols] bool array"""
        mask = np.asarray(mask, bool)
        arr = self.lib.fj_objects(mask.shape[0])
        for s, row in enumerate(mask):
            r = np.ascontiguousarray(row, np.uint8)
            self.lib.fj_set(arr, s, self.lib.fj_booleans(r.size, r.ctypes.data))
        return arr

    def read_doubles(self, obj):
        out = np.zeros(max(self.lib.fj_length(obj), 0))
        self.lib.fj_read_doubles(obj, out.ctypes.data)
        return out

    def read_ints(self, obj):
        out = np.zeros(max(self.lib.fj_length(obj), 0), np.int32)
        self.lib.fj_read_ints(obj, out.ctypes.data)
        return out

    def read_rows(self, obj):
        """double[][] -> [rows][cols]"""
        if not obj:
            return None
        return np.array([self.read_doubles(self.lib.fj_get(obj, s)) for s in range(self.lib.fj_length(obj))])

    def exception(self):
        """(class, message) of the pending exception, cleared; None without one."""
        cls, msg = C.create_string_buffer(128), C.create_string_buffer(512)
        if not self.lib.fj_exception(cls, 128, msg, 512):
            return None
        self.lib.fj_clear_exception()
        return cls.value.decode(), msg.value.decode()

    def violations(self):
        return self.lib.fj_violations()

    # ---- the natives
    def create_structure(self, n, p8, idx12, shunt_p=None, shunt_idx=None, **families):
        """StructureNative.createStructure with the closed-branch and shunt packs; `families` names other arrays of
        the reference's signature to pass non-empty (open_branch_indices, phase_control_side, ...)."""
        args = [None] * 32
        args[0] = self.doubles(p8)
        args[1] = self.ints(idx12)
        if shunt_p is not None:
            args[4] = self.doubles(shunt_p)
            args[5] = self.ints(shunt_idx)
        slot = {"open_branch_indices": 3, "hvdc_indices": 7, "load_model_indices": 23, "identity_rows": 31,
                "phase_control_side": 21}
        for k, v in families.items():
            args[slot[k]] = self.ints(v)
        return getattr(self.lib, STRUCT + "createStructure")(self.env, None, n, *args)

    def destroy_structure(self, h):
        getattr(self.lib, STRUCT + "destroyStructure")(self.env, None, h)

    def recommend_batch_size(self, h, max_cap, mode):
        return getattr(self.lib, STRUCT + "recommendBatchSize")(self.env, None, h, max_cap, mode)

    def create_batch_context(self, sh, n_scenarios, mode, rl, flags=None):
        args = [self.ints(rl["voltage_row"]), self.doubles(rl["min_q"]), self.doubles(rl["max_q"]),
                self.doubles(rl["load_q"]), self.doubles(rl["target_v"]),
                None if flags is None else self.bytes(flags), None, None, None]
        return getattr(self.lib, BATCH + "createBatchContext")(self.env, None, sh, n_scenarios, mode, *args, -1, 0.0,
                                                              None, None, None, 0, None, None, 0.0, 0.0)

    def destroy_batch_context(self, h):
        getattr(self.lib, BATCH + "destroyBatchContext")(self.env, None, h)

    def solve(self, ch, x0_base, target_base, base_row_mode, mask, row_mode_entries=None, target_override=None,
              seed_state=None, x0_per_scenario=None, max_iter=15, tol=1e-4, rl_tol=1e-4, max_outer=20,
              max_switch=3, pause_at_ended=0):
        """BatchSolveNative.solveBatchContext -> double[handed back][n + 3] as [rows][n + 3] (None on a throw)."""
        ti = tv = None
        if target_override is not None and len(target_override[0]):
            ti, tv = self.ints(target_override[0]), self.doubles(target_override[1])
        rme = None if row_mode_entries is None or len(row_mode_entries) == 0 else self.ints(row_mode_entries)
        a20 = [self.doubles(x0_base), self.doubles(target_base), rme, self.ints(base_row_mode),
               self.boolean_rows(mask), ti, tv] + [None] * 13
        tail = [None, None if x0_per_scenario is None else self.doubles(x0_per_scenario), None, None,
                None if seed_state is None else self.ints(seed_state), None, None]
        rows = getattr(self.lib, BATCH + "solveBatchContext")(
            self.env, None, ch, *a20, max_iter, tol, rl_tol, 0.0, 0.0, 0, 0.1, 0.174533, *tail,
            max_outer, max_switch, 0, 0, 0, None, None, pause_at_ended)
        return self.read_rows(rows)

    def begin_install(self, ch, slots):
        getattr(self.lib, BATCH + "beginBatchInstall")(self.env, None, ch, self.ints(slots))

    def handed_back_slots(self, ch):
        return self.read_ints(getattr(self.lib, BATCH + "readBatchHandedBackSlots")(self.env, None, ch))
