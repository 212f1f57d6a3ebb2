"""Writes tests/golden/closed_branch_vectors.npz: inputs and the 28 outputs of the reference's own
closed-branch formulas (native/cuda/equations/generated/closed_branch.cuh, updateClosedBranch), evaluated through
oracle/_ref/libref_formulas.so, which oracle/Makefile (`make ref`) builds from the reference tree.

Run in the build container (where the reference tree exists):  python tests/golden/make_closed_branch_vectors.py
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def main():
    lib = C.CDLL(os.path.join(REPO, "oracle", "_ref", "libref_formulas.so"))
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    lib.ref_closed_branch.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_double, dp]
    rng = np.random.default_rng(20260101)
    k = 256
    p8 = np.zeros((k, 8))
    xr = rng.uniform(0.01, 0.2, k)
    rr = xr / rng.uniform(3, 10, k)
    p8[:, 0] = 1 / np.hypot(rr, xr)
    p8[:, 1] = rng.uniform(0, 0.01, k)      # g1
    p8[:, 2] = rng.uniform(0, 0.01, k)      # g2
    p8[:, 3] = rng.uniform(0, 0.03, k)      # b1
    p8[:, 4] = rng.uniform(0, 0.03, k)      # b2
    p8[:, 5] = np.arctan2(rr, xr)           # ksi
    p8[:, 6] = rng.uniform(0.9, 1.1, k)     # r1
    p8[:, 7] = rng.uniform(-0.2, 0.2, k)    # a1
    state = np.column_stack([rng.uniform(0.85, 1.15, k), rng.uniform(-1, 1, k),
                             rng.uniform(0.85, 1.15, k), rng.uniform(-1, 1, k)])  # v1 ph1 v2 ph2
    out = np.zeros((k, 28))
    for i in range(k):
        o = np.zeros(28)
        lib.ref_closed_branch(np.ascontiguousarray(p8[i]), state[i, 0], state[i, 1], state[i, 2], state[i, 3], o)
        out[i] = o
    np.savez(os.path.join(HERE, "closed_branch_vectors.npz"), p8=p8, state=state, out=out)
    print("wrote", k, "vectors")


if __name__ == "__main__":
    main()
