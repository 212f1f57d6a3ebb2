"""Pins of the oracle's branch formulas: the reference's own formulas (golden vectors and, when built, the
oracle/_ref library), central finite differences, and the textbook complex-power form S = V conj(I)."""
import ctypes as C
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "closed_branch_vectors.npz"))


def test_oracle_matches_reference_golden_vectors():
    p8, st, out = GOLD["p8"], GOLD["state"], GOLD["out"]
    for i in range(p8.shape[0]):
        o = po.closed_branch(p8[i], *st[i])
        np.testing.assert_allclose(o, out[i], rtol=1e-13, atol=1e-13)


def test_oracle_matches_reference_library_when_built():
    path = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_formulas.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    lib = C.CDLL(path)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    lib.ref_closed_branch.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_double, dp]
    rng = np.random.default_rng(7)
    for _ in range(200):
        p8 = np.concatenate([rng.uniform(1, 100, 1), rng.uniform(0, .02, 4), rng.uniform(0, 0.4, 1),
                             rng.uniform(0.9, 1.1, 1), rng.uniform(-0.3, 0.3, 1)])
        s = (rng.uniform(0.8, 1.2), rng.uniform(-2, 2), rng.uniform(0.8, 1.2), rng.uniform(-2, 2))
        ref = np.zeros(28)
        lib.ref_closed_branch(np.ascontiguousarray(p8), *s, ref)
        np.testing.assert_allclose(po.closed_branch(p8, *s), ref, rtol=1e-13, atol=1e-12)


def _values(p8, v1, ph1, v2, ph2, a1=None, r1=None):
    p = p8.copy()
    if a1 is not None:
        p[7] = a1
    if r1 is not None:
        p[6] = r1
    o = po.closed_branch(p, v1, ph1, v2, ph2)
    return np.array([o[0], o[7], o[14], o[21]])


def test_derivatives_by_central_differences():
    # outputs per function: value index, then d/dv1, d/dv2, d/dph1, d/dph2, d/da1, d/dr1
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(50):
        p8 = np.concatenate([rng.uniform(5, 50, 1), rng.uniform(0, .02, 4), rng.uniform(0, 0.3, 1),
                             rng.uniform(0.9, 1.1, 1), rng.uniform(-0.2, 0.2, 1)])
        v1, ph1, v2, ph2 = rng.uniform(0.9, 1.1), rng.uniform(-1, 1), rng.uniform(0.9, 1.1), rng.uniform(-1, 1)
        o = po.closed_branch(p8, v1, ph1, v2, ph2)
        fd = {
            "v1": (_values(p8, v1 + h, ph1, v2, ph2) - _values(p8, v1 - h, ph1, v2, ph2)) / (2 * h),
            "v2": (_values(p8, v1, ph1, v2 + h, ph2) - _values(p8, v1, ph1, v2 - h, ph2)) / (2 * h),
            "ph1": (_values(p8, v1, ph1 + h, v2, ph2) - _values(p8, v1, ph1 - h, v2, ph2)) / (2 * h),
            "ph2": (_values(p8, v1, ph1, v2, ph2 + h) - _values(p8, v1, ph1, v2, ph2 - h)) / (2 * h),
            "a1": (_values(p8, v1, ph1, v2, ph2, a1=p8[7] + h) - _values(p8, v1, ph1, v2, ph2, a1=p8[7] - h)) / (2 * h),
            "r1": (_values(p8, v1, ph1, v2, ph2, r1=p8[6] + h) - _values(p8, v1, ph1, v2, ph2, r1=p8[6] - h)) / (2 * h),
        }
        for f in range(4):
            base = 7 * f
            for k, var in enumerate(["v1", "v2", "ph1", "ph2", "a1", "r1"]):
                assert abs(o[base + 1 + k] - fd[var][f]) < 1e-6 * max(1.0, abs(fd[var][f])), (f, var)


def test_values_equal_textbook_complex_power():
    """p1 + j q1 = V1 conj(I1) for a pi-branch with series admittance y e^{-j(pi/2 - ksi)}, shunt g + jb per side,
    ideal ratio r1 and phase a1 on side 1 (closed form, independent of the OLF expressions)."""
    rng = np.random.default_rng(11)
    for _ in range(100):
        r, x = rng.uniform(0.001, 0.05), rng.uniform(0.01, 0.2)
        g1, g2, b1, b2 = rng.uniform(0, 0.02, 4)
        rho, alpha = rng.uniform(0.9, 1.1), rng.uniform(-0.2, 0.2)
        v1, ph1, v2, ph2 = rng.uniform(0.9, 1.1), rng.uniform(-1, 1), rng.uniform(0.9, 1.1), rng.uniform(-1, 1)
        ys = 1 / complex(r, x)
        p8 = np.array([abs(ys), g1, g2, b1, b2, np.arctan2(r, x), rho, alpha])
        V1 = v1 * np.exp(1j * ph1)
        V2 = v2 * np.exp(1j * ph2)
        k = rho * np.exp(1j * alpha)          # ideal transformer on side 1: V1' = k V1
        V1p = k * V1
        I1p = (V1p - V2) * ys + V1p * complex(g1, b1)
        I2 = (V2 - V1p) * ys + V2 * complex(g2, b2)
        S1 = V1p * np.conj(I1p)               # lossless ideal transformer: same power on both of its sides
        S2 = V2 * np.conj(I2)
        o = po.closed_branch(p8, v1, ph1, v2, ph2)
        np.testing.assert_allclose([o[0], o[7], o[14], o[21]], [S1.real, S1.imag, S2.real, S2.imag], rtol=1e-11,
                                   atol=1e-11)
