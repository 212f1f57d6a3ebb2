"""The oracle on a network that needs pivoting (CPU). gridgen.Network.with_series_resonance puts a load-free bus i
between reactances x and -x beside a branch a-b. Bus i's own 2x2 Jacobian block is then singular at the flat start
while the Jacobian is not. An elimination without row pivoting that takes bus i first meets a zero pivot. The
oracle's sparse LU (threshold partial pivoting, oracle/oracle.c) must pivot past it.

Pins: the oracle's solve against LAPACK's dense partial-pivoting solve on that Jacobian; the converged state against
what the circuit fixes in closed form (a zero series impedance ties a and b to the same complex voltage, and the
current through the path is the one that makes bus i's injections zero).
"""
import numpy as np
import scipy.sparse as sp

import pyoracle as po
from olfgpu import gridgen

PAIRS_BESIDE = [0, 40]


def resonant(grid):
    base = grid("grid118")
    return base, base.with_series_resonance(PAIRS_BESIDE, x=0.05)


def test_series_resonance_network_shape(grid):
    base, net = resonant(grid)
    m = len(PAIRS_BESIDE)
    assert net.n_bus == base.n_bus + m and net.n_branch == base.n_branch + 2 * m
    new = base.n_bus + np.arange(m)
    assert (net.bus_type[new] == gridgen.PQ).all()
    assert (net.pload[new] == 0).all() and (net.qload[new] == 0).all() and (net.pgen[new] == 0).all()
    e = base.n_branch
    assert np.array_equal(net.bus1[e:e + m], base.bus1[PAIRS_BESIDE]) and np.array_equal(net.bus2[e:e + m], new)
    assert np.array_equal(net.bus1[e + m:], new) and np.array_equal(net.bus2[e + m:], base.bus2[PAIRS_BESIDE])
    assert np.allclose(net.x[e:e + m], 0.05) and np.allclose(net.x[e + m:], -0.05)
    assert (net.r[e:] == 0).all() and (net.b[e:] == 0).all()


def test_zero_pivot_block_and_pivoted_solve_matches_lapack(grid):
    _, net = resonant(grid)
    osys = po.OracleSystem(net)
    x0, tgt, md, _ = net.system()
    F, _, J = osys.fill(x0, md, tgt)
    A = sp.csr_matrix((J, osys.colidx, osys.rowptr), shape=(net.n, net.n)).toarray()
    for i in net.n_bus - np.arange(1, len(PAIRS_BESIDE) + 1):
        blk = A[2 * i:2 * i + 2, 2 * i:2 * i + 2]
        assert abs(np.linalg.det(blk)) < 1e-26, blk      # what an unpivoted block elimination of bus i meets
        assert np.abs(A[2 * i:2 * i + 2]).max() > 10.0    # while its rows are far from zero
    assert np.linalg.cond(A) < 1e6
    x = po.lu_solve(osys.rowptr, osys.colidx, J, -F)
    ref = np.linalg.solve(A, -F)
    assert np.abs(x - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(A @ x + F).max() <= 1e-12 * np.abs(F).max()


def test_converged_state_ties_the_resonant_path(grid):
    base, net = resonant(grid)
    osys = po.OracleSystem(net)
    x0, tgt, md, _ = net.system()
    x, _, _, _, res = osys.solve(x0, md, tgt)
    assert po.STATUS[res.status] == "CONVERGED"
    v = x[1::2] * np.exp(1j * x[0::2])
    for p, e in enumerate(PAIRS_BESIDE):
        a, b, i = base.bus1[e], base.bus2[e], base.n_bus + p
        # x + (-x) in series is a short: a and b at one complex voltage, up to the stopping tolerance's effect
        assert abs(v[a] - v[b]) < 1e-6, (a, b, v[a], v[b])
        # bus i draws nothing: the current in from a equals the current out to b, I = (Va - Vi) / (j x)
        ia = (v[a] - v[i]) / (1j * 0.05)
        ib = (v[i] - v[b]) / (-1j * 0.05)
        assert abs(ia - ib) < 1e-4 * max(1.0, abs(ia))
