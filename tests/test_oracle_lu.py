"""Pins of the oracle's sparse LU: dense partial-pivot solves (numpy) on Jacobian-like systems, forced
off-diagonal pivots, singular detection, and the minimum-degree order (a permutation, fill no worse than the
natural order)."""
import numpy as np
import pytest
import scipy.sparse as sp

import pyoracle as po


def _jacobian_like(osys, rng, scale=0.1):
    n = osys.n
    x = np.zeros(n)
    x[1::2] = 1 + rng.normal(0, 0.03, n // 2)
    x[0::2] = rng.normal(0, scale, n // 2)
    _, tgt, mode, _ = osys.net.system()   # slack angle row fixed: without it the Jacobian is singular
    _, _, J = osys.fill(x, mode, tgt)
    return J


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_lu_matches_dense_solve_on_jacobian(grid, seed):
    net = grid("grid118")
    osys = po.OracleSystem(net)
    rng = np.random.default_rng(seed)
    J = _jacobian_like(osys, rng)
    b = rng.normal(size=osys.n)
    x = po.lu_solve(osys.rowptr, osys.colidx, J, b)
    A = sp.csr_matrix((J, osys.colidx, osys.rowptr), shape=(osys.n, osys.n)).toarray()
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-10)


def test_lu_pivots_off_diagonal_when_the_diagonal_vanishes():
    rng = np.random.default_rng(5)
    n = 40
    A = rng.normal(size=(n, n)) * (rng.random((n, n)) < 0.15) + np.diag(rng.uniform(1, 2, n))
    A[3, 3] = 0.0
    A[3, 7] = 2.5
    A[7, 3] = 1.5
    A[10, 10] = 1e-9                       # far below the 1e-3 threshold of its column
    A[12, 10] = 3.0
    S = sp.csr_matrix(A)
    b = rng.normal(size=n)
    x = po.lu_solve(S.indptr, S.indices, S.data, b)
    np.testing.assert_allclose(A @ x, b, atol=1e-9)


def test_lu_reports_singular():
    A = sp.csr_matrix(np.array([[1.0, 2.0, 0], [2.0, 4.0, 0], [0, 0, 1.0]]))
    with pytest.raises(np.linalg.LinAlgError):
        po.lu_solve(A.indptr, A.indices, A.data, np.ones(3))


def test_min_degree_is_a_permutation_with_low_fill(grid):
    net = grid("grid2000")
    osys = po.OracleSystem(net)
    n = net.n_bus
    adj = sp.coo_matrix((np.ones(net.n_branch), (net.bus1, net.bus2)), shape=(n, n))
    adj = ((adj + adj.T) > 0).tocsr()
    adj.setdiag(0)
    adj.eliminate_zeros()
    perm = po.min_degree(adj.indptr.astype(np.int32), adj.indices.astype(np.int32))
    assert sorted(perm.tolist()) == list(range(n))

    def fill(order):
        # symbolic Cholesky fill by explicit elimination (boolean Gaussian elimination)
        pos = np.empty(n, int)
        pos[order] = np.arange(n)
        nb = [set(pos[adj.indices[adj.indptr[i]:adj.indptr[i + 1]]]) for i in order]
        total = 0
        for k in range(n):
            higher = {j for j in nb[k] if j > k}
            total += len(higher)
            for j in higher:
                nb[j] |= higher - {j}
        return total
    assert fill(perm) <= fill(np.arange(n))
    assert fill(perm) < 3 * (adj.nnz // 2)
