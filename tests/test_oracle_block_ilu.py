"""The oracle's block ILU(0) for variable-size blocks (lar_fill_preconditioner
and lar_perform_rhsub for block storage, lad_block/las_block_intf.c:1409-2143
and 3085-3476), pinned by what the mathematics fixes: the LU and inverse of a
block against numpy; ILU(0) factors reproduce the matrix on its pattern (which
fails for a dropped or misordered update); on a block chain ILU(0) is the exact
LU, so its solve is the direct solve; and the solve inverts the dense L U."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from oracle_lib import BlockSystem, P, oracle  # noqa: E402


@pytest.mark.parametrize("n", [1, 2, 4, 6, 18, 40])
def test_lun_invn_invert(n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + n * np.eye(n) * 0.1
    if n > 1:
        A[0, 0] = 0.0  # forces a row interchange
    a = np.ascontiguousarray(A.T.ravel()).copy()
    piv = np.zeros(n, np.int32)
    assert oracle().orc_lun(n, P(a), P(piv)) == 0
    if n > 1:
        assert piv[0] != 0
    oracle().orc_invn(n, P(a), P(piv))
    inv = a.reshape(n, n).T
    np.testing.assert_allclose(inv @ A, np.eye(n), atol=1e-11 * np.linalg.cond(A))
    np.testing.assert_allclose(inv, np.linalg.inv(A), rtol=0, atol=1e-12 * np.abs(np.linalg.inv(A)).max() * np.linalg.cond(A))


def test_lun_reports_a_singular_block():
    a = np.zeros(9)
    piv = np.zeros(3, np.int32)
    assert oracle().orc_lun(3, P(a), P(piv)) == 1


def random_system(ns, edges, seed, shift=4.0):
    """a structurally symmetric block matrix on the given undirected edges,
    diagonally dominant by shift"""
    rng = np.random.default_rng(seed)
    nb = len(ns)
    nbrs = [[] for _ in range(nb)]
    for a, b in edges:
        nbrs[a].append(b)
        nbrs[b].append(a)
    for i in range(nb):  # neighbour lists in no particular order
        rng.shuffle(nbrs[i])
    diag = [rng.normal(size=(ns[i], ns[i])) + shift * (1 + len(nbrs[i])) * np.eye(ns[i]) for i in range(nb)]
    off = [[rng.normal(size=(ns[i], ns[j])) for j in nbrs[i]] for i in range(nb)]
    return BlockSystem(ns, nbrs, diag, off)


# blocks 0-1-2 and 2-3-4 form triangles, so updates a_ij -= a_ik a_kj with
# k < j < i occur (fill inside the pattern), plus a chain and a pendant block
TRI_EDGES = [(0, 1), (0, 2), (1, 2), (2, 3), (2, 4), (3, 4), (4, 5), (5, 6), (1, 6), (3, 7)]
TRI_NS = [3, 1, 6, 4, 2, 6, 3, 5]


def test_ilu0_reproduces_the_matrix_on_its_pattern():
    S = random_system(TRI_NS, TRI_EDGES, seed=1)
    f, rc = S.ilu0()
    assert rc == 0
    L, U = S.factors_dense(f)
    A = S.dense()
    LU = L @ U
    mask = S.dense(np.ones_like(S.val)) != 0
    np.testing.assert_allclose(LU[mask], A[mask], rtol=0, atol=1e-12 * np.abs(A).max())
    # and the triangles make it incomplete: L U differs from A off the pattern
    assert np.abs(LU[~mask]).max() > 1e-3


def test_ilu0_of_a_block_chain_is_the_exact_factorisation():
    ns = [2, 5, 1, 6, 3, 4]
    S = random_system(ns, [(k, k + 1) for k in range(len(ns) - 1)], seed=2, shift=0.5)
    f, rc = S.ilu0()
    assert rc == 0
    b = np.random.default_rng(3).normal(size=int(S.xoff[-1]))
    x = S.ilu_solve(f, b)
    A = S.dense()
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=0,
                               atol=1e-12 * np.linalg.cond(A) * np.abs(x).max())


def test_ilu_solve_inverts_the_factors():
    S = random_system(TRI_NS, TRI_EDGES, seed=4)
    f, rc = S.ilu0()
    assert rc == 0
    L, U = S.factors_dense(f)
    b = np.random.default_rng(5).normal(size=int(S.xoff[-1]))
    v = S.ilu_solve(f, b)
    np.testing.assert_allclose(L @ (U @ v), b, rtol=0, atol=1e-12 * np.abs(b).max() * np.linalg.cond(L @ U))


def test_ilu_solve_in_place():
    S = random_system(TRI_NS, TRI_EDGES, seed=6)
    f, _ = S.ilu0()
    b = np.random.default_rng(7).normal(size=int(S.xoff[-1]))
    v = S.ilu_solve(f, b)
    w = b.copy()
    oracle().orc_block_ilu_solve(S.nb, P(S.ns), P(S.xoff), P(S.nptr), P(S.nbr), P(S.doff), P(S.ooff),
                                 P(f), P(w), P(w))
    assert np.array_equal(v, w)


def test_missing_mirrored_block_is_reported():
    ns = [2, 2]
    S = BlockSystem(ns, [[], [0]], [np.eye(2), np.eye(2)], [[], [np.ones((2, 2))]])
    _, rc = S.ilu0()
    assert rc == -1


def test_singular_diagonal_block_is_reported():
    ns = [2, 3]
    S = BlockSystem(ns, [[1], [0]], [np.eye(2), np.zeros((3, 3))], [[np.ones((2, 3))], [np.zeros((3, 2))]])
    _, rc = S.ilu0()
    assert rc == 2
