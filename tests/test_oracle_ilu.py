"""The block ILU restatement (oracle/ilu.py, after AS/CellILU.cs) pinned by what
defines an incomplete factorisation on a pattern: (L U)_ij = A_ij for every block
(i, j) of the pattern, and the exact LU where the pattern is full."""
import numpy as np
import pytest

import oracle as O
from oracle import ilu
from tests.helpers import log_uniform_nu, nodes_3d


def _blocks(n, k, seed=4):
    g = O.Grid(*nodes_3d(*n, jitter=0.3, seed=seed))
    op = O.Sip(g, k, log_uniform_nu(g.J, seed))
    A, _ = op.assemble()
    M = A.dense()
    bs = op.Np
    J = g.J
    blocks, rows = {}, []
    for i in range(J):
        r = []
        for j in range(J):
            B = M[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs]
            if np.any(B != 0):
                blocks[(i, j)] = B.copy()
                r.append(j)
        rows.append(r)
    return M, blocks, rows, bs


@pytest.mark.parametrize("p", [0, 1])
@pytest.mark.parametrize("k", [1, 2])
def test_lu_equals_a_on_the_pattern(p, k):
    M, blocks, rows, bs = _blocks((3, 3, 2), k)
    W, pat = ilu.factor(blocks, rows, p)
    LU = ilu.lu_product(W, pat, bs)
    for i, r in enumerate(pat):
        for j in r:
            s = (slice(i * bs, (i + 1) * bs), slice(j * bs, (j + 1) * bs))
            assert np.abs(LU[s] - M[s]).max() <= 1e-10 * np.abs(M).max()
    # ILU(1) fills more blocks than A holds, ILU(0) none
    nA = sum(len(r) for r in rows)
    nP = sum(len(r) for r in pat)
    assert (nP > nA) if p == 1 else (nP == nA)
    # outside the pattern L U differs from A (the dropped fill), for ILU(0) at least
    if p == 0:
        mask = np.ones_like(M, dtype=bool)
        for i, r in enumerate(pat):
            for j in r:
                mask[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = False
        assert np.abs(LU - M)[mask].max() > 1e-6 * np.abs(M).max()


def test_full_pattern_is_the_exact_solve():
    """2 x 2 x 1 cells: the ILU(1) pattern couples every pair, so L U = A and the
    solve is A^-1 b."""
    M, blocks, rows, bs = _blocks((2, 2, 1), 2)
    W, pat = ilu.factor(blocks, rows, 1)
    assert all(len(r) == 4 for r in pat)
    b = np.random.default_rng(3).standard_normal(M.shape[0])
    assert np.abs(ilu.solve(W, pat, b, bs) - np.linalg.solve(M, b)).max() <= 1e-10 * np.abs(np.linalg.solve(M, b)).max()


def test_solve_inverts_lu():
    M, blocks, rows, bs = _blocks((3, 3, 2), 2)
    W, pat = ilu.factor(blocks, rows, 1)
    b = np.random.default_rng(5).standard_normal(M.shape[0])
    x = ilu.solve(W, pat, b, bs)
    LU = ilu.lu_product(W, pat, bs)
    assert np.abs(LU @ x - b).max() <= 1e-10 * np.abs(b).max()
