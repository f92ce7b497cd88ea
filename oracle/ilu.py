"""Block ILU on the cell blocks (CellILU) restated in numpy (test infrastructure only).

Follows L3-solution/BoSSS.Solution.AdvancedSolvers/CellILU.cs:
* GetPattern (:1490-1560): the ILU(p) pattern is the block pattern of A^(p+1)
  (the ILU(0) pattern, the nonzero blocks of A, multiplied by itself p times);
* ComputeILU (:1229-1310): for k = 0 .. J-2: the diagonal block A_kk is factorised;
  for every i > k with (i, k) in the pattern, A_ik <- A_ik A_kk^-1 and, for every
  j > k with (i, j) in the pattern, A_ij <- A_ij - A_ik A_kj (blocks outside the
  pattern are dropped; blocks of the pattern that A does not hold start at zero);
* Solve (:182-250): X = U^-1 L^-1 B with L unit block lower (the scaled A_ik) and
  U block upper (the updated A_kj, j >= k).
Blocks are dense bs x bs numpy arrays keyed by (row, column).
"""
import numpy as np


def pattern(rows, p):
    """rows[i]: sorted block columns of A's row i; returns the rows of pattern(A^(p+1))."""
    pat = [set(r) for r in rows]
    base = [set(r) for r in rows]
    for _ in range(p):
        pat = [set().union(*(base[k] for k in pat[i])) for i in range(len(rows))]
    return [sorted(s) for s in pat]


def factor(blocks, rows, p):
    """blocks: {(i, j): bs x bs array} of A; returns (W, pat): the ILU(p) factors on the
    pattern (strictly lower blocks: L; diagonal and upper: U)."""
    pat = pattern(rows, p)
    J = len(rows)
    bs = next(iter(blocks.values())).shape[0]
    W = {}
    for i in range(J):
        for j in pat[i]:
            W[(i, j)] = blocks[(i, j)].copy() if (i, j) in blocks else np.zeros((bs, bs))
    colpat = [[] for _ in range(J)]
    for i in range(J):
        for k in pat[i]:
            colpat[k].append(i)
    for k in range(J - 1):
        Akk_inv = np.linalg.inv(W[(k, k)])
        for i in colpat[k]:
            if i <= k:
                continue
            W[(i, k)] = W[(i, k)] @ Akk_inv
            for j in pat[i]:
                if j >= k + 1 and (k, j) in W:
                    W[(i, j)] = W[(i, j)] - W[(i, k)] @ W[(k, j)]
    return W, pat


def solve(W, pat, b, bs):
    """x = U^-1 L^-1 b."""
    J = len(pat)
    y = np.array(b, dtype=np.float64).reshape(J, bs).copy()
    for i in range(J):
        for k in pat[i]:
            if k < i:
                y[i] -= W[(i, k)] @ y[k]
    x = np.zeros_like(y)
    for i in reversed(range(J)):
        s = y[i].copy()
        for j in pat[i]:
            if j > i:
                s -= W[(i, j)] @ x[j]
        x[i] = np.linalg.solve(W[(i, i)], s)
    return x.ravel()


def lu_product(W, pat, bs):
    """(L U) as a dense matrix, L unit block lower, U block upper."""
    J = len(pat)
    L = np.eye(J * bs)
    U = np.zeros((J * bs, J * bs))
    for (i, j), B in W.items():
        if j < i:
            L[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = B
        else:
            U[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = B
    return L @ U
