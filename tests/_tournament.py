"""CPU model of the LU with tournament pivoting (test helper).

Panels of 64 columns.  A panel's rows are cut into blocks of 128; each block is
factored with partial pivoting (largest |re|+|im|, the lowest row on ties) and
keeps the 64 rows it took, in order.  Lists of sibling nodes are stacked (left
first) over their rows' values at the start of the panel and factored again,
which keeps 64; a node without a sibling passes its list up.  The root's rows,
in the order its factorisation took them, are the panel's pivot rows.  The
interchanges follow LAPACK's convention (row k0 + t swapped with piv[k0 + t], in
order); L11 U11 is the unpivoted LU of the chosen rows, L21 = A21 U11^{-1},
U12 = L11^{-1} A12, A22 -= L21 U12.
"""
from __future__ import annotations

import numpy as np

NB = 64
ROWS = 128


def cabs1(x):
    return np.abs(x.real) + np.abs(x.imag)


def gepp_choose(B: np.ndarray, nsel: int) -> list:
    """Rows of B (local indices) taken by partial pivoting, in order, at most nsel."""
    B = np.array(B, dtype=np.complex128, copy=True)
    r, nb = B.shape
    live = np.ones(r, dtype=bool)
    out = []
    for j in range(min(nb, nsel)):
        a = np.where(live, np.maximum(cabs1(B[:, j]), 2.0 ** -1020), -1.0)
        p = int(np.argmax(a))              # first index of the maximum
        if a[p] < 0:
            break
        out.append(p)
        live[p] = False
        piv = B[p, j]
        den = piv.real ** 2 + piv.imag ** 2
        l = B[:, j] * (np.conj(piv) / den) if den > 0 else np.zeros(r, dtype=np.complex128)
        B[:, j + 1:] -= np.outer(l, B[p, j + 1:])
    return out


def tournament(P: np.ndarray) -> list:
    """The panel's pivot rows (indices into P's rows), in order."""
    m, nb = P.shape
    G = (m + ROWS - 1) // ROWS
    lists = []
    for g in range(G):
        rows = list(range(g * ROWS, min(m, (g + 1) * ROWS)))
        lists.append([rows[i] for i in gepp_choose(P[rows], nb)])
    while len(lists) > 1:
        nxt = []
        for a in range(0, len(lists), 2):
            if a + 1 == len(lists):
                nxt.append(lists[a])           # no sibling: passes up unchanged
                continue
            rows = lists[a] + lists[a + 1]
            nxt.append([rows[i] for i in gepp_choose(P[rows], nb)])
        lists = nxt
    return lists[0]


def interchanges(sel: list, k0: int) -> list:
    """LAPACK-order pivots for rows sel[t] (absolute) to land at k0 + t."""
    pos = list(range(k0, k0 + len(sel)))
    row_at = {}                                # position -> row, for positions that changed
    where = {}                                 # row -> position, for rows that moved
    piv = []
    for t, r in enumerate(sel):
        p = where.get(r, r)
        piv.append(p)
        a, b = pos[t], p
        ra, rb = row_at.get(a, a), row_at.get(b, b)
        row_at[a], row_at[b] = rb, ra
        where[rb], where[ra] = a, b
    return piv


def calu(A: np.ndarray):
    """Packed LU (L unit lower below the diagonal, U on and above) and pivots."""
    A = np.array(A, dtype=np.complex128, copy=True)
    n = A.shape[0]
    piv = np.zeros(n, dtype=np.int64)
    for k0 in range(0, n, NB):
        nb = min(NB, n - k0)
        c1 = k0 + nb
        sel = [k0 + i for i in tournament(A[k0:, k0:c1])]
        p = interchanges(sel, k0)
        for t in range(nb):
            piv[k0 + t] = p[t]
            A[[k0 + t, p[t]]] = A[[p[t], k0 + t]]
        # unpivoted LU of the top block
        T = A[k0:c1, k0:c1]
        for j in range(nb):
            T[j + 1:, j] /= T[j, j]
            T[j + 1:, j + 1:] -= np.outer(T[j + 1:, j], T[j, j + 1:])
        U11 = np.triu(T)
        L11 = np.tril(T, -1) + np.eye(nb)
        if c1 < n:
            A[c1:, k0:c1] = np.linalg.solve(U11.T, A[c1:, k0:c1].T).T
            A[k0:c1, c1:] = np.linalg.solve(L11, A[k0:c1, c1:])
            A[c1:, c1:] -= A[c1:, k0:c1] @ A[k0:c1, c1:]
    return A, piv


def apply_pivots(A: np.ndarray, piv) -> np.ndarray:
    PA = np.array(A, copy=True)
    for j in range(len(piv)):
        PA[[j, piv[j]]] = PA[[piv[j], j]]
    return PA


def backward_error(A: np.ndarray, F: np.ndarray, piv) -> float:
    """||P A - L U||_F / ||A||_F for packed factors F (row-major, as A)."""
    n = A.shape[0]
    F = np.asarray(F, dtype=np.complex128)
    L, U = np.tril(F, -1) + np.eye(n), np.triu(F)
    return float(np.linalg.norm(apply_pivots(A, piv) - L @ U) / np.linalg.norm(A))
