"""The CPU model of the LU with tournament pivoting (tests/_tournament.py) that the
GPU tests compare against: within one block of 128 rows it is partial pivoting
(LAPACK's pivots, izamax's |re|+|im| rule); over many blocks its factors
reproduce P A to roundoff; its interchanges put the chosen rows on top."""
import numpy as np
import scipy.linalg as sla

import _tournament as T


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))


def test_one_block_is_partial_pivoting():
    for n in (5, 64, 100, 128):
        A = _rand(n, n)
        F, piv = T.calu(A)
        lu, p = sla.lu_factor(A)
        assert np.array_equal(piv, p), n
        assert np.abs(F - lu).max() <= 1e-12 * np.abs(lu).max()


def test_interchanges_bring_the_chosen_rows_on_top():
    rng = np.random.default_rng(3)
    for _ in range(200):
        n, k0 = 300, int(rng.integers(0, 200))
        nb = int(rng.integers(1, 65))
        sel = list(rng.choice(np.arange(k0, n), size=nb, replace=False))
        piv = T.interchanges(sel, k0)
        rows = np.arange(n)
        for t, p in enumerate(piv):
            assert k0 + t <= p < n
            rows[[k0 + t, p]] = rows[[p, k0 + t]]
        assert list(rows[k0:k0 + nb]) == sel


def test_many_blocks_backward_error():
    for n, seed in ((300, 1), (700, 2)):
        A = _rand(n, seed)
        F, piv = T.calu(A)
        assert T.backward_error(A, F, piv) <= 1e-16 * n, n   # n u
        assert np.abs(np.tril(F, -1)).max() < 10.0        # multipliers above 1, but modest


def test_tree_choice_on_a_planted_panel():
    """A panel whose dominant rows are planted in different blocks: the tournament
    keeps them, in the order partial pivoting over the whole panel would take them
    when each dominates its own column."""
    rng = np.random.default_rng(5)
    m, nb = 1000, 64
    P = 1e-3 * (rng.normal(size=(m, nb)) + 1j * rng.normal(size=(m, nb)))
    rows = rng.choice(m, size=nb, replace=False)
    for j, r in enumerate(rows):
        P[r, j] = 10.0 + j
    assert T.tournament(P) == list(rows)
