"""LU with tournament pivoting on the GPU (lu_tournament = 1): the pivots of the CPU
model (tests/_tournament.py) exactly and its factors to rounding; backward error
of FP64 and FP32 factors at orders where the tree has up to seven levels; two-level
blocking, look-ahead, concurrent LUs and repeated runs give the same factors; a
zero pivot is reported."""
import threading

import numpy as np
import pytest

import bielcc_amd as B
import _tournament as T

pytestmark = pytest.mark.gpu
def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))


def _factor(ctx, A, fp32=False, **tuning):
    n = A.shape[0]
    tuning = {"lu_tournament": 1, **tuning}
    old = {k: ctx.get_tuning(k) for k in tuning}
    try:
        for k, v in tuning.items():
            ctx.set_tuning(k, v)
        dA = ctx.to_device(np.ascontiguousarray(A.T))
        piv = ctx.empty(4 * n)
        if fp32:
            dF = ctx.empty(n * n * 8)
            ctx.lu32_factor(dA, n, n, dF, n, piv)
            F = dF.to_numpy((n, n), np.complex64).T.astype(np.complex128)
            dF.free()
        else:
            ctx.lu_factor(dA, n, n, piv)
            F = dA.to_numpy((n, n), np.complex128).T
        p = piv.to_numpy((n,), np.int32)
        dA.free()
        piv.free()
        return F, p
    finally:
        for k, v in old.items():
            ctx.set_tuning(k, v)


def _probe_backward(A, F, piv, nprobe=4, seed=0):
    """max over random z of ||P A z - L (U z)|| / (||A||_F ||z||): O(n^2) per probe."""
    n = A.shape[0]
    rng = np.random.default_rng(seed)
    PA = T.apply_pivots(A, piv)
    L, U = np.tril(F, -1) + np.eye(n), np.triu(F)
    nA = np.linalg.norm(A)
    worst = 0.0
    for _ in range(nprobe):
        z = rng.normal(size=n) + 1j * rng.normal(size=n)
        worst = max(worst, np.linalg.norm(PA @ z - L @ (U @ z)) / (nA * np.linalg.norm(z)))
    return worst


@pytest.mark.parametrize("n", [64, 65, 200, 300, 1000])
def test_pivots_and_factors_of_the_cpu_model(ctx, n):
    A = _rand(n, 100 + n)
    F, piv = _factor(ctx, A)
    Fm, pm = T.calu(A)
    assert np.array_equal(piv, pm), (n, np.argwhere(piv != pm)[:5].ravel().tolist())
    assert np.abs(F - Fm).max() <= 1e-12 * np.abs(Fm).max()
    assert T.backward_error(A, F, piv) <= 1e-16 * max(n, 100)


@pytest.mark.parametrize("n", [1000, 2500, 6500])
def test_backward_error_fp64_and_fp32(ctx, n):
    """Panels of up to 6,500 rows: 51 blocks, a tree of seven levels."""
    A = _rand(n, 200 + n)
    F, piv = _factor(ctx, A)
    assert _probe_backward(A, F, piv) <= 1e-16 * n
    F32, piv32 = _factor(ctx, A, fp32=True)
    assert _probe_backward(A, F32, piv32) <= 1e-7 * n
    assert np.abs(np.tril(F, -1)).max() < 20.0


def test_solve_and_refinement(ctx):
    n = 3000
    rng = np.random.default_rng(31)
    A = _rand(n, 31)
    b = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    old = ctx.get_tuning("lu_tournament")
    ctx.set_tuning("lu_tournament", 1)
    try:
        dA = ctx.to_device(np.ascontiguousarray(A.T))
        dB = ctx.to_device(np.ascontiguousarray(b.T))
        piv = ctx.empty(4 * n)
        ctx.lu_factor(dA, n, n, piv)
        ctx.lu_solve(dA, n, n, piv, dB, 3, n)
        X = dB.to_numpy((3, n), np.complex128).T
        assert np.linalg.norm(A @ X - b) / (np.linalg.norm(A) * np.linalg.norm(X)) <= 1e-16 * n   # n u
        dA = ctx.to_device(np.ascontiguousarray(A.T))
        dB = ctx.to_device(np.ascontiguousarray(b.T))
        dF = ctx.empty(n * n * 8)
        ctx.lu32_factor(dA, n, n, dF, n, piv)
        it, res = ctx.refine_solve(dA, n, n, dF, n, piv, dB, 3, n, 1e-12, 30)
        X = dB.to_numpy((3, n), np.complex128).T
        assert res <= 1e-12 and np.linalg.norm(A @ X - b) / np.linalg.norm(b) <= 1e-12
    finally:
        ctx.set_tuning("lu_tournament", old)


@pytest.mark.parametrize("n", [700, 2500])
def test_two_level_and_look_ahead_pick_the_same_pivots(ctx, n):
    """Outer blocks of 128 and 512 columns, and the one-level loop with the next panel
    on the side stream, take the pivots of the in-order one-level LU (the tree reads
    only the panel's columns) and give its factors to rounding."""
    A = _rand(n, 300 + n)
    ref = _factor(ctx, A, lu_outer_nb=64, lu_look_ahead=0)
    for kw in ({"lu_outer_nb": 128}, {"lu_outer_nb": 512}, {"lu_outer_nb": 64, "lu_look_ahead": 1}):
        F, piv = _factor(ctx, A, **kw)
        assert np.array_equal(piv, ref[1]), kw
        assert np.abs(F - ref[0]).max() <= 1e-14 * n * np.abs(ref[0]).max(), kw


def test_repeatable_and_the_same_alone_and_concurrent():
    """No workgroup of the tree waits for another: four LUs on four contexts at once
    give, bit for bit, what each gives alone, which is what a repeated run gives."""
    ctxs = [B.Context(0) for _ in range(4)]
    try:
        mats = [_rand(n, 400 + n) for n in (1500, 2600, 3300, 4100)]
        alone = [_factor(c, A) for c, A in zip(ctxs, mats)]
        again = _factor(ctxs[0], mats[0])
        assert np.array_equal(again[0], alone[0][0]) and np.array_equal(again[1], alone[0][1])
        out = [None] * 4

        def run(t):
            out[t] = _factor(ctxs[t], mats[t])

        th = [threading.Thread(target=run, args=(t,)) for t in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=300)
        for a, b in zip(alone, out):
            assert b is not None and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    finally:
        for c in ctxs:
            c.close()


def test_zero_pivot_reported(ctx):
    for n in (10, 300):
        A = np.eye(n, dtype=complex)
        A[:, 4] = 0
        old = ctx.get_tuning("lu_tournament")
        ctx.set_tuning("lu_tournament", 1)
        try:
            dA = ctx.to_device(np.ascontiguousarray(A.T))
            piv = ctx.empty(4 * n)
            with pytest.raises(B.BielccError) as e:
                ctx.lu_factor(dA, n, n, piv)
            assert e.value.code == 5
        finally:
            ctx.set_tuning("lu_tournament", old)


@pytest.mark.parametrize("n", [1500, 3500])
def test_two_level_look_ahead(ctx, n):
    """lu_look_ahead2: the next outer block factored on the side stream (tournament
    panels) during the bulk of the rank-outer update, its interchanges outside the
    block applied afterwards.  With tournament panels throughout it takes the pivots
    of the LU without look-ahead; with partial pivoting elsewhere it reaches the same
    backward error; FP64 and FP32 (FP16 split updates)."""
    A = _rand(n, 500 + n)
    for outer in (128, 512):
        ref = _factor(ctx, A, lu_outer_nb=outer, lu_look_ahead2=0)
        la = _factor(ctx, A, lu_outer_nb=outer, lu_look_ahead2=1)
        assert np.array_equal(la[1], ref[1]), outer
        assert np.abs(la[0] - ref[0]).max() <= 1e-14 * n * np.abs(ref[0]).max(), outer
        mixed = _factor(ctx, A, lu_outer_nb=outer, lu_look_ahead2=1, lu_tournament=0)
        assert _probe_backward(A, *mixed) <= 1e-16 * n, outer
        m32 = _factor(ctx, A, fp32=True, lu_outer_nb=outer, lu_look_ahead2=1, lu_tournament=0)
        assert _probe_backward(A, *m32) <= 1e-7 * n, outer
