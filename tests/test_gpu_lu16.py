"""FP32 LU whose rank-512 updates run on the FP16 matrix cores (lu32_gemm16):
split operands (1) give factors with the backward error of the complex FP32 GEMM's,
FP16 operands alone (2) give FP16-sized errors that the FP64 refinement still
removes; both refine to a 1e-12 residual."""
import numpy as np
import pytest

import _tournament as T

pytestmark = pytest.mark.gpu


def _rand(n, seed, shift=0.0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + shift * np.eye(n)


def _lu32(ctx, A, **tuning):
    n = A.shape[0]
    old = {k: ctx.get_tuning(k) for k in tuning}
    try:
        for k, v in tuning.items():
            ctx.set_tuning(k, v)
        dA = ctx.to_device(np.ascontiguousarray(A.T))
        dF, piv = ctx.empty(n * n * 8), ctx.empty(4 * n)
        ctx.lu32_factor(dA, n, n, dF, n, piv)
        F = dF.to_numpy((n, n), np.complex64).T.astype(np.complex128)
        p = piv.to_numpy((n,), np.int32)
        return F, p, (dA, dF, piv)
    finally:
        for k, v in old.items():
            ctx.set_tuning(k, v)


def _probe(A, F, piv, seed=0):
    n = A.shape[0]
    rng = np.random.default_rng(seed)
    PA = T.apply_pivots(A, piv)
    L, U = np.tril(F, -1) + np.eye(n), np.triu(F)
    z = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    return np.linalg.norm(PA @ z - L @ (U @ z)) / (np.linalg.norm(A) * np.linalg.norm(z))


@pytest.mark.parametrize("n", [1500, 3000])
def test_split_fp16_updates_match_fp32_accuracy(ctx, n):
    A = _rand(n, 600 + n)
    ref = _lu32(ctx, A, lu_outer_nb=512, lu32_gemm16=0)
    spl = _lu32(ctx, A, lu_outer_nb=512, lu32_gemm16=1)
    h16 = _lu32(ctx, A, lu_outer_nb=512, lu32_gemm16=2)
    e_ref, e_spl, e_h16 = _probe(A, *ref[:2]), _probe(A, *spl[:2]), _probe(A, *h16[:2])
    print(n, e_ref, e_spl, e_h16)
    assert e_spl <= 2.0 * e_ref + 1e-9
    assert e_ref <= 1e-7 * n
    assert e_h16 <= 1e-3 and e_h16 > e_spl   # FP16 products: errors of FP16 size


@pytest.mark.parametrize("mode", [1, 2])
def test_refinement_on_fp16_updated_factors(ctx, mode):
    n = 3000
    A = _rand(n, 700, shift=2 * np.sqrt(n))
    rng = np.random.default_rng(701)
    b = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    F, p, (dA, dF, piv) = _lu32(ctx, A, lu_outer_nb=512, lu32_gemm16=mode)
    dB = ctx.to_device(np.ascontiguousarray(b.T))
    it, res = ctx.refine_solve(dA, n, n, dF, n, piv, dB, 2, n, 1e-12, 40)
    X = dB.to_numpy((2, n), np.complex128).T
    assert res <= 1e-12 and np.linalg.norm(A @ X - b) / np.linalg.norm(b) <= 1e-12
    assert it <= (4 if mode == 1 else 12), it


def test_scaling_of_small_and_large_entries(ctx):
    """Operand maxima far from 1 (1e-12 and 1e12 times a random matrix): the power-of-two
    scaling keeps the FP16 parts in range, so the backward error is unchanged."""
    n = 1500
    A = _rand(n, 800)
    base = _probe(A, *_lu32(ctx, A, lu_outer_nb=512, lu32_gemm16=1)[:2])
    for s in (1e-12, 1e12):
        As = s * A
        e = _probe(As, *_lu32(ctx, As, lu_outer_nb=512, lu32_gemm16=1)[:2])
        assert np.isfinite(e) and e <= 2.0 * base + 1e-9, (s, e, base)
