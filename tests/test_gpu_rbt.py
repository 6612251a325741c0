"""LU without interchanges of a randomly butterfly-transformed matrix (bielcc_lu_factor_rbt,
bielcc_rbt_solve).  The factors are checked against the definition of the transform,
U = diag(B', B'') B with B = (1/sqrt 2) [R S; R -S], built here as dense matrices from the
diagonals the library drew: ||U^T A V - L U|| / ||U^T A V|| at roundoff.  The solve is
checked by its residual against A -- at least as small as that of LAPACK's solution of the
same system (partial pivoting, numpy) -- and by its distance to LAPACK's solution, on random matrices
and on collocation matrices of the library's own assembly; orders that are not multiples of
4 leave their last rows and columns untransformed."""
import numpy as np
import pytest
from scipy.linalg import block_diag

import bielcc_amd as B

pytestmark = pytest.mark.gpu


def _bf(R, S):
    return np.block([[np.diag(R), np.diag(S)], [np.diag(R), -np.diag(S)]]) / np.sqrt(2)


def _dense(d, n):
    """U (n x n) of one transform from its 2 n4 diagonals: outer d[:n4], inner d[n4:]."""
    n4 = n & ~3
    U = np.eye(n)
    if n4:
        h, q = n4 // 2, n4 // 4
        d0, d1 = d[:n4], d[n4:]
        U[:n4, :n4] = block_diag(_bf(d1[:q], d1[q:h]), _bf(d1[h:h + q], d1[h + q:])) @ _bf(d0[:h], d0[h:])
    return U


def _factor(ctx, A, seed=7):
    n = A.shape[0]
    dA = ctx.to_device(np.ascontiguousarray(A.T))
    dF, piv, rbt = ctx.empty(n * n * 16), ctx.empty(4 * n), ctx.empty(32 * n)
    ctx.lu_factor_rbt(dA, n, n, dF, n, piv, rbt, seed=seed)
    F = dF.to_numpy((n, n), np.complex128).T
    d = rbt.to_numpy((4 * (n & ~3),), np.float64)
    return F, piv.to_numpy((n,), np.int32), d, (dA, dF, piv, rbt)


@pytest.mark.parametrize("n", [1000, 1003, 2048])
def test_factors_of_the_transformed_matrix(ctx, n):
    rng = np.random.default_rng(100 + n)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    F, piv, d, _ = _factor(ctx, A)
    n4 = n & ~3
    U, V = _dense(d[:2 * n4], n), _dense(d[2 * n4:], n)
    assert np.all(np.abs(np.log(d)) <= 0.05 + 1e-15) and np.array_equal(piv, np.arange(n))
    T = U.T @ A @ V
    L, Uf = np.tril(F, -1) + np.eye(n), np.triu(F)
    err = np.linalg.norm(T - L @ Uf) / np.linalg.norm(T)
    assert err <= 1e-15 * n, err


def test_same_seed_same_bits(ctx):
    rng = np.random.default_rng(3)
    n = 700
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    F1, _, d1, _ = _factor(ctx, A, seed=11)
    F2, _, d2, _ = _factor(ctx, A, seed=11)
    F3, _, d3, _ = _factor(ctx, A, seed=12)
    assert np.array_equal(F1, F2) and np.array_equal(d1, d2)
    assert not np.array_equal(d1, d3)


def _lapack_residual(A, b):
    """Relative residual of LAPACK's solution (partial pivoting, no refinement)."""
    x = np.linalg.solve(A, b)
    return float((np.linalg.norm(b - A @ x, axis=0) / np.linalg.norm(b, axis=0)).max()), x


def _solve(ctx, A, b, tol, seed=7):
    n, nr = A.shape[0], b.shape[1]
    F, piv, d, (dA, dF, dp, rbt) = _factor(ctx, A, seed)
    dB = ctx.to_device(np.ascontiguousarray(b.T))
    it, res = ctx.rbt_solve(dA, n, n, dF, n, dp, rbt, dB, nr, n, tol=tol, max_iter=10)
    return dB.to_numpy((nr, n), np.complex128).T, it, res


@pytest.mark.parametrize("n", [1500, 1501])
def test_solve_random_system(ctx, n):
    rng = np.random.default_rng(200 + n)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    b = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    r_lapack, xt = _lapack_residual(A, b)
    X, it, res = _solve(ctx, A, b, tol=r_lapack)   # at least as small a residual as LAPACK's solution
    assert res <= r_lapack and it <= 4, (it, res, r_lapack)
    assert (np.linalg.norm(A @ X - b, axis=0) / np.linalg.norm(b, axis=0)).max() <= 1.5 * r_lapack
    assert np.linalg.norm(X - xt) / np.linalg.norm(xt) <= 1e-15 * np.linalg.cond(A)


@pytest.mark.parametrize("shape,k", [("sphere", 2.0), ("torus", 4.0), ("ellipsoid", 1.0)])
def test_solve_collocation_matrix(ctx, shape, k):
    """The K collocation matrix of the library's assembly (FP64), plane-wave right-hand
    sides: the refined solution against LAPACK's, to the conditioning of the matrix."""
    verts = {"sphere": lambda: B.mesh.sphere_quads(16, 32, seed=1),
             "torus": lambda: B.mesh.torus_quads(40, 16, R=1.0, r=0.3),
             "ellipsoid": lambda: B.mesh.ellipsoid_quads(16, 32, axes=(2.0, 0.5, 0.5), seed=4)}[shape]()
    m = ctx.mesh(verts)
    Z = m.assemble(B.num_param(), k)
    n = Z.shape[0]
    rng = np.random.default_rng(5)
    b = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    r_lapack, xt = _lapack_residual(Z, b)
    X, it, res = _solve(ctx, Z, b, tol=r_lapack)
    assert res <= r_lapack and it <= 4, (it, res, r_lapack)
    cond = np.linalg.cond(Z)
    assert np.linalg.norm(X - xt) / np.linalg.norm(xt) <= 1e-15 * cond, (np.linalg.norm(X - xt), cond)
    m.close()


def test_zero_matrix_reports_the_zero_pivot(ctx):
    n = 64
    with pytest.raises(B.BielccError) as e:
        _factor(ctx, np.zeros((n, n), np.complex128))
    assert "column 1" in str(e.value)


def _factor32(ctx, A, seed=7):
    n = A.shape[0]
    dA = ctx.to_device(np.ascontiguousarray(A.T))
    dF, piv, rbt = ctx.empty(n * n * 8), ctx.empty(4 * n), ctx.empty(32 * n)
    ctx.lu32_factor_rbt(dA, n, n, dF, n, piv, rbt, seed=seed)
    return dA, dF, piv, rbt


@pytest.mark.parametrize("n", [1000, 1003])
def test_fp32_factors_of_the_transformed_matrix(ctx, n):
    """FP32 factors of U^T A V: backward error of FP32 size, as the pivoting FP32 LU's."""
    rng = np.random.default_rng(300 + n)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    dA, dF, piv, rbt = _factor32(ctx, A)
    F = dF.to_numpy((n, n), np.complex64).T.astype(np.complex128)
    d = rbt.to_numpy((4 * (n & ~3),), np.float64)
    n4 = n & ~3
    T = _dense(d[:2 * n4], n).T @ A @ _dense(d[2 * n4:], n)
    err = np.linalg.norm(T - (np.tril(F, -1) + np.eye(n)) @ np.triu(F)) / np.linalg.norm(T)
    assert err <= 1e-7 * n, err
    assert np.array_equal(piv.to_numpy((n,), np.int32), np.arange(n))


@pytest.mark.parametrize("mode", [1, 0])
@pytest.mark.parametrize("shape", ["random", "sphere"])
def test_fp32_refinement_on_transformed_factors(ctx, mode, shape):
    """FGMRES (1) and classic steps (0) on the FP32 butterfly factors reach a residual at
    most LAPACK's, with the solution within the conditioning of LAPACK's."""
    rng = np.random.default_rng(17)
    if shape == "random":
        n = 1500
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        m = None
    else:
        m = ctx.mesh(B.mesh.sphere_quads(16, 32, seed=1))
        A = m.assemble(B.num_param(), 2.0)
        n = A.shape[0]
    b = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    r_lapack, xt = _lapack_residual(A, b)
    dA, dF, piv, rbt = _factor32(ctx, A)
    dB = ctx.to_device(np.ascontiguousarray(b.T))
    old = ctx.get_tuning("refine_mode")
    ctx.set_tuning("refine_mode", mode)
    try:
        it, res = ctx.refine_solve(dA, n, n, dF, n, piv, dB, 2, n, tol=max(r_lapack, 1e-13), max_iter=40, rbt=rbt)
    finally:
        ctx.set_tuning("refine_mode", old)
    X = dB.to_numpy((2, n), np.complex128).T
    assert res <= max(r_lapack, 1e-13), (it, res, r_lapack)
    assert np.linalg.norm(X - xt) / np.linalg.norm(xt) <= 1e-15 * np.linalg.cond(A)
    if m is not None:
        m.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_concurrent_factorisations_match_alone(prec):
    """Four contexts factor at once on four host threads (orders 6,000-9,000: up to 141
    workgroups per panel launch each, more than the device holds together), so workgroups
    of a panel start late while others have finished: the factors equal those of the same
    factorisations run alone, bit for bit.  (Every workgroup reads the diagonal block
    unfactored; the last one to finish writes it.)"""
    import threading
    ns = [9000, 8000, 7000, 6000]
    rng = np.random.default_rng(41)
    mats = [rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) for n in ns]
    ctxs = [B.Context(0) for _ in ns]
    item = 8 if prec == "fp32" else 16

    def run(i, out):
        c, n = ctxs[i], ns[i]
        dA = c.to_device(np.ascontiguousarray(mats[i].T))
        dF, piv, rbt = c.empty(n * n * item), c.empty(4 * n), c.empty(32 * n)
        if prec == "fp32":
            c.lu32_factor_rbt(dA, n, n, dF, n, piv, rbt, seed=5)
        else:
            c.lu_factor_rbt(dA, n, n, dF, n, piv, rbt, seed=5)
        c.sync()
        out[i] = dF.to_numpy((n * n * item // 8,), np.float64)
        for d in (dA, dF, piv, rbt):
            d.free()

    alone = {}
    for i in range(len(ns)):
        run(i, alone)
    together = {}
    th = [threading.Thread(target=run, args=(i, together)) for i in range(len(ns))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for c in ctxs:
        c.close()
    for i in range(len(ns)):
        assert np.array_equal(alone[i], together[i]), ns[i]
