"""Batched (many current sets) and mixed-precision forms.

- Batched FP64 field == one current set at a time, for every pass width
  (1e-12; near pairs go through per-pair operator matrices, so the
  agreement is to rounding, not bitwise).
- Mixed precision (FP32 far-pair node sums): field within 1e-6 of the FP64
  oracle; matrix blocks within 1e-6 of the FP64 matrix.
- FP32 LU + FP64 refinement: the FP64 solution to the stated residual, and a
  clean error when the matrix is too ill-conditioned for FP32 factors.
"""
import numpy as np
import pytest

import bielcc_amd as B
import oracle as O
from _util import block_metric, golden, np_dict, rowwise_rel

pytestmark = pytest.mark.gpu


def _sphere(nt=10, nph=20, seed=5):
    return B.mesh.sphere_quads(nt, nph, jitter=0.005, seed=seed)


def _near_points(rng, m_pts, radius):
    p = rng.normal(size=(m_pts, 3))
    return radius * p / np.linalg.norm(p, axis=1, keepdims=True)


@pytest.mark.parametrize("near_mode,k,with_jm", [(0, 3.0, False), (1, 2.0 + 0.3j, True), (0, 2.5, True)])
def test_batched_field_equals_one_set_at_a_time(ctx, near_mode, k, with_jm):
    rng = np.random.default_rng(31)
    verts = _sphere()
    n = verts.shape[0]
    nrhs = 5
    jE = rng.normal(size=(nrhs, n, 3)) + 1j * rng.normal(size=(nrhs, n, 3))
    jM = rng.normal(size=(nrhs, n, 3)) + 1j * rng.normal(size=(nrhs, n, 3)) if with_jm else None
    pts = np.concatenate([_near_points(rng, 24, 1.04), _near_points(rng, 24, 2.5)])
    kv = rng.normal(size=(nrhs, 3))
    ev = rng.normal(size=(nrhs, 3))
    npp = B.num_param()
    m = ctx.mesh(verts)
    ref_E, ref_H = [], []
    for r in range(nrhs):
        E1, H1 = m.field(npp, k, jE[r], pts, jM=None if jM is None else jM[r], multE=0.7 - 0.2j, multH=1.3j,
                         incident=(kv[r], ev[r]), near_mode=near_mode)
        ref_E.append(E1)
        ref_H.append(H1)
    for chunk in (1, 2, 4):
        ctx.set_tuning("field_rchunk", chunk)
        E, H = m.field_batch(npp, k, jE, pts, jM=jM, multE=0.7 - 0.2j, multH=1.3j, incident=(kv, ev),
                             near_mode=near_mode)
        for r in range(nrhs):
            assert rowwise_rel(E[r], ref_E[r]).max() <= 1e-12, (chunk, r)
            assert rowwise_rel(H[r], ref_H[r]).max() <= 1e-12, (chunk, r)
    ctx.set_tuning("field_rchunk", 4)
    m.close()


def test_tuning_keys(ctx):
    assert ctx.get_tuning("field_wg") == 2048
    ctx.set_tuning("field_wg", 512)
    assert ctx.get_tuning("field_wg") == 512
    ctx.set_tuning("field_wg", 2048)
    with pytest.raises(B.BielccError):
        ctx.set_tuning("field_rchunk", 3)
    with pytest.raises(B.BielccError):
        ctx.set_tuning("no_such_key", 1)


@pytest.mark.parametrize("k,general", [(2.0, False), (2.0 + 0.4j, False), (9.0, True)])
def test_mixed_field_vs_fp64_oracle(ctx, orc, monkeypatch, k, general):
    """FP32 far sums: 1e-6 relative to the FP64 oracle, near points included."""
    rng = np.random.default_rng(32)
    verts = _sphere(12, 24, seed=7)
    n = verts.shape[0]
    jE = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    pts = np.concatenate([_near_points(rng, 16, 1.05), _near_points(rng, 48, 3.0)])
    Eo, Ho = orc.field_eh(verts, jE, k, O.num_param(), pts)
    if general:   # the kernels' general sincos branch (k * diam beyond the reduced range)
        monkeypatch.setenv("BIELCC_GENERAL_SINCOS", "1")
    c = B.Context(0) if general else ctx
    m = c.mesh(verts)
    E, H = m.field_batch(B.num_param(), k, jE[None], pts, precision=B.PREC_MIXED)
    err_e, err_h = rowwise_rel(E[0], Eo).max(), rowwise_rel(H[0], Ho).max()
    print(f"mixed field k={k}: max rel err E {err_e:.2e} H {err_h:.2e}")
    assert err_e <= 1e-6 and err_h <= 1e-6
    m.close()
    if general:
        c.close()


@pytest.mark.parametrize("which,k", [(0, 2.0), (1, 2.0), (0, 3.0 + 0.5j)])
def test_mixed_matrix_vs_fp64(ctx, which, k):
    verts = _sphere(12, 24, seed=9)
    m = ctx.mesh(verts)
    npp = B.num_param()
    Z64 = m.assemble(npp, k, which=which)
    Z32 = m.assemble(npp, k, which=which, precision=B.PREC_MIXED)
    err = block_metric(Z32, Z64).max()
    print(f"mixed matrix which={which} k={k}: max block err {err:.2e}")
    assert err <= 1e-6
    m.close()


def test_refine_random_system(ctx):
    rng = np.random.default_rng(33)
    n, nrhs = 300, 3
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 30 * np.eye(n)
    Bm = rng.normal(size=(n, nrhs)) + 1j * rng.normal(size=(n, nrhs))
    m = ctx.mesh(_sphere(4, 8))
    X, it, res = m.solve_mixed(A, Bm, tol=1e-13)
    Xr = np.linalg.solve(A, Bm)
    assert np.linalg.norm(X - Xr) <= 1e-12 * np.linalg.norm(Xr)
    assert res <= 1e-13 and 1 <= it <= 8
    m.close()


@pytest.mark.parametrize("fname", ["config1.npz", "config1_tri.npz"])
def test_refine_on_collocation_matrix(ctx, fname):
    """The collocation system of the reference's example mesh: refinement on FP32
    factors reaches the FP64 LU solution (cond ~ 5.6e4 / 2.3e3)."""
    g = golden(fname)
    m = ctx.mesh(g["verts"])
    npp = B.num_param(**np_dict(g["num_param"]))
    Z = m.assemble(npp, float(g["k"]))
    rhs = m.rhs_planewave(g["kvec"], g["e0"])
    X, it, res = m.solve_mixed(Z, rhs, tol=1e-13)
    Xr = m.solve(Z, rhs)
    rel = np.linalg.norm(X - Xr) / np.linalg.norm(Xr)
    print(f"{fname}: {it} corrections, residual {res:.2e}, |X - X_lu| / |X_lu| = {rel:.2e}")
    assert res <= 1e-13
    assert rel <= 1e-10
    m.close()


def test_refine_reports_ill_conditioned(ctx):
    """cond ~ 1e11: FP32 factors cannot drive the residual down; the call fails
    with BIELCC_ERR_UNSUPPORTED instead of returning a wrong answer silently."""
    rng = np.random.default_rng(34)
    n = 200
    q1, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    q2, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    A = (q1 * np.logspace(0, -11, n)) @ q2
    Bm = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    m = ctx.mesh(_sphere(4, 8))
    with pytest.raises(B.BielccError) as ei:
        m.solve_mixed(A, Bm, tol=1e-12, max_iter=20)
    assert ei.value.code in (4, 5)
    m.close()
