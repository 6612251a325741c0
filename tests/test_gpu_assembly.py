"""Collocation matrix, RHS, LU solve and currents on the GPU against the
reference's recorded outputs (config 1: 512-quad sphere and its 960-triangle
split, ka = 2) and the oracle.  Bars: blocks 1e-12 (per-block metric),
currents within the conditioning-limited bound stated per test."""
import numpy as np
import pytest

import bielcc_amd as B
import oracle as O
from _util import golden, np_dict, rel_err, rowwise_rel, block_metric

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["config1.npz", "config1_tri.npz"])
def c1(request, ctx):
    g = golden(request.param)
    m = ctx.mesh(g["verts"])
    npp = B.num_param(**np_dict(g["num_param"]))
    Z = m.assemble(npp, float(g["k"]), which=0)
    yield g, m, npp, Z
    m.close()


def test_matrix_vs_reference_samples(c1):
    g, m, npp, Z = c1
    n = g["verts"].shape[0]
    si, sj = g["sample_i"], g["sample_j"]
    blocks = np.stack([Z[2 * si + a, 2 * sj + b] for a in range(2) for b in range(2)], axis=1)
    bm = np.sqrt((np.abs(blocks - g["blocks"]) ** 2).sum(1)) / np.sqrt((np.abs(g["blocks"]) ** 2).sum(1))
    assert bm.max() <= 1e-12, bm.max()
    diag = np.stack([Z[2 * np.arange(n) + a, 2 * np.arange(n) + b] for a in range(2) for b in range(2)], axis=1)
    assert rel_err(diag, g["diag"]) <= 1e-12
    assert np.linalg.norm(Z @ g["row_w"] - g["row_sums"]) <= 1e-12 * np.linalg.norm(g["row_sums"])
    assert np.linalg.norm(g["row_w"] @ Z - g["col_sums"]) <= 1e-12 * np.linalg.norm(g["col_sums"])


def test_matrix_vs_oracle_every_block(c1, orc):
    g, m, npp, Z = c1
    Zo = orc.assemble(g["verts"], O.num_param(**np_dict(g["num_param"])), float(g["k"]), which=0)
    bm = block_metric(Z, Zo)
    assert bm.max() <= 1e-12, bm.max()


def test_r_matrix(c1, orc):
    g, m, npp, _ = c1
    ZR = m.assemble(npp, float(g["k"]), which=1)
    si, sj = g["sample_i"], g["sample_j"]
    rb = np.stack([ZR[2 * si + a, 2 * sj + b] for a in range(2) for b in range(2)], axis=1)
    assert rel_err(rb, g["rblocks"]) <= 1e-12
    assert np.linalg.norm(ZR @ g["row_w"] - g["rrow_sums"]) <= 1e-12 * np.linalg.norm(g["rrow_sums"])
    # self blocks are exactly zero (R_Rot_Colloc at the centroid, R.h:94-97)
    n = g["verts"].shape[0]
    assert np.all(ZR[2 * np.arange(n), 2 * np.arange(n)] == 0)


def test_kept_near_list_follows_rows_and_rule(ctx, orc):
    """The mesh keeps the near-pair list of its last assembly: the same rows again give
    the same matrix bit for bit; other rows, then all rows again, and a wider near
    rule (analytic_calc 4, more exact pairs) each match the oracle as on a fresh list."""
    g = golden("config1.npz")
    verts = g["verts"][:240]
    m = ctx.mesh(verts)
    k = 2.5
    npp = B.num_param()
    Z1 = m.assemble(npp, k)
    assert np.array_equal(m.assemble(npp, k), Z1)
    Zp = m.assemble(npp, k, row0=100, nrows=80)
    assert np.array_equal(Zp, Z1[100:180])
    assert np.array_equal(m.assemble(npp, k), Z1)
    assert block_metric(Z1, orc.assemble(verts, O.num_param(), k)).max() <= 1e-12
    npw = B.num_param(analytic_calc=4.0)
    Zw = m.assemble(npw, k)
    assert block_metric(Zw, orc.assemble(verts, O.num_param(analytic_calc=4.0), k)).max() <= 1e-12
    assert not np.array_equal(Zw, Z1)
    m.close()


def test_partial_rows_and_complex_k(ctx, orc):
    g = golden("config1.npz")
    verts = g["verts"][:200]
    m = ctx.mesh(verts)
    k = 2.0 + 0.25j
    npp = B.num_param()
    Zfull = m.assemble(npp, k)
    Zpart = m.assemble(npp, k, row0=40, nrows=60)
    assert np.array_equal(Zpart, Zfull[40:100])
    Zo = orc.assemble(verts, O.num_param(), k, row0=40, nrows=60)
    assert block_metric(Zpart, Zo).max() <= 1e-12
    m.close()


def test_adaptive_parameters_generic_path(ctx, orc):
    """p_max > 1 (the examples' settings, main0_K.cpp:40-73) runs the exact per-pair kernel."""
    verts = B.mesh.sphere_quads(6, 12, jitter=0.005, seed=9)
    m = ctx.mesh(verts)
    npd = dict(n_start=6, n_start_seg=8, p_max=3, p_max_seg=2, eps=1e-5)
    Z = m.assemble(B.num_param(**npd), 1.5)
    Zo = orc.assemble(verts, O.num_param(**npd), 1.5)
    assert block_metric(Z, Zo).max() <= 1e-12
    m.close()


def test_rhs_and_currents(c1):
    g, m, npp, Z = c1
    rhs = m.rhs_planewave(g["kvec"], g["e0"])[:, 0]
    assert rel_err(rhs, g["rhs"]) <= 1e-15
    j = m.currents(g["b"])[0]
    assert rel_err(j, g["jE"]) <= 1e-15


def test_lu_solve_currents(c1):
    """Currents from the GPU LU of the GPU matrix vs LAPACK on the reference
    matrix.  The reference has no solver (parity of the solve is against
    LAPACK); the bound is cond(Z) * 1e-15 with cond recorded in the fixture."""
    g, m, npp, Z = c1
    b = m.solve(Z, g["rhs"])[:, 0]
    cond = float(g["cond"])
    err = np.linalg.norm(b - g["b"]) / np.linalg.norm(g["b"])
    assert err <= max(1e-12, 2e-15 * cond), (err, cond)
    res = np.linalg.norm(Z @ b - g["rhs"]) / np.linalg.norm(g["rhs"])
    assert res <= 1e-13, res


def test_lu_random_multirhs(ctx):
    rng = np.random.default_rng(3)
    for n in (1, 63, 64, 65, 700):
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 4 * np.sqrt(n) * np.eye(n)
        Bm = rng.normal(size=(n, 5)) + 1j * rng.normal(size=(n, 5))
        X = ctx_solve(ctx, A, Bm)
        Xr = np.linalg.solve(A, Bm)
        assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) <= 1e-13, n


@pytest.mark.parametrize("nrhs", [1, 3, 8, 9, 64])
def test_lu_solve_sync_free_and_blocked_agree(ctx, nrhs):
    """FP64 solves: one right-hand side goes through the one-kernel-per-triangle
    solve, more through rocBLAS trsm or the blocked TRSM + GEMM loop (trsv = 0 sends
    one there too); every path against numpy, ragged last block."""
    rng = np.random.default_rng(4)
    for n in (1, 64, 200, 1000):
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 4 * np.sqrt(n) * np.eye(n)
        Bm = rng.normal(size=(n, nrhs)) + 1j * rng.normal(size=(n, nrhs))
        Xr = np.linalg.solve(A, Bm)
        out = []
        for trsv, lib in ((1, 1), (0, 1), (0, 0)):
            ctx.set_tuning("trsv", trsv)
            ctx.set_tuning("trsm_lib", lib)
            X = ctx_solve(ctx, A, Bm)
            assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) <= 1e-13, (n, trsv, lib)
            out.append(X)
        ctx.set_tuning("trsv", 1)
        ctx.set_tuning("trsm_lib", 1)
        for X in out[1:]:
            assert np.linalg.norm(out[0] - X) <= 1e-13 * np.linalg.norm(Xr)


def ctx_solve(ctx, A, Bm):
    n = A.shape[0]
    dA = ctx.to_device(np.ascontiguousarray(A.T))
    dB = ctx.to_device(np.ascontiguousarray(Bm.T))
    piv = ctx.empty(4 * n)
    ctx.lu_factor(dA, n, n, piv)
    ctx.lu_solve(dA, n, n, piv, dB, Bm.shape[1], n)
    return dB.to_numpy((Bm.shape[1], n), np.complex128).T


def test_lu_singular_reports(ctx):
    n = 10
    A = np.eye(n, dtype=complex)
    A[:, 4] = 0
    dA = ctx.to_device(np.ascontiguousarray(A.T))
    piv = ctx.empty(4 * n)
    with pytest.raises(B.BielccError) as e:
        ctx.lu_factor(dA, n, n, piv)
    assert e.value.code == 5


def test_bad_arguments(ctx):
    m = ctx.mesh(golden("config1.npz")["verts"][:10])
    with pytest.raises(B.BielccError) as e:
        m.assemble(B.num_param(), 1.0, row0=1, nrows=2)
    assert e.value.code == 1
    with pytest.raises(B.BielccError):
        m.assemble(B.num_param(n_start=0), 1.0)
    with pytest.raises(ValueError):
        ctx.mesh(np.zeros((3, 5, 3)))
    m.close()


def test_sync_free_solve_recovers_from_an_expired_wait(ctx):
    """A wait of the one-kernel triangular solve that gives up (here forced from
    the host with a one-poll limit, so every kernel still drains normally) is
    redone on the blocked path: lu_solve and the FP32 refinement still return
    the solution."""
    rng = np.random.default_rng(6)
    n = 1000
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 4 * np.sqrt(n) * np.eye(n)
    Bm = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    Xr = np.linalg.solve(A, Bm)
    ctx.set_tuning("trsv_spin", 1)
    try:
        X = ctx_solve(ctx, A, Bm)
        assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) <= 1e-13
        dA = ctx.to_device(np.ascontiguousarray(A.T))
        dB = ctx.to_device(np.ascontiguousarray(Bm.T))
        dF, piv = ctx.empty(n * n * 8), ctx.empty(4 * n)
        ctx.lu32_factor(dA, n, n, dF, n, piv)
        it, res = ctx.refine_solve(dA, n, n, dF, n, piv, dB, 3, n, 1e-13, 30)
        X = dB.to_numpy((3, n), np.complex128).T
        assert res <= 1e-13 and np.linalg.norm(X - Xr) / np.linalg.norm(Xr) <= 1e-12
    finally:
        ctx.set_tuning("trsv_spin", 1 << 20)


def _fp32_first_solve(ctx, A, Bm, trsv):
    """X0 = F^-1 B of the refinement (FP32 factors, one FP32 solve: no correction
    allowed, so the refinement stops with code 4 and leaves X0), with the factors
    and pivots it used."""
    n, nr = A.shape[0], Bm.shape[1]
    dA = ctx.to_device(np.ascontiguousarray(A.T))
    dB = ctx.to_device(np.ascontiguousarray(Bm.T))
    dF, piv = ctx.empty(n * n * 8), ctx.empty(4 * n)
    ctx.lu32_factor(dA, n, n, dF, n, piv)
    ctx.set_tuning("trsv", trsv)
    try:
        with pytest.raises(B.BielccError) as e:
            ctx.refine_solve(dA, n, n, dF, n, piv, dB, nr, n, 1e-30, 0)
        assert e.value.code == 4
    finally:
        ctx.set_tuning("trsv", 1)
    return (dB.to_numpy((nr, n), np.complex128).T, dF.to_numpy((n, n), np.complex64).T,
            piv.to_numpy((n,), np.int32))


@pytest.mark.parametrize("nrhs", [8, 9, 12, 16, 17])
def test_fp32_solve_up_to_16_rhs_in_one_kernel(ctx, nrhs):
    """FP32 solves of up to 16 right-hand sides take the one-kernel solve (8: the
    8-wide instantiation, 9-16 the 16-wide one; 17: rocBLAS trsm).  Its X0 against the
    same FP32 factors and pivots substituted in FP64, and against the trsm path; ragged
    last blocks."""
    assert ctx.get_tuning("trsv_max32") == 16
    _fp32_solves_against_fp64_substitution(ctx, nrhs)


def _fp32_solves_against_fp64_substitution(ctx, nrhs):
    import scipy.linalg as sl
    rng = np.random.default_rng(9)
    for n in (64, 700, 1000):
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 4 * np.sqrt(n) * np.eye(n)
        Bm = rng.normal(size=(n, nrhs)) + 1j * rng.normal(size=(n, nrhs))
        X1, F, piv = _fp32_first_solve(ctx, A, Bm, 1)
        X0, F0, piv0 = _fp32_first_solve(ctx, A, Bm, 0)
        assert np.array_equal(F, F0) and np.array_equal(piv, piv0)
        Y = Bm.copy()
        for j in range(n):   # the factorisation's interchanges, in order
            Y[[j, piv[j]]] = Y[[piv[j], j]]
        F64 = F.astype(np.complex128)
        Y = sl.solve_triangular(F64, Y, lower=True, unit_diagonal=True)
        Xf = sl.solve_triangular(F64, Y, lower=False)
        for X in (X1, X0):
            assert np.linalg.norm(X - Xf, axis=0).max() / np.linalg.norm(Xf, axis=0).min() <= 1e-5, n
        assert np.linalg.norm(X1 - X0) <= 1e-5 * np.linalg.norm(Xf)
        # and refined to 1e-12 on both paths
        for trsv in (1, 0):
            ctx.set_tuning("trsv", trsv)
            dA = ctx.to_device(np.ascontiguousarray(A.T))
            dB = ctx.to_device(np.ascontiguousarray(Bm.T))
            dF, dp = ctx.empty(n * n * 8), ctx.empty(4 * n)
            ctx.lu32_factor(dA, n, n, dF, n, dp)
            it, res = ctx.refine_solve(dA, n, n, dF, n, dp, dB, nrhs, n, 1e-12, 10)
            X = dB.to_numpy((nrhs, n), np.complex128).T
            assert res <= 1e-12 and np.linalg.norm(A @ X - Bm) / np.linalg.norm(Bm) <= 1e-12, (n, trsv)
        ctx.set_tuning("trsv", 1)


@pytest.mark.parametrize("n", [65, 300, 1000])
def test_fused_panel_matches_two_launch_panel(ctx, n):
    """One launch per panel column (next column's pivot by the last workgroup)
    gives bit for bit the factors and pivots of the two-launch panel, FP64 and FP32."""
    rng = np.random.default_rng(7)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    out = {}
    for fused in (1, 0):
        ctx.set_tuning("lu_fused_panel", fused)
        dA = ctx.to_device(np.ascontiguousarray(A.T))
        piv = ctx.empty(4 * n)
        ctx.lu_factor(dA, n, n, piv)
        dF, piv32 = ctx.empty(n * n * 8), ctx.empty(4 * n)
        ctx.lu32_factor(ctx.to_device(np.ascontiguousarray(A.T)), n, n, dF, n, piv32)
        out[fused] = (dA.to_numpy((n, n), np.complex128), piv.to_numpy((n,), np.int32),
                      dF.to_numpy((n, n), np.complex64), piv32.to_numpy((n,), np.int32))
    ctx.set_tuning("lu_fused_panel", 2)
    for a, b in zip(out[1], out[0]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [2500, 6500])
def test_lu_residual_at_panel_sizes(ctx, n):
    """FP64 and FP32 LU at orders where the panel kernels run hundreds of
    workgroups (2,500: fused panel steps; 6,500: two launches per column):
    the FP64 solve's residual, and FP32 factors refined to 1e-12."""
    rng = np.random.default_rng(8)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 2 * np.sqrt(n) * np.eye(n)
    b = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    X = ctx_solve(ctx, A, b)
    assert np.linalg.norm(A @ X - b) / np.linalg.norm(b) <= 1e-13
    dA = ctx.to_device(np.ascontiguousarray(A.T))
    dB = ctx.to_device(np.ascontiguousarray(b.T))
    dF, piv = ctx.empty(n * n * 8), ctx.empty(4 * n)
    ctx.lu32_factor(dA, n, n, dF, n, piv)
    it, res = ctx.refine_solve(dA, n, n, dF, n, piv, dB, 2, n, 1e-12, 10)
    X = dB.to_numpy((2, n), np.complex128).T
    assert it <= 4 and np.linalg.norm(A @ X - b) / np.linalg.norm(b) <= 1e-12


@pytest.mark.parametrize("n", [300, 1000, 2500])
def test_look_ahead_matches_in_order_lu(ctx, n):
    """The look-ahead LU (next panel factored on the side stream while the rest of
    the trailing update runs) picks the pivots of the in-order LU and gives its
    factors to rounding (the trailing GEMM is split in two calls, so sums may round
    differently; the bound grows with the n elimination steps), FP64 and FP32."""
    rng = np.random.default_rng(9)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    out = {}
    ctx.set_tuning("lu_outer_nb", 64)   # the look-ahead belongs to the one-level loop
    try:
        for la in (1, 0):
            ctx.set_tuning("lu_look_ahead", la)
            dA = ctx.to_device(np.ascontiguousarray(A.T))
            piv = ctx.empty(4 * n)
            ctx.lu_factor(dA, n, n, piv)
            dF, piv32 = ctx.empty(n * n * 8), ctx.empty(4 * n)
            ctx.lu32_factor(ctx.to_device(np.ascontiguousarray(A.T)), n, n, dF, n, piv32)
            out[la] = (dA.to_numpy((n, n), np.complex128), piv.to_numpy((n,), np.int32),
                       dF.to_numpy((n, n), np.complex64), piv32.to_numpy((n,), np.int32))
    finally:
        ctx.set_tuning("lu_look_ahead", 2)
        ctx.set_tuning("lu_outer_nb", 0)
    assert np.array_equal(out[1][1], out[0][1]) and np.array_equal(out[1][3], out[0][3])
    scale = np.abs(out[0][0]).max()
    assert np.abs(out[1][0] - out[0][0]).max() <= 1e-14 * n * scale
    scale = np.abs(out[0][2]).max()
    assert np.abs(out[1][2] - out[0][2]).max() <= 1e-6 * n * scale


def _factors(ctx, A, fp32=True, **tuning):
    n = A.shape[0]
    old = {k: ctx.get_tuning(k) for k in tuning}
    try:
        for k, v in tuning.items():
            ctx.set_tuning(k, v)
        dA = ctx.to_device(np.ascontiguousarray(A.T))
        piv = ctx.empty(4 * n)
        ctx.lu_factor(dA, n, n, piv)
        out = (dA.to_numpy((n, n), np.complex128), piv.to_numpy((n,), np.int32))
        dA.free()
        if not fp32:
            return out
        dF, piv32 = ctx.empty(n * n * 8), ctx.empty(4 * n)
        ctx.lu32_factor(ctx.to_device(np.ascontiguousarray(A.T)), n, n, dF, n, piv32)
        return out + (dF.to_numpy((n, n), np.complex64), piv32.to_numpy((n,), np.int32))
    finally:
        for k, v in old.items():
            ctx.set_tuning(k, v)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_lu_u12_triangle_modes(ctx, mode):
    """The factorisation's U12 triangles through the library's LDS sweep (0), rocBLAS
    trsm (1) and the lane-substitution kernel (2): each gives FP64 and FP32 factors
    whose backward error ||P A - L U|| / ||A|| is at roundoff level; ragged last panel."""
    rng = np.random.default_rng(22)
    for n in (300, 1000, 2000):
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        LU, piv, F, piv32 = _factors(ctx, A, lu_trsm_lib=mode)
        for fac, p, tol in ((LU.T, piv, 1e-13), (F.T.astype(np.complex128), piv32, 1e-4)):
            PA = A.copy()
            for j in range(n):   # the factorisation's interchanges, in order
                PA[[j, p[j]]] = PA[[p[j], j]]
            L = np.tril(fac, -1) + np.eye(n)
            err = np.linalg.norm(PA - L @ np.triu(fac)) / np.linalg.norm(A)
            assert err <= tol, (n, mode, err)


@pytest.mark.parametrize("n", [64, 65, 300, 1000, 2500])
def test_whole_panel_kernel_matches_column_kernels(ctx, n):
    """The whole-panel launch (panel rows in LDS across workgroups, one grid
    barrier per column) gives bit for bit the factors and pivots of the
    per-column launches, FP64 and FP32 (same arithmetic, same order)."""
    rng = np.random.default_rng(11)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    a = _factors(ctx, A, lu_panel_lds=1, lu_look_ahead=0)
    b = _factors(ctx, A, lu_panel_lds=0, lu_look_ahead=0)
    for name, x, y in zip(("LU64", "piv64", "LU32", "piv32"), a, b):
        d = np.abs(x.astype(np.complex128) - y.astype(np.complex128))
        where = np.argwhere(d > 0)
        assert np.array_equal(x, y), (name, d.max(), len(where), where[:8].tolist(),
                                      np.array_equal(a[1], b[1]), np.array_equal(a[3], b[3]))


@pytest.mark.parametrize("n", [300, 1000, 2500])
def test_fp32_whole_panel_two_rows_per_lane(ctx, n):
    """FP32 whole panels with 128 rows per workgroup, two per lane (by default
    used only for panels of more than 512 x 64 = 32,768 rows), forced here at
    small orders: the same FP32 factors and pivots, bit for bit, as the
    per-column launches, and a backward error ||P A - L U|| / ||A|| at FP32
    roundoff."""
    rng = np.random.default_rng(12)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    a = _factors(ctx, A, lu_panel_lds=2, lu_panel_rows32=128, lu_look_ahead=0)
    b = _factors(ctx, A, lu_panel_lds=0, lu_look_ahead=0)
    assert ctx.get_tuning("lu_panel_rows32") == 0
    assert np.array_equal(a[3], b[3]) and np.array_equal(a[2], b[2])
    F, p = a[2].T.astype(np.complex128), a[3]
    PA = A.copy()
    for j in range(n):
        PA[[j, p[j]]] = PA[[p[j], j]]
    err = np.linalg.norm(PA - (np.tril(F, -1) + np.eye(n)) @ np.triu(F)) / np.linalg.norm(A)
    assert err <= 1e-4, err


@pytest.mark.parametrize("n", [65, 1000, 2500])
def test_fp64_whole_panel_two_rows_per_lane(ctx, n):
    """FP64 whole panels with 128 rows per workgroup (lu_panel_rows64 = 128: one
    workgroup per compute unit, half the barrier arrivals): the FP64 factors and
    pivots of the per-column launches, bit for bit, and the interchange map that the
    panel kernel's last workgroup writes gives the same matrix as the map kernel's."""
    rng = np.random.default_rng(13)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    a = _factors(ctx, A, lu_panel_lds=2, lu_panel_rows64=128, lu_look_ahead=0)
    b = _factors(ctx, A, lu_panel_lds=0, lu_look_ahead=0)
    assert ctx.get_tuning("lu_panel_rows64") == 64
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[0], b[0])


def test_whole_panel_abort_falls_back_to_one_workgroup(ctx):
    """A zero wait bound makes the whole-panel launches abort at their first
    barrier (forced from the host; every wait still ends).  The last workgroup
    then factors the panel alone: same factors and pivots, bit for bit, and the
    abort counter moves."""
    n = 700
    rng = np.random.default_rng(12)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    a = _factors(ctx, A, lu_panel_lds=0, lu_look_ahead=0)
    before = ctx.get_tuning("lu_panel_aborts")
    b = _factors(ctx, A, lu_panel_lds=1, lu_look_ahead=0, lu_panel_wait_us=0)
    assert ctx.get_tuning("lu_panel_aborts") > max(before, 0)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_whole_panel_concurrent_lus():
    """Four contexts force the whole-panel kernel (lu_panel_lds = 2: every LU uses it,
    whatever else runs) and factor at once on four host threads.  The contexts are
    created late in the process, so their buffers come from recycled device memory:
    the barrier words must be zeroed before any launch reads them.  Each result
    equals, bit for bit, the same LU run alone, and no launch aborts."""
    import threading
    n = 3000
    rng = np.random.default_rng(13)
    mats = [rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 0.1 * t * np.eye(n) for t in range(4)]
    ctxs = [B.Context(0) for _ in range(4)]
    try:
        alone = [_factors(c, A, lu_panel_lds=2) for c, A in zip(ctxs, mats)]
        out = [None] * 4
        errs = []

        def run(t):
            try:
                out[t] = _factors(ctxs[t], mats[t], lu_panel_lds=2)
            except BaseException as exc:   # noqa: BLE001 - reported below
                errs.append(exc)

        th = [threading.Thread(target=run, args=(t,)) for t in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errs, errs
        for a, b in zip(alone, out):
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
        assert all(c.get_tuning("lu_panel_aborts") == 0 for c in ctxs)
    finally:
        for c in ctxs:
            c.close()


def test_concurrent_per_column_lus_match_alone():
    """Four contexts run per-column LUs at once on four host threads (the fused
    panel step's last workgroup reads rows other workgroups have just updated,
    on any XCD, while other streams load the memory system): every factor and
    pivot equals, bit for bit, the same LU run alone.  FP64 and FP32, order
    3,000 (fused steps) and 7,000 (two launches per column)."""
    import threading
    rng = np.random.default_rng(15)
    for n in (3000, 7000):
        mats = [rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) for _ in range(4)]
        ctxs = [B.Context(0) for _ in range(4)]
        try:
            alone = [_factors(c, A, lu_panel_lds=0) for c, A in zip(ctxs, mats)]
            for rep in range(2):
                out = [None] * 4

                def run(t):
                    out[t] = _factors(ctxs[t], mats[t], lu_panel_lds=0)

                th = [threading.Thread(target=run, args=(t,)) for t in range(4)]
                for t in th:
                    t.start()
                for t in th:
                    t.join()
                for a, b in zip(alone, out):
                    for x, y in zip(a, b):
                        assert np.array_equal(x, y), (n, rep)
        finally:
            for c in ctxs:
                c.close()


def test_lus_admitted_to_the_whole_panel_kernel_at_a_later_panel():
    """Four LUs of order 9,000 start at once with the default admission: 141 workgroups
    each against a budget of two per compute unit (512), so at least one starts on the
    per-column kernels and is admitted to the whole-panel kernel at a later, shorter
    panel.  FP64 and FP32 factors and pivots equal, bit for bit, the same LUs run alone
    (both panel paths give the same bits)."""
    import hashlib
    import threading
    n = 9000
    ctxs = [B.Context(0) for _ in range(4)]
    try:
        if ctxs[0].get_tuning("n_cu") != 256:
            pytest.skip("sized for 256 compute units")
        rng = np.random.default_rng(21)
        dA0 = []
        for t in range(4):
            A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
            dA0.append(ctxs[t].to_device(np.ascontiguousarray(A.T)))
            del A
        dA = [c.empty(n * n * 16) for c in ctxs]
        dF = [c.empty(n * n * 8) for c in ctxs]
        piv = [c.empty(4 * n) for c in ctxs]
        piv32 = [c.empty(4 * n) for c in ctxs]

        def factor(t):
            c = ctxs[t]
            c.copy(dA[t], dA0[t], n * n * 16)
            c.lu_factor(dA[t], n, n, piv[t])
            c.lu32_factor(dA0[t], n, n, dF[t], n, piv32[t])

        def digest(t):
            h = hashlib.blake2b()
            for d, shape, dt in ((dA[t], (n, n), np.complex128), (piv[t], (n,), np.int32),
                                 (dF[t], (n, n), np.complex64), (piv32[t], (n,), np.int32)):
                h.update(d.to_numpy(shape, dt).tobytes())
            return h.hexdigest()

        alone = []
        for t in range(4):
            factor(t)
            alone.append(digest(t))
        late0 = sum(c.get_tuning("lu_panel_late_admits") for c in ctxs)
        start = threading.Barrier(4)
        errs = []

        def run(t):
            try:
                start.wait()
                factor(t)
            except BaseException as exc:   # noqa: BLE001 - reported below
                errs.append(exc)

        th = [threading.Thread(target=run, args=(t,)) for t in range(4)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        assert not errs, errs
        assert sum(c.get_tuning("lu_panel_late_admits") for c in ctxs) > late0
        for t in range(4):
            assert digest(t) == alone[t], t
    finally:
        for c in ctxs:
            c.close()


def test_whole_panel_kernel_two_workgroups_per_cu(ctx):
    """Order 17,000 in FP64: the first panels need 266 workgroups of 64 rows, more
    than one per compute unit; factors and pivots equal the per-column path's bit
    for bit (the interchanges of both paths go through the same row gather)."""
    n = 17000
    rng = np.random.default_rng(16)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    a = _factors(ctx, A, fp32=False, lu_panel_lds=1, lu_look_ahead=0)
    b = _factors(ctx, A, fp32=False, lu_panel_lds=0, lu_look_ahead=0)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[0], b[0])


@pytest.mark.parametrize("n", [700, 2500])
def test_two_level_lu_matches_one_level(ctx, n):
    """Outer blocks of 128, 256 and 512 columns (panel updates kept inside the block,
    one update of that rank for the rest) pick the FP64 pivots of the one-level LU
    and give its FP64 factors to rounding.  Both precisions' factors reproduce P A
    to a backward error of n * u * max|A| (LAPACK-style bound; u = 1e-15 FP64,
    1e-6 FP32).  FP32 pivots may differ: the blockings round the triangles'
    columns differently (library TRSM), and near-ties then resolve differently."""
    rng = np.random.default_rng(17)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))

    def backward(F, piv):
        F = F.T.astype(np.complex128)
        L, U = np.tril(F, -1) + np.eye(n), np.triu(F)
        PA = A.copy()
        for j in range(n):
            PA[[j, piv[j]]] = PA[[piv[j], j]]
        return np.abs(PA - L @ U).max() / np.abs(A).max()

    ref = _factors(ctx, A, lu_outer_nb=64)
    for nbo in (128, 256, 512):
        out = _factors(ctx, A, lu_outer_nb=nbo)
        assert np.array_equal(out[1], ref[1])
        assert np.abs(out[0] - ref[0]).max() <= 1e-14 * n * np.abs(ref[0]).max()
        assert backward(out[0], out[1]) <= n * 1e-15
        assert backward(out[2], out[3]) <= n * 1e-6


@pytest.mark.parametrize("mode", [1, 0])
def test_refine_forward_error_stop(ctx, mode):
    """refine_ferr_exp = e: the refinement goes on past the residual target (classic
    steps, mode 0: until the estimated forward error, cond estimated from the
    residual's contraction times the residual, is below 10^-e; FGMRES, mode 1: to
    the FP64 floor); the solution then matches the FP64 solve at least as closely,
    with at least as many corrections."""
    n = 600
    rng = np.random.default_rng(18)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 2 * np.sqrt(n) * np.eye(n)
    b = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    Xr = np.linalg.solve(A, b)
    dA = ctx.to_device(np.ascontiguousarray(A.T))
    dF, piv = ctx.empty(n * n * 8), ctx.empty(4 * n)
    ctx.lu32_factor(dA, n, n, dF, n, piv)
    out = {}
    ctx.set_tuning("refine_mode", mode)
    try:
        for e in (0, 11):
            ctx.set_tuning("refine_ferr_exp", e)
            dB = ctx.to_device(np.ascontiguousarray(b.T))
            it, res = ctx.refine_solve(dA, n, n, dF, n, piv, dB, 3, n, 1e-8, 30)
            X = dB.to_numpy((3, n), np.complex128).T
            out[e] = (it, res, np.linalg.norm(X - Xr) / np.linalg.norm(Xr))
    finally:
        ctx.set_tuning("refine_ferr_exp", 0)
        ctx.set_tuning("refine_mode", 1)
    assert ctx.get_tuning("refine_ferr_exp") == 0
    assert out[11][0] >= out[0][0] and out[11][1] <= out[0][1]
    assert out[11][2] <= max(out[0][2], 1e-13) and out[11][2] <= 1e-11


def _graded(n, smin, seed):
    """Random complex matrix with its two smallest singular values set to 3 smin and
    smin times the largest (cond(A) = 1 / smin)."""
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    U, s, Vh = np.linalg.svd(A)
    s[-1], s[-2] = s[0] * smin, s[0] * smin * 3
    return (U * s) @ Vh, rng


def test_fgmres_converges_where_classic_refinement_fails(ctx):
    """cond(A) = 1e7: cond u32 is near 1, so classic corrections on the FP32 factors
    do not shrink the residual (code 4) or take more steps, while the default
    FGMRES refinement (basis of 20) reaches tol 1e-9, below the residual of the
    FP64 direct solve of the same system, and matches that solve to cond u64; a
    zero right-hand side stays zero."""
    n = 400
    A, rng = _graded(n, 1e-7, 3)
    b = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    b[:, 1] = 0
    Xr = np.linalg.solve(A, b)
    r64 = (np.linalg.norm(A @ Xr - b, axis=0)[[0, 2]] / np.linalg.norm(b, axis=0)[[0, 2]]).max()
    assert r64 > 1e-9   # the direct FP64 solve itself stays above tol
    dA = ctx.to_device(np.ascontiguousarray(A.T))
    dF, piv = ctx.empty(n * n * 8), ctx.empty(4 * n)
    ctx.lu32_factor(dA, n, n, dF, n, piv)
    assert ctx.get_tuning("refine_mode") == 1 and ctx.get_tuning("refine_restart") == 20
    dB = ctx.to_device(np.ascontiguousarray(b.T))
    it, res = ctx.refine_solve(dA, n, n, dF, n, piv, dB, 3, n, 1e-9, 60)
    X = dB.to_numpy((3, n), np.complex128).T
    rel = np.linalg.norm(A @ X - b, axis=0)[[0, 2]] / np.linalg.norm(b, axis=0)[[0, 2]]
    assert res <= 1e-9 and rel.max() <= 2e-9 and 1 <= it <= 60
    assert not np.any(X[:, 1])
    assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) <= 1e-7
    ctx.set_tuning("refine_mode", 0)
    try:
        dB = ctx.to_device(np.ascontiguousarray(b.T))
        try:
            it0, _ = ctx.refine_solve(dA, n, n, dF, n, piv, dB, 3, n, 1e-9, 60)
            assert it0 > it
        except B.BielccError as err:
            assert err.code == 4
    finally:
        ctx.set_tuning("refine_mode", 1)


def _two_prod(a, b):
    """Exact products a b = p + e of float64 arrays (Dekker's splitting, no FMA)."""
    def split(x):
        c = 134217729.0 * x
        hi = c - (c - x)
        return hi, x - hi
    p = a * b
    ah, al = split(a)
    bh, bl = split(b)
    return p, ((ah * bh - p) + ah * bl + al * bh) + al * bl


def _exact_residual(A, X, Bm):
    """B - A X rounded once to FP64: every product split exactly, each sum by math.fsum."""
    import math
    n, nr = X.shape
    R = np.empty((n, nr), complex)
    for c in range(nr):
        terms_re, terms_im = [], []
        for (a, x, sgn, im) in ((A.real, X[:, c].real, -1, 0), (A.imag, X[:, c].imag, 1, 0),
                                (A.real, X[:, c].imag, -1, 1), (A.imag, X[:, c].real, -1, 1)):
            p, e = _two_prod(a, x[None, :])
            (terms_im if im else terms_re).append((sgn * p, sgn * e))
        for r in range(n):
            re = [Bm[r, c].real] + [t[r, j] for pe in terms_re for t in pe for j in range(n)]
            ii = [Bm[r, c].imag] + [t[r, j] for pe in terms_im for t in pe for j in range(n)]
            R[r, c] = complex(math.fsum(re), math.fsum(ii))
    return R


def test_residual_doubled_is_exactly_rounded(ctx):
    """bielcc_residual with doubled sums against the exactly rounded B - A X (products
    split exactly, sums by math.fsum) where B = fl(A X) + a few ulps, so that the true
    residual is tiny next to |A||X|: doubled, within a few ulps of the exact residual
    (several column chunks, a second group of right-hand sides); FP64 sums, wrong by
    far more than the residual itself."""
    rng = np.random.default_rng(31)
    n, nr = 200, 9
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    X = rng.normal(size=(n, nr)) + 1j * rng.normal(size=(n, nr))
    Bm = A @ X
    Bm = Bm + 4e-16 * np.abs(Bm) * (rng.normal(size=Bm.shape) + 1j * rng.normal(size=Bm.shape))
    Rx = _exact_residual(A, X, Bm)
    dA = ctx.to_device(np.ascontiguousarray(A.T))
    dX = ctx.to_device(np.ascontiguousarray(X.T))
    dB = ctx.to_device(np.ascontiguousarray(Bm.T))
    out = {}
    for doubled in (True, False):
        dR = ctx.empty(n * nr * 16)
        ctx.residual(dA, n, n, dX, n, dB, n, nr, dR, n, doubled=doubled)
        out[doubled] = dR.to_numpy((nr, n), np.complex128).T
    scale = np.abs(Rx).max()
    assert np.abs(out[True] - Rx).max() <= 1e-12 * scale
    assert np.abs(out[False] - Rx).max() >= 1e-2 * scale
    with pytest.raises(B.BielccError):
        ctx.residual(dA, n, n - 1, dX, n, dB, n, nr, ctx.empty(16 * n * nr), n)


def test_fgmres_doubled_residual_accuracy(ctx):
    """cond(A) = 1e6 with a forward-error request: FGMRES with FP64 residuals stops
    at their rounding floor; switching to doubled residuals there (the default) takes
    the solution on to about u64 of the exact one (reference: LAPACK FP64 LU refined
    with 80-bit residuals), ten times closer than LAPACK's own FP64 solve."""
    n = 400
    A, rng = _graded(n, 1e-6, 5)
    b = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    L = np.longdouble

    def resid_ext(X):
        Ar, Ai, Xr, Xi = A.real.astype(L), A.imag.astype(L), X.real.astype(L), X.imag.astype(L)
        return ((b.real.astype(L) - (Ar @ Xr - Ai @ Xi)).astype(float)
                + 1j * (b.imag.astype(L) - (Ar @ Xi + Ai @ Xr)).astype(float))
    import scipy.linalg as sl
    lu = sl.lu_factor(A)
    X64 = sl.lu_solve(lu, b)
    Xt = X64.copy()
    for _ in range(8):
        Xt = Xt + sl.lu_solve(lu, resid_ext(Xt))
    err = lambda X: np.linalg.norm(X - Xt) / np.linalg.norm(Xt)
    dA = ctx.to_device(np.ascontiguousarray(A.T))
    dF, piv = ctx.empty(n * n * 8), ctx.empty(4 * n)
    ctx.lu32_factor(dA, n, n, dF, n, piv)
    assert ctx.get_tuning("refine_resid_dd") == 1
    out = {}
    ctx.set_tuning("refine_ferr_exp", 7)
    try:
        for mode in (0, 1):
            ctx.set_tuning("refine_resid_dd", mode)
            dB = ctx.to_device(np.ascontiguousarray(b.T))
            it, res = ctx.refine_solve(dA, n, n, dF, n, piv, dB, 2, n, 1e-10, 60)
            out[mode] = err(dB.to_numpy((2, n), np.complex128).T)
    finally:
        ctx.set_tuning("refine_ferr_exp", 0)
        ctx.set_tuning("refine_resid_dd", 1)
    assert out[1] <= 1e-12 and out[1] <= 0.1 * err(X64) and out[1] <= 0.1 * out[0], (out, err(X64))


def test_fgmres_restart_sizes(ctx):
    """Bases of 1, 2 and 20 vectors per cycle (1: each cycle is one classic step
    preconditioned and minimised along one direction) converge on a well-conditioned
    system to 1e-12 with the same solution; the tunables reject sizes outside 1..64
    and modes other than 0 and 1."""
    n = 500
    rng = np.random.default_rng(21)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + np.sqrt(n) * np.eye(n)
    b = rng.normal(size=(n, 4)) + 1j * rng.normal(size=(n, 4))
    Xr = np.linalg.solve(A, b)
    dA = ctx.to_device(np.ascontiguousarray(A.T))
    dF, piv = ctx.empty(n * n * 8), ctx.empty(4 * n)
    ctx.lu32_factor(dA, n, n, dF, n, piv)
    try:
        for m in (1, 2, 20):
            ctx.set_tuning("refine_restart", m)
            dB = ctx.to_device(np.ascontiguousarray(b.T))
            it, res = ctx.refine_solve(dA, n, n, dF, n, piv, dB, 4, n, 1e-12, 30)
            X = dB.to_numpy((4, n), np.complex128).T
            assert res <= 1e-12 and np.linalg.norm(X - Xr) / np.linalg.norm(Xr) <= 1e-11, m
        for key, bad in (("refine_restart", 0), ("refine_restart", 65), ("refine_mode", 2)):
            with pytest.raises(B.BielccError):
                ctx.set_tuning(key, bad)
    finally:
        ctx.set_tuning("refine_restart", 20)
