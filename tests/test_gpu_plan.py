"""Plans: several meshes per persistent launch, operator-matrix fields, graphs.

- Assembly plan over uneven meshes (quads and triangles, real and complex k,
  FP64 and mixed) == bielcc_assemble_ex per mesh.  Each matrix entry is one
  (point, cell) pair evaluated by the same tile code; the tile loops may fuse
  multiply-adds differently in the two kernels, so the bar is 1e-13 per block
  (the near blocks, contraction-free, match bitwise).
- Field plan (operator matrices + complex GEMM) == bielcc_field_batch per mesh,
  with near pairs (both near rules), incident field and H, to 1e-12 per point:
  only the order of the cross-cell sums differs.  The mixed form (FP32 node
  sums) is held to its own bar, 1e-6 per point against the FP64 field: its
  FP32 roundings depend on how the compiler fuses the node loop.
- A workspace too small for one item splits it into point ranges and groups of
  launches; the result does not change.
- Captured plans replay from a HIP graph: several replays with new currents and
  eager work in between, each equal to the eager result.
"""
import numpy as np
import pytest

import bielcc_amd as B
import oracle as O
from _util import block_metric, rowwise_rel

pytestmark = pytest.mark.gpu


def _meshes():
    sph = B.mesh.sphere_quads(8, 16, jitter=0.005, seed=3)
    ell = B.mesh.triangulate(B.mesh.ellipsoid_quads(6, 12, axes=(1.2, 0.8, 0.6), seed=4))
    tor = B.mesh.torus_quads(24, 8, R=1.0, r=0.3)
    return [sph, ell, tor]


def _points(rng, m, lo, hi):
    p = rng.normal(size=(m, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    return p * rng.uniform(lo, hi, size=(m, 1))


@pytest.mark.parametrize("precision", [B.PREC_FP64, B.PREC_MIXED])
def test_assemble_plan_equals_per_mesh(ctx, precision):
    npp = B.num_param()
    verts = _meshes()
    ks = [2.0, 3.5 + 0.25j, 6.0]
    meshes = [ctx.mesh(v) for v in verts]
    Zs, items = [], []
    for m, k in zip(meshes, ks):
        n2 = 2 * m.n
        Z = ctx.empty(n2 * n2 * 16)
        Zs.append(Z)
        items.append((m, k, Z, n2))
    ctx.set_tuning("plan_asm_cells", 40)          # ragged last cell chunk on every mesh
    plan = ctx.plan_assemble(items, npp, precision=precision)
    assert plan.info("segments") == 3 and plan.info("tiles") > 3
    plan.run()
    ctx.sync()
    for (m, k, Z, n2) in items:
        got = Z.to_numpy((n2, n2), np.complex128).T
        ref = m.assemble(npp, k, precision=precision)
        err = block_metric(got, ref).max()
        assert err <= 1e-13, (m.n, k, err)
    # a graph replay writes the same matrices again
    for (_, _, Z, n2) in items:
        ctx.upload(Z, np.zeros((n2, n2), np.complex128))
    plan.capture()
    plan.run()
    ctx.sync()
    for (m, k, Z, n2) in items:
        err = block_metric(Z.to_numpy((n2, n2), np.complex128).T, m.assemble(npp, k, precision=precision)).max()
        assert err <= 1e-13, (m.n, k, err)
    plan.close()
    ctx.set_tuning("plan_asm_cells", 64)
    for m in meshes:
        m.close()


def test_assemble_plan_against_oracle(ctx, orc):
    """One small mesh through a plan against the C restatement of the reference."""
    verts = B.mesh.sphere_quads(5, 10, jitter=0.005, seed=9)
    m = ctx.mesh(verts)
    n2 = 2 * m.n
    Z = ctx.empty(n2 * n2 * 16)
    plan = ctx.plan_assemble([(m, 2.5, Z, n2)], B.num_param())
    plan.run()
    ctx.sync()
    got = Z.to_numpy((n2, n2), np.complex128).T
    ref = orc.assemble(verts, O.num_param(), 2.5)
    assert np.abs(got - ref).max() / np.abs(ref).max() <= 1e-12
    plan.close()
    m.close()


def _field_setup(ctx, rng, precision, near_mode, with_h=True):
    verts = _meshes()
    ks = [2.0, 3.0 + 0.2j, 4.5]
    nrhs = [5, 3, 9]
    Ms = [300, 41, 517]
    meshes = [ctx.mesh(v) for v in verts]
    items, host = [], []
    for m, k, nr, M in zip(meshes, ks, nrhs, Ms):
        jE = rng.normal(size=(nr, m.n, 3)) + 1j * rng.normal(size=(nr, m.n, 3))
        pts = np.concatenate([_points(rng, M // 4, 1.0, 1.3), _points(rng, M - M // 4, 2.0, 4.0)])
        kv = rng.normal(size=(nr, 3))
        ev = rng.normal(size=(nr, 3))
        inc = np.concatenate([kv, ev], axis=1)
        d = dict(mesh=m, k=k, nrhs=nr, M=M, jE=ctx.to_device(jE), pts=ctx.to_device(pts), inc=ctx.to_device(inc),
                 E=ctx.empty(nr * M * 48), H=ctx.empty(nr * M * 48) if with_h else None, multE=0.7 - 0.2j)
        items.append(d)
        host.append(dict(jE=jE, pts=pts, inc=inc))
    return meshes, items, host


def _field_ref(m, it, h, npp, near_mode, precision):
    return m.field_batch(npp, it["k"], h["jE"], h["pts"], multE=it["multE"], incident=(h["inc"][:, :3], h["inc"][:, 3:]),
                         near_mode=near_mode, precision=precision)


def _check_field(items, host, npp, near_mode, precision):
    tol = 1e-12 if precision == B.PREC_FP64 else 1e-6
    for it, h in zip(items, host):
        m, nr, M = it["mesh"], it["nrhs"], it["M"]
        E = it["E"].to_numpy((nr, M, 3), np.complex128)
        Eref, Href = _field_ref(m, it, h, npp, near_mode, B.PREC_FP64)
        assert rowwise_rel(E.reshape(-1, 3), Eref.reshape(-1, 3)).max() <= tol, m.n
        if it["H"] is not None:
            H = it["H"].to_numpy((nr, M, 3), np.complex128)
            assert rowwise_rel(H.reshape(-1, 3), Href.reshape(-1, 3)).max() <= tol, m.n


@pytest.mark.parametrize("precision,near_mode", [(B.PREC_FP64, 0), (B.PREC_FP64, 1), (B.PREC_MIXED, 0)])
def test_field_plan_equals_field_batch(ctx, precision, near_mode):
    rng = np.random.default_rng(11)
    npp = B.num_param()
    meshes, items, host = _field_setup(ctx, rng, precision, near_mode)
    ctx.set_tuning("plan_field_cells", 48)
    plan = ctx.plan_field(items, npp, near_mode=near_mode, precision=precision)
    assert plan.info("near_pairs") > 0 and plan.info("launches") == 1
    plan.run()
    ctx.sync()
    _check_field(items, host, npp, near_mode, precision)
    plan.close()
    ctx.set_tuning("plan_field_cells", 64)
    for m in meshes:
        m.close()


def test_field_plan_split_workspace(ctx):
    """A 16 MiB cap splits the items into point ranges and several launches."""
    rng = np.random.default_rng(12)
    npp = B.num_param()
    meshes, items, host = _field_setup(ctx, rng, B.PREC_FP64, 0)
    ctx.set_tuning("plan_ws_mb", 16)
    plan = ctx.plan_field(items, npp)
    ctx.set_tuning("plan_ws_mb", 24576)
    assert plan.info("launches") > 1 and plan.info("segments") > 3
    assert plan.info("workspace_bytes") <= 16 << 20
    plan.run()
    ctx.sync()
    _check_field(items, host, npp, 0, B.PREC_FP64)
    plan.close()
    for m in meshes:
        m.close()


def test_field_plan_graph_replays(ctx):
    """Capture once, then replay with new currents and incidences; eager work
    (another mesh's field) runs between the replays."""
    rng = np.random.default_rng(13)
    npp = B.num_param()
    meshes, items, host = _field_setup(ctx, rng, B.PREC_MIXED, 0)
    plan = ctx.plan_field(items, npp, precision=B.PREC_MIXED)
    plan.capture()
    assert plan.info("captured") == 1
    other = ctx.mesh(B.mesh.sphere_quads(6, 12, seed=8))
    for rep in range(3):
        for it, h in zip(items, host):
            h["jE"] = rng.normal(size=h["jE"].shape) + 1j * rng.normal(size=h["jE"].shape)
            h["inc"] = rng.normal(size=h["inc"].shape)
            ctx.upload(it["jE"], h["jE"])
            ctx.upload(it["inc"], h["inc"])
        plan.run()
        other.field(npp, 2.0, rng.normal(size=(other.n, 3)) + 0j, _points(rng, 64, 2.0, 3.0))
        ctx.sync()
        _check_field(items, host, npp, 0, B.PREC_MIXED)
    plan.close()
    other.close()
    for m in meshes:
        m.close()


def test_mixed_field_plan_switches_to_fp64_sums(ctx):
    """Two copies of a sphere 1e-4 apart carry currents J and -J: their fields cancel to
    about 1e-4 of either, beyond what FP32 node sums resolve, so the FP32 error control
    recomputes every point.  After that run the mixed plan sums in FP64: from then on its
    results, eager and from a graph captured again, are those of the FP64 plan bit for bit.
    With mixed_probe_pct = 0 the plan keeps FP32 sums (and the error control)."""
    rng = np.random.default_rng(31)
    npp = B.num_param()
    v = B.mesh.sphere_quads(8, 16, seed=3)
    m = ctx.mesh(np.concatenate([v, v + np.array([1e-4, 0.0, 0.0])]))
    nr, M, k = 4, 600, 2.0
    J = rng.normal(size=(nr, v.shape[0], 3)) + 1j * rng.normal(size=(nr, v.shape[0], 3))
    jE = np.concatenate([J, -J], axis=1)
    pts = _points(rng, M, 2.0, 4.0)

    def items():
        return [dict(mesh=m, k=k, nrhs=nr, M=M, jE=ctx.to_device(jE), pts=ctx.to_device(pts), inc=None,
                     E=ctx.empty(nr * M * 48), H=ctx.empty(nr * M * 48), multE=1.0)]

    def fields(it):
        ctx.sync()
        return it["E"].to_numpy((nr, M, 3), np.complex128), it["H"].to_numpy((nr, M, 3), np.complex128)

    i64 = items()
    p64 = ctx.plan_field(i64, npp, precision=B.PREC_FP64)
    p64.run()
    E64, H64 = fields(i64[0])
    assert ctx.get_tuning("mixed_probe_pct") == 25
    for capture in (False, True):
        it = items()
        plan = ctx.plan_field(it, npp, precision=B.PREC_MIXED)
        if capture:
            plan.capture()
        fixed = ctx.get_tuning("mixed_fixup_points")
        plan.run()   # FP32 sums, every point recomputed
        E, H = fields(it[0])
        assert ctx.get_tuning("mixed_fixup_points") - fixed == M
        assert plan.info("fp64_sums") == 1 and plan.info("captured") == (1 if capture else 0)
        assert rowwise_rel(E.reshape(-1, 3), E64.reshape(-1, 3)).max() <= 1e-6
        for _ in range(2):
            ctx.upload(it[0]["E"], np.zeros((nr, M, 3), np.complex128))
            fixed = ctx.get_tuning("mixed_fixup_points")
            plan.run()
            E, H = fields(it[0])
            assert ctx.get_tuning("mixed_fixup_points") == fixed
            assert np.array_equal(E, E64) and np.array_equal(H, H64)
        plan.close()
    ctx.set_tuning("mixed_probe_pct", 0)
    try:
        it = items()
        plan = ctx.plan_field(it, npp, precision=B.PREC_MIXED)
        for _ in range(2):
            fixed = ctx.get_tuning("mixed_fixup_points")
            plan.run()
            E, H = fields(it[0])
            assert ctx.get_tuning("mixed_fixup_points") - fixed == M and plan.info("fp64_sums") == 0
            assert rowwise_rel(E.reshape(-1, 3), E64.reshape(-1, 3)).max() <= 1e-6
        plan.close()
    finally:
        ctx.set_tuning("mixed_probe_pct", 25)
    p64.close()
    m.close()


def test_field_plan_without_h_and_without_incidence(ctx):
    rng = np.random.default_rng(14)
    npp = B.num_param()
    m = ctx.mesh(B.mesh.sphere_quads(7, 14, seed=2))
    nr, M = 4, 200
    jE = rng.normal(size=(nr, m.n, 3)) + 1j * rng.normal(size=(nr, m.n, 3))
    pts = _points(rng, M, 2.0, 3.0)
    dE = ctx.empty(nr * M * 48)
    plan = ctx.plan_field([dict(mesh=m, k=3.0, nrhs=nr, M=M, jE=ctx.to_device(jE), pts=ctx.to_device(pts), E=dE)],
                          npp)
    plan.run()
    ctx.sync()
    E = dE.to_numpy((nr, M, 3), np.complex128)
    Eref, _ = m.field_batch(npp, 3.0, jE, pts)
    assert rowwise_rel(E.reshape(-1, 3), Eref.reshape(-1, 3)).max() <= 1e-12
    plan.close()
    m.close()


def test_plan_argument_errors(ctx):
    npp = B.num_param()
    m = ctx.mesh(B.mesh.sphere_quads(4, 8, seed=1))
    Z = ctx.empty(16)
    with pytest.raises(B.BielccError):
        ctx.plan_assemble([(m, 2.0, Z, 1)], npp)            # ld < 2N
    bad = B.num_param(p_max=2)
    with pytest.raises(B.BielccError):
        ctx.plan_assemble([(m, 2.0, Z, 2 * m.n)], bad)      # adaptive passes: not a tile plan
    with pytest.raises(B.BielccError):
        ctx.set_tuning("plan_wg_per_cu", 0)
    m.close()


@pytest.mark.parametrize("precision,near_mode,with_h", [(B.PREC_FP64, 0, True), (B.PREC_FP64, 1, False),
                                                        (B.PREC_MIXED, 0, True)])
def test_field_batch_many_sets_through_operator_matrices(ctx, precision, near_mode, with_h):
    """bielcc_field_batch with at least opmat_min_rhs current sets takes the
    operator-matrix path; it equals the pass-per-4-sets path (FP64: 1e-12 per
    point; mixed: 1e-6 against FP64), chunked or not."""
    rng = np.random.default_rng(15)
    npp = B.num_param()
    m = ctx.mesh(B.mesh.triangulate(B.mesh.sphere_quads(7, 14, seed=6)))
    nr, M = 11, 700
    jE = rng.normal(size=(nr, m.n, 3)) + 1j * rng.normal(size=(nr, m.n, 3))
    pts = np.concatenate([_points(rng, 200, 1.0, 1.2), _points(rng, M - 200, 2.0, 4.0)])
    inc = (rng.normal(size=(nr, 3)), rng.normal(size=(nr, 3)))
    ctx.set_tuning("opmat_min_rhs", 1000)
    Er, Hr = m.field_batch(npp, 2.5, jE, pts, multE=0.4 + 0.9j, incident=inc, near_mode=near_mode)
    tol = 1e-12 if precision == B.PREC_FP64 else 1e-6
    for ws_mb in (24576, 16):          # 16 MiB: several point chunks through one workspace
        ctx.set_tuning("opmat_min_rhs", 9)
        ctx.set_tuning("plan_ws_mb", ws_mb)
        E, H = m.field_batch(npp, 2.5, jE, pts, multE=0.4 + 0.9j, incident=inc, near_mode=near_mode,
                             precision=precision)
        ctx.set_tuning("plan_ws_mb", 24576)
        assert rowwise_rel(E.reshape(-1, 3), Er.reshape(-1, 3)).max() <= tol, ws_mb
        if with_h:
            assert rowwise_rel(H.reshape(-1, 3), Hr.reshape(-1, 3)).max() <= tol, ws_mb
    m.close()
