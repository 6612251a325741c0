"""Scattered field on the GPU against the reference's GetFieldE_Ideal output
(config 1) and the oracle (near field, magnetic currents, the example's
near rule, lossy media).  Bar: 1e-10 relative per point."""
import numpy as np
import pytest

import bielcc_amd as B
import oracle as O
from _util import golden, np_dict, rowwise_rel

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.mark.parametrize("fname", ["config1.npz", "config1_tri.npz"])
def test_get_field_e_vs_reference(ctx, fname):
    g = golden(fname)
    m = ctx.mesh(g["verts"])
    npp = B.num_param(**np_dict(g["num_param"]))
    E = m.get_field_e(g["jE"], 1.0, g["kvec"], g["e0"], npp, float(g["k"]), g["pts"])
    assert rowwise_rel(E, g["E"]).max() <= TOL
    Ef, Hf = m.field(npp, float(g["k"]), g["jE"], g["pts"], incident=(g["kvec"], g["e0"]))
    assert rowwise_rel(Ef, g["E"]).max() <= TOL
    assert rowwise_rel(Hf, g["H"]).max() <= TOL
    m.close()


def _near_points(rng, m_pts, radius):
    p = rng.normal(size=(m_pts, 3))
    return radius * p / np.linalg.norm(p, axis=1, keepdims=True)


@pytest.mark.parametrize("near_mode", [0, 1])
def test_near_field_vs_oracle(ctx, orc, near_mode):
    rng = np.random.default_rng(21)
    verts = B.mesh.sphere_quads(10, 20, jitter=0.005, seed=5)
    n = verts.shape[0]
    jE = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    jM = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    pts = np.concatenate([_near_points(rng, 24, 1.04), _near_points(rng, 8, 0.97), _near_points(rng, 8, 2.5)])
    k = 3.0
    npp = B.num_param()
    m = ctx.mesh(verts)
    E, H = m.field(npp, k, jE, pts, jM=jM, multE=0.3 + 1j, multH=-0.5j, near_mode=near_mode)
    Eo, Ho = orc.field_eh(verts, jE, k, O.num_param(), pts, jM=jM, multE=0.3 + 1j, multH=-0.5j, near_mode=near_mode)
    assert rowwise_rel(E, Eo).max() <= TOL
    assert rowwise_rel(H, Ho).max() <= TOL
    m.close()


@pytest.mark.parametrize("near_mode", [0, 1])
def test_near_list_shortcut_at_its_boundary(ctx, orc, near_mode):
    """The near-pair scan is skipped by workgroups of 256 points that all lie beyond the
    sphere around the cell centres plus the largest near distance (3 grid steps, both
    rules).  One workgroup of points just inside that boundary (scanned, near pairs among
    them) and one just beyond it (skipped): the field at every point equals the oracle's."""
    rng = np.random.default_rng(23)
    verts = B.mesh.sphere_quads(10, 20, jitter=0.005, seed=5)
    n = verts.shape[0]
    m = ctx.mesh(verts)
    c = m.tables()["center"]
    bc = 0.5 * (c.min(axis=0) + c.max(axis=0))
    lim = np.linalg.norm(c - bc, axis=1).max() + 3.0 * m.grid_step()
    d = rng.normal(size=(512, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.concatenate([np.sort(rng.uniform(lim - 0.25, lim - 1e-3, 256)), np.sort(rng.uniform(lim + 1e-3, lim + 0.25, 256))])
    pts = bc + r[:, None] * d
    jE = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    k = 3.0
    E, H = m.field(B.num_param(), k, jE, pts, near_mode=near_mode)
    Eo, Ho = orc.field_eh(verts, jE, k, O.num_param(), pts, near_mode=near_mode)
    assert rowwise_rel(E, Eo).max() <= TOL
    assert rowwise_rel(H, Ho).max() <= TOL
    m.close()


def test_triangles_complex_k_vs_oracle(ctx, orc):
    rng = np.random.default_rng(22)
    verts = B.mesh.triangulate(B.mesh.sphere_quads(8, 16, jitter=0.005, seed=6))
    n = verts.shape[0]
    jE = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    pts = np.concatenate([_near_points(rng, 16, 1.05), _near_points(rng, 16, 3.0)])
    k = 2.5 + 0.4j
    m = ctx.mesh(verts)
    E, H = m.field(B.num_param(), k, jE, pts, incident=((0., 0., 2.5), (1., 0., 0.)))
    Eo, Ho = orc.field_eh(verts, jE, k, O.num_param(), pts, incident=((0., 0., 2.5), (1., 0., 0.)))
    assert rowwise_rel(E, Eo).max() <= TOL
    assert rowwise_rel(H, Ho).max() <= TOL
    m.close()


def test_reduced_phase_and_general_sincos_paths(ctx, orc, monkeypatch):
    """The tile kernels' reduced-range e^{ikr} (one quadrant per cell) and the
    general sincos give the oracle's values to the same bar."""
    verts = B.mesh.sphere_quads(16, 32, jitter=0.005, seed=1)
    n = verts.shape[0]
    rng = np.random.default_rng(24)
    jE = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    pts = B.mesh.sphere_points(40, 3.0, seed=5)
    k = 2.0
    npp = B.num_param()
    Eo, Ho = orc.field_eh(verts, jE, k, O.num_param(), pts)
    Zo = orc.assemble(verts[:60], O.num_param(), k)
    monkeypatch.setenv("BIELCC_GENERAL_SINCOS", "1")
    cg = B.Context(0)
    for c in (ctx, cg):
        m = c.mesh(verts)
        E, H = m.field(npp, k, jE, pts)
        assert rowwise_rel(E, Eo).max() <= TOL
        assert rowwise_rel(H, Ho).max() <= TOL
        m2 = c.mesh(verts[:60])
        Z = m2.assemble(npp, k)
        from _util import block_metric
        assert block_metric(Z, Zo).max() <= 1e-12
        m.close(); m2.close()
    cg.close()


@pytest.mark.parametrize("t_max", [0.2, 0.4, 0.9])
def test_phase_modes_vs_oracle(ctx, orc, t_max):
    """k chosen so that |k| max|y - c| (the largest node phase offset from a cell's centre)
    is 0.2, 0.4 or 0.9: the centre-phase series on |t| <= 0.25 (phase mode 3) and on
    |t| <= 0.5 (mode 2), and the quadrant reduction (mode 1).  Field and matrix against the
    oracle."""
    verts = B.mesh.sphere_quads(16, 32, jitter=0.005, seed=1)
    n = verts.shape[0]
    m = ctx.mesh(verts)
    cen = m.tables()["center"]
    max_rad = np.linalg.norm(verts - cen[:, None, :], axis=2).max()
    k = t_max / max_rad
    rng = np.random.default_rng(25)
    jE = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    pts = B.mesh.sphere_points(40, 3.0, seed=5)
    E, H = m.field(B.num_param(), k, jE, pts)
    Eo, Ho = orc.field_eh(verts, jE, k, O.num_param(), pts)
    assert rowwise_rel(E, Eo).max() <= TOL
    assert rowwise_rel(H, Ho).max() <= TOL
    m.close()
    m2 = ctx.mesh(verts[:60])
    from _util import block_metric
    assert block_metric(m2.assemble(B.num_param(), k), orc.assemble(verts[:60], O.num_param(), k)).max() <= 1e-12
    m2.close()


def test_field_linearity_in_currents(ctx):
    """E and H are linear in (jE, jM): a size-independent property."""
    rng = np.random.default_rng(23)
    verts = B.mesh.sphere_quads(8, 16, seed=8)
    n = verts.shape[0]
    j1 = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    j2 = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    pts = B.mesh.sphere_points(64, 2.0, seed=3)
    m = ctx.mesh(verts)
    npp = B.num_param()
    E1, H1 = m.field(npp, 2.0, j1, pts)
    E2, H2 = m.field(npp, 2.0, j2, pts)
    E3, H3 = m.field(npp, 2.0, 2.0 * j1 - 0.5j * j2, pts)
    assert rowwise_rel(E3, 2.0 * E1 - 0.5j * E2).max() <= 1e-12
    assert rowwise_rel(H3, 2.0 * H1 - 0.5j * H2).max() <= 1e-12
    m.close()


@pytest.mark.parametrize("t_max,auto64", [(0.2, True), (0.9, False)])
def test_mixed_far_sums_rule(ctx, t_max, auto64):
    """mixed_far32 = 0 (the library default): a mixed field call with real k whose node phases use
    the centre-phase series (|k| max|y - c| <= 0.5) runs with FP64 far sums, the FP64 field bit for
    bit; beyond that (quadrant reduction) it keeps the FP32 sums.  mixed_far32 = 1 keeps the FP32
    sums everywhere.  The FP32 error control is off here (mixed_fixup = 0), so FP32 sums show as
    a difference from the FP64 field (set-wise within 1e-5 for these random currents)."""
    verts = B.mesh.sphere_quads(16, 32, jitter=0.005, seed=1)
    n = verts.shape[0]
    m = ctx.mesh(verts)
    cen = m.tables()["center"]
    k = t_max / np.linalg.norm(verts - cen[:, None, :], axis=2).max()
    rng = np.random.default_rng(26)
    jE = rng.normal(size=(2, n, 3)) + 1j * rng.normal(size=(2, n, 3))
    pts = B.mesh.sphere_points(300, 3.0, seed=6)
    npp = B.num_param()
    E64, H64 = m.field_batch(npp, k, jE, pts)
    old, oldfix = ctx.get_tuning("mixed_far32"), ctx.get_tuning("mixed_fixup")
    try:
        ctx.set_tuning("mixed_fixup", 0)
        ctx.set_tuning("mixed_far32", 0)
        a0 = ctx.get_tuning("mixed_auto_fp64")
        Ea, Ha = m.field_batch(npp, k, jE, pts, precision=B.PREC_MIXED)
        assert ctx.get_tuning("mixed_auto_fp64") == a0 + (1 if auto64 else 0)
        ctx.set_tuning("mixed_far32", 1)
        E32, H32 = m.field_batch(npp, k, jE, pts, precision=B.PREC_MIXED)
        assert ctx.get_tuning("mixed_auto_fp64") == a0 + (1 if auto64 else 0)
    finally:
        ctx.set_tuning("mixed_far32", old)
        ctx.set_tuning("mixed_fixup", oldfix)
    assert np.array_equal(Ea, E64) == auto64 and np.array_equal(Ha, H64) == auto64
    assert not np.array_equal(E32, E64) and not np.array_equal(H32, H64)
    for X, Y in ((Ea, E64), (Ha, H64), (E32, E64), (H32, H64)):
        assert np.linalg.norm(X - Y) <= 1e-5 * np.linalg.norm(Y)
    m.close()
