"""bielcc_field_noalloc (the field driver with its work memory in a fixed
per-GPU device array, no allocation) against bielcc_field on the same inputs:
two calls in a row, the second with a smaller input, each within 1e-10 per
point (bit for bit when the cell split is the same); inputs beyond its scratch
and meshes without node tables are rejected with an error, not run."""
import numpy as np
import pytest

import bielcc_amd as B

pytestmark = pytest.mark.gpu


def _near_and_far(rng, m_near, m_far, r_near, r_far):
    p = rng.normal(size=(m_near + m_far, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    p[:m_near] *= r_near
    p[m_near:] *= r_far
    return p


def _both(ctx, m, npp, k, jE, jM, pts, incident, multE, multH):
    M = len(pts)
    dp, dje = ctx.to_device(pts), ctx.to_device(jE)
    djm = ctx.to_device(jM) if jM is not None else None
    out = []
    for use_noalloc in (False, True):
        dE, dH = ctx.empty(M * 48), ctx.empty(M * 48)
        if use_noalloc:
            m.field_noalloc_into(npp, k, dje, dp, M, dE, dH, djm, multE, multH, incident)
        else:
            m.field_into(npp, k, dje, dp, M, dE, dH, djm, multE, multH, incident)
        ctx.sync()
        out.append((dE.to_numpy((M, 3), np.complex128), dH.to_numpy((M, 3), np.complex128)))
    return out


def _rel(a, b):
    nb = np.linalg.norm(b, axis=1)
    return float((np.linalg.norm(a - b, axis=1) / np.where(nb == 0, 1.0, nb)).max())


def test_noalloc_field_matches_field_twice_then_smaller():
    ctx = B.Context(0)
    rng = np.random.default_rng(91)
    npp = B.num_param()
    big = B.mesh.sphere_quads(24, 48, jitter=0.005, seed=3)          # 1,152 cells
    small = B.mesh.triangulate(B.mesh.sphere_quads(8, 16, jitter=0.005, seed=4))
    cases = [(big, 3000, 600, 2.5, None, True), (small, 400, 150, 1.7 + 0.1j, True, False)]
    for verts, m_far, m_near, k, with_jm, with_inc in cases:
        m = ctx.mesh(verts)
        n = verts.shape[0]
        jE = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
        jM = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3)) if with_jm else None
        pts = _near_and_far(rng, m_near, m_far, 1.02, 3.0)
        inc = ((0.0, 0.0, 2.5), (1.0, 0.0, 0.0)) if with_inc else None
        # the mesh's node tables exist only after one call of the allocating driver
        with pytest.raises(B.BielccError) as e:
            dp, dje = ctx.to_device(pts[:8]), ctx.to_device(jE)
            m.field_noalloc_into(npp, k, dje, dp, 8, ctx.empty(8 * 48), ctx.empty(8 * 48))
        assert e.value.code == 4 and "node" in str(e.value)
        (Ea, Ha), (Eb, Hb) = _both(ctx, m, npp, k, jE, jM, pts, inc, 0.3 + 1j, -0.5j)
        assert np.isfinite(Eb).all() and np.isfinite(Hb).all()
        assert _rel(Eb, Ea) <= 1e-10 and _rel(Hb, Ha) <= 1e-10
        m.close()
    ctx.close()


def test_noalloc_field_rejects_inputs_beyond_its_scratch():
    ctx = B.Context(0)
    verts = B.mesh.sphere_quads(6, 12, jitter=0.005, seed=5)
    m = ctx.mesh(verts)
    npp = B.num_param()
    n = verts.shape[0]
    jE = np.ones((n, 3), dtype=np.complex128)
    M = 700_000                                        # 108 B per point > 64 MiB
    dp = ctx.to_device(np.tile([0.0, 0.0, 3.0], (M, 1)))
    dE, dH = ctx.empty(M * 48), ctx.empty(M * 48)
    dje = ctx.to_device(jE)
    m.field_into(npp, 1.0, dje, dp, 1, dE, dH)           # builds the node tables
    ctx.sync()
    with pytest.raises(B.BielccError) as e:
        m.field_noalloc_into(npp, 1.0, dje, dp, M, dE, dH)
    assert e.value.code == 4 and "64 MiB" in str(e.value)
    m.close()
    ctx.close()
