"""Serving loop on the GPU: every query's result is the same served alone
and in a concurrent mix (three workers, Poisson arrivals, an HBM budget small
enough to force evictions); the mixed form agrees with the FP64 form within
1e-6; one query's FP64 result (total E and H) against the oracle chain (1e-10)."""
import numpy as np
import pytest

import bielcc_amd as B
import oracle as O
from bielcc_amd import serve
from _util import rowwise_rel

pytestmark = pytest.mark.gpu


def _pool():
    m = B.mesh
    return {"sphere": m.sphere_quads(6, 12, jitter=0.005, seed=1),
            "ellipsoid": m.ellipsoid_quads(6, 12, axes=(1.2, 0.8, 0.6), seed=2),
            "torus_tri": m.triangulate(m.torus_quads(12, 6, R=1.0, r=0.35)),
            "cube": m.cube_quads(3, side=1.5)}


def _queries(seed=71, n=12):
    rng = np.random.default_rng(seed)
    names = list(_pool().keys())
    ks = [1.5, 2.0, 3.0]
    out = []
    for i in range(n):
        nrhs = int(rng.integers(1, 5))
        d = rng.normal(size=(nrhs, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        p = np.cross(d, rng.normal(size=(nrhs, 3)))
        p /= np.linalg.norm(p, axis=1, keepdims=True)
        M = int(rng.integers(40, 160))
        pts = B.mesh.sphere_points(M, float(rng.uniform(2.5, 3.5)), seed=int(rng.integers(1 << 30)))
        out.append(serve.Query(names[int(rng.integers(len(names)))], ks[int(rng.integers(3))], d, p, pts, qid=i))
    return out


def _copy(q):
    return serve.Query(q.mesh, q.k, q.dirs, q.pols, q.points, qid=q.qid)


def _alone(form, queries):
    out = {}
    with serve.Server(_pool(), form=form, workers=1, hbm_budget=4e9) as s:
        for q in queries:
            s.submit(_copy(q))
            out[q.qid] = s.wait([q.qid])[q.qid]
    return out


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_same_result_alone_and_in_the_mix(form):
    qs = _queries()
    alone = _alone(form, qs)
    # budget: room for about two factor sets plus workspaces, so the mix evicts
    with serve.Server(_pool(), form=form, workers=3, hbm_budget=5e6) as s:
        run = serve.poisson_run(s, [_copy(q) for q in qs], qps=200.0, seed=3)
    st = run["stats"]
    print(form, {k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()})
    assert st["p50_s"] <= st["p99_s"] and st["solves"] == sum(q.nrhs for q in qs)
    for q in qs:
        a, b = alone[q.qid], run["results"][q.qid]
        assert np.abs(a.E - b.E).max() <= 1e-12 * np.abs(a.E).max(), q.qid
        assert np.abs(a.H - b.H).max() <= 1e-12 * np.abs(a.H).max(), q.qid


def test_mixed_form_within_1e6_of_fp64():
    qs = _queries(seed=72, n=6)
    r64 = _alone("fp64", qs)
    r32 = _alone("mixed", qs)
    for q in qs:
        Es64 = r64[q.qid].E
        Es32 = r32[q.qid].E
        err = max(rowwise_rel(Es32[r], Es64[r]).max() for r in range(q.nrhs))
        errh = max(rowwise_rel(r32[q.qid].H[r], r64[q.qid].H[r]).max() for r in range(q.nrhs))
        assert err <= 1e-6 and errh <= 1e-6, (q.qid, err, errh)
        assert r32[q.qid].refine_resid <= 1e-10


def test_fp64_query_against_the_oracle_chain(orc):
    q = _queries(seed=73, n=1)[0]
    res = _alone("fp64", [q])[q.qid]
    verts = _pool()[q.mesh]
    npp = O.num_param()
    Z = orc.assemble(verts, npp, q.k)
    for r in range(q.nrhs):
        kv = q.k * q.dirs[r]
        rhs = orc.rhs_planewave(verts, kv, q.pols[r])
        b = orc.lu_solve(Z, rhs)[:, 0]
        j = orc.currents_by_basis(verts, b)
        Eo, Ho = orc.field_eh(verts, j, q.k, npp, q.points, incident=(kv, q.pols[r]))
        # total H in normalised units: R[j'] / multE (multE = i/k) plus PlaneWaveFieldH (omega = k, mu = 1)
        Ht = -1j * q.k * Ho + orc.plane_wave_h(q.points, kv, q.pols[r], q.k / (2 * np.pi), mu=1.0)
        assert rowwise_rel(res.E[r], Eo).max() <= 1e-10
        assert rowwise_rel(res.H[r], Ht).max() <= 1e-10


def test_mixed_form_falls_back_to_fp64_factors_when_refinement_stalls():
    """An unreachable refinement tolerance stalls every refinement: the entry
    gets FP64 factors once and its queries match the FP64 form."""
    qs = _queries(seed=74, n=4)
    r64 = _alone("fp64", qs)
    out = {}
    with serve.Server(_pool(), form="mixed", workers=1, hbm_budget=4e9, refine_tol=1e-30) as s:
        for q in qs:
            s.submit(_copy(q))
            out[q.qid] = s.wait([q.qid])[q.qid]
    for q in qs:
        assert out[q.qid].fp64_fallback
        for r in range(q.nrhs):
            assert rowwise_rel(out[q.qid].E[r][:, :], r64[q.qid].E[r]).max() <= 1e-6
