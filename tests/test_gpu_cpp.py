"""The drop-in C++ layer (bielcc_amd/cpp/bielcc_hip.hpp): a reference-style
program (reference names and signatures) run on the GPU, checked against the
oracle: operators 1e-12, field after assembly + LU 1e-10."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from bielcc_amd import build, mesh
from _util import GOLDEN

pytestmark = pytest.mark.gpu

RUT1 = np.array([[0., 0., 0.], [2., 0., 0.], [3., 2., 0.], [0., 2., 0.]])
RUT2 = np.array([[0., 2., 0.], [0., 0., 0.], [2., 0., 0.]])


def _run():
    out = subprocess.run([build.EXAMPLE, os.path.join(GOLDEN, "sphere_quad.dat")], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    rec = {}
    pts, E = [], []
    for line in out.stdout.splitlines():
        tok = line.split()
        key, vals = tok[0], [float(t) for t in tok[1:]]
        if key == "point":
            pts.append(vals)
        elif key == "E":
            E.append(np.array(vals[0::2]) + 1j * np.array(vals[1::2]))
        else:
            rec[key] = np.array(vals)
    return rec, np.array(pts), np.array(E)


def _c(v):
    return v[0::2] + 1j * v[1::2]


def test_cpp_dropin_program(orc):
    rec, pts, E = _run()
    npp = O.num_param()
    j = np.array([1, 0, 0], dtype=complex)
    x, xf = np.array([2., 1.5, 0.]), np.array([10., 10., 10.])
    cases = {"K_RotRot_Near_Smooth": (0, RUT1, x), "K_RotRot_HS": (1, RUT1, x), "K_RotRot_HS_Colloc": (2, RUT1, x),
             "K_RotRot_Far": (4, RUT1, xf), "R_Rot_Smooth": (5, RUT2, x), "R_Rot_Colloc": (6, RUT2, xf)}
    for key, (op, cell, pt) in cases.items():
        ref = orc.op(op, cell, pt, j, npp, 1.0)
        assert np.linalg.norm(_c(rec[key]) - ref) <= 1e-12 * np.linalg.norm(ref), key
    kp = O.kernel_param(0.5, 1e-6, 1.0)
    s = orc.integral_pnt(4, RUT1, x, kp, O.IntegralParam(1, 8, 3, 1e-6))[0]
    assert abs(_c(rec["IntegralUniversalPnt"])[0] - s) <= 1e-12 * abs(s)
    assert rec["smoothR_after"][0] == kp.smoothR
    assert abs(rec["integral1Divr"][0] - orc.integral1divr(RUT1, x)) <= 1e-13 * abs(orc.integral1divr(RUT1, x))
    # F(y) forms and PlaneWaveFieldH (vacuum: mu = mu0)
    assert abs(rec["area_rut1"][0] - 5.0) <= 1e-14
    kv, e0 = np.array([0.3, -1.2, 0.5]), np.array([1.0, 0.5, 0.0])
    fy = orc.integral_y(1, RUT1, e0, O.kernel_param(1e-16, 1e-16, 0.0, vec_dbl=kv), O.IntegralParam(3, 8, 1, 1e-6))
    assert np.linalg.norm(_c(rec["IntegralUniversal_PlaneWaveFieldE"]) - fy) <= 1e-12 * np.linalg.norm(fy)
    h = orc.plane_wave_h(x[None], kv, e0, 0.2, mu=4e-7 * np.pi)[0]
    assert np.linalg.norm(_c(rec["PlaneWaveFieldH"]) - h) <= 1e-12 * np.linalg.norm(h)
    xs = np.stack([0.001 * np.arange(2000), np.full(2000, 0.5), np.full(2000, -0.25)], axis=1)
    es = orc.plane_wave_e(xs, kv, e0).sum(axis=0)
    assert np.linalg.norm(_c(rec["PlaneWaveFieldE_sum2000"]) - es) <= 1e-12 * np.linalg.norm(es)
    # whole problem on the reference's own example mesh
    verts = mesh.read_dat(os.path.join(GOLDEN, "sphere_quad.dat"), 4)
    assert int(rec["cells"][0]) == verts.shape[0]
    assert rec["grid_step"][0] == max(orc.cellfill(v)["diam"] for v in verts)
    Z = orc.assemble(verts, npp, 2.0)
    rhs = orc.rhs_planewave(verts, (0, 0, 2.0), (1, 0, 0))
    b = orc.lu_solve(Z, rhs)[:, 0]
    jE = orc.currents_by_basis(verts, b)
    Eo = orc.get_field_e(verts, jE, 1.0, (0, 0, 2.0), (1, 0, 0), 2.0, npp, pts)
    err = np.linalg.norm(E - Eo, axis=1) / np.linalg.norm(Eo, axis=1)
    cond = np.linalg.cond(Z)
    assert err.max() <= max(1e-10, 1e-15 * cond), (err.max(), cond)
