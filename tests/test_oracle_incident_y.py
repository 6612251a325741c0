"""The oracle's PlaneWaveFieldH and F(y) integrators (IntegralUniversal over
quads and triangles, the F(y) segment form) against vectors recorded from the
reference, and against what the mathematics fixes:
- the constant integrand over the examples' cells gives their areas (rut1 = 5,
  main0_integrals.cpp:24-27; rut2 + rut3 tile rut1) and a segment's length;
- over a parallelogram split into n x n sub-parallelograms (and a segment into
  n pieces) the midpoint sum of e^{i k.y} has the closed form
  |u x v| M(k.u) M(k.v) e^{i k.y0},  M(a) = e^{ia/2n} (e^{ia} - 1) / (n (e^{ia/n} - 1)),
  so a wrong node, weight or partition fails it;
- k = 0 gives area x e0; k along a flat cell's normal gives e^{i k.y0} area e0;
- PlaneWaveFieldH = (k x e0) e^{ik.x} / (omega mu) and k.H = 0, e0.H = 0 (transverse)."""
import numpy as np
import pytest

import oracle as O
from _util import golden

RUT1 = np.array([[0., 0., 0.], [2., 0., 0.], [3., 2., 0.], [0., 2., 0.]])
RUT2 = np.array([[0., 2., 0.], [0., 0., 0.], [2., 0., 0.]])
RUT3 = np.array([[3., 2., 0.], [2., 0., 0.], [0., 2., 0.]])


def _kp(ch, R, Rmin, kvec, kt=0, k=1.5):
    return O.kernel_param(R, Rmin, k if kt == 0 else k - 0.3j, vec_dbl=kvec)


def test_plane_wave_h_vs_reference(orc):
    g = golden("incident_y.npz")
    H = np.array([orc.plane_wave_h(g["h_pts"][i:i + 1], g["h_kvec"][i], g["h_e0"][i], g["h_freq"][i],
                                   mu=g["h_mu"][i])[0] for i in range(len(g["h_H"]))])
    assert np.abs(H - g["h_H"]).max() <= 1e-15 * np.abs(g["h_H"]).max()
    # mu = mu0 mu_r (MaterialParam::CalcMu)
    assert np.allclose(g["h_mu"], 4e-7 * np.pi * g["h_mu_r"], rtol=1e-15, atol=0)


def test_plane_wave_h_is_the_transverse_plane_wave(orc):
    rng = np.random.default_rng(5)
    k = rng.normal(size=3) * 3
    e0 = np.cross(k, rng.normal(size=3))
    pts = rng.normal(size=(20, 3))
    for mu in (1.0, 2.0 + 0.5j):
        H = orc.plane_wave_h(pts, k, e0, 7.0, mu=mu)
        ph = np.exp(1j * pts @ k)
        expect = np.cross(k, e0)[None, :] * ph[:, None] / (2 * np.pi * 7.0 * mu)
        assert np.abs(H - expect).max() <= 1e-14 * np.abs(expect).max()
        assert np.abs(H @ k).max() <= 1e-13 * np.abs(H).max() * np.linalg.norm(k)
        assert np.abs(H @ e0).max() <= 1e-13 * np.abs(H).max() * np.linalg.norm(e0)


def test_integral_y_vs_reference(orc):
    g = golden("incident_y.npz")
    for i in range(len(g["y_f"])):
        nv = int(g["y_nv"][i])
        cell = g["y_cell"][i][:nv if nv > 2 else 2]
        R, Rmin = g["y_kp"][i]
        idim, n0, pm, eps = g["y_ip"][i]
        kp = _kp(orc, R, Rmin, g["y_kvec"][i], int(g["y_kt"][i]))
        ip = O.IntegralParam(int(idim), int(n0), int(pm), float(eps))
        r = orc.integral_y(int(g["y_f"][i]), cell, g["y_e0"][i], kp, ip)
        ref = g["y_res"][i]
        assert np.abs(r - ref).max() <= 1e-14 * max(np.abs(ref).max(), 1e-300), i
        assert kp.smoothR == g["y_R_out"][i], i


def test_unit_integrand_gives_areas_and_lengths(orc):
    ip = O.IntegralParam(1, 8, 1, 1e-6)
    a = [orc.integral_y(0, c, (0, 0, 0), _kp(orc, 1e-16, 1e-16, (0, 0, 0)), ip)[0].real for c in (RUT1, RUT2, RUT3)]
    assert abs(a[0] - 5.0) <= 1e-14 and abs(a[1] - 2.0) <= 1e-14 and abs(a[2] - 3.0) <= 1e-14
    seg = np.array([[1., 1., 1.], [4., 5., 1.]])
    assert abs(orc.integral_y(0, seg, (0, 0, 0), _kp(orc, 1e-16, 1e-16, (0, 0, 0)), ip)[0].real - 5.0) <= 1e-14


def _M(a, n):
    if a == 0:
        return 1.0
    return np.exp(1j * a / (2 * n)) * (np.exp(1j * a) - 1) / (n * (np.exp(1j * a / n) - 1))


@pytest.mark.parametrize("n", [5, 8, 16])
def test_plane_wave_midpoint_sums_closed_form(orc, n):
    rng = np.random.default_rng(n)
    y0 = rng.normal(size=3)
    u, v = np.array([2.0, 0.3, -0.2]), np.array([0.7, 1.5, 0.4])
    par = np.array([y0, y0 + u, y0 + u + v, y0 + v])     # a parallelogram (bilinear map is affine)
    k = rng.normal(size=3) * 2.5
    e0 = np.array([0.3, -1.0, 0.5])
    ip = O.IntegralParam(3, n, 1, 1e-6)
    r = orc.integral_y(1, par, e0, _kp(orc, 1e-16, 1e-16, k), ip)
    area = np.linalg.norm(np.cross(u, v))
    # the rule's sub-cell (i, j) spans p in [i/n, (i+1)/n] along cell[1]->cell[2] (v), q along cell[0]->cell[1] (u)
    expect = area * _M(k @ u, n) * _M(k @ v, n) * np.exp(1j * k @ y0) * e0
    assert np.abs(r - expect).max() <= 1e-13 * np.abs(expect).max()
    seg = np.array([y0, y0 + u])
    rs = orc.integral_y(1, seg, e0, _kp(orc, 1e-16, 1e-16, k), O.IntegralParam(3, n, 1, 1e-6))
    expect = np.linalg.norm(u) * _M(k @ u, n) * np.exp(1j * k @ y0) * e0
    assert np.abs(rs - expect).max() <= 1e-13 * np.abs(expect).max()


def test_plane_wave_special_directions(orc):
    e0 = np.array([1.0, 0.5, 0.0])
    ip = O.IntegralParam(3, 8, 1, 1e-6)
    for cell, area in ((RUT1 + [0, 0, 0.7], 5.0), (RUT3 + [0, 0, 0.7], 3.0)):
        r0 = orc.integral_y(1, cell, e0, _kp(orc, 1e-16, 1e-16, (0, 0, 0)), ip)
        assert np.abs(r0 - area * e0).max() <= 1e-14
        kz = 2.3
        rn = orc.integral_y(1, cell, e0, _kp(orc, 1e-16, 1e-16, (0, 0, kz)), ip)
        assert np.abs(rn - np.exp(1j * kz * 0.7) * area * e0).max() <= 1e-14
