"""ctypes bindings for the two CPU checkers built by oracle/Makefile.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker, never by the product package.

``Checker("oracle")`` wraps liboracle.so (the C restatement, functions
``orc_*``); ``Checker("ref")`` wraps _ref/libbielcc_ref.so (the reference
compiled from its own sources, functions ``ref_*``).  Both expose the same
Python methods, taking/returning numpy arrays; complex values are numpy
complex128.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_DP = C.POINTER(C.c_double)


class NumParam(C.Structure):
    _fields_ = [("eps", C.c_double), ("rs", C.c_double), ("rs_seg", C.c_double),
                ("analytic_calc", C.c_double), ("n_start", C.c_int), ("n_start_seg", C.c_int),
                ("p_max", C.c_int), ("p_max_seg", C.c_int)]


class KernelParam(C.Structure):
    _fields_ = [("smoothR", C.c_double), ("smoothRMin", C.c_double), ("k_re", C.c_double),
                ("k_im", C.c_double), ("vec_dbl", C.c_double * 3), ("n", C.c_double * 3),
                ("current", C.c_double * 6)]


class IntegralParam(C.Structure):
    _fields_ = [("idim", C.c_int), ("n_start", C.c_int), ("p_max", C.c_int), ("eps", C.c_double)]


# Defaults of Calculation_Constants (electrodynamics/constants.h:17-27)
DEFAULT_NUM_PARAM = dict(eps=1e-6, rs=2.0, rs_seg=2.0, analytic_calc=3.0,
                         n_start=8, n_start_seg=10, p_max=1, p_max_seg=1)

KERNELS = ["F_SimplePot_L", "F_GradSimplePot_L", "F_DoublePot_L", "F_VectorPot_L",
           "F_SimplePot_H", "nuF_SimplePot_H", "F_SimplePot_Hm1", "F_GradSimplePot_H",
           "F_GradSimplePot_Hm1", "F_DoublePot_H", "nF_DoublePot_H", "F_RotRotSimplePot_H"]
OPS = ["K_RotRot_Near_Smooth", "K_RotRot_HS", "K_RotRot_HS_Colloc", "K_RotRot_HS2_Colloc",
       "K_RotRot_Far", "R_Rot_Smooth", "R_Rot_Colloc", "R_Rot_Near"]


def num_param(**kw) -> NumParam:
    d = dict(DEFAULT_NUM_PARAM)
    d.update(kw)
    return NumParam(**d)


def kernel_param(smoothR, smoothRMin, k, vec_dbl=(0, 0, 0), n=(0, 0, 0), current=(0, 0, 0)) -> KernelParam:
    kp = KernelParam()
    kp.smoothR, kp.smoothRMin = smoothR, smoothRMin
    k = complex(k)
    kp.k_re, kp.k_im = k.real, k.imag
    for i in range(3):
        kp.vec_dbl[i] = vec_dbl[i]
        kp.n[i] = n[i]
        c = complex(current[i])
        kp.current[2 * i], kp.current[2 * i + 1] = c.real, c.imag
    return kp


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_DP)


def _cvec(z, n):
    z = np.asarray(z, dtype=np.complex128).reshape(-1)
    out = np.empty(2 * z.size)
    out[0::2], out[1::2] = z.real, z.imag
    return out


def _to_c(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.complex128)


class Checker:
    def __init__(self, which: str = "oracle"):
        self.which = which
        if which == "oracle":
            path, self.pre = os.path.join(HERE, "liboracle.so"), "orc_"
        elif which == "ref":
            path, self.pre = os.path.join(HERE, "_ref", "libbielcc_ref.so"), "ref_"
        else:
            raise ValueError(which)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle {'ref' if which == 'ref' else 'oracle'}`)")
        self.lib = C.CDLL(path)
        self._kt = [] if which == "oracle" else [C.c_int(1)]
        getattr(self.lib, self.pre + "integral1divr").restype = C.c_double

    def _f(self, name):
        # attribute access caches the function object, so restype settings persist
        return getattr(self.lib, self.pre + name)

    # -- geometry
    def cellfill(self, cell):
        cell, pc = _d(cell)
        out = np.zeros(20)
        rc = self._f("cellfill")(cell.shape[0], pc, out.ctypes.data_as(_DP))
        assert rc == 0
        return dict(rkt=out[0:3], norm=out[3:6], tau1=out[6:9], tau2=out[9:12], area=out[12],
                    diam=out[13], norm_func=out[14:17], center=out[17:20])

    # -- kernels and integrators (kp is mutated like the reference's KernelParam&)
    def kernel(self, f, x, y, kp):
        x, px = _d(x); y, py = _d(y)
        out = np.zeros(6)
        args = [C.c_int(f)] + self._kt + [px, py, C.byref(kp), out.ctypes.data_as(_DP)]
        assert self._f("kernel")(*args) == 0
        return _to_c(out)

    def integral_pnt(self, f, cell, x, kp, ip):
        cell, pc = _d(cell); x, px = _d(x)
        out = np.zeros(6)
        args = [C.c_int(f)] + self._kt + [C.c_int(cell.shape[0]), pc, px, C.byref(kp), C.byref(ip),
                                          out.ctypes.data_as(_DP)]
        assert self._f("integral_pnt")(*args) == 0
        return _to_c(out)

    def integral_seg(self, f, a, b, x, kp, ip):
        a, pa = _d(a); b, pb = _d(b); x, px = _d(x)
        out = np.zeros(6)
        args = [C.c_int(f)] + self._kt + [pa, pb, px, C.byref(kp), C.byref(ip), out.ctypes.data_as(_DP)]
        assert self._f("integral_seg")(*args) == 0
        return _to_c(out)

    def k0(self, variant, f, cell, x, j, kp, ip):
        cell, pc = _d(cell); x, px = _d(x)
        jj, pj = _d(_cvec(j, 3))
        out = np.zeros(6)
        args = [C.c_int(variant), C.c_int(f)] + self._kt + [C.c_int(cell.shape[0]), pc, px, pj,
                                                             C.byref(kp), C.byref(ip), out.ctypes.data_as(_DP)]
        assert self._f("k0")(*args) == 0
        return _to_c(out)

    def k0_hs(self, cell, x, j, ip, analyticR, k):
        cell, pc = _d(cell); x, px = _d(x)
        jj, pj = _d(_cvec(j, 3))
        out = np.zeros(6)
        k = complex(k)
        args = self._kt + [C.c_int(cell.shape[0]), pc, px, pj, C.byref(ip), C.c_double(analyticR),
                           C.c_double(k.real), C.c_double(k.imag), out.ctypes.data_as(_DP)]
        assert self._f("k0_hs")(*args) == 0
        return _to_c(out)

    def integral1divr(self, cell, x):
        cell, pc = _d(cell); x, px = _d(x)
        return self._f("integral1divr")(C.c_int(cell.shape[0]), pc, px)

    def integral_grad1divr_seg(self, a, b, x):
        a, pa = _d(a); b, pb = _d(b); x, px = _d(x)
        out = np.zeros(3)
        self._f("integral_grad1divr_seg")(pa, pb, px, out.ctypes.data_as(_DP))
        return out

    def surf_hs(self, cell, x, analyticR, ip, k):
        cell, pc = _d(cell); x, px = _d(x)
        out = np.zeros(2)
        k = complex(k)
        args = self._kt + [C.c_int(cell.shape[0]), pc, px, C.c_double(analyticR), C.byref(ip),
                           C.c_double(k.real), C.c_double(k.imag), out.ctypes.data_as(_DP)]
        assert self._f("surf_hs")(*args) == 0
        return complex(out[0], out[1])

    def seg_hs(self, a, b, x, analyticR, ip, k):
        a, pa = _d(a); b, pb = _d(b); x, px = _d(x)
        out = np.zeros(6)
        k = complex(k)
        args = self._kt + [pa, pb, px, C.c_double(analyticR), C.byref(ip), C.c_double(k.real),
                           C.c_double(k.imag), out.ctypes.data_as(_DP)]
        assert self._f("seg_hs")(*args) == 0
        return _to_c(out)

    # -- operators
    def op(self, op, cell, x, j, np_, k):
        cell, pc = _d(cell); x, px = _d(x)
        jj, pj = _d(_cvec(j, 3))
        out = np.zeros(6)
        k = complex(k)
        args = [C.c_int(op)] + self._kt + [C.c_int(cell.shape[0]), pc, px, pj, C.byref(np_),
                                           C.c_double(k.real), C.c_double(k.imag), out.ctypes.data_as(_DP)]
        assert self._f("op")(*args) == 0
        return _to_c(out)

    def op_pairs(self, op, verts, j, pts, np_, k):
        verts, pv = _d(verts); pts, pp = _d(pts)
        jj, pj = _d(_cvec(j, 3 * verts.shape[0]))
        npts, ncell = pts.shape[0], verts.shape[0]
        out = np.zeros(npts * ncell * 6)
        k = complex(k)
        args = [C.c_int(op)] + self._kt + [C.c_int(verts.shape[1]), C.c_int(ncell), pv, pj, C.c_int(npts), pp,
                                           C.byref(np_), C.c_double(k.real), C.c_double(k.imag),
                                           out.ctypes.data_as(_DP)]
        assert self._f("op_pairs")(*args) == 0
        return _to_c(out).reshape(npts, ncell, 3)

    # -- drivers
    def get_field_e(self, verts, jE, mult, kvec, e0, k, np_, pts, tmpdir="/tmp"):
        verts, pv = _d(verts); pts, pp = _d(pts)
        jj, pj = _d(_cvec(jE, 3 * verts.shape[0]))
        kv, pk = _d(kvec); e, pe = _d(e0)
        out = np.zeros(pts.shape[0] * 6)
        mult, k = complex(mult), complex(k)
        args = self._kt + [C.c_int(verts.shape[1]), C.c_int(verts.shape[0]), pv, pj,
                           C.c_double(mult.real), C.c_double(mult.imag), pk, pe,
                           C.c_double(k.real), C.c_double(k.imag), C.byref(np_), C.c_int(pts.shape[0]),
                           pp, out.ctypes.data_as(_DP)]
        if self.which == "ref":
            args.append(C.c_char_p(tmpdir.encode()))
        rc = self._f("get_field_e")(*args)
        assert rc == 0, rc
        return _to_c(out).reshape(-1, 3)

    def plane_wave_e(self, pts, kvec, e0):
        pts, pp = _d(pts); kv, pk = _d(kvec); e, pe = _d(e0)
        out = np.zeros(pts.shape[0] * 6)
        self._f("plane_wave_e")(C.c_int(pts.shape[0]), pp, pk, pe, out.ctypes.data_as(_DP))
        return _to_c(out).reshape(-1, 3)

    def plane_wave_h(self, pts, kvec, e0, freq, mu=None, mu_r=1.0):
        """PlaneWaveFieldH at pts [n][3] -> [n][3] complex.  The reference's form
        takes a MaterialParam of relative permeability mu_r (mu = mu0 mu_r after
        Calc(freq)) and returns (H, mu); the oracle takes mu itself (complex allowed)."""
        pts, pp = _d(pts); kv, pk = _d(kvec); e, pe = _d(e0)
        out = np.zeros(pts.shape[0] * 6)
        if self.which == "ref":
            mu_out = C.c_double(0.)
            self._f("plane_wave_h")(C.c_int(pts.shape[0]), pp, pk, pe, C.c_double(freq), C.c_double(mu_r),
                                    out.ctypes.data_as(_DP), C.byref(mu_out))
            return _to_c(out).reshape(-1, 3), mu_out.value
        mu = complex(mu)
        self._f("plane_wave_h")(C.c_int(pts.shape[0]), pp, pk, pe, C.c_double(freq), C.c_double(mu.real),
                                C.c_double(mu.imag), out.ctypes.data_as(_DP))
        return _to_c(out).reshape(-1, 3)

    def integral_y(self, f, cell, e0, kp, ip):
        """IntegralUniversal F(y) forms over a cell [nv][3] (nv = 4, 3) or the F(y)
        segment form over [2][3]; f 0 = the constant 1, 1 = PlaneWaveFieldE
        (kp.vec_dbl = wave vector).  kp is mutated like the reference's."""
        cell, pc = _d(cell); e, pe = _d(e0)
        out = np.zeros(6)
        args = [C.c_int(f)] + self._kt + [C.c_int(cell.shape[0]), pc, pe, C.byref(kp), C.byref(ip),
                                          out.ctypes.data_as(_DP)]
        assert self._f("integral_y")(*args) == 0
        return _to_c(out)

    # -- oracle-only drivers of this library's own loops
    def field_eh(self, verts, jE, k, np_, pts, jM=None, multE=1.0, multH=1.0, incident=None,
                 near_mode=0):
        assert self.which == "oracle"
        verts, pv = _d(verts); pts, pp = _d(pts)
        jj, pj = _d(_cvec(jE, 3 * verts.shape[0]))
        if jM is not None:
            jm, pm = _d(_cvec(jM, 3 * verts.shape[0]))
        else:
            pm = None
        if incident is not None:
            kv, pk = _d(incident[0]); e, pe = _d(incident[1])
        else:
            pk = pe = None
        E = np.zeros(pts.shape[0] * 6); H = np.zeros(pts.shape[0] * 6)
        multE, multH, k = complex(multE), complex(multH), complex(k)
        rc = self._f("field_eh")(C.c_int(verts.shape[1]), C.c_int(verts.shape[0]), pv, pj, pm,
                                 C.c_double(multE.real), C.c_double(multE.imag),
                                 C.c_double(multH.real), C.c_double(multH.imag),
                                 C.c_int(incident is not None), pk, pe, C.c_double(k.real),
                                 C.c_double(k.imag), C.byref(np_), C.c_int(near_mode),
                                 C.c_int(pts.shape[0]), pp, E.ctypes.data_as(_DP), H.ctypes.data_as(_DP))
        assert rc == 0
        return _to_c(E).reshape(-1, 3), _to_c(H).reshape(-1, 3)

    def far_pattern(self, verts, jE, k, dirs, jM=None, multE=1.0, multH=1.0, n_start=8):
        """Far-field pattern (E_inf, H_inf) [ndir][3] at unit directions dirs."""
        assert self.which == "oracle"
        verts, pv = _d(verts); dirs, pd = _d(dirs)
        jj, pj = _d(_cvec(jE, 3 * verts.shape[0]))
        if jM is not None:
            jm, pm = _d(_cvec(jM, 3 * verts.shape[0]))
        else:
            pm = None
        E = np.zeros(dirs.shape[0] * 6); H = np.zeros(dirs.shape[0] * 6)
        multE, multH, k = complex(multE), complex(multH), complex(k)
        rc = self._f("far_pattern")(C.c_int(verts.shape[1]), C.c_int(verts.shape[0]), pv, pj, pm,
                                    C.c_double(multE.real), C.c_double(multE.imag), C.c_double(multH.real),
                                    C.c_double(multH.imag), C.c_double(k.real), C.c_double(k.imag),
                                    C.c_int(n_start), C.c_int(dirs.shape[0]), pd, E.ctypes.data_as(_DP),
                                    H.ctypes.data_as(_DP))
        assert rc == 0
        return _to_c(E).reshape(-1, 3), _to_c(H).reshape(-1, 3)

    def assemble(self, verts, np_, k, which=0, row0=0, nrows=None):
        assert self.which == "oracle"
        verts, pv = _d(verts)
        n2 = 2 * verts.shape[0]
        nrows = n2 - row0 if nrows is None else nrows
        Z = np.zeros((n2, nrows, 2))          # column-major (nrows x n2) with ld = nrows
        k = complex(k)
        rc = self._f("assemble")(C.c_int(which), C.c_int(verts.shape[1]), C.c_int(verts.shape[0]), pv,
                                 C.byref(np_), C.c_double(k.real), C.c_double(k.imag), C.c_int(row0),
                                 C.c_int(nrows), Z.ctypes.data_as(_DP), C.c_longlong(nrows))
        assert rc == 0
        return np.ascontiguousarray((Z[..., 0] + 1j * Z[..., 1]).T)   # (nrows, n2)

    def rhs_planewave(self, verts, kvec, e0):
        assert self.which == "oracle"
        verts, pv = _d(verts); kv, pk = _d(kvec); e, pe = _d(e0)
        out = np.zeros(4 * verts.shape[0])
        self._f("rhs_planewave")(C.c_int(verts.shape[1]), C.c_int(verts.shape[0]), pv, pk, pe,
                                 out.ctypes.data_as(_DP))
        return _to_c(out)

    def currents_by_basis(self, verts, b):
        assert self.which == "oracle"
        verts, pv = _d(verts)
        bb, pb = _d(_cvec(b, 2 * verts.shape[0]))
        out = np.zeros(6 * verts.shape[0])
        self._f("currents_by_basis")(C.c_int(verts.shape[1]), C.c_int(verts.shape[0]), pv, pb,
                                     out.ctypes.data_as(_DP))
        return _to_c(out).reshape(-1, 3)

    def lu_solve(self, Z, rhs):
        assert self.which == "oracle"
        n = Z.shape[0]
        A = np.asfortranarray(Z.astype(np.complex128))
        B = np.array(np.asarray(rhs, dtype=np.complex128).reshape(n, -1), order="F", copy=True)
        rc = self._f("lu_solve")(C.c_int(n), A.ctypes.data_as(_DP), C.c_longlong(n),
                                 B.ctypes.data_as(_DP), C.c_int(B.shape[1]), C.c_longlong(n))
        assert rc == 0, rc
        return B

    # -- reference-only helpers
    def ref_currents_by_basis(self, tau, b):
        assert self.which == "ref"
        t, pt = _d(tau)
        bb, pb = _d(_cvec(b, 2 * t.shape[0]))
        out = np.zeros(6 * t.shape[0])
        self._f("currents_by_basis")(C.c_int(t.shape[0]), pt, pb, out.ctypes.data_as(_DP))
        return _to_c(out).reshape(-1, 3)

    def ref_read_mesh(self, path, nv, cap=100000):
        assert self.which == "ref"
        verts = np.zeros(cap * nv * 3); rkt = np.zeros(cap * 3); nrm = np.zeros(cap * 3)
        tau = np.zeros(cap * 6); gs = C.c_double(0.); n = C.c_int(0)
        rc = self._f("read_mesh")(C.c_char_p(path.encode()), C.c_int(nv), C.c_int(cap), C.byref(n),
                                  verts.ctypes.data_as(_DP), rkt.ctypes.data_as(_DP),
                                  nrm.ctypes.data_as(_DP), tau.ctypes.data_as(_DP), C.byref(gs))
        assert rc == 0, rc
        n = n.value
        return dict(verts=verts[:n * nv * 3].reshape(n, nv, 3), rkt=rkt[:3 * n].reshape(n, 3),
                    norm=nrm[:3 * n].reshape(n, 3), tau=tau[:6 * n].reshape(n, 2, 3), grid_step=gs.value)
