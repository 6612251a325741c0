"""Python host layer over the C ABI: a context, device buffers, meshes and the
batched drivers.  Arrays cross as numpy (host) or `DeviceArray` (device);
complex data are numpy complex128 on the host and interleaved doubles on the
device.  Every call goes to the HIP library; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import IntegralParam, KernelParam, NumParam

__all__ = ["BielccError", "Context", "DeviceArray", "Mesh", "NumParam", "KernelParam", "IntegralParam",
           "num_param", "kernel_param", "OPS", "KERNELS", "DEFAULT_NUM_PARAM"]

# far-pair quadrature precision (include/bielcc_hip.h BIELCC_PREC_*)
PREC_FP64 = 0
PREC_MIXED = 1

OPS = {"K_RotRot_Near_Smooth": 0, "K_RotRot_HS": 1, "K_RotRot_HS_Colloc": 2, "K_RotRot_HS2_Colloc": 3,
       "K_RotRot_Far": 4, "R_Rot_Smooth": 5, "R_Rot_Colloc": 6, "R_Rot_Near": 7}
KERNELS = {"F_SimplePot_L": 0, "F_GradSimplePot_L": 1, "F_DoublePot_L": 2, "F_VectorPot_L": 3,
           "F_SimplePot_H": 4, "nuF_SimplePot_H": 5, "F_SimplePot_Hm1": 6, "F_GradSimplePot_H": 7,
           "F_GradSimplePot_Hm1": 8, "F_DoublePot_H": 9, "nF_DoublePot_H": 10, "F_RotRotSimplePot_H": 11}
# Calculation_Constants (electrodynamics/constants.h:17-27)
DEFAULT_NUM_PARAM = dict(eps=1e-6, rs=2.0, rs_seg=2.0, analytic_calc=3.0, n_start=8, n_start_seg=10,
                         p_max=1, p_max_seg=1)


class BielccError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bielcc error {code}: {msg}")
        self.code = code


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


def _as_real(a) -> np.ndarray:
    """numpy array -> contiguous float64 view (complex -> interleaved)."""
    a = np.asarray(a)
    if np.iscomplexobj(a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        return a.view(np.float64)
    return np.ascontiguousarray(a, dtype=np.float64)


class DeviceArray:
    """A device allocation owned by a Context (freed on close/GC)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx._check(ctx.lib.bielcc_malloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, ctx: "Context", a) -> "DeviceArray":
        r = _as_real(a) if not (isinstance(a, np.ndarray) and a.dtype == np.int32) else np.ascontiguousarray(a)
        d = cls(ctx, max(r.nbytes, 1))
        if r.nbytes:
            ctx._check(ctx.lib.bielcc_memcpy(ctx.h, C.c_void_p(d.ptr), r.ctypes.data_as(C.c_void_p), r.nbytes, 0))
        return d

    def to_numpy(self, shape, dtype=np.float64, offset: int = 0) -> np.ndarray:
        """Copy of the bytes [offset, offset + size) as a host array."""
        out = np.empty(shape, dtype=dtype)
        if offset < 0 or offset + out.nbytes > self.nbytes:
            raise ValueError("range outside the device buffer")
        if out.nbytes:
            self.ctx._check(self.ctx.lib.bielcc_memcpy(self.ctx.h, out.ctypes.data_as(C.c_void_p),
                                                       C.c_void_p(self.ptr + int(offset)), out.nbytes, 1))
        return out

    def free(self):
        if self.ptr is not None and self.ctx.h is not None:
            self.ctx.lib.bielcc_free(self.ctx.h, C.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _p(x):
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return C.c_void_p(x.ptr)
    if isinstance(x, int):
        return C.c_void_p(x)
    raise TypeError(f"expected a DeviceArray or device pointer, got {type(x)}")


def _vec3(v):
    a = (C.c_double * 3)(*[float(t) for t in v])
    return C.cast(a, C.POINTER(C.c_double)), a


class Context:
    """A bielcc_ctx: one HIP device, one stream.

    ``Context.default_tuning`` (key -> value) is applied to every context
    created afterwards in this process, e.g. by a server's workers."""

    default_tuning: dict = {}

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.bielcc_ctx_create(int(device), C.byref(h))
        if rc != 0:
            raise BielccError(rc, f"bielcc_ctx_create(device={device}) failed (no HIP device?)")
        self.h = h
        for k, v in Context.default_tuning.items():
            self.set_tuning(k, v)

    def close(self):
        if self.h is not None:
            self.lib.bielcc_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            buf = C.create_string_buffer(512)
            self.lib.bielcc_last_error(self.h, buf, 512)
            raise BielccError(rc, buf.value.decode(errors="replace"))

    def sync(self):
        self._check(self.lib.bielcc_sync(self.h))

    def stream(self) -> int:
        return self.lib.bielcc_stream(self.h)

    # -- memory
    def upload(self, dst: DeviceArray, a) -> None:
        """Host array into the front of an existing device buffer (synchronous)."""
        a = np.ascontiguousarray(a)
        if a.nbytes > dst.nbytes:
            raise ValueError("array larger than the device buffer")
        self._check(self.lib.bielcc_memcpy(self.h, C.c_void_p(dst.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, 0))

    def copy(self, dst: DeviceArray, src: DeviceArray, nbytes: int) -> None:
        """Device-to-device copy of nbytes (synchronous)."""
        self._check(self.lib.bielcc_memcpy(self.h, C.c_void_p(dst.ptr), C.c_void_p(src.ptr), int(nbytes), 2))

    def to_device(self, a) -> DeviceArray:
        return DeviceArray.from_numpy(self, a)

    def empty(self, nbytes: int) -> DeviceArray:
        return DeviceArray(self, nbytes)

    # -- meshes
    def mesh(self, verts) -> "Mesh":
        return Mesh(self, verts)

    # -- plans (several meshes per launch)
    def plan_assemble(self, items, np_: NumParam, which=0, precision=PREC_FP64) -> Plan:
        """items: [(mesh, k, Z DeviceArray, ld)]."""
        arr = (_lib.AssembleItem * len(items))()
        for a, (m, k, Z, ld) in zip(arr, items):
            k = complex(k)
            a.mesh_id, a.k_re, a.k_im, a.Z_d, a.ld = m.id, k.real, k.imag, _p(Z), int(ld)
        h = C.c_void_p()
        self._check(self.lib.bielcc_plan_assemble(self.h, len(items), arr, int(which), C.byref(np_),
                                                  int(precision), C.byref(h)))
        return Plan(self, h, (arr, items))

    def plan_field(self, items, np_: NumParam, near_mode=0, precision=PREC_FP64) -> Plan:
        """items: [dict(mesh, k, nrhs, jE, M, pts, E, H=None, inc=None, multE=1)] with device buffers
        (jE [nrhs][N][3], inc [nrhs][6] (kvec, e0), pts [M][3], E/H [nrhs][M][3])."""
        arr = (_lib.FieldItem * len(items))()
        for a, it in zip(arr, items):
            k, mE = complex(it["k"]), complex(it.get("multE", 1.0))
            a.mesh_id, a.k_re, a.k_im, a.nrhs = it["mesh"].id, k.real, k.imag, int(it["nrhs"])
            a.jE_d, a.inc_d, a.multE_re, a.multE_im = _p(it["jE"]), _p(it.get("inc")), mE.real, mE.imag
            a.M, a.pts_d, a.E_d, a.H_d = int(it["M"]), _p(it["pts"]), _p(it["E"]), _p(it.get("H"))
        h = C.c_void_p()
        self._check(self.lib.bielcc_plan_field(self.h, len(items), arr, C.byref(np_), int(near_mode),
                                               int(precision), C.byref(h)))
        return Plan(self, h, (arr, items))

    # -- single functions, batched (host arrays in, host arrays out)
    def op_batch(self, op, cells, x, j, np_: NumParam, k) -> np.ndarray:
        op = OPS[op] if isinstance(op, str) else int(op)
        cells = np.ascontiguousarray(cells, dtype=np.float64)
        n, nv = cells.shape[0], cells.shape[1]
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(n, 3)
        j = np.ascontiguousarray(j, dtype=np.complex128).reshape(n, 3)
        dc, dx, dj = self.to_device(cells), self.to_device(x), self.to_device(j)
        dr = self.empty(n * 48)
        k = complex(k)
        self._check(self.lib.bielcc_op_batch(self.h, op, nv, n, _p(dc), _p(dx), _p(dj), C.byref(np_), k.real,
                                             k.imag, _p(dr)))
        return dr.to_numpy((n, 3), np.complex128)

    def kernel_batch(self, f, x, y, kp: KernelParam) -> np.ndarray:
        f = KERNELS[f] if isinstance(f, str) else int(f)
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]
        dx, dy, dr = self.to_device(x), self.to_device(y), self.empty(n * 48)
        self._check(self.lib.bielcc_kernel_batch(self.h, f, n, _p(dx), _p(dy), C.byref(kp), _p(dr)))
        return dr.to_numpy((n, 3), np.complex128)

    def integral_batch(self, f, cells, x, kp: KernelParam, ip: IntegralParam):
        """cells [n][nv][3] (nv = 4, 3) or segments [n][2][3]; returns (res [n][3], smoothR [n])."""
        f = KERNELS[f] if isinstance(f, str) else int(f)
        cells = np.ascontiguousarray(cells, dtype=np.float64)
        n, nv = cells.shape[0], cells.shape[1]
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(n, 3)
        dc, dx, dr, ds = self.to_device(cells), self.to_device(x), self.empty(n * 48), self.empty(n * 8)
        self._check(self.lib.bielcc_integral_batch(self.h, f, nv, n, _p(dc), _p(dx), C.byref(kp), C.byref(ip),
                                                   _p(dr), _p(ds)))
        return dr.to_numpy((n, 3), np.complex128), ds.to_numpy((n,))

    def integral_y_batch(self, f, cells, kp: KernelParam, ip: IntegralParam, e0=(0.0, 0.0, 0.0)):
        """IntegralUniversal F(y) forms over cells [n][nv][3] (nv = 4, 3) or the F(y)
        segment form over [n][2][3]; f 0 = the constant 1, 1 = PlaneWaveFieldE
        (kp.vec_dbl = wave vector, e0 = polarisation).  Returns (res [n][3], smoothR [n])."""
        cells = np.ascontiguousarray(cells, dtype=np.float64)
        n, nv = cells.shape[0], cells.shape[1]
        dc, dr, ds = self.to_device(cells), self.empty(max(n, 1) * 48), self.empty(max(n, 1) * 8)
        ev, _a = _vec3(e0)
        self._check(self.lib.bielcc_integral_y_batch(self.h, int(f), nv, n, _p(dc), C.byref(kp), ev, C.byref(ip),
                                                     _p(dr), _p(ds)))
        return dr.to_numpy((n, 3), np.complex128), ds.to_numpy((n,))

    def k0_batch(self, variant, f, cells, x, j, kp: KernelParam, ip: IntegralParam, analyticR=0.0):
        f = KERNELS[f] if isinstance(f, str) else int(f)
        cells = np.ascontiguousarray(cells, dtype=np.float64)
        n, nv = cells.shape[0], cells.shape[1]
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(n, 3)
        j = np.ascontiguousarray(j, dtype=np.complex128).reshape(n, 3)
        dc, dx, dj = self.to_device(cells), self.to_device(x), self.to_device(j)
        dr, ds = self.empty(n * 48), self.empty(n * 8)
        self._check(self.lib.bielcc_k0_batch(self.h, int(variant), f, nv, n, _p(dc), _p(dx), _p(dj), C.byref(kp),
                                             C.byref(ip), float(analyticR), _p(dr), _p(ds)))
        return dr.to_numpy((n, 3), np.complex128), ds.to_numpy((n,))

    def analytic_batch(self, which, cells, x, ip: IntegralParam | None = None, analyticR=0.0, k=0.0):
        cells = np.ascontiguousarray(cells, dtype=np.float64)
        n, nv = cells.shape[0], cells.shape[1]
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(n, 3)
        dc, dx, dr = self.to_device(cells), self.to_device(x), self.empty(n * 48)
        k = complex(k)
        ipp = C.byref(ip) if ip is not None else None
        self._check(self.lib.bielcc_analytic_batch(self.h, int(which), nv, n, _p(dc), _p(dx), ipp,
                                                   float(analyticR), k.real, k.imag, _p(dr)))
        return dr.to_numpy((n, 3), np.complex128)

    def aver_batch(self, which, x, y, eps) -> np.ndarray:
        """Averaging functions (kernel_lib.h:560-634): which 0 Ker_AverDiv, 1
        Ker_AverDiv_NearEdge, 2 Ker_BasicAverFunc (column 0).  Returns [n][3]."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]
        dx, dy, dr = self.to_device(x), self.to_device(y), self.empty(max(n, 1) * 24)
        self._check(self.lib.bielcc_aver_batch(self.h, int(which), n, _p(dx), _p(dy), float(eps), _p(dr)))
        return dr.to_numpy((n, 3))

    def plane_wave_e(self, pts, kvec, e0) -> np.ndarray:
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        n = pts.shape[0]
        dp, dr = self.to_device(pts), self.empty(n * 48)
        kv, _a = _vec3(kvec)
        ev, _b = _vec3(e0)
        self._check(self.lib.bielcc_plane_wave_e(self.h, n, _p(dp), kv, ev, _p(dr)))
        return dr.to_numpy((n, 3), np.complex128)

    def plane_wave_h(self, pts, kvec, e0, freq, mu=1.0) -> np.ndarray:
        """PlaneWaveFieldH at pts [n][3]; mu the absolute permeability (real or complex)."""
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        n = pts.shape[0]
        dp, dr = self.to_device(pts), self.empty(max(n, 1) * 48)
        kv, _a = _vec3(kvec)
        ev, _b = _vec3(e0)
        mu = complex(mu)
        self._check(self.lib.bielcc_plane_wave_h(self.h, n, _p(dp), kv, ev, float(freq), mu.real, mu.imag, _p(dr)))
        return dr.to_numpy((n, 3), np.complex128)

    def add_incident_h(self, M: int, nrhs: int, pts: DeviceArray, inc, freq, H: DeviceArray, mu=1.0, scale=1.0):
        """H[r][p] = scale H[r][p] + PlaneWaveFieldH(x_p) of incidence r (device H [nrhs][M][3],
        device pts [M][3], host inc [nrhs][6] = (kvec, e0))."""
        inc = np.ascontiguousarray(inc, dtype=np.float64).reshape(nrhs, 6)
        mu, scale = complex(mu), complex(scale)
        self._check(self.lib.bielcc_field_add_incident_h(self.h, int(M), int(nrhs), _p(pts),
                                                         inc.ctypes.data_as(C.POINTER(C.c_double)), float(freq),
                                                         mu.real, mu.imag, scale.real, scale.imag, _p(H)))

    # -- LU (device buffers)
    def lu_factor(self, A: DeviceArray, n: int, lda: int, piv: DeviceArray) -> None:
        info = C.c_int(0)
        self._check(self.lib.bielcc_lu_factor(self.h, int(n), _p(A), int(lda), _p(piv), C.byref(info)))

    def lu_solve(self, A: DeviceArray, n: int, lda: int, piv: DeviceArray, B: DeviceArray, nrhs: int,
                 ldb: int) -> None:
        self._check(self.lib.bielcc_lu_solve(self.h, int(n), _p(A), int(lda), _p(piv), int(nrhs), _p(B),
                                             int(ldb)))

    def lu32_factor(self, A: DeviceArray, n: int, lda: int, F: DeviceArray, ldf: int, piv: DeviceArray) -> None:
        """FP32 factors of the FP64 matrix A into F (complex64, n x n, ldf); A is kept."""
        info = C.c_int(0)
        self._check(self.lib.bielcc_lu32_factor(self.h, int(n), _p(A), int(lda), _p(F), int(ldf), _p(piv),
                                                C.byref(info)))

    def refine_solve(self, A: DeviceArray, n: int, lda: int, F: DeviceArray, ldf: int, piv: DeviceArray,
                     B: DeviceArray, nrhs: int, ldb: int, tol=1e-10, max_iter=30, rbt: DeviceArray = None):
        """FP64 solution of A X = B on FP32 factors (FGMRES preconditioned by them, or
        classic refinement with tunable refine_mode = 0); returns (corrections, residual).
        rbt: the butterflies when F came from lu32_factor_rbt."""
        it, res = C.c_int(0), C.c_double(0.)
        if rbt is None:
            rc = self.lib.bielcc_refine_solve(self.h, int(n), _p(A), int(lda), _p(F), int(ldf), _p(piv), int(nrhs),
                                              _p(B), int(ldb), float(tol), int(max_iter), C.byref(it), C.byref(res))
        else:
            rc = self.lib.bielcc_refine_solve_rbt(self.h, int(n), _p(A), int(lda), _p(F), int(ldf), _p(piv), _p(rbt),
                                                  int(nrhs), _p(B), int(ldb), float(tol), int(max_iter), C.byref(it),
                                                  C.byref(res))
        self._check(rc)
        return it.value, res.value

    def lu32_factor_rbt(self, A: DeviceArray, n: int, lda: int, F: DeviceArray, ldf: int, piv: DeviceArray,
                        rbt: DeviceArray, seed: int = 1) -> None:
        """FP32 factors (complex64) of U^T A V without interchanges (butterflies as
        lu_factor_rbt, diagonals into rbt, 32 n bytes); A is kept."""
        info = C.c_int(0)
        self._check(self.lib.bielcc_lu32_factor_rbt(self.h, int(n), _p(A), int(lda), _p(F), int(ldf), _p(piv),
                                                    _p(rbt), int(seed), C.byref(info)))

    def lu_factor_rbt(self, A: DeviceArray, n: int, lda: int, F: DeviceArray, ldf: int, piv: DeviceArray,
                      rbt: DeviceArray, seed: int = 1) -> None:
        """F = U^T A V (random butterflies from seed; their diagonals into rbt, 32 n bytes)
        factored without interchanges (piv: identity); A is kept."""
        info = C.c_int(0)
        self._check(self.lib.bielcc_lu_factor_rbt(self.h, int(n), _p(A), int(lda), _p(F), int(ldf), _p(piv), _p(rbt),
                                                  int(seed), C.byref(info)))

    def rbt_solve(self, A: DeviceArray, n: int, lda: int, F: DeviceArray, ldf: int, piv: DeviceArray,
                  rbt: DeviceArray, B: DeviceArray, nrhs: int, ldb: int, tol=1e-12, max_iter=10):
        """A X = B (X over B) on lu_factor_rbt's factors, refined against A; returns
        (corrections, relative residual)."""
        it, res = C.c_int(0), C.c_double(0.)
        self._check(self.lib.bielcc_rbt_solve(self.h, int(n), _p(A), int(lda), _p(F), int(ldf), _p(piv), _p(rbt),
                                              int(nrhs), _p(B), int(ldb), float(tol), int(max_iter), C.byref(it),
                                              C.byref(res)))
        return it.value, res.value

    def residual(self, A: DeviceArray, n: int, lda: int, X: DeviceArray, ldx: int, B: DeviceArray, ldb: int,
                 nrhs: int, R: DeviceArray, ldr: int, doubled: bool = False) -> None:
        """R = B - A X on the device, sums in FP64 or (doubled) in twice its precision."""
        self._check(self.lib.bielcc_residual(self.h, int(n), _p(A), int(lda), _p(X), int(ldx), _p(B), int(ldb),
                                             int(nrhs), int(bool(doubled)), _p(R), int(ldr)))

    def set_tuning(self, key: str, value: int):
        self._check(self.lib.bielcc_set_tuning(self.h, key.encode(), int(value)))

    def get_tuning(self, key: str) -> int:
        v = C.c_int(0)
        self._check(self.lib.bielcc_get_tuning(self.h, key.encode(), C.byref(v)))
        return v.value

    # -- timing
    def timing_enable(self, on=True):
        self._check(self.lib.bielcc_timing_enable(self.h, int(bool(on))))

    def timing_reset(self):
        self._check(self.lib.bielcc_timing_reset(self.h))

    def timing(self, name: str):
        ms, cnt = C.c_double(0.), C.c_int(0)
        self._check(self.lib.bielcc_timing(self.h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


class Plan:
    """A bielcc_plan: the matrices or fields of several meshes, enqueued with
    run() (no host synchronisation, no allocation), optionally replayed from a
    HIP graph after capture().  The item buffers must outlive the plan; their
    contents may change between runs."""

    def __init__(self, ctx: "Context", handle, keep):
        self.ctx = ctx
        self.h = handle
        self._keep = keep          # buffers and descriptors the plan points into

    def run(self):
        self.ctx._check(self.ctx.lib.bielcc_plan_run(self.ctx.h, self.h))

    def capture(self):
        self.ctx._check(self.ctx.lib.bielcc_plan_capture(self.ctx.h, self.h))

    def info(self, key: str) -> float:
        v = C.c_double(0.)
        self.ctx._check(self.ctx.lib.bielcc_plan_info(self.ctx.h, self.h, key.encode(), C.byref(v)))
        return v.value

    def close(self):
        if self.h is not None and self.ctx.h is not None:
            self.ctx.lib.bielcc_plan_destroy(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mesh:
    """A device-resident surface mesh (the SurfMesh of mesh/surf_mesh.h)."""

    def __init__(self, ctx: Context, verts):
        self.ctx = ctx
        v = np.ascontiguousarray(verts, dtype=np.float64)
        if v.ndim != 3 or v.shape[2] != 3 or v.shape[1] not in (3, 4):
            raise ValueError("verts must be [n][3|4][3]")
        self.n, self.nv = v.shape[0], v.shape[1]
        mid = C.c_int(0)
        ctx._check(ctx.lib.bielcc_mesh_create(ctx.h, self.n, self.nv, v.ctypes.data_as(C.POINTER(C.c_double)),
                                              C.byref(mid)))
        self.id = mid.value

    def close(self):
        if self.id is not None and self.ctx.h is not None:
            self.ctx.lib.bielcc_mesh_destroy(self.ctx.h, self.id)
        self.id = None

    def tables(self) -> dict:
        out = np.empty((self.n, 20))
        self.ctx._check(self.ctx.lib.bielcc_mesh_tables(self.ctx.h, self.id,
                                                        out.ctypes.data_as(C.POINTER(C.c_double))))
        return dict(rkt=out[:, 0:3], norm=out[:, 3:6], tau1=out[:, 6:9], tau2=out[:, 9:12], area=out[:, 12],
                    diam=out[:, 13], norm_func=out[:, 14:17], center=out[:, 17:20])

    def grid_step(self) -> float:
        g = C.c_double(0.)
        self.ctx._check(self.ctx.lib.bielcc_mesh_grid_step(self.ctx.h, self.id, C.byref(g)))
        return g.value

    # device-buffer drivers
    def assemble_into(self, Z: DeviceArray, ld: int, np_: NumParam, k, which=0, row0=0, nrows=None,
                      precision=PREC_FP64):
        nrows = 2 * self.n - row0 if nrows is None else nrows
        k = complex(k)
        if precision == PREC_FP64:
            self.ctx._check(self.ctx.lib.bielcc_assemble(self.ctx.h, self.id, int(which), C.byref(np_), k.real,
                                                         k.imag, int(row0), int(nrows), _p(Z), int(ld)))
        else:
            self.ctx._check(self.ctx.lib.bielcc_assemble_ex(self.ctx.h, self.id, int(which), C.byref(np_), k.real,
                                                            k.imag, int(row0), int(nrows), int(precision), _p(Z),
                                                            int(ld)))

    def rhs_planewave_into(self, B: DeviceArray, ldb: int, kvec, e0):
        kvec = np.ascontiguousarray(kvec, dtype=np.float64).reshape(-1, 3)
        e0 = np.ascontiguousarray(e0, dtype=np.float64).reshape(-1, 3)
        self.ctx._check(self.ctx.lib.bielcc_rhs_planewave(self.ctx.h, self.id, kvec.shape[0],
                                                          kvec.ctypes.data_as(C.POINTER(C.c_double)),
                                                          e0.ctypes.data_as(C.POINTER(C.c_double)), _p(B),
                                                          int(ldb)))

    def currents_into(self, b: DeviceArray, ldb: int, nrhs: int, j: DeviceArray):
        self.ctx._check(self.ctx.lib.bielcc_currents_by_basis(self.ctx.h, self.id, int(nrhs), _p(b), int(ldb),
                                                              _p(j)))

    def field_into(self, np_: NumParam, k, jE, pts, M, E, H=None, jM=None, multE=1.0, multH=1.0,
                   incident=None, near_mode=0):
        k, multE, multH = complex(k), complex(multE), complex(multH)
        if incident is not None:
            kv, _a = _vec3(incident[0])
            ev, _b = _vec3(incident[1])
        else:
            kv = ev = None
        self.ctx._check(self.ctx.lib.bielcc_field(self.ctx.h, self.id, C.byref(np_), k.real, k.imag, _p(jE),
                                                  _p(jM), multE.real, multE.imag, multH.real, multH.imag, kv, ev,
                                                  int(near_mode), int(M), _p(pts), _p(E), _p(H)))

    def field_noalloc_into(self, np_: NumParam, k, jE, pts, M, E, H=None, jM=None, multE=1.0, multH=1.0,
                           incident=None, near_mode=0):
        """field_into through bielcc_field_noalloc: no device allocation (work memory in the
        library's fixed per-GPU scratch; the mesh's node tables must exist)."""
        k, multE, multH = complex(k), complex(multE), complex(multH)
        if incident is not None:
            kv, _a = _vec3(incident[0])
            ev, _b = _vec3(incident[1])
        else:
            kv = ev = None
        self.ctx._check(self.ctx.lib.bielcc_field_noalloc(self.ctx.h, self.id, C.byref(np_), k.real, k.imag, _p(jE),
                                                          _p(jM), multE.real, multE.imag, multH.real, multH.imag,
                                                          kv, ev, int(near_mode), int(M), _p(pts), _p(E), _p(H)))

    def field_batch_into(self, np_: NumParam, k, nrhs: int, jE, pts, M, E, H=None, jM=None, multE=1.0,
                         multH=1.0, inc=None, near_mode=0, precision=PREC_FP64):
        """nrhs current sets (device [nrhs][N][3]) -> E, H (device [nrhs][M][3]); inc: host [nrhs][6] or None."""
        k, multE, multH = complex(k), complex(multE), complex(multH)
        incp = None
        if inc is not None:
            inc = np.ascontiguousarray(inc, dtype=np.float64).reshape(nrhs, 6)
            incp = inc.ctypes.data_as(C.POINTER(C.c_double))
        self.ctx._check(self.ctx.lib.bielcc_field_batch(self.ctx.h, self.id, C.byref(np_), k.real, k.imag,
                                                        int(nrhs), _p(jE), _p(jM), multE.real, multE.imag,
                                                        multH.real, multH.imag, incp, int(near_mode),
                                                        int(precision), int(M), _p(pts), _p(E), _p(H)))

    # host convenience wrappers
    def assemble(self, np_: NumParam, k, which=0, row0=0, nrows=None, precision=PREC_FP64) -> np.ndarray:
        nrows = 2 * self.n - row0 if nrows is None else nrows
        Z = self.ctx.empty(nrows * 2 * self.n * 16)
        self.assemble_into(Z, nrows, np_, k, which, row0, nrows, precision)
        return Z.to_numpy((2 * self.n, nrows), np.complex128).T   # column-major -> (rows, cols)

    def rhs_planewave(self, kvec, e0) -> np.ndarray:
        kvec = np.asarray(kvec, dtype=np.float64).reshape(-1, 3)
        nrhs = kvec.shape[0]
        B = self.ctx.empty(nrhs * 2 * self.n * 16)
        self.rhs_planewave_into(B, 2 * self.n, kvec, e0)
        return B.to_numpy((nrhs, 2 * self.n), np.complex128).T

    def currents(self, b) -> np.ndarray:
        b = np.asarray(b, dtype=np.complex128).reshape(2 * self.n, -1, order="F")
        nrhs = b.shape[1]
        db = self.ctx.to_device(np.asfortranarray(b).T.copy())
        dj = self.ctx.empty(nrhs * self.n * 48)
        self.currents_into(db, 2 * self.n, nrhs, dj)
        return dj.to_numpy((nrhs, self.n, 3), np.complex128)

    def field(self, np_: NumParam, k, jE, pts, jM=None, multE=1.0, multH=1.0, incident=None, near_mode=0):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        M = pts.shape[0]
        dp = self.ctx.to_device(pts)
        dje = self.ctx.to_device(np.asarray(jE, dtype=np.complex128).reshape(self.n, 3))
        djm = self.ctx.to_device(np.asarray(jM, dtype=np.complex128).reshape(self.n, 3)) if jM is not None else None
        dE, dH = self.ctx.empty(M * 48), self.ctx.empty(M * 48)
        self.field_into(np_, k, dje, dp, M, dE, dH, djm, multE, multH, incident, near_mode)
        return dE.to_numpy((M, 3), np.complex128), dH.to_numpy((M, 3), np.complex128)

    def field_batch(self, np_: NumParam, k, jE, pts, jM=None, multE=1.0, multH=1.0, incident=None, near_mode=0,
                    precision=PREC_FP64):
        """jE (and jM): [nrhs][N][3]; incident: None or (kvec, e0) with [3] or [nrhs][3] rows.
        Returns E, H as [nrhs][M][3]."""
        jE = np.asarray(jE, dtype=np.complex128).reshape(-1, self.n, 3)
        nrhs = jE.shape[0]
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        M = pts.shape[0]
        inc = None
        if incident is not None:
            kv = np.broadcast_to(np.asarray(incident[0], dtype=np.float64).reshape(-1, 3), (nrhs, 3))
            ev = np.broadcast_to(np.asarray(incident[1], dtype=np.float64).reshape(-1, 3), (nrhs, 3))
            inc = np.concatenate([kv, ev], axis=1)
        dp, dje = self.ctx.to_device(pts), self.ctx.to_device(jE)
        djm = self.ctx.to_device(np.asarray(jM, dtype=np.complex128).reshape(nrhs, self.n, 3)) \
            if jM is not None else None
        dE, dH = self.ctx.empty(nrhs * M * 48), self.ctx.empty(nrhs * M * 48)
        self.field_batch_into(np_, k, nrhs, dje, dp, M, dE, dH, djm, multE, multH, inc, near_mode, precision)
        return dE.to_numpy((nrhs, M, 3), np.complex128), dH.to_numpy((nrhs, M, 3), np.complex128)

    def far_pattern_into(self, np_: NumParam, k, nrhs: int, jE, dirs, ndir, Einf, Hinf=None, jM=None,
                         multE=1.0, multH=1.0, precision=PREC_FP64):
        k, multE, multH = complex(k), complex(multE), complex(multH)
        self.ctx._check(self.ctx.lib.bielcc_far_pattern(self.ctx.h, self.id, C.byref(np_), k.real, k.imag,
                                                        int(nrhs), _p(jE), _p(jM), multE.real, multE.imag,
                                                        multH.real, multH.imag, int(precision), int(ndir),
                                                        _p(dirs), _p(Einf), _p(Hinf)))

    def far_pattern(self, np_: NumParam, k, jE, dirs, jM=None, multE=1.0, multH=1.0, precision=PREC_FP64):
        """Far-field pattern at unit directions dirs [ndir][3]; jE (jM) [N][3] or
        [nrhs][N][3].  Returns (E_inf, H_inf) shaped like the currents' sets."""
        jE = np.asarray(jE, dtype=np.complex128)
        single = jE.ndim == 2
        jE = jE.reshape(-1, self.n, 3)
        nrhs = jE.shape[0]
        dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        nd = dirs.shape[0]
        dd, dje = self.ctx.to_device(dirs), self.ctx.to_device(jE)
        djm = self.ctx.to_device(np.asarray(jM, dtype=np.complex128).reshape(nrhs, self.n, 3)) \
            if jM is not None else None
        dE, dH = self.ctx.empty(nrhs * nd * 48), self.ctx.empty(nrhs * nd * 48)
        self.far_pattern_into(np_, k, nrhs, dje, dd, nd, dE, dH, djm, multE, multH, precision)
        E, H = dE.to_numpy((nrhs, nd, 3), np.complex128), dH.to_numpy((nrhs, nd, 3), np.complex128)
        return (E[0], H[0]) if single else (E, H)

    def get_field_e(self, jE, mult, kvec, e0, np_: NumParam, k, pts) -> np.ndarray:
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        M = pts.shape[0]
        dp = self.ctx.to_device(pts)
        dje = self.ctx.to_device(np.asarray(jE, dtype=np.complex128).reshape(self.n, 3))
        dE = self.ctx.empty(M * 48)
        mult, k = complex(mult), complex(k)
        kv, _a = _vec3(kvec)
        ev, _b = _vec3(e0)
        self.ctx._check(self.ctx.lib.bielcc_get_field_e(self.ctx.h, self.id, _p(dje), mult.real, mult.imag, kv, ev,
                                                        C.byref(np_), k.real, k.imag, M, _p(dp), _p(dE)))
        return dE.to_numpy((M, 3), np.complex128)

    def solve(self, Z_host, rhs_host) -> np.ndarray:
        """LU-solve on the device (host matrices in/out; for tests and small cases)."""
        n = Z_host.shape[0]
        B = np.asarray(rhs_host, dtype=np.complex128).reshape(n, -1, order="F")
        dA = self.ctx.to_device(np.asfortranarray(Z_host).T.copy())
        dB = self.ctx.to_device(np.asfortranarray(B).T.copy())
        piv = self.ctx.empty(n * 4)
        self.ctx.lu_factor(dA, n, n, piv)
        self.ctx.lu_solve(dA, n, n, piv, dB, B.shape[1], n)
        return dB.to_numpy((B.shape[1], n), np.complex128).T

    def solve_mixed(self, Z_host, rhs_host, tol=1e-12, max_iter=30):
        """FP32 LU + FP64 refinement on the device; returns (X, corrections, residual)."""
        n = Z_host.shape[0]
        B = np.asarray(rhs_host, dtype=np.complex128).reshape(n, -1, order="F")
        dA = self.ctx.to_device(np.asfortranarray(Z_host).T.copy())
        dB = self.ctx.to_device(np.asfortranarray(B).T.copy())
        dF = self.ctx.empty(n * n * 8)
        piv = self.ctx.empty(n * 4)
        self.ctx.lu32_factor(dA, n, n, dF, n, piv)
        it, res = self.ctx.refine_solve(dA, n, n, dF, n, piv, dB, B.shape[1], n, tol, max_iter)
        return dB.to_numpy((B.shape[1], n), np.complex128).T, it, res
