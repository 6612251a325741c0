"""Record golden input/output vectors of the reference (BIELcc compiled from
its own sources into oracle/_ref/ by `make -C oracle ref`).

Run in the build container only (the reference is not on the GPU host):
    python tests/golden/make_golden.py
Writes tests/golden/*.npz.  Seeds are fixed; every file stores its inputs
next to the reference's outputs.  The reference's default numerical
parameters are Calculation_Constants (electrodynamics/constants.h:17-27).
"""
from __future__ import annotations

import gzip
import os
import sys

import numpy as np
import scipy.linalg
from ctypes import c_int as C_int

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from bielcc_amd import mesh  # noqa: E402

REF_DATA = "/root/reference/examples/data/surface_mesh"

# cells of examples/ED_Operators/main0_K.cpp:101-112 and main0_integrals.cpp
RUT1 = np.array([[0., 0., 0.], [2., 0., 0.], [3., 2., 0.], [0., 2., 0.]])
RUT2 = np.array([[0., 2., 0.], [0., 0., 0.], [2., 0., 0.]])
RUT3 = np.array([[3., 2., 0.], [2., 0., 0.], [0., 2., 0.]])

# NumParam sets: defaults; main0_K.cpp:152-169 (x10 splits, p_max 6); main0_R.cpp:110-127
NP_SETS = [
    dict(eps=1e-6, rs=2.0, rs_seg=2.0, analytic_calc=3.0, n_start=8, n_start_seg=10, p_max=1, p_max_seg=1),
    dict(eps=1e-6, rs=2.0, rs_seg=2.0, analytic_calc=3.0, n_start=8, n_start_seg=10, p_max=3, p_max_seg=3),
    dict(eps=1e-5, rs=0.5, rs_seg=0.5, analytic_calc=3.0, n_start=20, n_start_seg=30, p_max=2, p_max_seg=2),
    dict(eps=1e-6, rs=3.0, rs_seg=3.0, analytic_calc=3.0, n_start=160, n_start_seg=400, p_max=1, p_max_seg=1),
    dict(eps=1e-6, rs=2.0, rs_seg=2.0, analytic_calc=3.0, n_start=5, n_start_seg=7, p_max=1, p_max_seg=1),
]


def case_cells():
    q = mesh.sphere_quads(8, 16, jitter=0.005, seed=7)
    t = mesh.triangulate(q)
    quads = [RUT1, q[0], q[37], q[-3]]            # q[0] is a north-pole quad (v0 == v1)
    tris = [RUT2, RUT3, t[3], t[50]]
    return quads, tris


def case_points(cell, rng):
    c = cell.mean(axis=0)
    d = np.max(np.linalg.norm(cell - c, axis=1))
    if len(cell) == 4:
        n = np.cross(cell[2] - cell[0], cell[3] - cell[1])
    else:
        n = np.cross(cell[1] - cell[0], cell[2] - cell[0])
    n /= np.linalg.norm(n)
    on = 0.3 * cell[0] + 0.3 * cell[1] + 0.4 * cell[2]                  # on the cell
    pts = [c, on, c + 0.05 * d * n, c + 0.5 * d * n + 0.2 * (cell[1] - cell[0]),
           c + 2.0 * d * rng.normal(size=3), c + 10.0 * d * n, np.array([2., 1.5, 0.]),
           np.array([10., 10., 10.])]
    return [np.asarray(p, dtype=np.float64) for p in pts]


def gen_ops(ref):
    rng = np.random.default_rng(11)
    quads, tris = case_cells()
    rec = dict(cells4=[], cells3=[], x4=[], x3=[], j4=[], j3=[], np4=[], np3=[], k4=[], k3=[], op4=[], op3=[],
               res4=[], res3=[])
    for nv, cells in ((4, quads), (3, tris)):
        for cell in cells:
            for x in case_points(cell, rng):
                for ki, k in enumerate([1.0, 2.0, 2.0 + 0.3j]):
                    j = (rng.normal(size=3) + 1j * rng.normal(size=3)) if ki else np.array([1., 1., 1.])
                    for npi, npd in enumerate(NP_SETS):
                        if npi == 3 and (ki != 0 or not np.allclose(x, [2., 1.5, 0.])):
                            continue      # the expensive set only on the examples' own case
                        npp = O.num_param(**npd)
                        for op in range(8):
                            r = ref.op(op, cell, x, j, npp, k)
                            rec[f"cells{nv}"].append(cell); rec[f"x{nv}"].append(x); rec[f"j{nv}"].append(j)
                            rec[f"np{nv}"].append(npi); rec[f"k{nv}"].append(k); rec[f"op{nv}"].append(op)
                            rec[f"res{nv}"].append(r)
    out = {k: np.asarray(v) for k, v in rec.items()}
    out["np_sets"] = np.array([[d[k] for k in ("eps", "rs", "rs_seg", "analytic_calc", "n_start", "n_start_seg",
                                                "p_max", "p_max_seg")] for d in NP_SETS])
    np.savez_compressed(os.path.join(HERE, "ops.npz"), **out)
    print("ops:", len(out["op4"]) + len(out["op3"]), "cases")


def gen_integrators(ref):
    rng = np.random.default_rng(12)
    quads, tris = case_cells()
    KP = [(1e-16, 1e-16, 1, 8), (0.5, 1e-6, 1, 8), (0.5, 1e-3, 3, 6), (2.0, 0.1, 4, 5)]  # R, Rmin, p_max, n
    rec = {k: [] for k in ["kind", "f", "nv", "cell", "x", "j", "kp", "ip", "k", "res", "R_out"]}
    for nv, cells in ((4, quads), (3, tris)):
        for cell in cells[:3]:
            for x in case_points(cell, rng)[:6]:
                for k in (1.5, 2.0 - 0.4j):
                    j = rng.normal(size=3) + 1j * rng.normal(size=3)
                    for (R, Rmin, pm, n0) in KP:
                        for f in range(12):
                            def mk():
                                return O.kernel_param(R, Rmin, k, vec_dbl=(0.3, -0.2, 0.5), n=(0.0, 0.6, 0.8),
                                                      current=(1 + 1j, 0.5, -0.2j))
                            for idim in (1, 3):
                                ip = O.IntegralParam(idim, n0, pm, 1e-6)
                                kp = mk(); r = ref.integral_pnt(f, cell, x, kp, ip)
                                rec["kind"].append(0); rec["f"].append(f); rec["nv"].append(nv)
                                rec["cell"].append(np.pad(cell, ((0, 4 - nv), (0, 0)))); rec["x"].append(x)
                                rec["j"].append(j); rec["kp"].append([R, Rmin]); rec["ip"].append([idim, n0, pm, 1e-6])
                                rec["k"].append(k); rec["res"].append(r); rec["R_out"].append(kp.smoothR)
                            ip = O.IntegralParam(3, n0 + 2, pm, 1e-6)
                            kp = mk(); r = ref.integral_seg(f, cell[0], cell[2], x, kp, ip)
                            seg = np.zeros((4, 3)); seg[0], seg[1] = cell[0], cell[2]
                            rec["kind"].append(1); rec["f"].append(f); rec["nv"].append(2); rec["cell"].append(seg)
                            rec["x"].append(x); rec["j"].append(j); rec["kp"].append([R, Rmin])
                            rec["ip"].append([3, n0 + 2, pm, 1e-6]); rec["k"].append(k); rec["res"].append(r)
                            rec["R_out"].append(kp.smoothR)
                            if f < 4:
                                continue   # K0 with a real integrand needs a real current
                            for var in (0, 1):
                                kp = mk(); r = ref.k0(var, f, cell, x, j, kp, ip)
                                rec["kind"].append(2 + var); rec["f"].append(f); rec["nv"].append(nv)
                                rec["cell"].append(np.pad(cell, ((0, 4 - nv), (0, 0)))); rec["x"].append(x)
                                rec["j"].append(j); rec["kp"].append([R, Rmin]); rec["ip"].append([3, n0 + 2, pm, 1e-6])
                                rec["k"].append(k); rec["res"].append(r); rec["R_out"].append(kp.smoothR)
    out = {k: np.asarray(v) for k, v in rec.items()}
    np.savez_compressed(os.path.join(HERE, "integrators.npz"), **out)
    print("integrators:", len(out["f"]), "cases")


def gen_kernels_analytic(ref):
    rng = np.random.default_rng(13)
    rec = {k: [] for k in ["f", "x", "y", "kp", "k", "res"]}
    pts = [rng.normal(size=3) for _ in range(12)] + [np.array([0.3, 0.1, 0.2])]
    for f in range(12):
        for a in range(len(pts)):
            x = pts[a]
            for dy in (1e-9, 1e-5, 0.01, 0.3, 2.0):
                y = x + dy * rng.normal(size=3)
                for (R, Rmin) in ((1e-16, 1e-16), (0.2, 1e-6), (5.0, 0.1)):
                    for k in (1.0, 3.0 + 0.5j, 1e-5):
                        kp = O.kernel_param(R, Rmin, k, vec_dbl=(0.3, -0.2, 0.5), n=(0.0, 0.6, 0.8),
                                            current=(1 + 1j, 0.5, -0.2j))
                        rec["f"].append(f); rec["x"].append(x); rec["y"].append(y); rec["kp"].append([R, Rmin])
                        rec["k"].append(k); rec["res"].append(ref.kernel(f, x, y, kp))
    out = {k: np.asarray(v) for k, v in rec.items()}
    # analytic integrals
    quads, tris = case_cells()
    an = {k: [] for k in ["which", "nv", "cell", "x", "k", "res"]}
    for nv, cells in ((4, quads), (3, tris)):
        for cell in cells:
            for x in case_points(cell, rng):
                for k in (1.0, 2.5 + 0.2j):
                    ip = O.IntegralParam(1, 8, 1, 1e-6)
                    pad = np.pad(cell, ((0, 4 - nv), (0, 0)))
                    an["which"].append(0); an["nv"].append(nv); an["cell"].append(pad); an["x"].append(x)
                    an["k"].append(k); an["res"].append([ref.integral1divr(cell, x), 0, 0])
                    an["which"].append(2); an["nv"].append(nv); an["cell"].append(pad); an["x"].append(x)
                    an["k"].append(k); an["res"].append([ref.surf_hs(cell, x, 3.0, ip, k), 0, 0])
                    seg = np.zeros((4, 3)); seg[0], seg[1] = cell[0], cell[1]
                    g = ref.integral_grad1divr_seg(cell[0], cell[1], x)
                    an["which"].append(1); an["nv"].append(2); an["cell"].append(seg); an["x"].append(x)
                    an["k"].append(k); an["res"].append(g.astype(complex))
                    ip3 = O.IntegralParam(3, 10, 2, 1e-6)
                    an["which"].append(3); an["nv"].append(2); an["cell"].append(seg); an["x"].append(x)
                    an["k"].append(k); an["res"].append(ref.seg_hs(cell[0], cell[1], x, 3.0, ip3, k))
                    an["which"].append(4); an["nv"].append(nv); an["cell"].append(pad); an["x"].append(x)
                    an["k"].append(k)
                    an["res"].append(ref.k0_hs(cell, x, np.array([1 + 0.5j, -0.3, 0.2j]), ip3, 3.0, k))
    for key, v in an.items():
        out["an_" + key] = np.asarray(v)
    # cell tables
    ct = {k: [] for k in ["nv", "cell", "tab"]}
    for nv, cells in ((4, quads), (3, tris)):
        for cell in cells:
            d = ref.cellfill(cell)
            ct["nv"].append(nv); ct["cell"].append(np.pad(cell, ((0, 4 - nv), (0, 0))))
            ct["tab"].append(np.concatenate([d["rkt"], d["norm"], d["tau1"], d["tau2"], [d["area"], d["diam"]],
                                             d["norm_func"], d["center"]]))
    for key, v in ct.items():
        out["cf_" + key] = np.asarray(v)
    np.savez_compressed(os.path.join(HERE, "kernels.npz"), **out)
    print("kernels:", len(out["f"]), "analytic:", len(out["an_which"]))


def ref_matrix(ref, verts, k, npd, op=2):
    """Z[2i+a, 2j+b] = tau_i^a . Op(tau_j^b, rkt_i, cell_j) from the reference's operator."""
    n = verts.shape[0]
    tabs = [ref.cellfill(c) for c in verts]
    rkt = np.array([t["rkt"] for t in tabs])
    tau = np.array([[t["tau1"], t["tau2"]] for t in tabs])       # [n][2][3]
    npp = O.num_param(**npd)
    Z = np.empty((2 * n, 2 * n), dtype=complex)
    for b in range(2):
        res = ref.op_pairs(op, verts, tau[:, b, :], rkt, npp, k)   # [i][j][3]
        for a in range(2):
            t = tau[:, a, :]
            z = (res[:, :, 0] * t[:, None, 0] + res[:, :, 1] * t[:, None, 1]) + res[:, :, 2] * t[:, None, 2]
            Z[a::2, b::2] = z
    return Z, rkt, tau


def gen_config1(ref, nv=4):
    """Config 1: PEC sphere r = 1, 16x32 theta-phi quads, +-0.5% jitter (seed 1),
    ka = 2, E along x propagating +z, 1,000 points on r = 3 (seed 2)."""
    k = 2.0
    verts = mesh.sphere_quads(16, 32, radius=1.0, jitter=0.005, seed=1)
    if nv == 3:
        verts = mesh.triangulate(verts)
    n = verts.shape[0]
    npd = NP_SETS[0]
    Z, rkt, tau = ref_matrix(ref, verts, k, npd, op=2)
    ZR, _, _ = ref_matrix(ref, verts, k, npd, op=6)
    kvec, e0 = np.array([0., 0., k]), np.array([1., 0., 0.])
    einc = ref.plane_wave_e(rkt, kvec, e0)
    rhs = np.empty(2 * n, dtype=complex)
    for a in range(2):
        t = tau[:, a, :]
        rhs[a::2] = -((einc[:, 0] * t[:, 0] + einc[:, 1] * t[:, 1]) + einc[:, 2] * t[:, 2])
    lu = scipy.linalg.lu_factor(Z)
    b = scipy.linalg.lu_solve(lu, rhs)
    jE = ref.ref_currents_by_basis(tau.reshape(n, 6), b)
    pts = mesh.sphere_points(1000, 3.0, seed=2)
    E = ref.get_field_e(verts, jE, 1.0, kvec, e0, k, O.num_param(**npd), pts)
    Einc = ref.plane_wave_e(pts, kvec, e0)
    H = ref.op_pairs(6, verts, jE, pts, O.num_param(**npd), k).sum(axis=1)
    rng = np.random.default_rng(5)
    w = rng.normal(size=2 * n) + 1j * rng.normal(size=2 * n)
    sample = rng.choice(n * n, size=4000, replace=False)
    si, sj = sample // n, sample % n
    blocks = np.stack([Z[2 * si + a, 2 * sj + b] for a in range(2) for b in range(2)], axis=1)
    rblocks = np.stack([ZR[2 * si + a, 2 * sj + b] for a in range(2) for b in range(2)], axis=1)
    diag = np.stack([Z[2 * np.arange(n) + a, 2 * np.arange(n) + b] for a in range(2) for b in range(2)], axis=1)
    name = "config1.npz" if nv == 4 else "config1_tri.npz"
    np.savez_compressed(os.path.join(HERE, name), verts=verts, k=k, kvec=kvec, e0=e0, pts=pts,
                        num_param=np.array([npd[kk] for kk in ("eps", "rs", "rs_seg", "analytic_calc", "n_start",
                                                               "n_start_seg", "p_max", "p_max_seg")]),
                        sample_i=si, sample_j=sj, blocks=blocks, rblocks=rblocks, diag=diag, row_w=w,
                        row_sums=Z @ w, rrow_sums=ZR @ w, col_sums=w @ Z, rhs=rhs, b=b, jE=jE, E=E, Einc=Einc,
                        H=H, Znorm=np.linalg.norm(Z), cond=np.linalg.cond(Z))
    print(name, "n =", n, "cond(Z) =", np.linalg.cond(Z))


def gen_mesh_files(ref):
    out = {}
    for fname, nv in (("sphere_quad.dat", 4), ("sphere_triangle.dat", 3), ("sphere_quad.vtk", 4)):
        d = ref.ref_read_mesh(os.path.join(REF_DATA, fname), nv)
        key = fname.replace(".", "_")
        for kk, v in d.items():
            out[f"{key}_{kk}"] = np.asarray(v)
    np.savez_compressed(os.path.join(HERE, "mesh_files.npz"), **out)
    print("mesh files:", [k for k in out if k.endswith("_verts")])


def writer_values(rng, n):
    """points and complex 3-vectors with the formatting edge cases of ostream's
    default (%g, 6 digits): exponents, exact zeros, negative zero, rounding."""
    pts = rng.normal(size=(n, 3)) * np.array([1.0, 1e-7, 12345.678])
    v = (rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))) * np.logspace(-12, 6, n)[:, None]
    v[0] = [0.0, -0.0 + 1j * 0.0, 1.0 - 1j]
    v[1] = [1234567.0 + 0.5j, -0.000123456789 - 1e-300j, 3.0 + 4.0j]
    return pts, v


def gen_writers(ref):
    """Files written by the reference's own writers (IO_VTK::WritePntsWithField_C,
    IO_AERECO::WriteVecs_C and ::WriteField_C) for byte-level comparison."""
    import ctypes as C
    rng = np.random.default_rng(61)
    out = os.path.join(HERE, "writers")
    os.makedirs(out, exist_ok=True)
    pts, v = writer_values(rng, 9)
    vals = np.ascontiguousarray(np.stack([v.real, v.imag], -1).reshape(-1))
    P = np.ascontiguousarray(pts.reshape(-1))
    dp = C.POINTER(C.c_double)
    lib = ref.lib
    assert lib.ref_write_vtk_field_c(os.path.join(out, "field_c.vtk").encode(), 9, P.ctypes.data_as(dp),
                                     vals.ctypes.data_as(dp)) == 0
    assert lib.ref_write_aereco_vecs_c(os.path.join(out, "vecs_real.gv").encode(),
                                       os.path.join(out, "vecs_image.gv").encode(), 9,
                                       vals.ctypes.data_as(dp)) == 0
    # WriteField_C writes (n/2)^2 rows: n = 4 is the largest size it reads in bounds
    assert lib.ref_write_aereco_field_c(os.path.join(out, "field_real.gv").encode(),
                                        os.path.join(out, "field_image.gv").encode(),
                                        os.path.join(out, "field_abs.gdr").encode(), 4,
                                        vals.ctypes.data_as(dp)) == 0
    np.savez(os.path.join(out, "inputs.npz"), pts=pts, vals=v)
    print("writers:", sorted(os.listdir(out)))


RWG_MESHES = (("sphere_quad", 4), ("sphere_triangle", 3), ("pole_sphere", 4), ("open_face", 4),
              ("torus_tri", 3))


def rwg_inputs(tmp):
    """Mesh files of the RwgSurfMesh cases: the reference's two example meshes,
    plus written ones: a sphere with degenerate pole quads (pseudo-edges), one
    open cube face (boundary edges) and a triangulated torus."""
    paths = {"sphere_quad": os.path.join(REF_DATA, "sphere_quad.dat"),
             "sphere_triangle": os.path.join(REF_DATA, "sphere_triangle.dat")}
    extra = {"pole_sphere": mesh.sphere_quads(5, 10, jitter=0.005, seed=4),
             "open_face": mesh.cube_quads(4)[:16],
             "torus_tri": mesh.triangulate(mesh.torus_quads(12, 6, R=1.0, r=0.3))}
    for name, v in extra.items():
        paths[name] = os.path.join(tmp, name + ".dat")
        mesh.write_dat(paths[name], v)
    return paths


def gen_rwg(ref):
    """RwgSurfMesh topology arrays and the reference's VTK export of each case."""
    import ctypes as C
    import tempfile
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    out, wdir = {}, os.path.join(HERE, "writers")
    with tempfile.TemporaryDirectory() as tmp:
        paths = rwg_inputs(tmp)
        for name, nv in RWG_MESHES:
            cap = 200000
            ncell, counts = C.c_int(0), (C.c_int * 3)()
            arrs = dict(points=np.zeros(3 * cap), cell_points=np.zeros(cap, np.int32),
                        point_num_cells=np.zeros(cap, np.int32), point_num_edges=np.zeros(cap, np.int32),
                        edges=np.zeros(2 * cap, np.int32), edge_num_cells=np.zeros(cap, np.int32),
                        cell_edges=np.zeros(cap, np.int32), cell_dir=np.zeros(cap, np.int32),
                        edge_mid=np.zeros(3 * cap), edge_ort=np.zeros(3 * cap),
                        nghb_off=np.zeros(cap + 1, np.int32), nghb=np.zeros(4 * cap, np.int32))
            ptr = lambda a: a.ctypes.data_as(dp if a.dtype == np.float64 else ip)
            vtk = os.path.join(wdir, f"rwg_{name}.vtk")
            rc = ref.lib.ref_rwg(paths[name].encode(), nv, cap, C.byref(ncell), counts,
                                 *[ptr(arrs[k]) for k in ("points", "cell_points", "point_num_cells",
                                                          "point_num_edges", "edges", "edge_num_cells",
                                                          "cell_edges", "cell_dir", "edge_mid", "edge_ort",
                                                          "nghb_off", "nghb")], vtk.encode())
            assert rc == 0, (name, rc)
            with open(vtk, "rb") as fi, gzip.GzipFile(vtk + ".gz", "wb", mtime=0) as fo:
                fo.write(fi.read())
            os.remove(vtk)
            n, npt, ne, ng = ncell.value, counts[0], counts[1], counts[2]
            d = dict(points=arrs["points"][:3 * npt].reshape(npt, 3),
                     cell_points=arrs["cell_points"][:n * nv].reshape(n, nv),
                     point_num_cells=arrs["point_num_cells"][:npt], point_num_edges=arrs["point_num_edges"][:npt],
                     edges=arrs["edges"][:2 * ne].reshape(ne, 2), edge_num_cells=arrs["edge_num_cells"][:ne],
                     cell_edges=arrs["cell_edges"][:n * nv].reshape(n, nv),
                     cell_dir=arrs["cell_dir"][:n * nv].reshape(n, nv),
                     edge_mid=arrs["edge_mid"][:3 * ne].reshape(ne, 3), edge_ort=arrs["edge_ort"][:3 * ne].reshape(ne, 3),
                     nghb_off=arrs["nghb_off"][:n * nv + 1], nghb=arrs["nghb"][:ng])
            rd = ref.ref_read_mesh(paths[name], nv)
            d["verts"], d["norm"] = np.asarray(rd["verts"]), np.asarray(rd["norm"])
            for k, v in d.items():
                out[f"{name}_{k}"] = v
            print("rwg", name, "cells", n, "points", npt, "edges", ne, "neighbour entries", ng)
    np.savez_compressed(os.path.join(HERE, "rwg.npz"), **out)


def gen_aver(ref):
    """Ker_BasicAverFunc (kernel_lib.h:626-634) at seeded pairs inside and outside eps."""
    import ctypes as C
    rng = np.random.default_rng(71)
    dp = C.POINTER(C.c_double)
    ref.lib.ref_basic_aver.restype = C.c_double
    ref.lib.ref_basic_aver.argtypes = [dp, dp, C.c_double]
    x = rng.normal(size=(400, 3))
    eps = rng.uniform(0.05, 2.0, 400)
    y = x + rng.normal(size=(400, 3)) * eps[:, None] * rng.uniform(0.1, 1.0, (400, 1))
    y[0] = x[0]                                             # r = 0
    y[1] = x[1] + np.array([eps[1], 0.0, 0.0])              # r = eps exactly
    res = np.array([ref.lib.ref_basic_aver(np.ascontiguousarray(x[i]).ctypes.data_as(dp),
                                           np.ascontiguousarray(y[i]).ctypes.data_as(dp), float(eps[i]))
                    for i in range(400)])
    np.savez_compressed(os.path.join(HERE, "aver.npz"), x=x, y=y, eps=eps, basic=res)
    print("aver: inside eps", int((res > 0).sum()), "of", len(res))


def gen_incident_y(ref):
    """PlaneWaveFieldH (incident_field.h:40-59) through a MaterialParam, and the F(y)
    integrators: IntegralUniversal over quads and triangles (integral_universal.h:252-431)
    and the F(y) segment form (integral_universal_seg.h:120-179), with the constant
    integrand of examples/integrators/main0_integrals.cpp:77-81 and PlaneWaveFieldE."""
    rng = np.random.default_rng(81)
    out = {}
    pts = rng.normal(size=(300, 3)) * 3.0
    kv = rng.normal(size=(300, 3)) * 4.0
    e0 = np.cross(kv, rng.normal(size=(300, 3)))
    e0 /= np.linalg.norm(e0, axis=1, keepdims=True)
    freq = rng.uniform(1e6, 1e10, 300)
    mu_r = rng.uniform(0.5, 4.0, 300)
    H, mu = [], []
    for i in range(300):
        h, m = ref.plane_wave_h(pts[i:i + 1], kv[i], e0[i], float(freq[i]), mu_r=float(mu_r[i]))
        H.append(h[0]); mu.append(m)
    out.update(h_pts=pts, h_kvec=kv, h_e0=e0, h_freq=freq, h_mu_r=mu_r, h_mu=np.array(mu), h_H=np.array(H))
    quads, tris = case_cells()
    KP = [(1e-16, 1e-16, 1, 8), (0.5, 1e-6, 1, 8), (0.5, 1e-3, 3, 6), (2.0, 0.1, 4, 5)]  # R, Rmin, p_max, n
    rec = {k: [] for k in ["f", "nv", "kt", "cell", "kvec", "e0", "kp", "ip", "res", "R_out"]}
    segs = [np.array([c[0], c[2]]) for c in quads[:2] + tris[:2]] + [np.array([[0., 0., 0.], [3., 4., 0.]])]
    for nv, cells in ((4, quads), (3, tris), (2, segs)):
        for cell in cells:
            for kt in (0, 1):
                for (R, Rmin, pm, n0) in KP:
                    for f in (0, 1):
                        for idim in ((1, 3) if f == 1 else (1,)):
                            kvec = rng.normal(size=3) * 2.0
                            e = rng.normal(size=3)
                            kp = O.kernel_param(R, Rmin, 1.5 if kt == 0 else 1.5 - 0.3j, vec_dbl=kvec)
                            ip = O.IntegralParam(idim, n0, pm, 1e-6)
                            ref._kt = [C_int(kt)]
                            r = ref.integral_y(f, cell, e, kp, ip)
                            rec["f"].append(f); rec["nv"].append(nv); rec["kt"].append(kt)
                            rec["cell"].append(np.pad(cell, ((0, 4 - len(cell)), (0, 0))))
                            rec["kvec"].append(kvec); rec["e0"].append(e); rec["kp"].append([R, Rmin])
                            rec["ip"].append([idim, n0, pm, 1e-6]); rec["res"].append(r)
                            rec["R_out"].append(kp.smoothR)
    ref._kt = [C_int(1)]
    out.update({"y_" + k: np.asarray(v) for k, v in rec.items()})
    np.savez_compressed(os.path.join(HERE, "incident_y.npz"), **out)
    print("incident_y:", len(H), "PlaneWaveFieldH points,", len(rec["f"]), "F(y) integrals")


if __name__ == "__main__":
    ref = O.Checker("ref")
    which = sys.argv[1:] or ["ops", "integrators", "kernels", "config1", "config1_tri", "mesh_files", "writers",
                             "rwg", "aver", "incident_y"]
    if "ops" in which:
        gen_ops(ref)
    if "integrators" in which:
        gen_integrators(ref)
    if "kernels" in which:
        gen_kernels_analytic(ref)
    if "config1" in which:
        gen_config1(ref, 4)
    if "config1_tri" in which:
        gen_config1(ref, 3)
    if "mesh_files" in which:
        gen_mesh_files(ref)
    if "writers" in which:
        gen_writers(ref)
    if "rwg" in which:
        gen_rwg(ref)
    if "aver" in which:
        gen_aver(ref)
    if "incident_y" in which:
        gen_incident_y(ref)
