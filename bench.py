#!/usr/bin/env python3
"""Benchmark of the Stratton-Chu K/R path on one MI355X.

Workload (default, --config sphere20k): perfectly conducting sphere of radius
1, 100 x 200 theta-phi quads (20,000 cells, +-0.5% radial jitter, seed 1),
ka = 10.  One step = the full K collocation matrix (40,000 x 40,000 complex
FP64, 25.6 GB in HBM, K_RotRot_HS_Colloc per (point, cell, basis)) plus the
scattered E and H at 100,000 points on a 200 x 500 theta-phi grid of radius 5
(K_RotRot_Far + R_Rot_Colloc per (point, cell)).  Inputs are resident in HBM
before timing; the currents come from one LU solve in the warm-up.

Metric: cell-pair interactions per second (assembly pairs + field pairs) over
the timed steps.  Also reported: the LU factorisation and solve times, the
roofline of the dominant kernel (field tile kernel) and a CPU baseline of the
reference compiled from its own sources (or the oracle when that library is
absent), timed on a bounded sample of the same workload.

The mixed-precision form (FP32 LU on the matrix cores, FP64 iterative
refinement to a 1e-10 residual, FP32 far-pair node sums in the field) is timed
once on the same workload after the FP64 steps: its assembly, factorisation,
refinement and field times, one cold solve end to end against the FP64 form,
and its error against the FP64 results (currents and field).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# nominal algorithmic FP64 flops per cell pair (n_start = 8, n_start_seg = 10,
# flop weights: add/sub/mul 1, fma 2, sqrt 16, div 12, sincos 46)
FLOPS_FIELD_PAIR = 12544      # fused E (K_RotRot_Far) + H (R_Rot_Colloc), far pair
FLOPS_ASM_PAIR = 10256        # K_RotRot_HS_Colloc block, far quad pair
PEAK_FP64_TFLOPS = 78.6       # MI355X FP64 vector (dense) peak
# Work the sphere20k field kernel (k_field_far<false, false, 3, 1>) executes per far pair, counted by
# rocprofv3 --pmc: FP64 flop = (2 FMA + MUL + ADD) x 64 lanes from SQ_INSTS_VALU_{FMA,MUL,ADD}_F64, and
# VALU lane instructions = SQ_INSTS_VALU x 64, each per launch / 2e9 pairs.  The nominal convention above
# prices sincos, sqrt and division by weight where the kernel spends a short series and one v_rsq_f64,
# so it counts 1.86x the executed flop; the roofline uses the executed work.
EXEC_FLOP_FIELD_PAIR = 6757
VALU_LANE_INSTR_FIELD_PAIR = 4472
VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4   # wave64 VALU instructions: 1,024 SIMDs x 2.4 GHz, one per 4 cycles
# HBM bytes moved per k_field_far launch on sphere20k: rocprofv3 --pmc FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00
# (scale factors from a 1 GiB copy / fill in the same run); the kernel is VALU-bound, the bytes are L2 re-reads
FIELD_TRAFFIC_BYTES = {"sphere20k": 1.0565e9}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fp64_calls(ctx):
    """Mixed field calls that ran with FP64 far sums: chosen by the sample or by the phase rule."""
    return ctx.get_tuning("mixed_probe_fp64") + ctx.get_tuning("mixed_auto_fp64")


def workload(name):
    from bielcc_amd import mesh
    if name == "sphere20k":
        return dict(verts=mesh.sphere_quads(100, 200, radius=1.0, jitter=0.005, seed=1), k=10.0,
                    pts=mesh.grid_points(200, 500, 5.0), desc="sphere r=1, 100x200 quads (20,000 cells), ka=10, "
                                                                  "field at 200x500 points r=5")
    if name == "sphere512":
        return dict(verts=mesh.sphere_quads(16, 32, radius=1.0, jitter=0.005, seed=1), k=2.0,
                    pts=mesh.sphere_points(1000, 3.0, seed=2), desc="sphere r=1, 16x32 quads (512 cells), ka=2, "
                                                                    "field at 1,000 points r=3")
    raise SystemExit(f"unknown config {name}")


def cpu_baseline(verts, pts, k, budget_s=20.0):
    """Time the reference (or the oracle) on a bounded sample of the same
    workload: field pairs (K_RotRot_Far + R_Rot_Colloc) for a sample of
    points over all cells, and assembly pairs (K_RotRot_HS_Colloc for both
    basis vectors) for a sample of collocation points over all cells."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    kind = "reference"
    try:
        chk = O.Checker("ref")
    except (FileNotFoundError, OSError):
        chk, kind = O.Checker("oracle"), "port"
    npp = O.num_param()
    n = verts.shape[0]
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    rng = np.random.default_rng(0)
    j = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    tabs = [chk.cellfill(v) for v in verts]
    rkt = np.array([t["rkt"] for t in tabs])
    tau = np.array([[t["tau1"], t["tau2"]] for t in tabs])
    fsel = pts[rng.permutation(len(pts))]
    asel = rkt[rng.permutation(n)]

    def field_pts(m):
        t0 = time.perf_counter()
        chk.op_pairs(4, verts, j, fsel[:m], npp, k)
        chk.op_pairs(6, verts, j, fsel[:m], npp, k)
        return time.perf_counter() - t0

    def asm_pts(m):
        t0 = time.perf_counter()
        for b in range(2):
            chk.op_pairs(2, verts, tau[:, b, :], asel[:m], npp, k)
        return time.perf_counter() - t0

    # calibrate on a sample that keeps every thread busy, then scale to the budget
    cal = 2 * threads
    per_f = field_pts(cal) / cal
    per_a = asm_pts(cal) / cal
    m_f = int(max(cal, min(len(pts), 0.6 * budget_s / per_f)))
    m_a = int(max(cal, min(n, 0.4 * budget_s / per_a)))
    tf = field_pts(m_f)
    ta = asm_pts(m_a)
    pairs = (m_f + m_a) * n
    return {"value": pairs / (tf + ta), "unit": "cell-pair interactions/s", "cores": threads, "kind": kind,
            "sample": f"{m_f} field points x {n} cells (K_RotRot_Far + R_Rot_Colloc) and {m_a} collocation "
                      f"points x {n} cells x 2 basis vectors (K_RotRot_HS_Colloc); OpenMP over points; "
                      f"{tf + ta:.1f} s of CPU work",
            "field_pairs_per_s": m_f * n / tf, "assembly_pairs_per_s": m_a * n / ta}


def c1_full():
    """The 512-cell sphere (ka = 2, 1,000 points at r = 3) end to end on the host
    cores through the compiled reference's own routines -- every matrix entry from
    K_RotRot_HS_Colloc (collocation point x cell x basis vector), the right-hand side
    from PlaneWaveFieldE, SurfaceCurrents_byBasis, GetFieldE_Ideal for E and
    R_Rot_Colloc for H; the solve by LAPACK (numpy), which the reference lacks --
    timed phase by phase, beside the same chain on the GPU (the library's assembly,
    LU, currents and field drivers), with the GPU's currents and total E against
    the host chain's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import bielcc_amd as B
    kind = "reference"
    try:
        chk = O.Checker("ref")
    except (FileNotFoundError, OSError):
        chk, kind = O.Checker("oracle"), "port"
    w = workload("sphere512")
    verts, k, pts = w["verts"], w["k"], w["pts"]
    n, M = verts.shape[0], pts.shape[0]
    n2 = 2 * n
    kvec, e0 = np.array([0.0, 0.0, k]), np.array([1.0, 0.0, 0.0])
    npp = O.num_param()
    cpu = {}
    t0 = time.perf_counter()
    tabs = [chk.cellfill(v) for v in verts]
    rkt = np.array([t["rkt"] for t in tabs])
    tau = np.array([[t["tau1"], t["tau2"]] for t in tabs])
    Z = np.empty((n2, n2), complex)
    for b in range(2):
        Ob = chk.op_pairs(2, verts, tau[:, b, :], rkt, npp, k)    # [point i][cell j][3]
        for a in range(2):
            Z[a::2, b::2] = np.einsum("ijg,ig->ij", Ob, tau[:, a, :])
    einc = chk.plane_wave_e(rkt, kvec, e0)
    rhs = -np.einsum("iag,ig->ia", tau, einc).reshape(n2)
    cpu["assemble"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    b = np.linalg.solve(Z, rhs)
    cpu["solve"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    j = chk.ref_currents_by_basis(tau, b) if kind == "reference" else chk.currents_by_basis(verts, b)
    cpu["currents"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    E = chk.get_field_e(verts, j, 1.0, kvec, e0, k, npp, pts)
    cpu["field_e"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    chk.op_pairs(6, verts, j, pts, npp, k).sum(axis=1)
    cpu["field_h"] = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)

    ctx = B.Context(0)
    gnp = B.num_param()
    gpu = {}
    m = ctx.mesh(verts)
    dZ, dpiv, dB, dJ = ctx.empty(n2 * n2 * 16), ctx.empty(n2 * 4), ctx.empty(n2 * 16), ctx.empty(n * 48)
    dP, dE, dH = ctx.to_device(pts), ctx.empty(M * 48), ctx.empty(M * 48)
    for rep in range(2):    # the first pass builds tables and grows work buffers
        t0 = time.perf_counter()
        m.assemble_into(dZ, n2, gnp, k)
        m.rhs_planewave_into(dB, n2, kvec, e0)
        ctx.sync()
        gpu["assemble"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        ctx.lu_factor(dZ, n2, n2, dpiv)
        ctx.lu_solve(dZ, n2, n2, dpiv, dB, 1, n2)
        ctx.sync()
        gpu["solve"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        m.currents_into(dB, n2, 1, dJ)
        ctx.sync()
        gpu["currents"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        m.field_into(gnp, k, dJ, dP, M, dE, dH)
        ctx.sync()
        gpu["field_e_h"] = time.perf_counter() - t0
    jg = dJ.to_numpy((n, 3), np.complex128)
    Eg = m.get_field_e(jg, 1.0, kvec, e0, gnp, k, pts)
    m.close()
    ctx.close()
    out = {"workload": w["desc"] + ", plane wave k = 2 z, E along x", "cpu_kind": kind, "cpu_cores": threads,
           "cpu_seconds": {k_: round(v, 4) for k_, v in cpu.items()}, "cpu_total_s": round(sum(cpu.values()), 4),
           "gpu_seconds": {k_: round(v, 5) for k_, v in gpu.items()}, "gpu_total_s": round(sum(gpu.values()), 5),
           "currents_rel_err": float(np.linalg.norm(jg - j) / np.linalg.norm(j)),
           "E_rel_err": float(np.linalg.norm(Eg - E) / np.linalg.norm(E))}
    out["speedup"] = round(out["cpu_total_s"] / out["gpu_total_s"], 1)
    log(f"[bench] C1 in full: CPU ({kind}, {threads} threads) {out['cpu_total_s']:.2f} s, GPU "
        f"{out['gpu_total_s'] * 1e3:.1f} ms; currents {out['currents_rel_err']:.1e}, E {out['E_rel_err']:.1e}")
    return out


def far_directions(n_theta=181, n_phi=360):
    """Unit directions on the 1-degree theta x phi grid (theta 0..180, phi 0..359)."""
    th = np.deg2rad(np.arange(n_theta))[:, None]
    ph = np.deg2rad(np.arange(n_phi))[None, :]
    d = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th) * np.ones_like(ph)], axis=-1)
    return np.ascontiguousarray(d.reshape(-1, 3))


def _errs(a, b):
    """(max over rows of |a-b| / max(|b|, 1e-3 max |b|), ||a-b|| / ||b||) for [n][3] complex rows."""
    nb = np.linalg.norm(b, axis=1)
    pt = float((np.linalg.norm(a - b, axis=1) / np.maximum(nb, 1e-3 * nb.max())).max())
    return pt, float(np.linalg.norm(a - b) / np.linalg.norm(b))


def near_far_errors(ctx, m, npp, k, dJ64, dJ32, near_pts, dirs):
    """Scattered E and H at near_pts and the far-field pattern at dirs, FP64
    form (FP64 currents, FP64 sums) against the mixed form (refined currents,
    FP32 far-pair node sums); errors and times of both."""
    import bielcc_amd as B
    out = {}
    Mn, nd = near_pts.shape[0], dirs.shape[0]
    dP, dD = ctx.to_device(near_pts), ctx.to_device(dirs)
    dE, dH = ctx.empty(Mn * 48), ctx.empty(Mn * 48)
    res = {}
    for tag, dJ, prec in (("fp64", dJ64, B.PREC_FP64), ("mixed", dJ32, B.PREC_MIXED)):
        probed, fixed = fp64_calls(ctx), ctx.get_tuning("mixed_fixup_points")
        ctx.sync()
        t0 = time.perf_counter()
        m.field_batch_into(npp, k, 1, dJ, dP, Mn, dE, dH, precision=prec)
        ctx.sync()
        t1 = time.perf_counter()
        res[tag] = (dE.to_numpy((Mn, 3), np.complex128), dH.to_numpy((Mn, 3), np.complex128))
        out[f"near_field_{tag}_s"] = round(t1 - t0, 4)
    # mixed call: far sums chosen by the sample, and points recomputed with FP64 sums
    out["near_field_mixed_far_sums"] = "FP64" if fp64_calls(ctx) > probed else "FP32"
    out["near_field_mixed_points_recomputed"] = ctx.get_tuning("mixed_fixup_points") - fixed
    eE, eH = _errs(res["mixed"][0], res["fp64"][0]), _errs(res["mixed"][1], res["fp64"][1])
    out.update({"near_points": Mn, "err_near_E_point_max": eE[0], "err_near_E_set": eE[1],
                "err_near_H_point_max": eH[0], "err_near_H_set": eH[1]})
    dE.free(); dH.free(); dP.free()
    dF, dG = ctx.empty(nd * 48), ctx.empty(nd * 48)
    for tag, dJ, prec in (("fp64", dJ64, B.PREC_FP64), ("mixed", dJ32, B.PREC_MIXED)):
        ctx.sync()
        t0 = time.perf_counter()
        m.far_pattern_into(npp, k, 1, dJ, dD, nd, dF, dG, precision=prec)
        ctx.sync()
        t1 = time.perf_counter()
        res[tag] = (dF.to_numpy((nd, 3), np.complex128), dG.to_numpy((nd, 3), np.complex128))
        out[f"far_pattern_{tag}_s"] = round(t1 - t0, 4)
    eE, eH = _errs(res["mixed"][0], res["fp64"][0]), _errs(res["mixed"][1], res["fp64"][1])
    out.update({"far_directions": nd, "err_far_E_point_max": eE[0], "err_far_E_set": eE[1],
                "err_far_H_point_max": eH[0], "err_far_H_set": eH[1]})
    dF.free(); dG.free(); dD.free()
    return out


def cube_precision(k=10.0, n_face=60):
    """Precision form on the cube of side 2 (6 x 60 x 60 quads): FP64 LU form
    against FP32 LU + FP64 refinement to 1e-10, currents, near field at 10,000
    points on r = 1.5 (partly inside the cube and close to its faces: exact near
    pairs) and the far-field pattern on the 1-degree grid."""
    import bielcc_amd as B
    from bielcc_amd import mesh
    verts = mesh.cube_quads(n_face, side=2.0)
    ctx = B.Context(0)
    m = ctx.mesh(verts)
    npp = B.num_param()
    n, n2 = m.n, 2 * m.n
    kvec, e0 = (0.0, 0.0, k), (1.0, 0.0, 0.0)

    def timed(fn):
        ctx.sync()
        t0 = time.perf_counter()
        fn()
        ctx.sync()
        return time.perf_counter() - t0
    dZ, dF, dpiv = ctx.empty(n2 * n2 * 16), ctx.empty(n2 * n2 * 8), ctx.empty(n2 * 4)
    dB, dJ32, dJ64 = ctx.empty(n2 * 16), ctx.empty(n * 48), ctx.empty(n * 48)
    t_asm = timed(lambda: m.assemble_into(dZ, n2, npp, k))
    _warm_lu32(ctx)   # the FP32 path's kernels loaded before the cold solve is timed
    t_lu32 = timed(lambda: ctx.lu32_factor(dZ, n2, n2, dF, n2, dpiv))
    m.rhs_planewave_into(dB, n2, kvec, e0)
    res, fallback = {}, False
    try:
        t_ref = timed(lambda: res.update(zip(("iters", "resid"), ctx.refine_solve(dZ, n2, n2, dF, n2, dpiv, dB, 1, n2,
                                                                                   1e-10, 30))))
    except B.BielccError as err:
        if err.code != 4:
            raise
        fallback, t_ref = True, float("nan")
    m.currents_into(dB, n2, 1, dJ32)
    dF.free()
    # FP64 form: LU in place of the matrix (assembled again after, if needed)
    t_lu64 = timed(lambda: ctx.lu_factor(dZ, n2, n2, dpiv))
    m.rhs_planewave_into(dB, n2, kvec, e0)
    t_s64 = timed(lambda: ctx.lu_solve(dZ, n2, n2, dpiv, dB, 1, n2))
    m.currents_into(dB, n2, 1, dJ64)
    dZ.free()
    j32, j64 = dJ32.to_numpy((n, 3), np.complex128), dJ64.to_numpy((n, 3), np.complex128)
    out = {"workload": f"cube of side 2, 6 x {n_face} x {n_face} quads ({n} cells, {n2} unknowns), ka = {k}",
           "assembly_s": round(t_asm, 4), "lu32_factor_s": round(t_lu32, 3), "refine_s": round(t_ref, 4),
           "refine_corrections": res.get("iters"), "refine_residual": res.get("resid"),
           "refine_stalled_fp64_needed": fallback,
           "lu64_factor_s": round(t_lu64, 3), "lu64_solve_s": round(t_s64, 4),
           "err_currents_rel": float(np.linalg.norm(j32 - j64) / np.linalg.norm(j64))}
    out.update(near_far_errors(ctx, m, npp, k, dJ64, dJ32, mesh.sphere_points(10000, 1.5, seed=21), far_directions()))
    log(f"[bench] cube precision: LU32 {t_lu32:.2f} s + refine {t_ref:.3f} s ({res.get('iters')} corrections) vs "
        f"LU64 {t_lu64:.2f} s; currents {out['err_currents_rel']:.1e}, near E {out['err_near_E_set']:.1e}, "
        f"far E {out['err_far_E_set']:.1e}")
    for d in (dpiv, dB, dJ32, dJ64):
        d.free()
    m.close()
    ctx.close()
    return out


FLOPS_ASM_PAIR_TRI = 9166    # K_RotRot_HS_Colloc block, far triangle pair (30 contour nodes)


def triangle_variant(steps=2, lu=False):
    """The 20,000-quad sphere with every quad split into two triangles ((0,1,2) and
    (2,3,0)); the zero-area triangle of each of the 400 pole quads is dropped
    (CellTriangle::CellFill divides by the normal's length), leaving 39,600 cells
    and 79,200 unknowns: the full K matrix (100.4 GB) and E and H at the
    100,000 points of radius 5, per step.  The field uses seeded random
    currents (its cost does not depend on them); with lu=True the FP64 and FP32
    LU of the matrix are timed once as well."""
    import bielcc_amd as B
    from bielcc_amd import mesh
    quads = mesh.sphere_quads(100, 200, radius=1.0, jitter=0.005, seed=1)
    verts = mesh.triangulate(quads)
    dropped = 2 * quads.shape[0] - verts.shape[0]
    pts = mesh.grid_points(200, 500, 5.0)
    k = 10.0
    n, M = verts.shape[0], pts.shape[0]
    n2 = 2 * n
    ctx = B.Context(0)
    m = ctx.mesh(verts)
    npp = B.num_param()
    rng = np.random.default_rng(7)
    dJ = ctx.to_device(rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3)))
    dP = ctx.to_device(pts)
    dE, dH = ctx.empty(M * 48), ctx.empty(M * 48)
    dZ = ctx.empty(n2 * n2 * 16)
    log(f"[bench] triangle variant: {n} triangles ({dropped} zero-area pole triangles dropped), {n2} unknowns, "
        f"matrix {n2 * n2 * 16 / 1e9:.1f} GB")
    m.assemble_into(dZ, n2, npp, k)
    m.field_into(npp, k, dJ, dP, M, dE, dH)
    ctx.sync()
    ctx.timing_enable(True)
    ctx.timing_reset()
    t0 = time.perf_counter()
    for s in range(steps):
        m.assemble_into(dZ, n2, npp, k)
        m.field_into(npp, k, dJ, dP, M, dE, dH)
        ctx.sync()
        log(f"[bench] triangle step {s} done ({time.perf_counter() - t0:.2f} s)")
    elapsed = time.perf_counter() - t0
    fld_ms, fld_cnt = ctx.timing("field_far")
    asm_ms, asm_cnt = ctx.timing("assemble_far")
    ctx.timing_enable(False)
    E = dE.to_numpy((M, 3), np.complex128)
    assert np.all(np.isfinite(E)), "non-finite field"
    fld_s, asm_s = fld_ms / max(fld_cnt, 1) / 1e3, asm_ms / max(asm_cnt, 1) / 1e3
    fld_tf = M * n * FLOPS_FIELD_PAIR / fld_s / 1e12
    asm_tf = n * n * FLOPS_ASM_PAIR_TRI / asm_s / 1e12
    out = {"workload": f"sphere r=1, 100x200 quads split into triangles: {n} cells ({dropped} zero-area pole "
                       f"triangles dropped), {n2} unknowns, ka=10, field at 200x500 points r=5",
           "cells": n, "dropped_triangles": dropped, "unknowns": n2, "steps": steps,
           "pairs_per_s": (n * n + M * n) * steps / elapsed, "ms_per_step": elapsed / steps * 1e3,
           "field_kernel": {"kernel": "k_field_far (triangles)", "launch_ms": round(fld_s * 1e3, 3),
                            "nominal_tflops": round(fld_tf, 3), "nominal_frac": round(fld_tf / PEAK_FP64_TFLOPS, 4),
                            "pairs_per_s": M * n / fld_s},
           "assembly_kernel": {"kernel": "k_assemble_far (triangles)", "launch_ms": round(asm_s * 1e3, 3),
                               "nominal_tflops": round(asm_tf, 3), "nominal_frac": round(asm_tf / PEAK_FP64_TFLOPS, 4),
                               "pairs_per_s": n * n / asm_s, "flop_per_pair": FLOPS_ASM_PAIR_TRI}}
    log(f"[bench] triangle variant: {out['pairs_per_s']:.3e} pairs/s; field kernel {fld_s * 1e3:.1f} ms "
        f"({fld_tf:.1f} TFLOP/s), assembly kernel {asm_s * 1e3:.1f} ms ({asm_tf:.1f} TFLOP/s)")
    if lu:
        dpiv = ctx.empty(n2 * 4)
        dF = ctx.empty(n2 * n2 * 8)
        ctx.sync()
        t1 = time.perf_counter()
        ctx.lu32_factor(dZ, n2, n2, dF, n2, dpiv)
        ctx.sync()
        t32 = time.perf_counter() - t1
        dF.free()
        t1 = time.perf_counter()
        ctx.lu_factor(dZ, n2, n2, dpiv)
        ctx.sync()
        t64 = time.perf_counter() - t1
        flop = 8.0 / 3.0 * n2 ** 3
        out["lu"] = {"lu64_s": round(t64, 3), "lu64_tflops": round(flop / t64 / 1e12, 2), "lu32_s": round(t32, 3),
                     "lu32_tflops": round(flop / t32 / 1e12, 2)}
        log(f"[bench] triangle variant LU: FP64 {t64:.2f} s, FP32 {t32:.2f} s")
        dpiv.free()
    for d in (dZ, dJ, dP, dE, dH):
        d.free()
    m.close()
    ctx.close()
    return out


def _warm_lu32(ctx, n=1536):
    """An FP32 and an FP64 LU of a small random matrix in two-level blocking (outer blocks
    of 512), so that the LUs' kernels, the FP16 matrix-core GEMMs of the FP32 LU among
    them, are loaded before anything is timed."""
    rng = np.random.default_rng(3)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    old = ctx.get_tuning("lu_outer_nb")
    ctx.set_tuning("lu_outer_nb", 512)
    try:
        dA, dF, dp = ctx.to_device(np.ascontiguousarray(A.T)), ctx.empty(n * n * 8), ctx.empty(4 * n)
        rbt = ctx.empty(32 * n)
        ctx.lu32_factor(dA, n, n, dF, n, dp)
        ctx.lu32_factor_rbt(dA, n, n, dF, n, dp, rbt)
        ctx.lu_factor(dA, n, n, dp)
        ctx.sync()
        for d in (dA, dF, dp, rbt):
            d.free()
    finally:
        ctx.set_tuning("lu_outer_nb", old)


def mixed_form(ctx, m, npp, k, kvec, e0, dP, M, dJ64, dE64, dH64, lu64_s, solve64_s, step64_s):
    """One cold solve in the mixed-precision form on the workload: FP64 matrix,
    FP32 LU on the matrix cores, FP64 refinement to a 1e-10 residual, currents,
    mixed field (FP32 far-pair sums with the error control, or FP64 sums when a
    1,024-point sample finds more than a quarter of the points beyond the FP32
    error model; which one is reported, and the FP32 sums are also timed alone).
    Errors are against the FP64 form's currents
    and field (which match the reference to 1e-12 / 1e-10).  FP32 far sums in
    the matrix itself are left out: they shift the currents by cond(Z) x 1e-7
    (4e-5 on this workload) and do not shorten the assembly, whose K pairs are
    bounded by the FP64 edge contours."""
    import bielcc_amd as B
    n = m.n
    n2 = 2 * n

    def timed(fn):
        ctx.sync()
        t0 = time.perf_counter()
        fn()
        ctx.sync()
        return time.perf_counter() - t0

    # The FP64 form's kernels are loaded by its warm-up; load the FP32 path's (its FP16
    # matrix-core GEMMs are loaded by the library on first use, about 1 s once per process)
    # with a small FP32 LU that takes that path, before the cold solve is timed.
    _warm_lu32(ctx)
    dZ = ctx.empty(n2 * n2 * 16)
    dF = ctx.empty(n2 * n2 * 8)
    dpiv = ctx.empty(n2 * 4)
    dB = ctx.empty(n2 * 16)
    dJ = ctx.empty(n * 48)
    dE, dH = ctx.empty(M * 48), ctx.empty(M * 48)
    rbt = ctx.empty(32 * n2)
    t_asm = timed(lambda: m.assemble_into(dZ, n2, npp, k))
    t_lu = timed(lambda: ctx.lu32_factor(dZ, n2, n2, dF, n2, dpiv))
    m.rhs_planewave_into(dB, n2, kvec, e0)
    res = {}
    t_ref = timed(lambda: res.update(zip(("iters", "resid"), ctx.refine_solve(dZ, n2, n2, dF, n2, dpiv, dB, 1, n2,
                                                                               1e-10, 30))))
    m.currents_into(dB, n2, 1, dJ)
    probed = fp64_calls(ctx)
    t_fld = timed(lambda: m.field_batch_into(npp, k, 1, dJ, dP, M, dE, dH, precision=B.PREC_MIXED))
    field_sums = "FP64" if fp64_calls(ctx) > probed else "FP32"
    log(f"[bench] mixed: assembly {t_asm:.3f} s, LU32 {t_lu:.2f} s, refine {t_ref:.3f} s "
        f"({res['iters']} corrections, residual {res['resid']:.1e}), field {t_fld:.3f} s ({field_sums} far sums)")
    j64, j32 = dJ64.to_numpy((n, 3), np.complex128), dJ.to_numpy((n, 3), np.complex128)
    E64, E32 = dE64.to_numpy((M, 3), np.complex128), dE.to_numpy((M, 3), np.complex128)
    H64, H32 = dH64.to_numpy((M, 3), np.complex128), dH.to_numpy((M, 3), np.complex128)
    # the FP32 far sums alone (the sample that chooses the sums switched off): mixed field of
    # the FP64 currents, FP32 sums plus the recomputation of the points they miss
    ctx.set_tuning("mixed_probe_pct", 0)
    ctx.set_tuning("mixed_far32", 1)
    fixed0 = ctx.get_tuning("mixed_fixup_points")
    t_fld32 = timed(lambda: m.field_batch_into(npp, k, 1, dJ64, dP, M, dE, dH, precision=B.PREC_MIXED))
    fixed = ctx.get_tuning("mixed_fixup_points") - fixed0
    ctx.set_tuning("mixed_far32", 0)
    ctx.set_tuning("mixed_probe_pct", 25)
    Eq, Hq = dE.to_numpy((M, 3), np.complex128), dH.to_numpy((M, 3), np.complex128)
    from bielcc_amd import mesh
    nearfar = near_far_errors(ctx, m, npp, k, dJ64, dJ, mesh.sphere_points(10000, 1.5, seed=21), far_directions())
    rowrel = lambda a, b: float((np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)).max())
    glob = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    # for reference: the FP32 factors without interchanges of the butterfly-transformed matrix
    t_lu_rbt = timed(lambda: ctx.lu32_factor_rbt(dZ, n2, n2, dF, n2, dpiv, rbt))
    m.rhs_planewave_into(dB, n2, kvec, e0)
    res_rbt = {}
    try:
        t_ref_rbt = timed(lambda: res_rbt.update(zip(("iters", "resid"), ctx.refine_solve(
            dZ, n2, n2, dF, n2, dpiv, dB, 1, n2, 1e-10, 60, rbt=rbt))))
    except B.BielccError as err:
        if err.code != 4:
            raise
        t_ref_rbt, res_rbt = float("nan"), {"iters": -1, "resid": float("nan")}
    for d in (dZ, dF, dpiv, dB, dJ, dE, dH, rbt):
        d.free()
    cold64 = step64_s + lu64_s + solve64_s   # FP64 assembly + field (one timed step) + LU + solve
    cold32 = t_asm + t_lu + t_ref + t_fld
    return {"form": "FP64 matrix; FP32 LU (rank-512 updates as FP16 GEMMs of split operands) + FP64 FGMRES "
                    "refinement to 1e-10; mixed far-pair field sums (near pairs, contours, cross-cell sums FP64)",
            "assembly_s": round(t_asm, 4), "lu32_factor_s": round(t_lu, 3),
            "butterfly_lu32_s": round(t_lu_rbt, 3), "butterfly_refine_s": round(t_ref_rbt, 4),
            "butterfly_refine_corrections": res_rbt["iters"],
            "lu32_tflops": round(8.0 / 3.0 * n2 ** 3 / t_lu / 1e12, 2),
            "refine_s": round(t_ref, 4), "refine_corrections": res["iters"], "refine_residual": res["resid"],
            "field_s": round(t_fld, 4), "field_pairs_per_s": M * n / t_fld, "field_far_sums": field_sums,
            "field_s_fp32_sums": round(t_fld32, 4), "fp32_sums_points_recomputed": fixed,
            "cold_solve_s": round(cold32, 3), "cold_solve_s_fp64": round(cold64, 3),
            "cold_solves_per_s_ratio": round(cold64 / cold32, 3),
            "err_currents_rel": glob(j32, j64),
            "err_field_E_max_rel": rowrel(E32, E64), "err_field_H_max_rel": rowrel(H32, H64),
            "err_field_E_rel": glob(E32, E64), "err_field_H_rel": glob(H32, H64),
            "err_field_E_max_rel_fp32_sums": rowrel(Eq, E64), "err_field_H_max_rel_fp32_sums": rowrel(Hq, H64),
            "near_r1.5_and_far_pattern": nearfar}


def _lu_probe_error(dF, n, piv, AX, X, dtype, itemsize, strip=2048):
    """||P A X - L (U X)|| / ||P A X|| for the device factors dF (n x n column-major,
    read back in column strips) and the pivots piv: a sampled backward error."""
    y = np.zeros_like(X)
    for c0 in range(0, n, strip):
        c1 = min(n, c0 + strip)
        S = dF.to_numpy((c1 - c0, n), dtype, offset=c0 * n * itemsize).T.astype(np.complex128)
        y += np.triu(S, -c0) @ X[c0:c1]
    z = y.copy()
    for c0 in range(0, n, strip):
        c1 = min(n, c0 + strip)
        S = dF.to_numpy((c1 - c0, n), dtype, offset=c0 * n * itemsize).T.astype(np.complex128)
        z += np.tril(S, -c0 - 1) @ y[c0:c1]
    PAX = AX.copy()
    for j in range(n):
        p = int(piv[j])
        if p != j:
            PAX[[j, p]] = PAX[[p, j]]
    return float(np.linalg.norm(PAX - z) / np.linalg.norm(PAX))


def lu_only(n, reps=3):
    """FP64 and FP32 LU of a seeded random complex n x n matrix (entries N(0,1) + i N(0,1)),
    each factored reps times from the same copy; the fastest of each is reported, with
    the factors' sampled backward error ||P A X - L U X|| / ||P A X|| on 4 random columns X.
    Also the FP64 factorisation without interchanges of the randomly butterfly-transformed
    matrix (lu_factor_rbt), and its refined solve of A X = (A X) against X."""
    import bielcc_amd as B
    ctx = B.Context(0)
    rng = np.random.default_rng(1)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    X = rng.normal(size=(n, 4)) + 1j * rng.normal(size=(n, 4))
    AX = A @ X
    dA0 = ctx.to_device(np.ascontiguousarray(A.T))
    del A
    dA, piv, piv32, dF = ctx.empty(n * n * 16), ctx.empty(4 * n), ctx.empty(4 * n), ctx.empty(n * n * 8)

    def timed(fn):
        ctx.sync()
        t0 = time.perf_counter()
        fn()
        ctx.sync()
        return time.perf_counter() - t0
    t64, t32 = [], []
    for _ in range(reps):
        ctx.copy(dA, dA0, n * n * 16)
        ctx.sync()
        t64.append(timed(lambda: ctx.lu_factor(dA, n, n, piv)))
        t32.append(timed(lambda: ctx.lu32_factor(dA0, n, n, dF, n, piv32)))
    flop = 8.0 / 3.0 * n ** 3
    be64 = _lu_probe_error(dA, n, piv.to_numpy((n,), np.int32), AX, X, np.complex128, 16)
    be32 = _lu_probe_error(dF, n, piv32.to_numpy((n,), np.int32), AX, X, np.complex64, 8)
    dF.free()
    # FP64 factors without interchanges of the butterfly-transformed matrix (no grid barrier
    # per column), then the solve of A X = A X_true refined against A
    dR, rbt = ctx.empty(n * n * 16), ctx.empty(32 * n)
    trbt = [timed(lambda: ctx.lu_factor_rbt(dA0, n, n, dR, n, piv, rbt, seed=1)) for _ in range(reps)]
    dB = ctx.to_device(np.ascontiguousarray(AX.T))
    it_rbt, res_rbt = ctx.rbt_solve(dA0, n, n, dR, n, piv, rbt, dB, 4, n, tol=1e-13, max_iter=10)
    err_rbt = float(np.linalg.norm(dB.to_numpy((4, n), np.complex128).T - X) / np.linalg.norm(X))
    dR.free()
    dR = ctx.empty(n * n * 8)
    t32rbt = [timed(lambda: ctx.lu32_factor_rbt(dA0, n, n, dR, n, piv, rbt, seed=1)) for _ in range(reps)]
    out = {"n": n, "lu64_s": round(min(t64), 4), "lu64_tflops": round(flop / min(t64) / 1e12, 2),
           "lu32_s": round(min(t32), 4), "lu32_tflops": round(flop / min(t32) / 1e12, 2),
           "backward_err_probe_fp64": be64, "backward_err_probe_fp32": be32,
           "lu64_rbt_s": round(min(trbt), 4), "lu64_rbt_tflops": round(flop / min(trbt) / 1e12, 2),
           "lu32_rbt_s": round(min(t32rbt), 4),
           "rbt_solve_corrections": it_rbt, "rbt_solve_residual": res_rbt, "rbt_solve_error": err_rbt,
           "panel_aborts": ctx.get_tuning("lu_panel_aborts")}
    log(f"[bench] LU n={n}: FP64 {out['lu64_s']} s ({out['lu64_tflops']} TFLOP/s, backward {be64:.1e}), "
        f"FP32 {out['lu32_s']} s ({out['lu32_tflops']} TFLOP/s, backward {be32:.1e}), "
        f"FP64 butterfly + no interchanges {out['lu64_rbt_s']} s (solve: {it_rbt} corrections, residual "
        f"{res_rbt:.1e}, error {err_rbt:.1e}), FP32 butterfly {out['lu32_rbt_s']} s")
    for d in (dA0, dA, piv, piv32, dR, rbt, dB):
        d.free()
    ctx.close()
    return out


def solve_only(n, reps=3):
    """Triangular solves on the factors of a seeded random complex n x n system
    (N(0,1) + i N(0,1) entries plus 2 sqrt(n) on the diagonal): FP64 solves (lu_solve)
    and FP32-factor refinements to a 1e-12 residual (refine_solve), for 1 to 16
    right-hand sides, through the one-kernel solve (library defaults: FP64 for one
    right-hand side, FP32 for up to 16) and through rocBLAS trsm (tunable trsv = 0).
    Fastest of reps each."""
    import bielcc_amd as B
    ctx = B.Context(0)
    rng = np.random.default_rng(2)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 2 * np.sqrt(n) * np.eye(n)
    dA = ctx.to_device(np.ascontiguousarray(A.T))
    del A
    dL, piv, dF, piv32 = ctx.empty(n * n * 16), ctx.empty(4 * n), ctx.empty(n * n * 8), ctx.empty(4 * n)
    ctx.copy(dL, dA, n * n * 16)
    ctx.lu_factor(dL, n, n, piv)
    ctx.lu32_factor(dA, n, n, dF, n, piv32)
    out = {"n": n, "fp64_solve_ms": {}, "fp32_refine_ms": {}, "refine_corrections": {}}
    for nrhs in (1, 2, 8, 16):
        Bm = rng.normal(size=(nrhs, n)) + 1j * rng.normal(size=(nrhs, n))
        dB0, dB = ctx.to_device(np.ascontiguousarray(Bm)), ctx.empty(n * nrhs * 16)
        for trsv in (1, 0):
            ctx.set_tuning("trsv", trsv)
            t64, t32 = [], []
            for _ in range(reps):
                ctx.copy(dB, dB0, n * nrhs * 16)
                ctx.sync()
                t0 = time.perf_counter()
                ctx.lu_solve(dL, n, n, piv, dB, nrhs, n)
                ctx.sync()
                t64.append(time.perf_counter() - t0)
                ctx.copy(dB, dB0, n * nrhs * 16)
                ctx.sync()
                t0 = time.perf_counter()
                it, res = ctx.refine_solve(dA, n, n, dF, n, piv32, dB, nrhs, n, 1e-12, 30)
                ctx.sync()
                t32.append(time.perf_counter() - t0)
            key = f"{nrhs} {'default' if trsv else 'trsm'}"
            out["fp64_solve_ms"][key] = round(min(t64) * 1e3, 3)
            out["fp32_refine_ms"][key] = round(min(t32) * 1e3, 3)
            out["refine_corrections"][key] = it
            log(f"[bench] n={n} nrhs={nrhs} {'default' if trsv else 'trsm'}: FP64 solve "
                f"{out['fp64_solve_ms'][key]} ms, FP32 refinement {out['fp32_refine_ms'][key]} ms ({it} corrections)")
        ctx.set_tuning("trsv", 1)
        dB0.free()
        dB.free()
    for d in (dA, dL, piv, dF, piv32):
        d.free()
    ctx.close()
    return out


def serving_pool(seed=5, n=16, lo=1000, hi=30000):
    """Seeded pool of closed meshes: spheres, triaxial ellipsoids (semi-axes in
    [0.5, 1.5]), tori (R = 1, r in [0.2, 0.4]) and cubes of side 2, with
    geometrically spaced target sizes lo..hi cells dealt to the kinds in a seeded
    order (so every kind gets small and large meshes)."""
    from bielcc_amd import mesh
    rng = np.random.default_rng(seed)
    targets = np.geomspace(lo, hi, n)[rng.permutation(n)]
    pool = {}
    for i in range(n):
        kind = ("sphere", "ellipsoid", "torus", "cube")[i % 4]
        target = int(targets[i])
        if kind == "sphere":
            nt = max(4, int(round(np.sqrt(target / 2))))
            v = mesh.sphere_quads(nt, 2 * nt, jitter=0.005, seed=100 + i)
        elif kind == "ellipsoid":
            nt = max(4, int(round(np.sqrt(target / 2))))
            v = mesh.ellipsoid_quads(nt, 2 * nt, axes=tuple(rng.uniform(0.5, 1.5, 3)), seed=100 + i)
        elif kind == "torus":
            nv = max(4, int(round(np.sqrt(target / 3))))
            v = mesh.torus_quads(3 * nv, nv, R=1.0, r=float(rng.uniform(0.2, 0.4)))
        else:
            v = mesh.cube_quads(max(2, int(round(np.sqrt(target / 6)))), side=2.0)
        pool[f"{kind}{i}"] = v
    return pool


def serving_queries(pool, nq, seed=6, m_lo=1000, m_hi=100000):
    """ka uniform in [1, 15] quantised to 8 levels, 1-16 incidences uniform on
    the sphere with orthogonal polarisations, m_lo..m_hi points uniform on a
    sphere of radius uniform in [3, 5]."""
    from bielcc_amd import serve
    rng = np.random.default_rng(seed)
    names = sorted(pool)
    levels = 1.0 + 14.0 * (np.arange(8) + 0.5) / 8
    qs = []
    for i in range(nq):
        nrhs = int(rng.integers(1, 17))
        d = rng.normal(size=(nrhs, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        p = np.cross(d, rng.normal(size=(nrhs, 3)))
        p /= np.linalg.norm(p, axis=1, keepdims=True)
        M = int(rng.integers(m_lo, m_hi + 1))
        rad = float(rng.uniform(3.0, 5.0))
        z, ph = 2.0 * rng.random(M) - 1.0, 2.0 * np.pi * rng.random(M)
        sz = np.sqrt(1.0 - z * z)
        pts = rad * np.stack([sz * np.cos(ph), sz * np.sin(ph), z], axis=1)
        qs.append(serve.Query(names[int(rng.integers(len(names)))], float(levels[int(rng.integers(8))]), d, p, pts,
                              qid=i))
    return qs


def plain_fp64_run(pool, queries):
    """Queries one at a time in the plain FP64 form: assemble, FP64 LU, solve,
    currents, FP64 field (passes of 4 current sets).  Meshes stay uploaded."""
    import bielcc_amd as B
    ctx = B.Context(0)
    npp = B.num_param()
    ctx.set_tuning("opmat_min_rhs", 1 << 20)
    meshes = {}
    phases = {}
    t0 = time.perf_counter()
    lat = []
    for qi, q in enumerate(queries):
        if qi % 4 == 0:
            log(f"[bench] plain FP64: query {qi}/{len(queries)}")
        ts = time.perf_counter()
        m = meshes.get(q.mesh) or meshes.setdefault(q.mesh, ctx.mesh(pool[q.mesh]))
        n2, nr, M = 2 * m.n, q.nrhs, q.points.shape[0]
        dZ, dpiv, dB = ctx.empty(n2 * n2 * 16), ctx.empty(n2 * 4), ctx.empty(n2 * nr * 16)
        dJ, dP, dE, dH = ctx.empty(m.n * nr * 48), ctx.to_device(q.points), ctx.empty(nr * M * 48), ctx.empty(nr * M * 48)
        t1 = time.perf_counter()
        m.assemble_into(dZ, n2, npp, q.k)
        ctx.sync()
        t2 = time.perf_counter()
        ctx.lu_factor(dZ, n2, n2, dpiv)
        t3 = time.perf_counter()
        kvec = q.k * np.asarray(q.dirs)
        m.rhs_planewave_into(dB, n2, kvec, q.pols)
        ctx.lu_solve(dZ, n2, n2, dpiv, dB, nr, n2)
        m.currents_into(dB, n2, nr, dJ)
        ctx.sync()
        t4 = time.perf_counter()
        m.field_batch_into(npp, q.k, nr, dJ, dP, M, dE, dH, inc=np.concatenate([kvec, q.pols], axis=1))
        dE.to_numpy((nr, M, 3), np.complex128)
        t5 = time.perf_counter()
        for k, v in (("setup", t1 - ts), ("assemble", t2 - t1), ("lu", t3 - t2), ("solve", t4 - t3),
                     ("field_and_copy", t5 - t4)):
            phases[k] = phases.get(k, 0.0) + v
        for d in (dZ, dpiv, dB, dJ, dP, dE, dH):
            d.free()
        lat.append(time.perf_counter() - ts)
    wall = time.perf_counter() - t0
    for m in meshes.values():
        m.close()
    ctx.close()
    solves = sum(q.nrhs for q in queries)
    log(f"[bench] plain FP64, one query at a time: {wall:.1f} s, {solves / wall:.1f} solves/s")
    return {"queries": len(queries), "solves": solves, "wall_s": round(wall, 3),
            "solves_per_s": round(solves / wall, 3), "mean_query_s": round(float(np.mean(lat)), 4),
            "seconds": {k: round(v, 3) for k, v in phases.items()}}


def _clock_ns():
    """Host clocks a profiler's kernel timestamps may be on (which one is found by
    trace_summary)."""
    return {"mono": time.monotonic_ns(), "boot": time.clock_gettime_ns(time.CLOCK_BOOTTIME)}


def trace_summary(trace_csv, windows_path, top=16):
    """Kernel activity of each recorded window of a run, from a rocprofv3 kernel trace
    (--kernel-trace --output-format csv): wall time, GPU busy time (union of the kernel
    intervals), the sum of kernel durations over it (kernels in flight on average), the
    kernels with the largest total time, and the whole-panel LU kernel's average duration
    by its number of workgroups."""
    import csv
    import collections
    import re
    win = json.load(open(windows_path))
    rows = []
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"],
                         int(r["Grid_Size_X"]) // max(1, int(r["Workgroup_Size_X"]))))
    rows.sort()
    lo, hi = rows[0][0], rows[-1][1]
    first = next(iter(win.values()))[0]
    key = min(first, key=lambda k: 0 if lo - 60e9 < first[k] < hi + 60e9 else 1)

    def short(n):
        n = re.sub(r"\(.*", "", n).replace("void ", "").replace("bielcc_impl::", "")
        return (n[:40] if n.startswith("Cijk_") else n)[:60]

    out = [f"{len(rows)} kernels in the trace"]
    for name, (a, b) in win.items():
        t0, t1 = a[key], b[key]
        ks = [r for r in rows if t0 <= r[0] < t1]
        if not ks:
            continue
        busy, cur_s, cur_e, tot = 0, None, None, 0
        for st, en, _, _ in ks:
            tot += en - st
            if cur_e is None or st > cur_e:
                if cur_e is not None:
                    busy += cur_e - cur_s
                cur_s, cur_e = st, en
            else:
                cur_e = max(cur_e, en)
        busy += cur_e - cur_s
        wall = t1 - t0
        out.append(f"\n{name}: wall {wall / 1e9:.3f} s, {len(ks)} kernels, GPU busy {busy / 1e9:.3f} s "
                   f"({busy / wall:.0%}), kernel time {tot / 1e9:.3f} s ({tot / busy:.2f} in flight on average)")
        out.append("| kernel | launches | total (s) | average (us) |")
        out.append("|---|---|---|---|")
        agg = collections.defaultdict(lambda: [0, 0])
        for st, en, n, _ in ks:
            x = agg[short(n)]
            x[0] += 1
            x[1] += en - st
        for n, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
            out.append(f"| `{n}` | {c} | {t / 1e9:.3f} | {t / c / 1e3:.1f} |")
        pb = collections.defaultdict(lambda: [0, 0])
        for st, en, n, g in ks:
            if "k_panel_lds" in n:   # the template argument, not the parameter list (double*)
                x = pb[(short(n).startswith("k_panel_lds<double"), (g + 31) // 32 * 32)]
                x[0] += 1
                x[1] += en - st
        for fp64 in (True, False):
            cells = [f"<={g}: {t / c / 1e3:.0f} us ({c})" for (d, g), (c, t) in sorted(pb.items()) if d == fp64]
            if cells:
                out.append(f"k_panel_lds<{'double' if fp64 else 'float'}> by workgroups: " + ", ".join(cells))
    return "\n".join(out)


def accuracy_probe(pool, q, refine_tol=1e-10):
    """One query solved three ways on one context: FP64 LU (the FP64 form), FP32 LU +
    FGMRES as served (refine_tol, forward-error request 1e-7), and an accurate
    reference (the FP64 LU refined by three classic steps with doubled-precision
    residuals, Dot2 sums).  Returns each solution's residual (doubled sums) and the
    scattered-E errors at the query's points against the reference's (FP64 field
    sums for all three, plus the served mixed field), per point with a floor of
    1e-3 of the largest scattered field, and set-wise."""
    import bielcc_amd as B
    ctx = B.Context(0)
    ctx.set_tuning("refine_ferr_exp", 7)
    npp = B.num_param()
    m = ctx.mesh(pool[q.mesh])
    n2, nr, M = 2 * m.n, q.nrhs, q.points.shape[0]
    kvec, pols = q.k * np.asarray(q.dirs), np.asarray(q.pols)
    inc = np.concatenate([kvec, pols], axis=1)
    dZ, dL = ctx.empty(n2 * n2 * 16), ctx.empty(n2 * n2 * 16)
    dF, dpiv, dp32 = ctx.empty(n2 * n2 * 8), ctx.empty(4 * n2), ctx.empty(4 * n2)
    dB0, dR = ctx.empty(n2 * nr * 16), ctx.empty(n2 * nr * 16)
    m.assemble_into(dZ, n2, npp, q.k)
    m.assemble_into(dL, n2, npp, q.k)
    m.rhs_planewave_into(dB0, n2, kvec, pols)
    b = dB0.to_numpy((nr, n2), np.complex128)
    ctx.lu_factor(dL, n2, n2, dpiv)
    ctx.lu32_factor(dZ, n2, n2, dF, n2, dp32)
    sols = {}
    d64 = ctx.to_device(b)
    ctx.lu_solve(dL, n2, n2, dpiv, d64, nr, n2)
    sols["fp64_lu"] = d64.to_numpy((nr, n2), np.complex128)
    dm = ctx.to_device(b)
    it, res = ctx.refine_solve(dZ, n2, n2, dF, n2, dp32, dm, nr, n2, refine_tol, 60)
    sols["mixed_fgmres"] = dm.to_numpy((nr, n2), np.complex128)
    xt = sols["fp64_lu"].copy()
    for _ in range(3):
        dX = ctx.to_device(xt)
        ctx.residual(dZ, n2, n2, dX, n2, dB0, n2, nr, dR, n2, doubled=True)
        ctx.lu_solve(dL, n2, n2, dpiv, dR, nr, n2)
        xt = xt + dR.to_numpy((nr, n2), np.complex128)
    sols["reference"] = xt
    out = {"mesh": q.mesh, "k": q.k, "nrhs": nr, "points": M, "unknowns": n2, "mixed_refine_corrections": it}
    dP = ctx.to_device(np.asarray(q.points))
    dJ, dE, dH = ctx.empty(m.n * nr * 48), ctx.empty(nr * M * 48), ctx.empty(nr * M * 48)
    fields = {}
    for name, x in sols.items():
        dX = ctx.to_device(x)
        ctx.residual(dZ, n2, n2, dX, n2, dB0, n2, nr, dR, n2, doubled=True)
        r = dR.to_numpy((nr, n2), np.complex128)
        out[f"resid_{name}"] = float((np.linalg.norm(r, axis=1) / np.linalg.norm(b, axis=1)).max())
        out[f"x_err_{name}"] = float(np.linalg.norm(x - xt) / np.linalg.norm(xt))
        m.currents_into(dX, n2, nr, dJ)
        for prec in ((B.PREC_FP64, B.PREC_MIXED) if name == "mixed_fgmres" else (B.PREC_FP64,)):
            m.field_batch_into(npp, q.k, nr, dJ, dP, M, dE, dH, inc=None, precision=prec)
            fields[(name, prec)] = dE.to_numpy((nr, M, 3), np.complex128)
    ref = fields[("reference", B.PREC_FP64)]
    for (name, prec), E in fields.items():
        if name == "reference":
            continue
        pt, st = 0.0, 0.0
        for r in range(nr):
            nes = np.linalg.norm(ref[r], axis=1)
            e = np.linalg.norm(E[r] - ref[r], axis=1) / np.maximum(nes, 1e-3 * nes.max())
            pt, st = max(pt, float(e.max())), max(st, float(np.linalg.norm(E[r] - ref[r]) / np.linalg.norm(ref[r])))
        tag = name + ("" if prec == B.PREC_FP64 else "_mixed_field")
        out[f"E_point_max_rel_{tag}"], out[f"E_set_max_rel_{tag}"] = pt, st
    m.close()
    ctx.close()
    log(f"[bench] accuracy probe {q.mesh} k={q.k}: " + ", ".join(f"{k} {v:.2e}" for k, v in out.items()
                                                                if k.startswith(("E_", "resid_", "x_err"))))
    return out


def serving_leg(nq=32, workers=4, refine_tol=1e-10, windows_path=None, nq_stream=24, plain_every=1):
    """The same query stream served in the FP64 form and in the mixed form
    (FP32 LU + FP64 refinement to 1e-10, FP32 far sums with FP64 recomputation
    of the points the FP32 error model does not clear), HBM budget 200 GB: once
    as a burst (every query queued at once: capacity in solves/s) and once as a
    Poisson stream (the first nq_stream queries) at 60% of the FP64 form's
    capacity (latency).  The plain FP64 form runs every plain_every-th query."""
    from bielcc_amd import serve
    pool = serving_pool()
    cells = [v.shape[0] for v in pool.values()]
    out = {"workload": f"{len(pool)} meshes ({min(cells)}-{max(cells)} cells: spheres, ellipsoids, tori, cubes), "
                       f"{nq} queries, ka in [1,15] on 8 levels, 1-16 incidences, 1,000-100,000 points at "
                       f"r in [3,5]; {workers} workers; HBM budget 200 GB; mixed form refines to a {refine_tol:g} residual "
                       f"and a 1e-7 estimated forward error"}
    res = {}
    rate = None
    windows = {}
    for mode in ("burst", "stream"):
        out[mode] = {}
        for form in ("fp64", "mixed"):
            qs = serving_queries(pool, nq)
            if mode == "stream":
                qs = qs[:nq_stream]
            qps = 1e6 if mode == "burst" else rate
            with serve.Server(pool, form=form, workers=workers, hbm_budget=200e9, refine_tol=refine_tol) as s:
                t0 = _clock_ns()
                run = serve.poisson_run(s, qs, qps=qps, seed=7,
                                        progress=lambda i, n: log(f"[bench] serving {mode} {form}: {i}/{n} "
                                                                  f"queries done"))
                windows[f"{mode} {form}"] = [t0, _clock_ns()]
            st = run["stats"]
            log(f"[bench] serving {mode} {form}: {st['wall_s']:.1f} s, {st['solves_per_s']:.1f} solves/s, "
                f"p50 {st['p50_s']:.3f} s, p99 {st['p99_s']:.3f} s, fallbacks {st['fp64_fallbacks']}")
            out[mode][form] = {k: (round(v, 5) if isinstance(v, float) else v) for k, v in st.items()}
            if mode == "burst":
                res[form] = run["results"]
        if mode == "burst":
            rate = 0.6 * nq / out["burst"]["fp64"]["wall_s"]
            out["stream_qps"] = round(rate, 3)
            out["stream_queries"] = min(nq, nq_stream)
    # error of the scattered field (the total field nearly vanishes at shadow points), split
    # by how the mixed server solved the query: FP32 factors + refinement, or FP64 factors
    # for systems FP32 factors cannot refine (there both servers run FP64 arithmetic and
    # differ by the system's own conditioning times rounding)
    qs = {q.qid: q for q in serving_queries(pool, nq)}
    errs = {"refined": [0.0, 0.0, 0], "fp64_factors": [0.0, 0.0, 0]}
    worst = worst_pt = None
    fp64_factor_queries = []
    for i, r64 in res["fp64"].items():
        r32, q = res["mixed"][i], qs[i]
        cat = errs["fp64_factors" if r32.fp64_fallback else "refined"]
        cat[2] += 1
        for r in range(r64.E.shape[0]):
            einc = q.pols[r][None, :] * np.exp(1j * q.k * (q.points @ q.dirs[r]))[:, None]
            es = r64.E[r] - einc
            nes = np.linalg.norm(es, axis=1)
            # per point, with a floor of 1e-3 of the set's largest scattered field (nulls)
            e = np.linalg.norm(r32.E[r] - r64.E[r], axis=1) / np.maximum(nes, 1e-3 * nes.max())
            g = float(np.linalg.norm(r32.E[r] - r64.E[r]) / np.linalg.norm(es))
            cat[0], cat[1] = max(cat[0], float(e.max())), max(cat[1], g)
            if r32.fp64_fallback and r == 0:
                fp64_factor_queries.append({"query": i, "mesh": q.mesh, "k": q.k, "nrhs": int(r64.E.shape[0]),
                                            "built_here": bool(r32.factored), "set_rel_err_first_set": g})
            if not r32.fp64_fallback and (worst is None or g > worst["set_rel_err"]):
                worst = {"query": i, "mesh": q.mesh, "k": q.k, "refine_resid": r32.refine_resid,
                         "refine_iters": r32.refine_iters, "set_rel_err": g}
            if not r32.fp64_fallback and (worst_pt is None or float(e.max()) > worst_pt[1]):
                worst_pt = (i, float(e.max()))
    out["err_scattered_E_refined"] = {"point_max_rel": errs["refined"][0], "set_max_rel": errs["refined"][1],
                                      "queries": errs["refined"][2]}
    out["err_scattered_E_fp64_factors"] = {"point_max_rel": errs["fp64_factors"][0],
                                           "set_max_rel": errs["fp64_factors"][1],
                                           "queries": errs["fp64_factors"][2]}
    out["worst_refined"] = worst
    out["fp64_factor_queries"] = fp64_factor_queries
    out["capacity_solves_per_s_ratio"] = round(out["burst"]["mixed"]["solves_per_s"] /
                                               out["burst"]["fp64"]["solves_per_s"], 3)
    # the plain FP64 GPU form on the same queries: each query on its own, in arrival
    # order, with its own assembly and FP64 LU (no factor reuse, no overlap, current
    # sets in passes of 4)
    t0 = _clock_ns()
    out["plain_fp64"] = plain_fp64_run(pool, serving_queries(pool, nq)[::plain_every])
    out["plain_fp64"]["sample"] = (f"all {nq} queries" if plain_every == 1 else
                                   f"every {plain_every}th query of the {nq}")
    windows["plain fp64"] = [t0, _clock_ns()]
    if windows_path:
        with open(windows_path, "w") as f:
            json.dump(windows, f)
    if worst_pt is not None:
        # which of the two differs from the exact solution: the refined query with the largest
        # point error against the FP64 form, solved again both ways beside an accurate reference
        out["accuracy_probe"] = accuracy_probe(pool, qs[worst_pt[0]], refine_tol)
        out["accuracy_probe"]["query"] = worst_pt[0]
    out["served_mixed_vs_plain_fp64_solves_per_s"] = round(out["burst"]["mixed"]["solves_per_s"] /
                                                           out["plain_fp64"]["solves_per_s"], 3)
    out["served_fp64_vs_plain_fp64_solves_per_s"] = round(out["burst"]["fp64"]["solves_per_s"] /
                                                          out["plain_fp64"]["solves_per_s"], 3)
    out["stream_p99_ratio"] = round(out["stream"]["fp64"]["p99_s"] / out["stream"]["mixed"]["p99_s"], 3)
    return out


def batched_workload(seed=4, n_mesh=12, n_dir=64, n_pts=2000):
    """12 seeded closed meshes: spheres, triaxial ellipsoids (semi-axes in
    [0.5, 1.5]) and tori (R = 1, r in [0.2, 0.4]) of 500-8,000 quads, every
    second one triangulated; ka uniform in [1, 12]; 64 incidence directions
    uniform on the sphere with orthogonal polarisations; 2,000 field points on a
    sphere of radius 3."""
    from bielcc_amd import mesh
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_mesh):
        kind = ("sphere", "ellipsoid", "torus")[i % 3]
        target = int(rng.uniform(500, 8000))
        if kind == "sphere":
            nt = max(4, int(round(np.sqrt(target / 2))))
            v = mesh.sphere_quads(nt, 2 * nt, jitter=0.005, seed=200 + i)
        elif kind == "ellipsoid":
            nt = max(4, int(round(np.sqrt(target / 2))))
            v = mesh.ellipsoid_quads(nt, 2 * nt, axes=tuple(rng.uniform(0.5, 1.5, 3)), seed=200 + i)
        else:
            nv = max(4, int(round(np.sqrt(target / 3))))
            v = mesh.torus_quads(3 * nv, nv, R=1.0, r=float(rng.uniform(0.2, 0.4)))
        tri = i % 2 == 1
        if tri:
            v = mesh.triangulate(v)
        k = float(rng.uniform(1.0, 12.0))
        d = rng.normal(size=(n_dir, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        pol = np.cross(d, rng.normal(size=(n_dir, 3)))
        pol /= np.linalg.norm(pol, axis=1, keepdims=True)
        pts = mesh.sphere_points(n_pts, 3.0, seed=300 + i)
        out.append(dict(kind=kind, tri=tri, verts=v, k=k, dirs=d, pols=pol, pts=pts))
    return out


def lpt_assign(weights, n):
    """Longest-processing-time assignment of jobs (by weight) to n workers."""
    loads, out = [0.0] * n, [[] for _ in range(n)]
    for i in sorted(range(len(weights)), key=lambda j: -weights[j]):
        t = min(range(n), key=lambda u: loads[u])
        out[t].append(i)
        loads[t] += weights[i]
    return out


def run_workers(ctxs, assign, fn):
    """fn(ctx, job) for every job of every worker, one host thread per worker
    context (the library releases the GIL); re-raises the first failure."""
    import threading
    errs = []

    def go(t):
        try:
            for i in assign[t]:
                fn(ctxs[t], i)
            ctxs[t].sync()
        except BaseException as exc:   # noqa: BLE001 - re-raised below
            errs.append(exc)
    th = [threading.Thread(target=go, args=(t,)) for t in range(len(ctxs))]
    for x in th:
        x.start()
    for x in th:
        x.join()
    if errs:
        raise errs[0]


def batched_leg(tune_cache=None, workers=4, solve=True):
    """Batched form on the 12-mesh workload, FP32 LU + FP64 refinement:
    assembly of all matrices in one persistent launch, FP32 LU and refinement
    (64 right-hand sides) per mesh, the field of all 768 current sets through
    one captured field plan (operator matrices + ZGEMM), with tile shapes from
    the autotuner.  The same batch in the FP64 form gives the reference for the
    errors and the time ratio; the field is also timed through the per-mesh
    driver (passes of 4 current sets) for comparison."""
    import bielcc_amd as B
    from bielcc_amd import tune as T
    work = batched_workload()
    ctx = B.Context(0)
    _warm_lu32(ctx)   # kernels of the FP32 LU path loaded before anything is timed
    npp = B.num_param()
    nd, M = work[0]["dirs"].shape[0], work[0]["pts"].shape[0]
    ms = [ctx.mesh(w["verts"]) for w in work]
    bufs = []
    for w, m in zip(work, ms):
        n, n2 = m.n, 2 * m.n
        inc = np.concatenate([w["k"] * w["dirs"], w["pols"]], axis=1)
        bufs.append(dict(Z=ctx.empty(n2 * n2 * 16), F=ctx.empty(n2 * n2 * 8), piv=ctx.empty(n2 * 4),
                         B=ctx.empty(n2 * nd * 16), J=ctx.empty(n * nd * 48), P=ctx.to_device(w["pts"]),
                         inc=ctx.to_device(inc), E=ctx.empty(nd * M * 48), H=ctx.empty(nd * M * 48)))
    cells = [m.n for m in ms]
    sig = f"batched12|cells{sum(cells)}|nd{nd}|M{M}"
    log(f"[bench] batched: {len(ms)} meshes, {min(cells)}-{max(cells)} cells ({sum(cells)} total), "
        f"{nd} incidences, {M} points each")

    def timed(fn):
        ctx.sync()
        t0 = time.perf_counter()
        fn()
        ctx.sync()
        return time.perf_counter() - t0

    # assembly: all matrices in one persistent launch, tile shape tuned
    aitems = [(m, w["k"], b["Z"], 2 * m.n) for w, m, b in zip(work, ms, bufs)]
    tune_a = T.tune(ctx, lambda: ctx.plan_assemble(aitems, npp), T.ASSEMBLE_SPACE, "plan_assemble",
                    sig + "|asm", repeats=1, cache_path=tune_cache, log=log)
    aplan = ctx.plan_assemble(aitems, npp)
    t_asm = timed(aplan.run)
    ctx.timing_enable(True)
    ctx.timing_reset()
    aplan.run()
    ctx.sync()
    kt_asm = ctx.timing("plan_assemble_tiles")[0]
    ctx.timing_enable(False)
    a_info = {k: aplan.info(k) for k in ("tiles", "grid", "launches", "pairs", "near_pairs")}
    aplan.close()
    if not solve:
        # counter runs of the plan kernels: synthetic currents, no factorisation
        rng = np.random.default_rng(1)
        for m, b in zip(ms, bufs):
            ctx.upload(b["J"], rng.normal(size=(nd, m.n, 3)) + 1j * rng.normal(size=(nd, m.n, 3)))
        fitems = [dict(mesh=m, k=w["k"], nrhs=nd, jE=b["J"], inc=b["inc"], M=M, pts=b["P"], E=b["E"], H=b["H"])
                  for w, m, b in zip(work, ms, bufs)]
        ctx.set_tuning("plan_ws_mb", 96 * 1024)
        fplan = ctx.plan_field(fitems, npp, precision=B.PREC_MIXED)
        t_f = [timed(fplan.run) for _ in range(3)]
        fplan.close()
        fplan64 = ctx.plan_field(fitems, npp, precision=B.PREC_FP64)
        t_f64 = [timed(fplan64.run) for _ in range(2)]
        fplan64.close()
        ctx.set_tuning("plan_ws_mb", 24576)
        for b in bufs:
            for d in b.values():
                d.free()
        for m in ms:
            m.close()
        ctx.close()
        return {"assembly_s": round(t_asm, 4), "field_mixed_s": [round(t, 4) for t in t_f],
                "field_fp64_s": [round(t, 4) for t in t_f64], "assembly_plan": a_info}
    # FP32 LU per mesh and refinement with 64 right-hand sides: the independent
    # systems run concurrently on worker contexts (own stream and BLAS handle each)
    wctx = [B.Context(0) for _ in range(workers)]
    assign = lpt_assign([float(m.n) ** 3 for m in ms], workers)
    wmesh = {}
    for t, idx in enumerate(assign):
        for i in idx:
            wmesh[i] = wctx[t].mesh(work[i]["verts"])
    t_lu = timed(lambda: run_workers(wctx, assign, lambda c, i: c.lu32_factor(
        bufs[i]["Z"], 2 * ms[i].n, 2 * ms[i].n, bufs[i]["F"], 2 * ms[i].n, bufs[i]["piv"])))
    refine, fallbacks = {}, []

    def solve_one(c, i):
        w, m, b = work[i], wmesh[i], bufs[i]
        n2 = 2 * m.n
        m.rhs_planewave_into(b["B"], n2, w["k"] * w["dirs"], w["pols"])
        try:
            refine[i] = c.refine_solve(b["Z"], n2, n2, b["F"], n2, b["piv"], b["B"], nd, n2, 1e-10, 30)
        except B.BielccError as err:
            if err.code != 4:      # refinement stalled: FP64 factors for this mesh (timed)
                raise
            fallbacks.append(i)
            lu = c.empty(n2 * n2 * 16)
            c.copy(lu, b["Z"], n2 * n2 * 16)
            c.lu_factor(lu, n2, n2, b["piv"])
            m.rhs_planewave_into(b["B"], n2, w["k"] * w["dirs"], w["pols"])
            c.lu_solve(lu, n2, n2, b["piv"], b["B"], nd, n2)
            c.sync()
            lu.free()
        m.currents_into(b["B"], n2, nd, b["J"])
    t_solve = timed(lambda: run_workers(wctx, assign, solve_one))
    refine = list(refine.values())
    # field of every current set: one field plan (mixed), tuned, captured, replayed
    fitems = [dict(mesh=m, k=w["k"], nrhs=nd, jE=b["J"], inc=b["inc"], M=M, pts=b["P"], E=b["E"], H=b["H"])
              for w, m, b in zip(work, ms, bufs)]
    ctx.set_tuning("plan_ws_mb", 96 * 1024)   # every mesh's operator matrices in one workspace: one launch
    tune_f = T.tune(ctx, lambda: ctx.plan_field(fitems, npp, precision=B.PREC_MIXED), T.FIELD_SPACE, "plan_field",
                    sig + "|field32", repeats=2, cache_path=tune_cache, log=log)
    fplan = ctx.plan_field(fitems, npp, precision=B.PREC_MIXED)
    fixed0 = ctx.get_tuning("mixed_fixup_points")
    t_field_eager = timed(fplan.run)   # first run: FP32 sums and their error control
    fixed_first = ctx.get_tuning("mixed_fixup_points") - fixed0
    ctx.timing_enable(True)
    ctx.timing_reset()
    fplan.run()              # eager, with HIP events around the tile launch and the GEMMs
    ctx.sync()
    kt = {key: ctx.timing(key)[0] for key in ("plan_field_tiles", "plan_field_gemm")}
    ctx.timing_enable(False)
    fplan.capture()
    reps = 3
    fixed0 = ctx.get_tuning("mixed_fixup_points")
    t_field = timed(lambda: [fplan.run() for _ in range(reps)]) / reps
    fixed = (ctx.get_tuning("mixed_fixup_points") - fixed0) // reps   # per run, of 12 x 2,000 points
    f_info = {k: fplan.info(k) for k in ("tiles", "grid", "launches", "pairs", "near_pairs", "workspace_bytes",
                                         "fp64_sums")}
    fplan.close()
    E32 = [b["E"].to_numpy((nd, M, 3), np.complex128) for b in bufs]
    H32 = [b["H"].to_numpy((nd, M, 3), np.complex128) for b in bufs]
    J32 = [b["J"].to_numpy((nd, m.n, 3), np.complex128) for m, b in zip(ms, bufs)]
    # the same fields through the per-mesh driver: one launch per mesh through operator
    # matrices (its default at 64 sets), and current sets in passes of 4
    def per_mesh():
        for w, m, b in zip(work, ms, bufs):
            m.field_batch_into(npp, w["k"], nd, b["J"], b["P"], M, b["E"], b["H"],
                               inc=np.concatenate([w["k"] * w["dirs"], w["pols"]], axis=1), precision=B.PREC_MIXED)
    t_field_opmat_mesh = timed(per_mesh)
    min_rhs = ctx.get_tuning("opmat_min_rhs")
    ctx.set_tuning("opmat_min_rhs", 1 << 20)
    t_field_driver = timed(per_mesh)
    ctx.set_tuning("opmat_min_rhs", min_rhs)
    # FP64 form of the same batch (same workers): FP64 LU of a copy of each matrix, FP64 solve, FP64 field plan
    for b in bufs:
        b["F"].free()
    tmp = [(c.empty(max((4 * ms[i].n ** 2 * 16 for i in idx), default=16)),
            c.empty(max((2 * ms[i].n * 4 for i in idx), default=4))) for c, idx in zip(wctx, assign)]

    def solve64_one(c, i):
        w, m, b = work[i], wmesh[i], bufs[i]
        n2 = 2 * m.n
        lu, piv = tmp[wctx.index(c)]
        c.copy(lu, b["Z"], n2 * n2 * 16)
        c.lu_factor(lu, n2, n2, piv)
        m.rhs_planewave_into(b["B"], n2, w["k"] * w["dirs"], w["pols"])
        c.lu_solve(lu, n2, n2, piv, b["B"], nd, n2)
        m.currents_into(b["B"], n2, nd, b["J"])
    t_lu64 = timed(lambda: run_workers(wctx, assign, solve64_one))
    t_solve64 = 0.0
    for lu, piv in tmp:
        lu.free()
        piv.free()
    for i, m in wmesh.items():
        m.close()
    panel_aborts = sum(max(c.get_tuning("lu_panel_aborts"), 0) for c in wctx)
    for c in wctx:
        c.close()

    fplan64 = ctx.plan_field(fitems, npp, precision=B.PREC_FP64)
    fplan64.run()
    t_field64 = timed(fplan64.run)
    fplan64.close()
    ctx.set_tuning("plan_ws_mb", 24576)
    e_cur, e_pt, e_set = 0.0, 0.0, 0.0
    per_mesh = []
    for i, (w, m, b) in enumerate(zip(work, ms, bufs)):
        e_pt_i = 0.0
        J64 = b["J"].to_numpy((nd, m.n, 3), np.complex128)
        E64 = b["E"].to_numpy((nd, M, 3), np.complex128)
        H64 = b["H"].to_numpy((nd, M, 3), np.complex128)
        e_cur = max(e_cur, float(np.linalg.norm(J32[i] - J64) / np.linalg.norm(J64)))
        for r in range(nd):
            einc = w["pols"][r][None, :] * np.exp(1j * w["k"] * (w["pts"] @ w["dirs"][r]))[:, None]
            for X32, X64, sub in ((E32[i][r], E64[r], einc), (H32[i][r], H64[r], 0.0)):
                sc = X64 - sub
                nsc = np.linalg.norm(sc, axis=1)
                e_pt_i = max(e_pt_i, float((np.linalg.norm(X32 - X64, axis=1) /
                                            np.maximum(nsc, 1e-3 * nsc.max())).max()))
                e_set = max(e_set, float(np.linalg.norm(X32 - X64) / np.linalg.norm(sc)))
        e_pt = max(e_pt, e_pt_i)
        per_mesh.append(float(f"{e_pt_i:.3e}"))
    for b in bufs:
        for d in b.values():
            if d.ptr:
                d.free()
    for m in ms:
        m.close()
    ctx.close()
    pairs_asm, pairs_fld = a_info["pairs"], f_info["pairs"]
    total32 = t_asm + t_lu + t_solve + t_field
    total64 = t_asm + t_lu64 + t_solve64 + t_field64
    solves = len(work) * nd
    res = {"workload": f"{len(work)} meshes (spheres, ellipsoids, tori; half triangulated), {min(cells)}-{max(cells)} "
                       f"cells, ka in [1,12], {nd} incidences and {M} points (r = 3) per mesh",
           "meshes": [{"kind": w["kind"], "triangles": w["tri"], "cells": m_n, "ka": round(w["k"], 3)}
                      for w, m_n in zip(work, cells)],
           "tuned": {"assemble": tune_a["best"], "field": tune_f["best"], "cached": tune_a["cached"] and tune_f["cached"],
                     "tuning_s": round(tune_a.get("tuning_s", 0) + tune_f.get("tuning_s", 0), 2)},
           "assembly_plan": a_info, "field_plan": f_info,
           "assembly_s": round(t_asm, 4), "lu32_s": round(t_lu, 3), "refine_solve_s": round(t_solve, 3),
           "refine": {"max_corrections": max((r[0] for r in refine), default=0),
                      "max_residual": max((r[1] for r in refine), default=0.0),
                      "fp64_fallback_meshes": fallbacks},
           "field_graph_replay_s": round(t_field, 4), "field_eager_s": round(t_field_eager, 4),
           "field_points_recomputed_fp64": fixed, "field_first_run_points_recomputed_fp64": fixed_first,
           "kernel_ms": {"k_assemble_multi": round(kt_asm, 3), "k_field_opmat": round(kt["plan_field_tiles"], 3),
                         "zgemm_all_meshes": round(kt["plan_field_gemm"], 3)},
           "k_field_opmat_tflops": round(f_info["pairs"] * FLOPS_FIELD_PAIR / kt["plan_field_tiles"] / 1e9, 2),
           "k_assemble_multi_pairs_per_s": a_info["pairs"] / (kt_asm / 1e3),
           "field_per_mesh_operator_matrix_s": round(t_field_opmat_mesh, 4),
           "field_per_mesh_passes_of_4_s": round(t_field_driver, 4),
           "field_plan_speedup_vs_passes": round(t_field_driver / t_field, 2),
           "assembly_pairs_per_s": pairs_asm / t_asm, "field_pair_sets_per_s": pairs_fld * nd / t_field,
           "total_s": round(total32, 3), "solves": solves, "solves_per_s": round(solves / total32, 2),
           "workers": workers, "lu_panel_aborts": panel_aborts,
           "fp64_form": {"lu64_and_solve64_s": round(t_lu64, 3), "field64_s": round(t_field64, 4),
                         "total_s": round(total64, 3), "solves_per_s": round(solves / total64, 2)},
           "solves_per_s_ratio_vs_fp64": round(total64 / total32, 3),
           "err_currents_rel_max": e_cur, "err_field_point_max_rel": e_pt, "err_field_set_max_rel": e_set,
           "err_field_point_max_rel_per_mesh": per_mesh}
    log(f"[bench] batched: assembly {t_asm:.3f} s, LU32 {t_lu:.2f} s, refine+solve {t_solve:.2f} s, field "
        f"{t_field * 1e3:.1f} ms (graph; eager {t_field_eager * 1e3:.1f} ms; per mesh {t_field_opmat_mesh * 1e3:.1f} "
        f"ms, passes of 4 {t_field_driver * 1e3:.1f} ms); FP64 form {total64:.2f} s vs {total32:.2f} s; "
        f"field err {e_set:.2e} set-wise")
    return res


def copy_calibration(nbytes=1 << 30):
    """A plain device copy of known size (torch), so that rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE can be scaled against it in the same run."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: {nbytes} bytes copied")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="sphere20k")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-lu", action="store_true", help="skip the LU (currents from the right-hand side)")
    ap.add_argument("--copy-calib", action="store_true", help="run a 1 GiB device copy first (counter scale)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision form")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-queries", type=int, default=32)
    ap.add_argument("--serve-stream", type=int, default=None, metavar="Q",
                    help="queries of the Poisson stream legs (default 12 in the full run, 24 with --serving-only)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 12-mesh batched form")
    ap.add_argument("--no-tri", action="store_true", help="skip the 39,600-triangle variant")
    ap.add_argument("--tri-only", type=int, default=0, metavar="STEPS",
                    help="run only the 39,600-triangle variant for STEPS steps")
    ap.add_argument("--tri-lu", action="store_true", help="with --tri-only: also time its FP64 and FP32 LU")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched form")
    ap.add_argument("--cube-only", action="store_true", help="run only the precision form on the cube")
    ap.add_argument("--batched-workers", type=int, default=4, help="worker streams of the batched form's solves")
    ap.add_argument("--batched-no-solve", action="store_true",
                    help="batched form: plan kernels only, on synthetic currents (counter runs)")
    ap.add_argument("--tune-cache", default=None, help="autotuner cache file (default: none, tune afresh)")
    ap.add_argument("--serving-only", action="store_true", help="run only the serving leg")
    ap.add_argument("--c1-only", action="store_true",
                    help="run only the 512-cell case in full on the host cores and on the GPU")
    ap.add_argument("--lu-only", type=int, default=0, metavar="N",
                    help="time only the FP64 and FP32 LU of a seeded random complex N x N matrix")
    ap.add_argument("--solve-only", type=int, default=0, metavar="N",
                    help="time only the FP64 solves and FP32 refinements of a seeded random N x N system")
    ap.add_argument("--serve-workers", type=int, default=4, help="worker streams of the served forms")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="library tunable for every context (repeatable), e.g. lu_look_ahead=0")
    ap.add_argument("--trace-windows", default=None, metavar="PATH",
                    help="serving leg: write the host-clock window of each run (burst, stream, plain) to PATH")
    ap.add_argument("--trace-summary", default=None, metavar="CSV",
                    help="print the kernel activity per window of a rocprofv3 kernel trace (needs --trace-windows)")
    ap.add_argument("--lu-panel", type=int, default=1, choices=(0, 1),
                    help="LU panels as one whole-panel launch (1) or per-column launches (0), every context")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the hot path is run on one MI355X; multi-GPU sharding is a design "
              "note only (row blocks of the matrix, observation points of the field).", file=sys.stderr)
        sys.exit(2)

    if args.trace_summary:
        print(trace_summary(args.trace_summary, args.trace_windows))
        return
    import bielcc_amd as B
    B.Context.default_tuning["lu_panel_lds"] = args.lu_panel
    for kv in args.tune:
        key, _, val = kv.partition("=")
        B.Context.default_tuning[key] = int(val)
    if args.copy_calib:
        copy_calibration()
    if args.batched_only:
        print(json.dumps({"batched": batched_leg(args.tune_cache, args.batched_workers,
                                                 solve=not args.batched_no_solve)}), flush=True)
        return
    if args.tri_only:
        print(json.dumps({"triangle_variant": triangle_variant(args.tri_only, lu=args.tri_lu)}), flush=True)
        return
    if args.cube_only:
        print(json.dumps({"precision_form_cube": cube_precision()}), flush=True)
        return
    if args.c1_only:
        print(json.dumps({"c1_full": c1_full()}), flush=True)
        return
    if args.lu_only:
        print(json.dumps({"lu": lu_only(args.lu_only)}), flush=True)
        return
    if args.solve_only:
        print(json.dumps({"solve": solve_only(args.solve_only)}), flush=True)
        return
    if args.serving_only:
        print(json.dumps({"serving": serving_leg(nq=args.serve_queries, workers=args.serve_workers,
                                                 windows_path=args.trace_windows,
                                                 nq_stream=args.serve_stream or 24)}), flush=True)
        return
    w = workload(args.config)
    verts, k, pts = w["verts"], w["k"], w["pts"]
    n, M = verts.shape[0], pts.shape[0]
    n2 = 2 * n
    log(f"[bench] {w['desc']}: N={n} cells, {n2} unknowns, M={M} points")
    ctx = B.Context(0)
    m = ctx.mesh(verts)
    npp = B.num_param()
    dZ = ctx.empty(n2 * n2 * 16)
    dB = ctx.empty(n2 * 16)
    dpiv = ctx.empty(n2 * 4)
    dJ = ctx.empty(n * 48)
    dP = ctx.to_device(pts)
    dE = ctx.empty(M * 48)
    dH = ctx.empty(M * 48)
    kvec, e0 = (0.0, 0.0, k), (1.0, 0.0, 0.0)

    lu_s = solve_s = None
    # warm-up: assemble, LU, solve, currents, field
    for wi in range(max(1, args.warmup)):
        t0 = time.perf_counter()
        m.assemble_into(dZ, n2, npp, k)
        ctx.sync()
        log(f"[bench] warmup {wi}: assembly {time.perf_counter() - t0:.3f} s")
        if wi == 0:
            m.rhs_planewave_into(dB, n2, kvec, e0)
            if not args.no_lu:
                t0 = time.perf_counter()
                ctx.lu_factor(dZ, n2, n2, dpiv)
                ctx.sync()
                lu_s = time.perf_counter() - t0
                t0 = time.perf_counter()
                ctx.lu_solve(dZ, n2, n2, dpiv, dB, 1, n2)
                ctx.sync()
                solve_s = time.perf_counter() - t0
                log(f"[bench] LU factor {lu_s:.2f} s, solve {solve_s:.3f} s")
            m.currents_into(dB, n2, 1, dJ)
        t0 = time.perf_counter()
        m.field_into(npp, k, dJ, dP, M, dE, dH)
        ctx.sync()
        log(f"[bench] warmup {wi}: field {time.perf_counter() - t0:.3f} s")

    ctx.timing_enable(True)
    ctx.timing_reset()
    ctx.sync()
    t_start = time.perf_counter()
    for s in range(args.steps):
        m.assemble_into(dZ, n2, npp, k)
        m.field_into(npp, k, dJ, dP, M, dE, dH)
        ctx.sync()
        log(f"[bench] step {s} done ({time.perf_counter() - t_start:.2f} s)")
    ctx.sync()
    elapsed = time.perf_counter() - t_start
    fld_ms, fld_cnt = ctx.timing("field_far")
    asm_ms, asm_cnt = ctx.timing("assemble_far")
    near_ms, _ = ctx.timing("near")
    ctx.timing_enable(False)

    pairs_step = n * n + M * n
    value = pairs_step * args.steps / elapsed
    E = dE.to_numpy((M, 3), np.complex128)
    assert np.all(np.isfinite(E)), "non-finite field"
    # roofline of the dominant kernel (field tile kernel; every pair of this workload is far)
    fld_avg_s = fld_ms / max(fld_cnt, 1) / 1e3
    achieved = M * n * EXEC_FLOP_FIELD_PAIR / fld_avg_s / 1e12
    nominal = M * n * FLOPS_FIELD_PAIR / fld_avg_s / 1e12
    asm_avg_s = asm_ms / max(asm_cnt, 1) / 1e3
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
            "flop_per_pair": EXEC_FLOP_FIELD_PAIR, "flop_count": "executed FP64 flop (rocprofv3 counters)",
            "valu_issue_frac": round(M * n * VALU_LANE_INSTR_FIELD_PAIR / 64 / fld_avg_s / VALU_ISSUE_PER_S, 4),
            "nominal_achieved": round(nominal, 3), "nominal_frac": round(nominal / PEAK_FP64_TFLOPS, 4),
            "traffic": FIELD_TRAFFIC_BYTES.get(args.config),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "kernel": "k_field_far (fused K_RotRot_Far + R_Rot_Colloc, FP64)",
            "launch_ms": round(fld_avg_s * 1e3, 3),
            "assembly_kernel_nominal_tflops": round(n * n * FLOPS_ASM_PAIR / asm_avg_s / 1e12, 3) if asm_cnt else None,
            "assembly_launch_ms": round(asm_avg_s * 1e3, 3) if asm_cnt else None}
    out = {"metric": "cell-pair interactions/s (K,R assembly+field)", "value": value,
           "unit": "cell-pair interactions/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded jittered sphere mesh)",
           "config": {"workload": w["desc"], "cells": n, "unknowns": n2, "points": M, "ka": k,
                      "pairs_per_step": pairs_step, "assembly_pairs": n * n, "field_pairs": M * n},
           "roofline": roof}
    if lu_s is not None:
        out["solve"] = {"lu_factor_s": round(lu_s, 3), "lu_solve_s": round(solve_s, 4),
                        "lu_tflops": round(8.0 / 3.0 * n2 ** 3 / lu_s / 1e12, 2),
                        "solves_per_s_after_factor": round(1.0 / solve_s, 2)}
    out["timing_ms"] = {"field_far": fld_ms, "assemble_far": asm_ms, "near": near_ms}
    if lu_s is not None and not args.no_mixed:
        out["precision_form"] = mixed_form(ctx, m, npp, k, kvec, e0, dP, M, dJ, dE, dH, lu_s, solve_s,
                                           elapsed / args.steps)
    for d in (dZ, dB, dpiv):
        d.free()
    if not args.no_tri:
        out["triangle_variant"] = triangle_variant()
    if lu_s is not None and not args.no_mixed:
        out["precision_form_cube"] = cube_precision()
    if not args.no_lu:   # a mid-size factorisation: partial pivoting against the butterfly form
        out["lu_order_6000"] = lu_only(6000)
    if not args.no_batched:
        out["batched"] = batched_leg(args.tune_cache, args.batched_workers)
    if not args.no_serve:
        out["serving"] = serving_leg(nq=args.serve_queries, nq_stream=args.serve_stream or 12)
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(verts, pts, k, args.cpu_budget)
        log(f"[bench] cpu baseline {out['cpu_baseline']['value']:.3e} pairs/s")
        out["cpu_baseline"]["c1_full"] = c1_full()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
