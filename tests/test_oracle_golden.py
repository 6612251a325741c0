"""The C restatement (oracle/liboracle.so) against the reference's own outputs
recorded in tests/golden/ (tests/golden/make_golden.py).  CPU only."""
import numpy as np
import pytest

import oracle as O
from _util import golden, np_dict, rel_err, rowwise_rel, block_metric


def test_ops_match_reference(orc):
    g = golden("ops.npz")
    sets = g["np_sets"]
    worst = 0.0
    for nv in (4, 3):
        for i in range(len(g[f"op{nv}"])):
            npp = O.num_param(**np_dict(sets[g[f"np{nv}"][i]]))
            r = orc.op(int(g[f"op{nv}"][i]), g[f"cells{nv}"][i], g[f"x{nv}"][i], g[f"j{nv}"][i], npp, g[f"k{nv}"][i])
            worst = max(worst, rel_err(r, g[f"res{nv}"][i], floor=1e-300))
    assert worst <= 1e-13, worst


def test_kernels_match_reference(orc):
    g = golden("kernels.npz")
    for i in range(len(g["f"])):
        R, Rmin = g["kp"][i]
        kp = O.kernel_param(R, Rmin, g["k"][i], vec_dbl=(0.3, -0.2, 0.5), n=(0.0, 0.6, 0.8),
                            current=(1 + 1j, 0.5, -0.2j))
        r = orc.kernel(int(g["f"][i]), g["x"][i], g["y"][i], kp)
        assert rel_err(r, g["res"][i], floor=1e-300) <= 1e-14, i


def test_integrators_match_reference(orc):
    g = golden("integrators.npz")
    for i in range(len(g["f"])):
        R, Rmin = g["kp"][i]
        kp = O.kernel_param(R, Rmin, g["k"][i], vec_dbl=(0.3, -0.2, 0.5), n=(0.0, 0.6, 0.8),
                            current=(1 + 1j, 0.5, -0.2j))
        idim, n0, pm, eps = g["ip"][i]
        ip = O.IntegralParam(int(idim), int(n0), int(pm), float(eps))
        kind, f, nv = int(g["kind"][i]), int(g["f"][i]), int(g["nv"][i])
        cell = g["cell"][i][:nv] if nv > 2 else None
        if kind == 0:
            r = orc.integral_pnt(f, cell, g["x"][i], kp, ip)
        elif kind == 1:
            r = orc.integral_seg(f, g["cell"][i][0], g["cell"][i][1], g["x"][i], kp, ip)
        else:
            r = orc.k0(kind - 2, f, cell, g["x"][i], g["j"][i], kp, ip)
        assert rel_err(r, g["res"][i], floor=1e-300) <= 1e-13, (i, kind, f)
        assert kp.smoothR == g["R_out"][i], (i, "smoothR side effect differs")


def test_analytic_and_cells_match_reference(orc):
    g = golden("kernels.npz")
    for i in range(len(g["an_which"])):
        w, nv = int(g["an_which"][i]), int(g["an_nv"][i])
        cell, x, k = g["an_cell"][i], g["an_x"][i], g["an_k"][i]
        if w == 0:
            r = np.array([orc.integral1divr(cell[:nv], x), 0, 0])
        elif w == 1:
            r = orc.integral_grad1divr_seg(cell[0], cell[1], x)
        elif w == 2:
            r = np.array([orc.surf_hs(cell[:nv], x, 3.0, O.IntegralParam(1, 8, 1, 1e-6), k), 0, 0])
        elif w == 3:
            r = orc.seg_hs(cell[0], cell[1], x, 3.0, O.IntegralParam(3, 10, 2, 1e-6), k)
        else:
            r = orc.k0_hs(cell[:nv], x, np.array([1 + 0.5j, -0.3, 0.2j]), O.IntegralParam(3, 10, 2, 1e-6), 3.0, k)
        assert rel_err(r, g["an_res"][i], floor=1e-300) <= 1e-13, (i, w)
    for i in range(len(g["cf_nv"])):
        nv = int(g["cf_nv"][i])
        d = orc.cellfill(g["cf_cell"][i][:nv])
        tab = np.concatenate([d["rkt"], d["norm"], d["tau1"], d["tau2"], [d["area"], d["diam"]], d["norm_func"],
                              d["center"]])
        assert np.array_equal(tab, g["cf_tab"][i])


@pytest.mark.parametrize("fname", ["config1.npz", "config1_tri.npz"])
def test_config1_matrix_rows(orc, fname):
    """Sampled blocks, all diagonal blocks and row/column checksums of the
    collocation matrix, plus the R-matrix samples."""
    g = golden(fname)
    verts, k = g["verts"], float(g["k"])
    npp = O.num_param(**np_dict(g["num_param"]))
    Z = orc.assemble(verts, npp, k, which=0)
    n = verts.shape[0]
    si, sj = g["sample_i"], g["sample_j"]
    blocks = np.stack([Z[2 * si + a, 2 * sj + b] for a in range(2) for b in range(2)], axis=1)
    assert rel_err(blocks, g["blocks"]) <= 1e-13
    diag = np.stack([Z[2 * np.arange(n) + a, 2 * np.arange(n) + b] for a in range(2) for b in range(2)], axis=1)
    assert rel_err(diag, g["diag"]) <= 1e-13
    w = g["row_w"]
    assert np.max(rowwise_rel((Z @ w)[:, None], g["row_sums"][:, None])) <= 1e-12
    assert np.max(rowwise_rel((w @ Z)[:, None], g["col_sums"][:, None])) <= 1e-12
    ZR = orc.assemble(verts, npp, k, which=1)
    rblocks = np.stack([ZR[2 * si + a, 2 * sj + b] for a in range(2) for b in range(2)], axis=1)
    assert rel_err(rblocks, g["rblocks"]) <= 1e-13
    assert np.linalg.norm(Z @ w - g["row_sums"]) <= 1e-13 * np.linalg.norm(g["row_sums"])


@pytest.mark.parametrize("fname", ["config1.npz", "config1_tri.npz"])
def test_config1_rhs_currents_field(orc, fname):
    g = golden(fname)
    verts, k = g["verts"], float(g["k"])
    npp = O.num_param(**np_dict(g["num_param"]))
    rhs = orc.rhs_planewave(verts, g["kvec"], g["e0"])
    assert rel_err(rhs, g["rhs"]) <= 1e-15
    jE = orc.currents_by_basis(verts, g["b"])
    assert rel_err(jE, g["jE"]) <= 1e-15
    E = orc.get_field_e(verts, g["jE"], 1.0, g["kvec"], g["e0"], k, npp, g["pts"])
    assert np.max(rowwise_rel(E, g["E"])) <= 1e-13
    Ef, Hf = orc.field_eh(verts, g["jE"], k, npp, g["pts"], incident=(g["kvec"], g["e0"]))
    assert np.max(rowwise_rel(Ef, g["E"])) <= 1e-13
    assert np.max(rowwise_rel(Hf, g["H"])) <= 1e-12


def test_config1_lu_against_lapack(orc):
    """The oracle's own LU against the scipy/LAPACK currents recorded with the
    golden matrix; the bound follows cond(Z) (parity of the solve itself is
    pinned by LAPACK, not by the reference, which has no solver)."""
    g = golden("config1.npz")
    verts, k = g["verts"], float(g["k"])
    npp = O.num_param(**np_dict(g["num_param"]))
    Z = orc.assemble(verts, npp, k, which=0)
    b = orc.lu_solve(Z, g["rhs"])[:, 0]
    cond = float(g["cond"])
    assert np.linalg.norm(b - g["b"]) / np.linalg.norm(g["b"]) <= 1e-15 * cond * 10
    assert np.linalg.norm(Z @ b - g["rhs"]) / np.linalg.norm(g["rhs"]) <= 1e-13
