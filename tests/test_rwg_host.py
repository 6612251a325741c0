"""RWG topology (host code of the library) against RwgSurfMesh compiled from
the reference (mesh/rwg_surf_mesh.h:300-526) on its two example meshes and
three written ones: a sphere with degenerate pole quads (pseudo-edges), an
open cube face (boundary sides) and a triangulated torus.  Every array must
match entry for entry, and the VTK export byte for byte.  Host-only: no GPU."""
import gzip
import os

import numpy as np
import pytest

from bielcc_amd.rwg import RwgMesh
from _util import GOLDEN, golden

CASES = (("sphere_quad", 4), ("sphere_triangle", 3), ("pole_sphere", 4), ("open_face", 4), ("torus_tri", 3))
ARRAYS = ("points", "cell_points", "point_num_cells", "point_num_edges", "edges", "edge_num_cells", "cell_edges",
          "cell_dir", "edge_mid", "edge_ort", "nghb_off", "nghb")


@pytest.mark.parametrize("name,nv", CASES)
def test_topology_matches_reference(name, nv):
    g = golden("rwg.npz")
    r = RwgMesh(g[f"{name}_verts"])
    got = dict(points=r.points, cell_points=r.cell_points, point_num_cells=r.point_num_cells,
               point_num_edges=r.point_num_edges, edges=r.edges, edge_num_cells=r.edge_num_cells,
               cell_edges=r.cell_edges, cell_dir=r.cell_edge_dir, edge_mid=r.edge_mid, edge_ort=r.edge_ort,
               nghb_off=r.neighbour_offsets, nghb=r.neighbours)
    for key in ARRAYS:
        ref = g[f"{name}_{key}"]
        assert got[key].shape == ref.shape, key
        assert np.array_equal(got[key], ref), key
    # Euler characteristic of the generated closed surfaces (sphere 2, torus 0); a
    # boundary on the open face.  (The reference's triangle example is not one
    # closed surface: V - E + F = 297.)
    if name in ("sphere_quad", "pole_sphere", "torus_tri"):
        assert r.n_points - r.n_edges + r.n == (0 if name == "torus_tri" else 2)
    if name == "open_face":
        assert (r.edge_num_cells == 1).sum() == 16


@pytest.mark.parametrize("name,nv", CASES)
def test_vtk_export_byte_identical(tmp_path, name, nv):
    g = golden("rwg.npz")
    r = RwgMesh(g[f"{name}_verts"])
    out = tmp_path / "rwg.vtk"
    r.write_vtk(out, g[f"{name}_norm"])
    with gzip.open(os.path.join(GOLDEN, "writers", f"rwg_{name}.vtk.gz"), "rb") as f:
        assert out.read_bytes() == f.read()


def test_neighbours_and_pseudo_edges():
    g = golden("rwg.npz")
    r = RwgMesh(g["pole_sphere_verts"])
    # every pole quad has one pseudo-edge: edge -5, direction 0, no neighbour
    pseudo = r.cell_edges < 0
    assert pseudo.sum() == 20 and np.all(r.cell_edge_dir[pseudo] == 0)
    for c, v in zip(*np.nonzero(pseudo)):
        assert r.neighbours_of(c, v) == [-5]
    # on a closed surface every real side has exactly one neighbour, which sees it back
    for c in range(r.n):
        for v in range(r.nv):
            if r.cell_edges[c, v] >= 0:
                (o,) = r.neighbours_of(c, v)
                assert c in [r.neighbours_of(o, w)[0] for w in range(r.nv)]


def test_bad_input_rejected():
    with pytest.raises(ValueError):
        RwgMesh(np.zeros((2, 5, 3)))
    v = np.zeros((1, 3, 3))
    v[0, 0, 0] = np.nan
    with pytest.raises(ValueError):
        RwgMesh(v)
