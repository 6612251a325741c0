"""Host-side checks that need no GPU: the C-ABI library loads and exports
every declared symbol, it reports a missing device as an error, the seeded
generators, and the mesh readers against the reference's own reader."""
import ctypes
import os
import re

import numpy as np
import pytest

import bielcc_amd as B
from bielcc_amd import mesh, _lib
from _util import golden, GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"^\w[\w\s\*]*?\b(bielcc_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 30
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.declared_symbols()) <= set(names)


def test_no_device_is_an_error_not_a_crash():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    h = ctypes.c_void_p()
    assert lib.bielcc_ctx_create(0, ctypes.byref(h)) == 3
    with pytest.raises(B.BielccError):
        B.Context(0)


def test_mt19937_64_known_answers():
    r = mesh.MT19937_64()
    assert r.next() == 14514284786278117030
    for _ in range(9998):
        r.next()
    assert r.next() == 9981545732273789042     # the 10000th output, C++ [rand.predef]
    u = mesh.uniforms(mesh.MT19937_64(7), 1000)
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.05


def test_sphere_mesh_orientation_and_poles(orc):
    v = mesh.sphere_quads(16, 32, jitter=0.005, seed=1)
    assert v.shape == (512, 4, 3)
    for c in range(0, 512, 7):
        d = orc.cellfill(v[c])
        assert np.dot(d["norm"], d["rkt"]) > 0          # outward CellFill normal
    assert np.array_equal(v[0, 0], v[0, 1])              # north pole quads repeat vertex 0 = 1
    assert np.array_equal(v[-1, 2], v[-1, 3])            # south pole quads repeat 2 = 3
    r = np.linalg.norm(v.reshape(-1, 3), axis=1)
    assert r.min() >= 0.995 - 1e-12 and r.max() <= 1.005 + 1e-12
    t = mesh.triangulate(v)
    assert t.shape == (960, 3, 3)                        # 64 degenerate pole triangles dropped
    big = mesh.sphere_quads(100, 200, jitter=0.005, seed=1)
    assert big.shape == (20000, 4, 3) and mesh.triangulate(big).shape == (39600, 3, 3)


def test_other_generators(orc):
    cube = mesh.cube_quads(6, 2.0)
    assert cube.shape == (216, 4, 3)
    for c in range(0, 216, 5):
        d = orc.cellfill(cube[c])
        assert np.dot(d["norm"], d["rkt"]) > 0
    tor = mesh.torus_quads(12, 8, 1.0, 0.3)
    for c in range(0, 96, 5):
        d = orc.cellfill(tor[c])
        ring = d["rkt"].copy(); ring[2] = 0; ring = ring / np.linalg.norm(ring)
        assert np.dot(d["norm"], d["rkt"] - ring) > 0
    p = mesh.sphere_points(100, 3.0, seed=2)
    assert np.allclose(np.linalg.norm(p, axis=1), 3.0)


@pytest.mark.parametrize("fname,nv", [("sphere_quad.dat", 4), ("sphere_triangle.dat", 3), ("sphere_quad.vtk", 4)])
def test_mesh_readers_match_reference_reader(fname, nv):
    """The reference's own example meshes read by our readers equal what the
    reference's SurfMesh reader produced (recorded in mesh_files.npz)."""
    g = golden("mesh_files.npz")
    key = fname.replace(".", "_")
    v = mesh.read_mesh(os.path.join(GOLDEN, fname), nv)
    assert np.array_equal(v, g[key + "_verts"])


def test_dat_round_trip(tmp_path):
    v = mesh.sphere_quads(6, 12, seed=3)
    p = str(tmp_path / "m.dat")
    mesh.write_dat(p, v)
    assert np.array_equal(mesh.read_dat(p, 4), v)
