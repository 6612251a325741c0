"""Seeded synthetic surface meshes and observation point sets.

Meshes are plain vertex arrays ``verts[n_cells, nv, 3]`` (nv = 4 quads or
3 triangles) in the reference's cell convention: CellQuad::CellFill takes the
normal as (mid23 - mid01) x (mid30 - mid12) (mesh/cell.cpp:21-34), so the
vertex order below is chosen to make that normal point outward, and pole
quads repeat their pole vertex as vertices 0 and 1 (north) or 2 and 3
(south), as the reference's own sphere_quad.dat does.

Random variates come from our own MT19937-64 (the standard 64-bit Mersenne
Twister, Matsumoto & Nishimura 2004; the same engine as C++'s
std::mt19937_64) mapped to [0, 1) as u = (x >> 11) * 2**-53, so the inputs are
reproducible in any language without implementation-defined distributions.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = [
    "MT19937_64", "uniforms", "sphere_quads", "triangulate", "ellipsoid_quads",
    "torus_quads", "cube_quads", "sphere_points", "grid_points", "write_dat", "read_dat", "read_vtk", "read_mesh",
]

_MASK64 = (1 << 64) - 1


class MT19937_64:
    """The standard 64-bit Mersenne Twister (parameters of C++ [rand.predef])."""

    _N, _M = 312, 156
    _MATRIX_A = 0xB5026F5AA96619E9
    _UPPER, _LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed: int = 5489):
        mt = [0] * self._N
        mt[0] = seed & _MASK64
        for i in range(1, self._N):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & _MASK64
        self._mt = mt
        self._idx = self._N

    def _twist(self) -> None:
        mt, n, m = self._mt, self._N, self._M
        for i in range(n):
            y = (mt[i] & self._UPPER) | (mt[(i + 1) % n] & self._LOWER)
            v = mt[(i + m) % n] ^ (y >> 1)
            if y & 1:
                v ^= self._MATRIX_A
            mt[i] = v
        self._idx = 0

    def next(self) -> int:
        if self._idx >= self._N:
            self._twist()
        x = self._mt[self._idx]
        self._idx += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & _MASK64


def uniforms(rng: MT19937_64, n: int) -> np.ndarray:
    """n variates u = (x >> 11) * 2^-53 in [0, 1)."""
    return np.array([(rng.next() >> 11) * (2.0 ** -53) for _ in range(n)], dtype=np.float64)


def _sphere_nodes(n_theta: int, n_phi: int, radius: float, jitter: float, seed: int):
    rng = MT19937_64(seed)
    th = np.pi * np.arange(n_theta + 1) / n_theta
    ph = 2.0 * np.pi * np.arange(n_phi) / n_phi
    nodes = np.empty((n_theta + 1, n_phi, 3))
    for i in range(n_theta + 1):
        for j in range(n_phi):
            s = radius
            if 0 < i < n_theta and jitter:
                s = radius * (1.0 + jitter * (2.0 * ((rng.next() >> 11) * 2.0 ** -53) - 1.0))
            if i == 0:
                nodes[i, j] = (0.0, 0.0, radius)
            elif i == n_theta:
                nodes[i, j] = (0.0, 0.0, -radius)
            else:
                nodes[i, j] = (s * math.sin(th[i]) * math.cos(ph[j]),
                               s * math.sin(th[i]) * math.sin(ph[j]),
                               s * math.cos(th[i]))
    return nodes


def sphere_quads(n_theta: int, n_phi: int, radius: float = 1.0, jitter: float = 0.005,
                 seed: int = 1) -> np.ndarray:
    """Theta-phi quad mesh of a sphere, interior nodes radially jittered by
    (1 + jitter*(2u-1)); n_theta*n_phi cells, 2*n_phi of them pole quads."""
    nodes = _sphere_nodes(n_theta, n_phi, radius, jitter, seed)
    v = np.empty((n_theta, n_phi, 4, 3))
    for i in range(n_theta):
        for j in range(n_phi):
            j1 = (j + 1) % n_phi
            v[i, j, 0] = nodes[i, j1]
            v[i, j, 1] = nodes[i, j]
            v[i, j, 2] = nodes[i + 1, j]
            v[i, j, 3] = nodes[i + 1, j1]
    return v.reshape(-1, 4, 3)


def triangulate(quads: np.ndarray, drop_degenerate: bool = True) -> np.ndarray:
    """Split each quad (0,1,2,3) into (0,1,2) and (2,3,0) - the split the
    reference's analytic 1/r integral uses (integrators/integrals_analytic.cpp:218-231).
    Zero-area triangles (one per pole quad) are dropped: CellTriangle::CellFill
    divides by the normal's length (mesh/cell.cpp:79-84)."""
    t = np.concatenate([quads[:, [0, 1, 2]], quads[:, [2, 3, 0]]], axis=1).reshape(-1, 3, 3)
    if drop_degenerate:
        n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
        keep = np.linalg.norm(n, axis=1) > 0.0
        t = t[keep]
    return np.ascontiguousarray(t)


def ellipsoid_quads(n_theta: int, n_phi: int, axes=(1.0, 1.0, 1.0), seed: int = 1,
                    jitter: float = 0.0) -> np.ndarray:
    """Triaxial ellipsoid: the unit-sphere theta-phi mesh scaled per axis."""
    v = sphere_quads(n_theta, n_phi, 1.0, jitter, seed)
    return v * np.asarray(axes, dtype=np.float64)


def torus_quads(n_u: int, n_v: int, R: float = 1.0, r: float = 0.3) -> np.ndarray:
    """Torus (major radius R, minor r) quad mesh with outward normals."""
    u = 2.0 * np.pi * np.arange(n_u) / n_u
    w = 2.0 * np.pi * np.arange(n_v) / n_v
    def P(a, b):
        return np.array([(R + r * math.cos(b)) * math.cos(a),
                         (R + r * math.cos(b)) * math.sin(a),
                         r * math.sin(b)])
    v = np.empty((n_u, n_v, 4, 3))
    for i in range(n_u):
        for j in range(n_v):
            i1, j1 = (i + 1) % n_u, (j + 1) % n_v
            v[i, j, 0] = P(u[i], w[j])
            v[i, j, 1] = P(u[i1], w[j])
            v[i, j, 2] = P(u[i1], w[j1])
            v[i, j, 3] = P(u[i], w[j1])
    return v.reshape(-1, 4, 3)


def cube_quads(n: int, side: float = 2.0) -> np.ndarray:
    """Axis-aligned cube centred at the origin, n x n quads per face, outward normals."""
    h = side / 2.0
    g = np.linspace(-h, h, n + 1)
    faces = []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            a1, a2 = [(axis + 1) % 3, (axis + 2) % 3]
            q = np.empty((n, n, 4, 3))
            for i in range(n):
                for j in range(n):
                    corners = [(g[i], g[j]), (g[i + 1], g[j]), (g[i + 1], g[j + 1]), (g[i], g[j + 1])]
                    for c, (s1, s2) in enumerate(corners):
                        p = np.zeros(3)
                        p[axis] = sgn * h
                        p[a1] = s1
                        p[a2] = s2
                        q[i, j, c] = p
            q = q.reshape(-1, 4, 3)
            # CellFill normal = (mid23-mid01) x (mid30-mid12); flip order where it points inward
            m01 = 0.5 * (q[:, 0] + q[:, 1]); m23 = 0.5 * (q[:, 2] + q[:, 3])
            m12 = 0.5 * (q[:, 1] + q[:, 2]); m30 = 0.5 * (q[:, 3] + q[:, 0])
            nrm = np.cross(m23 - m01, m30 - m12)
            if nrm[0, axis] * sgn < 0:
                q = q[:, ::-1]
            faces.append(q)
    return np.ascontiguousarray(np.concatenate(faces))


def sphere_points(m: int, radius: float, seed: int = 2) -> np.ndarray:
    """m points uniform on a sphere: z = 2u1 - 1, phi = 2 pi u2."""
    rng = MT19937_64(seed)
    p = np.empty((m, 3))
    for i in range(m):
        z = 2.0 * ((rng.next() >> 11) * 2.0 ** -53) - 1.0
        phi = 2.0 * math.pi * ((rng.next() >> 11) * 2.0 ** -53)
        s = math.sqrt(max(0.0, 1.0 - z * z))
        p[i] = (radius * s * math.cos(phi), radius * s * math.sin(phi), radius * z)
    return p


def grid_points(n_theta: int, n_phi: int, radius: float) -> np.ndarray:
    """n_theta x n_phi theta-phi grid of points (cell-centred in theta)."""
    th = np.pi * (np.arange(n_theta) + 0.5) / n_theta
    ph = 2.0 * np.pi * np.arange(n_phi) / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    return np.ascontiguousarray(np.stack([radius * np.sin(T) * np.cos(P),
                                          radius * np.sin(T) * np.sin(P),
                                          radius * np.cos(T)], axis=-1).reshape(-1, 3))


def write_dat(path: str, verts: np.ndarray) -> None:
    """Write cells in the reference's .dat layout (mesh/surf_mesh.cpp:102-175):
    one object, one module, `sign nfm x x`, the vertices, then 0 tear lines."""
    n, nv, _ = verts.shape
    with open(path, "w") as f:
        f.write(f"1\n1\n1 {n} 0 0\n")
        for c in range(n):
            for v in range(nv):
                f.write("%.17g %.17g %.17g\n" % tuple(verts[c, v]))
        f.write("0\n")


def read_vtk(path: str, nv: int) -> np.ndarray:
    """Read an ASCII legacy VTK surface (POINTS + CELLS or POLYGONS) as the
    reference's reader does (mesh/surf_mesh.h:367-555, surf_mesh.cpp:195-294):
    cells with exactly nv vertices are kept in file order, others skipped."""
    points, polys = None, []
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        line = lines[i].strip()
        i += 1
        if not line or line.startswith("#"):
            continue
        tok = line.split()
        kw = tok[0]
        if kw == "POINTS":
            npts = int(tok[1])
            pts = []
            while len(pts) < npts:
                vals = lines[i].split()
                i += 1
                pts.append([float(v) for v in vals[:3]])
            points = np.asarray(pts, dtype=np.float64)
        elif kw in ("CELLS", "POLYGONS"):
            ncell = int(tok[1])
            polys = []
            for _ in range(ncell):
                vals = lines[i].split()
                i += 1
                n = int(vals[0])
                polys.append([int(v) for v in vals[1:1 + n]])
        elif kw == "CELL_TYPES":
            i += int(tok[1])
        elif kw in ("CELL_DATA", "POINT_DATA"):
            break
    if points is None or not polys:
        raise ValueError(f"{path}: no points or cells")
    cells = [points[p] for p in polys if len(p) == nv]
    for p in polys:
        if len(p) == nv and (min(p) < 0 or max(p) >= len(points)):
            raise ValueError(f"{path}: point index out of range")
    return np.ascontiguousarray(np.asarray(cells, dtype=np.float64).reshape(-1, nv, 3))


def read_mesh(path: str, nv: int) -> np.ndarray:
    """SurfMesh::ReadFromFile (mesh/surf_mesh.h:353-364): by file extension."""
    if path.endswith(".dat"):
        return read_dat(path, nv)
    if path.endswith(".vtk"):
        return read_vtk(path, nv)
    raise ValueError(f"unsupported mesh file type: {path}")


def read_dat(path: str, nv: int) -> np.ndarray:
    """Read a .dat surface mesh (mesh/surf_mesh.cpp:18-175): objects, modules
    (`sign nfm unused unused`), nv vertices per cell; tear lines are skipped."""
    with open(path) as f:
        tok = f.read().split()
    pos = 0
    def nxt():
        nonlocal pos
        pos += 1
        return tok[pos - 1]
    cells = []
    n_obj = int(nxt())
    for _ in range(n_obj):
        kmod = int(nxt())
        for _ in range(kmod):
            _sign, nfm = int(nxt()), int(nxt())
            nxt(); nxt()
            for _ in range(nfm):
                cells.append([[float(nxt()) for _ in range(3)] for _ in range(nv)])
    return np.asarray(cells, dtype=np.float64).reshape(-1, nv, 3)
