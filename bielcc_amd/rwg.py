"""RWG-type mesh topology (RwgSurfMesh, mesh/rwg_surf_mesh.h:300-526).

Built by the library's host code (bielcc_rwg_* in include/bielcc_hip.h); no
GPU is involved.  Arrays follow the reference's orders entry for entry:
unique points (1e-5 tolerance), unique edges (lower, higher point), cell
sides with their edge (-5 for a pseudo-edge) and direction (+1 / -1 / 0),
and per side the cells beside it (-5: none).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


class RwgMesh:
    def __init__(self, verts):
        v = np.ascontiguousarray(verts, dtype=np.float64)
        if v.ndim != 3 or v.shape[2] != 3 or v.shape[1] not in (3, 4):
            raise ValueError("verts must be [n][3|4][3]")
        self.lib = _lib.load()
        self.n, self.nv = v.shape[0], v.shape[1]
        h = C.c_void_p()
        rc = self.lib.bielcc_rwg_create(self.n, self.nv, v.ctypes.data_as(_DP), C.byref(h))
        if rc != 0:
            raise ValueError(f"bielcc_rwg_create failed ({rc})")
        self.h = h
        npt, ne, ng = C.c_int(), C.c_int(), C.c_int()
        self.lib.bielcc_rwg_sizes(h, C.byref(npt), C.byref(ne), C.byref(ng))
        n, nv = self.n, self.nv
        self.points = np.zeros((npt.value, 3))
        self.cell_points = np.zeros((n, nv), np.int32)
        self.point_num_cells = np.zeros(npt.value, np.int32)
        self.point_num_edges = np.zeros(npt.value, np.int32)
        self.edges = np.zeros((ne.value, 2), np.int32)
        self.edge_num_cells = np.zeros(ne.value, np.int32)
        self.cell_edges = np.zeros((n, nv), np.int32)
        self.cell_edge_dir = np.zeros((n, nv), np.int32)
        self.edge_mid = np.zeros((ne.value, 3))
        self.edge_ort = np.zeros((ne.value, 3))
        self.neighbour_offsets = np.zeros(n * nv + 1, np.int32)
        self.neighbours = np.zeros(ng.value, np.int32)
        arrs = (self.points, self.cell_points, self.point_num_cells, self.point_num_edges, self.edges,
                self.edge_num_cells, self.cell_edges, self.cell_edge_dir, self.edge_mid, self.edge_ort,
                self.neighbour_offsets, self.neighbours)
        self.lib.bielcc_rwg_get(h, *[a.ctypes.data_as(_DP if a.dtype == np.float64 else _IP) for a in arrs])

    @property
    def n_points(self) -> int:
        return self.points.shape[0]

    @property
    def n_edges(self) -> int:
        return self.edges.shape[0]

    def neighbours_of(self, cell: int, side: int) -> list:
        k = cell * self.nv + side
        return self.neighbours[self.neighbour_offsets[k]:self.neighbour_offsets[k + 1]].tolist()

    def write_vtk(self, path: str, cell_normals) -> None:
        """ExportRwgfMeshToVTK (rwg_surf_mesh.h:541-660); cell_normals [n][3] are
        the CellFill normals (Mesh.tables()["norm"])."""
        nrm = np.ascontiguousarray(cell_normals, dtype=np.float64).reshape(self.n, 3)
        rc = self.lib.bielcc_rwg_write_vtk(self.h, nrm.ctypes.data_as(_DP), str(path).encode())
        if rc != 0:
            raise OSError(f"bielcc_rwg_write_vtk failed ({rc}) for {path}")

    def close(self):
        if getattr(self, "h", None) is not None:
            self.lib.bielcc_rwg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
