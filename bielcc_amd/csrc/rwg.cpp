// RWG-type mesh topology (host side): unique points, unique edges, edge-cell
// incidence, cell sides with their direction, edge mid points and orts, and
// the cells beside each side -- the data RwgSurfMesh builds
// (mesh/rwg_surf_mesh.h:300-526) for div-conforming formulations and for
// adjacency lists -- plus its VTK export (rwg_surf_mesh.h:541-660).
//
// The arrays must match the reference's entry for entry, so the orders are the
// reference's:
//  - points: every cell vertex in cell order, sorted by the tolerance
//    lexicographic order of Point::operator< (mesh/point.cpp:51-63), then
//    runs of points closer than POINT_TOLERANCE = 1e-5
//    (geometry/tolerance_constants.h:12) merged into their first member;
//    std::sort and std::unique with those two comparisons give the same
//    sequence as the reference, which calls the same library algorithms on the
//    same input;
//  - a cell vertex maps to the first point within the tolerance;
//  - edges: (lower, higher) point pairs of the cell sides, sorted and unique;
//    a side whose two points coincide is a pseudo-edge (-5, direction 0).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>

#include "bielcc_hip.h"

namespace {

constexpr double kPointTol = 1e-5;
constexpr double kPointTolSq = kPointTol * kPointTol;
constexpr int kNone = -5;

struct P3 {
    double c[3];
};

bool p_less(const P3& a, const P3& b) {
    for (int d = 0; d < 3; d++)
        if (std::abs(a.c[d] - b.c[d]) > kPointTol) return a.c[d] < b.c[d];
    return false;
}

bool p_same(const P3& a, const P3& b) {
    const double dx = a.c[0] - b.c[0], dy = a.c[1] - b.c[1], dz = a.c[2] - b.c[2];
    return (dx * dx + dy * dy + dz * dz) < kPointTolSq;
}

}  // namespace

struct bielcc_rwg_s {
    int n = 0, nv = 0;
    std::vector<P3> points;
    std::vector<int> cell_points;                 // [n][nv]
    std::vector<int> point_num_cells;
    std::vector<std::vector<int>> point_cells;
    std::vector<std::vector<int>> point_edges;
    std::vector<std::array<int, 2>> edges;
    std::vector<int> edge_num_cells;
    std::vector<std::vector<int>> edge_cells;
    std::vector<int> cell_edges, cell_dir;        // [n][nv]
    std::vector<P3> edge_mid, edge_ort;
    std::vector<int> nghb_off, nghb;              // per (cell, side): the other cells on the side, or -5
};

namespace {

// first point (lowest index) within the tolerance of v: candidates from a
// grid of 1e-3 buckets around v, so a match is never missed
struct PointIndex {
    double h = 1e-3;
    std::unordered_map<long long, std::vector<int>> grid;
    static long long key(long long i, long long j, long long k) {
        return ((i & 0x1fffff) << 42) | ((j & 0x1fffff) << 21) | (k & 0x1fffff);
    }
    void build(const std::vector<P3>& pts) {
        for (int p = 0; p < (int)pts.size(); p++)
            grid[key((long long)std::floor(pts[p].c[0] / h), (long long)std::floor(pts[p].c[1] / h),
                     (long long)std::floor(pts[p].c[2] / h))]
                .push_back(p);
    }
    int find(const std::vector<P3>& pts, const P3& v) const {
        const long long i0 = (long long)std::floor(v.c[0] / h), j0 = (long long)std::floor(v.c[1] / h),
                        k0 = (long long)std::floor(v.c[2] / h);
        int best = -1;
        for (long long i = i0 - 1; i <= i0 + 1; i++)
            for (long long j = j0 - 1; j <= j0 + 1; j++)
                for (long long k = k0 - 1; k <= k0 + 1; k++) {
                    auto it = grid.find(key(i, j, k));
                    if (it == grid.end()) continue;
                    for (int p : it->second)
                        if ((best < 0 || p < best) && p_same(pts[p], v)) best = p;
                }
        return best;
    }
};

void build(bielcc_rwg_s& r, const double* verts) {
    const int n = r.n, nv = r.nv;
    // unique points
    r.points.resize((size_t)n * nv);
    for (size_t t = 0; t < r.points.size(); t++)
        for (int d = 0; d < 3; d++) r.points[t].c[d] = verts[3 * t + d];
    std::sort(r.points.begin(), r.points.end(), p_less);
    r.points.resize(std::unique(r.points.begin(), r.points.end(), p_same) - r.points.begin());
    const int np_ = (int)r.points.size();
    PointIndex idx;
    idx.build(r.points);
    // cell points; a point counts once per cell that holds it
    r.cell_points.assign((size_t)n * nv, 0);
    r.point_num_cells.assign(np_, 0);
    r.point_cells.assign(np_, {});
    for (int c = 0; c < n; c++) {
        for (int v = 0; v < nv; v++) {
            P3 q;
            for (int d = 0; d < 3; d++) q.c[d] = verts[((size_t)c * nv + v) * 3 + d];
            const int p = idx.find(r.points, q);
            r.cell_points[(size_t)c * nv + v] = p;
            bool first = true;
            for (int u = 0; u < v; u++)
                if (r.cell_points[(size_t)c * nv + u] == p) { first = false; break; }
            if (first) {
                r.point_num_cells[p]++;
                r.point_cells[p].push_back(c);
            }
        }
    }
    // unique edges
    for (int c = 0; c < n; c++)
        for (int v = 0; v < nv; v++) {
            const int a = r.cell_points[(size_t)c * nv + v], b = r.cell_points[(size_t)c * nv + (v + 1) % nv];
            if (a != b) r.edges.push_back({std::min(a, b), std::max(a, b)});
        }
    std::sort(r.edges.begin(), r.edges.end());
    r.edges.resize(std::unique(r.edges.begin(), r.edges.end()) - r.edges.begin());
    const int ne = (int)r.edges.size();
    r.edge_num_cells.assign(ne, 0);
    r.edge_cells.assign(ne, {});
    r.cell_edges.assign((size_t)n * nv, kNone);
    r.cell_dir.assign((size_t)n * nv, 0);
    for (int c = 0; c < n; c++)
        for (int v = 0; v < nv; v++) {
            const int a = r.cell_points[(size_t)c * nv + v], b = r.cell_points[(size_t)c * nv + (v + 1) % nv];
            if (a == b) continue;   // pseudo-edge
            const std::array<int, 2> key = {std::min(a, b), std::max(a, b)};
            const int e = (int)(std::lower_bound(r.edges.begin(), r.edges.end(), key) - r.edges.begin());
            r.cell_edges[(size_t)c * nv + v] = e;
            r.cell_dir[(size_t)c * nv + v] = a < b ? 1 : -1;
            r.edge_num_cells[e]++;
            r.edge_cells[e].push_back(c);
        }
    // edge mid points and orts: (A + B) * 0.5 and (B - A) * (1 / |B - A|)
    r.edge_mid.resize(ne);
    r.edge_ort.resize(ne);
    r.point_edges.assign(np_, {});
    for (int e = 0; e < ne; e++) {
        const P3 &A = r.points[r.edges[e][0]], &B = r.points[r.edges[e][1]];
        r.point_edges[r.edges[e][0]].push_back(e);
        r.point_edges[r.edges[e][1]].push_back(e);
        const double dl = std::sqrt((A.c[0] - B.c[0]) * (A.c[0] - B.c[0]) + (A.c[1] - B.c[1]) * (A.c[1] - B.c[1]) +
                                    (A.c[2] - B.c[2]) * (A.c[2] - B.c[2]));
        const double inv = 1. / dl;
        for (int d = 0; d < 3; d++) {
            r.edge_mid[e].c[d] = (A.c[d] + B.c[d]) * 0.5;
            r.edge_ort[e].c[d] = (B.c[d] - A.c[d]) * inv;
        }
    }
    // the cells beside each side: the side's edge cells without this cell (one
    // occurrence removed), or -5 for a pseudo-edge or a side no other cell shares
    r.nghb_off.assign((size_t)n * nv + 1, 0);
    for (int c = 0; c < n; c++)
        for (int v = 0; v < nv; v++) {
            const int e = r.cell_edges[(size_t)c * nv + v];
            std::vector<int> others;
            if (e >= 0) {
                others = r.edge_cells[e];
                auto it = std::find(others.begin(), others.end(), c);
                if (it != others.end()) others.erase(it);
            }
            if (others.empty()) others.push_back(kNone);
            r.nghb.insert(r.nghb.end(), others.begin(), others.end());
            r.nghb_off[(size_t)c * nv + v + 1] = (int)r.nghb.size();
        }
}

}  // namespace

extern "C" {

int bielcc_rwg_create(int n_cells, int verts_per_cell, const double* verts, bielcc_rwg* out) {
    if (!out) return BIELCC_ERR_ARG;
    *out = nullptr;
    if (n_cells < 1 || (verts_per_cell != 3 && verts_per_cell != 4) || !verts) return BIELCC_ERR_ARG;
    for (size_t t = 0; t < (size_t)n_cells * verts_per_cell * 3; t++)
        if (!std::isfinite(verts[t])) return BIELCC_ERR_ARG;
    bielcc_rwg r = new bielcc_rwg_s();
    r->n = n_cells;
    r->nv = verts_per_cell;
    try {
        build(*r, verts);
    } catch (...) {
        delete r;
        return BIELCC_ERR_OOM;
    }
    *out = r;
    return BIELCC_OK;
}

int bielcc_rwg_destroy(bielcc_rwg r) {
    delete r;
    return BIELCC_OK;
}

int bielcc_rwg_sizes(bielcc_rwg r, int* n_points, int* n_edges, int* n_neighbours) {
    if (!r) return BIELCC_ERR_ARG;
    if (n_points) *n_points = (int)r->points.size();
    if (n_edges) *n_edges = (int)r->edges.size();
    if (n_neighbours) *n_neighbours = (int)r->nghb.size();
    return BIELCC_OK;
}

int bielcc_rwg_get(bielcc_rwg r, double* points, int* cell_points, int* point_num_cells, int* point_num_edges,
                   int* edges, int* edge_num_cells, int* cell_edges, int* cell_edge_dir, double* edge_mid,
                   double* edge_ort, int* neighbour_offsets, int* neighbours) {
    if (!r) return BIELCC_ERR_ARG;
    const size_t np_ = r->points.size(), ne = r->edges.size(), cs = (size_t)r->n * r->nv;
    for (size_t i = 0; i < np_; i++) {
        if (points)
            for (int d = 0; d < 3; d++) points[3 * i + d] = r->points[i].c[d];
        if (point_num_cells) point_num_cells[i] = r->point_num_cells[i];
        if (point_num_edges) point_num_edges[i] = (int)r->point_edges[i].size();
    }
    for (size_t t = 0; t < cs; t++) {
        if (cell_points) cell_points[t] = r->cell_points[t];
        if (cell_edges) cell_edges[t] = r->cell_edges[t];
        if (cell_edge_dir) cell_edge_dir[t] = r->cell_dir[t];
    }
    for (size_t e = 0; e < ne; e++) {
        if (edges) { edges[2 * e] = r->edges[e][0]; edges[2 * e + 1] = r->edges[e][1]; }
        if (edge_num_cells) edge_num_cells[e] = r->edge_num_cells[e];
        for (int d = 0; d < 3; d++) {
            if (edge_mid) edge_mid[3 * e + d] = r->edge_mid[e].c[d];
            if (edge_ort) edge_ort[3 * e + d] = r->edge_ort[e].c[d];
        }
    }
    if (neighbour_offsets) std::copy(r->nghb_off.begin(), r->nghb_off.end(), neighbour_offsets);
    if (neighbours) std::copy(r->nghb.begin(), r->nghb.end(), neighbours);
    return BIELCC_OK;
}

int bielcc_rwg_write_vtk(bielcc_rwg r, const double* cell_normals, const char* path) {
    if (!r || !cell_normals || !path) return BIELCC_ERR_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) return BIELCC_ERR_ARG;
    const int n = r->n, nv = r->nv, np_ = (int)r->points.size();
    std::fprintf(f, "# vtk DataFile Version 3.0\nRwgSurfMesh export\nASCII\nDATASET UNSTRUCTURED_GRID\n\n");
    std::fprintf(f, "POINTS %d double\n", np_);
    for (const P3& p : r->points) std::fprintf(f, "%.16e %.16e %.16e\n", p.c[0], p.c[1], p.c[2]);
    std::fprintf(f, "\nCELLS %d %d\n", n, n * (nv + 1));
    for (int c = 0; c < n; c++) {
        std::fprintf(f, "%d", nv);
        for (int v = 0; v < nv; v++) std::fprintf(f, " %d", r->cell_points[(size_t)c * nv + v]);
        std::fprintf(f, "\n");
    }
    std::fprintf(f, "\nCELL_TYPES %d\n", n);
    for (int c = 0; c < n; c++) std::fprintf(f, "%d\n", nv == 3 ? 5 : 9);
    std::fprintf(f, "\nCELL_DATA %d\nSCALARS cell_id int 1\nLOOKUP_TABLE default\n", n);
    for (int c = 0; c < n; c++) std::fprintf(f, "%d\n", c);
    std::fprintf(f, "\nFIELD CellFields 3\nedge_indices %d %d int\n", nv, n);
    for (int c = 0; c < n; c++)
        for (int v = 0; v < nv; v++) std::fprintf(f, v + 1 < nv ? "%d " : "%d\n", r->cell_edges[(size_t)c * nv + v]);
    std::fprintf(f, "\nedge_dir_index %d %d int\n", nv, n);
    for (int c = 0; c < n; c++)
        for (int v = 0; v < nv; v++) std::fprintf(f, v + 1 < nv ? "%d " : "%d\n", r->cell_dir[(size_t)c * nv + v]);
    std::fprintf(f, "\ncell_normal 3 %d double\n", n);
    for (int c = 0; c < n; c++)
        std::fprintf(f, "%.16e %.16e %.16e\n", cell_normals[3 * c], cell_normals[3 * c + 1], cell_normals[3 * c + 2]);
    std::fprintf(f, "\nPOINT_DATA %d\nSCALARS point_num_cells int 1\nLOOKUP_TABLE default\n", np_);
    for (int p = 0; p < np_; p++) std::fprintf(f, "%d\n", r->point_num_cells[p]);
    std::fprintf(f, "\nSCALARS point_num_edges int 1\nLOOKUP_TABLE default\n");
    for (int p = 0; p < np_; p++) std::fprintf(f, "%d\n", (int)r->point_edges[p].size());
    std::fprintf(f, "\n");
    const bool ok = std::ferror(f) == 0;
    return (std::fclose(f) == 0 && ok) ? BIELCC_OK : BIELCC_ERR_ARG;
}

}  // extern "C"
