// Internal host-side declarations shared by the HIP translation units and the
// C-ABI implementation.  Not installed; the public surface is include/bielcc_hip.h.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

#include "bielcc_dev.h"

namespace bielcc_impl {

// Per-cell constant table, CT doubles per cell (row-major [n][CT]).
enum : int {
    CT_RKT = 0,      // collocation point (CellFill, mesh/cell.cpp:15-17 / :68-70)
    CT_NORM = 3,     // CellFill normal (midline cross product for quads)
    CT_TAU1 = 6,     // local basis tau^1
    CT_TAU2 = 9,     // local basis tau^2
    CT_AREA = 12,
    CT_DIAM = 13,    // get_diam
    CT_NFUNC = 14,   // norm_func (diagonal cross product) used by K0
    CT_CENTER = 17,  // get_center_mass
    CT_NU = 20,      // K0 edge normals nu_e, [4][3] (triangles: 3 used)
    CT = 32
};

struct DevMesh {
    int n = 0, nv = 0;
    double grid_step = 0.;   // max cell diameter (SurfMesh::GetGridStep, mesh/surf_mesh.h:322-328)
    double max_rad = 0.;     // max distance from a cell's centre to its vertices (bounds node offsets)
    double bc[3] = {0., 0., 0.}, br = 0.;   // a sphere around every cell centre (centre, radius)
    double* verts = nullptr; // [n][nv][3]
    double* ctab = nullptr;  // [n][CT]
    double* nodes = nullptr; // [n][nn][4]: sub-cell centroid (3) + weight
    double* segs = nullptr;  // [n][nv*ns][4]: edge node (3) + segment step length
    int nodes_n = 0, segs_n = 0;
    // cell-local FP32 copies for the mixed-precision kernels (built on first use)
    float4* nodes32 = nullptr;   // [n][nn]: node - centre (3), weight
    float* nodes32_yy = nullptr; // [n][nn]: |node - centre|^2
    int nodes32_n = 0;
    // near pairs of the last assembly (collocation points i0 .. i0 + ni - 1 against every
    // cell, rule analytic_calc): reused by the next assembly of the same rows and rule
    int* near_lq = nullptr;      // [near_tot] point index relative to i0
    int* near_lc = nullptr;      // [near_tot] cell
    long long near_tot = -1;     // -1: no list kept
    int near_i0 = -1, near_ni = -1;
    double near_ac = 0.;
};

// ---- tables (tables.hip)
hipError_t launch_cell_tables(hipStream_t s, int nv, int n, const double* verts, double* ctab);
hipError_t launch_node_tables(hipStream_t s, int nv, int n, const double* verts, int n_start, double* nodes);
hipError_t launch_seg_tables(hipStream_t s, int nv, int n, const double* verts, int n_seg, double* segs);
hipError_t launch_local32(hipStream_t s, int n, int cnt, const double* nodes, const double* ctab, float4* loc,
                          float* yy);

// ---- near lists (tables.hip)
// mode 0: near iff dist(q, center_c) < ac * diam_c
// mode 1: near iff !(dist(q, center_c) > thresh)
// far: the sphere around every cell centre (far[0..2] centre, far[3] radius) plus the largest
// near distance (far[4]: ac * max diam, or thresh); a workgroup whose points all lie beyond it
// skips the scan (no pair of theirs can be near).  null: no shortcut.
hipError_t launch_near_count(hipStream_t s, int nq, const double* q, int q_stride, int n, const double* ctab,
                             int mode, double ac, double thresh, int* counts, const double* far = nullptr);
hipError_t launch_scan(hipStream_t s, int n, const int* in, long long* out_excl, long long* total);
hipError_t launch_near_fill(hipStream_t s, int nq, const double* q, int q_stride, int n, const double* ctab,
                            int mode, double ac, double thresh, const long long* offs, int* list_q, int* list_c,
                            const double* far = nullptr);

// The far shortcut of a mesh for a near rule: {centre, radius, reach} (see launch_near_count).
struct NearFar {
    double v[5];
};
inline NearFar near_far(const DevMesh& m, int mode, double ac, double thresh) {
    return NearFar{{m.bc[0], m.bc[1], m.bc[2], m.br, mode == 0 ? ac * m.grid_step : thresh}};
}

// ---- field (field.hip)
struct FieldArgs {
    int nv, M, N, nn;
    const double* pts;      // [M][3]
    const double* ctab;
    const double* nodes;
    const float4* nodes32;  // cell-local FP32 nodes (prec = 1)
    const float* nodes32_yy;
    const double* jE;       // [nrhs][N][3] complex
    const double* jM;       // [nrhs][N][3] complex or null
    int nrhs;
    int rchunk;             // current sets per pass: 1, 2 or 4 (1 with jM)
    double k_re, k_im;
    int mode;               // near rule (see launch_near_count)
    double ac, thresh;
    int nsplit;
    double* part;           // [nsplit][nrhs][NACC][M]
    int fast_phase;         // node phase mode: 0 general sincos, 1 quadrant reduction, 2 and 3 centre-phase (phase.h)
    int prec;               // 0: FP64 node sums; 1: FP32 cell-local node sums
};
hipError_t launch_field_far(hipStream_t s, const FieldArgs& a);
// Cell splits of the field tile launch for a.M points and a.N cells: at least s_min,
// chosen against the last-round tail (tail_aware_splits) for the kernel instance of the
// launch's first pass.
int field_far_splits(const FieldArgs& a, int s_min);

// Workgroups of a 2-D tile launch (bx point blocks x s cell splits) run in rounds of
// `slots` = resident workgroups on the device; a last round that is only partly
// filled leaves compute units idle.  Among s_min .. min(max_s, 3 s_min) the split
// count that minimises the time model ceil(bx s / slots) / s (fewer on ties).  On
// sphere20k: field 6 splits (4.6 rounds) -> 13 (9.9 rounds): 425 -> 392 ms; assembly
// 26 splits (2.7 rounds) -> 29 (3.0 rounds): 85 -> 79 ms.
inline int tail_aware_splits(int bx, int s_min, int max_s, int slots) {
    s_min = s_min < 1 ? 1 : s_min;
    if (max_s < s_min) return max_s < 1 ? 1 : max_s;
    if (slots <= 0 || bx <= 0) return s_min;
    int best = s_min;
    double best_t = 1e300;
    const int hi = max_s < 3 * s_min ? max_s : 3 * s_min;
    for (int s = s_min; s <= hi; s++) {
        const long long rounds = ((long long)bx * s + slots - 1) / slots;
        const double t = (double)rounds / s;
        if (t < best_t * (1. - 1e-9)) { best_t = t; best = s; }
    }
    return best;
}
// Resident workgroups of `kernel` (block threads, no dynamic LDS) on the current device.
int resident_slots(const void* kernel, int block);
// as_matrix = 0: per pair, the operators applied to jE / jM (one current set);
// as_matrix = 1: per pair, the 3x3 operator matrices (any number of sets).
hipError_t launch_field_near(hipStream_t s, int nv, long long npairs, const int* lq, const int* lc,
                             const double* pts, const double* verts, const double* jE, const double* jM,
                             const bielcc_dev::NumP& np, double k_re, double k_im, int mode, int as_matrix,
                             double* out);
struct FinalizeArgs {
    int M, N, nsplit, nrhs;
    const double* part;
    const double* jE;        // [nrhs][N][3] (read with near matrices)
    const double* jM;        // null without magnetic currents
    const long long* offs;   // near-pair offsets per point, or null
    const int* lc;           // near-pair cell indices
    const double* near_out;
    int near_mat;            // near_out holds operator matrices
    double mEr, mEi, mHr, mHi;
    const double* inc;       // [nrhs][6] plane wave (kvec, e0) or null
    const double* pts;
    double* E;               // [nrhs][M][3] complex
    double* H;               // same, or null
};
hipError_t launch_field_finalize(hipStream_t s, const FinalizeArgs& f);

// ---- FP32 error control of the mixed field (field.hip)
struct FixArgs {
    int M, N, nrhs;
    const double* pts;       // [M][3]
    const double* ctab;
    const double* jE;        // [nrhs][N][3] complex
    const double* jM;        // or null
    double k_re;
    double mE2, mH2;         // |multE|^2, |multH|^2
    const double* inc;       // device [nrhs][6] (kvec, e0) or null: E minus it is the scattered field
    const double* E;         // [nrhs][M][3] complex (FP32 sums)
    const double* H;         // or null
    double* mx;              // [nrhs][2] max |E_scat|, |H| (zeroed by the launch)
    int* flags;              // [nrhs][M]
};
// flags, then the list of points that need FP64 sums (count zeroed by the launch)
hipError_t launch_fix_flags(hipStream_t s, const FixArgs& f, int* count, int* list);
hipError_t launch_fix_gather(hipStream_t s, int m, const int* list, const double* pts, double* sub);
hipError_t launch_fix_scatter(hipStream_t s, int m, int M, int nrhs, const int* list, const double* Es,
                              const double* Hs, double* E, double* H);

// ---- operator-matrix field (field.hip): one segment = one mesh and a range of M points
struct OpSeg {
    int M, N, nn, nv;
    const double* pts;       // [M][3]: the segment's first point
    const double* ctab;
    const double* nodes;
    const float4* nodes32;   // prec = 1
    const float* nodes32_yy;
    const double* verts;     // [N][nv][3] (near pairs)
    double k_re, k_im;
    int fast_phase;
    int mode;                // near rule (see launch_near_count)
    double ac, thresh;
    double* WK;              // [3N][3M] complex, column-major, ld 3M: K_RotRot_Far / K near, /4pi included
    double* WR;              // same for R_Rot_Colloc / R_Rot_Smooth, or null
};
// Persistent tile kernel over tiles (segment, first point, cell begin, cell end).
hipError_t launch_field_opmat(hipStream_t s, const OpSeg* segs_d, const int4* tiles_d, int ntiles, int grid, bool ck,
                              bool p32);
// Exact near pairs of one segment (lq = point within the segment) into WK / WR.
hipError_t launch_opmat_near(hipStream_t s, long long npairs, const int* lq, const int* lc, const OpSeg& sg,
                             const bielcc_dev::NumP& np, int mode);
// E[r][p] = plane wave inc[r] = (kvec, e0) at pts[p], or 0 when inc is null.
hipError_t launch_init_incident(hipStream_t s, int M, int nrhs, const double* pts, const double* inc, double* E);

// ---- far-field pattern (farfield.hip)
struct PatternArgs {
    int N, nn, ndir, nrhs, nsplit;
    const double* dirs;     // [ndir][3] unit vectors
    const double* ctab;
    const double* nodes;
    const float4* nodes32;  // cell-local FP32 nodes (prec = 1)
    const double* jE;       // [nrhs][N][3] complex
    const double* jM;       // same or null
    double k_re, k_im;
    int fast_phase;         // |k| * max diameter within the polynomial range
    int prec;
    double* part;           // [nsplit][nrhs][12][ndir]
    double* Einf;           // [nrhs][ndir][3] complex
    double* Hinf;           // same or null
};
hipError_t launch_far_pattern(hipStream_t s, const PatternArgs& a, double mEr, double mEi, double mHr, double mHi);

// ---- assembly (assemble.hip)
struct AssembleArgs {
    int nv, N, nn, ns;
    int which;              // 0: K_RotRot_HS_Colloc, 1: R_Rot_Colloc
    int i0, ni;             // collocation points i0 .. i0+ni-1 (rows 2*i0 .. 2*(i0+ni)-1)
    const double* ctab;
    const double* nodes;
    const double* segs;
    double k_re, k_im;
    double ac;
    double* Z;              // column-major, complex, leading dimension ld
    long long ld;
    int fast_phase;         // see FieldArgs::fast_phase
    int prec;               // 0: FP64; 1: FP32 surface sums (contours stay FP64)
    const float4* nodes32;  // cell-local FP32 surface nodes (prec = 1)
    const float* nodes32_yy;
    int wg_target;          // workgroups the launch aims for (cells are split to reach it)
};

// pi/4 + |k| * (max cell diameter) bounds every reduced node phase (phase.h)
bool fast_phase_ok(double k_re, double grid_step);
// node phase mode of the tile kernels: 2 (centre-phase reduction, short series) when
// 3 when |k| max_rad <= 0.25, 2 when <= 0.5 (centre phase), else 1 (quadrant reduction) when
// fast_phase_ok, else 0
int phase_mode(double k_re, double grid_step, double max_rad);
hipError_t launch_assemble_far(hipStream_t s, const AssembleArgs& a);
// Persistent far-pair kernel over several meshes: items_d[tile.x] with point
// block tile.y (256 collocation points) and cells [tile.z, tile.w).  All
// items share which and prec; ck = some item has a complex wave number.
hipError_t launch_assemble_multi(hipStream_t s, const AssembleArgs* items_d, const int4* tiles_d, int ntiles,
                                 int grid, int which, bool ck, bool p32);
hipError_t launch_assemble_generic(hipStream_t s, int nv, int which, long long npairs, const int* lq,
                                   const int* lc, int N, int i0, int ni, const double* verts, const double* ctab,
                                   const bielcc_dev::NumP& np, double k_re, double k_im, double* Z, long long ld);
hipError_t launch_rhs_planewave(hipStream_t s, int N, const double* ctab, int nrhs, const double* kvec_e0,
                                double* B, long long ldb);
hipError_t launch_currents(hipStream_t s, int N, const double* ctab, int nrhs, const double* b, long long ldb,
                           double* j);

// ---- generic batches (generic.hip)
hipError_t launch_op_batch(hipStream_t s, int op, int nv, int n, const double* cells, const double* x,
                           const double* j, const bielcc_dev::NumP& np, double k_re, double k_im, double* res);
// averaging functions: which 0 Ker_AverDiv, 1 Ker_AverDiv_NearEdge, 2 Ker_BasicAverFunc
hipError_t launch_aver_batch(hipStream_t s, int which, int n, const double* x, const double* y, double eps,
                             double* res);
hipError_t launch_kernel_batch(hipStream_t s, int f, int n, const double* x, const double* y,
                               const bielcc_dev::KP& kp, double* res);
hipError_t launch_integral_batch(hipStream_t s, int f, int nv, int n, const double* cells, const double* x,
                                 const bielcc_dev::KP& kp, int idim, int n_start, int p_max, double eps,
                                 double* res, double* r_out);
hipError_t launch_k0_batch(hipStream_t s, int variant, int f, int nv, int n, const double* cells,
                           const double* x, const double* j, const bielcc_dev::KP& kp, int n_start, int p_max,
                           double eps, double analyticR, double* res, double* r_out);
hipError_t launch_analytic_batch(hipStream_t s, int which, int nv, int n, const double* cells, const double* x,
                                 int n_start, int p_max, double eps, double analyticR, double k_re, double k_im,
                                 double* res);
hipError_t launch_plane_wave(hipStream_t s, int n, const double* pts, const double* kvec_e0, double* out);
// PlaneWaveFieldH parameters: ke = (kvec, e0), frequency, absolute permeability
struct PlaneWaveH {
    double ke[6];
    double freq, mu_re, mu_im;
};
hipError_t launch_plane_wave_h(hipStream_t s, int n, const double* pts, const PlaneWaveH& pw, double* out);
hipError_t launch_add_incident_h(hipStream_t s, int M, int nrhs, const double* pts, const double* inc, double freq,
                                 double mu_re, double mu_im, double s_re, double s_im, double* H);
hipError_t launch_integral_y_batch(hipStream_t s, int f, int nv, int n, const double* cells, const bielcc_dev::KP& kp,
                                   const double* e0, int idim, int n_start, int p_max, double eps, double* res,
                                   double* r_out);

// ---- LU (lu.hip)
struct LuWork;
LuWork* lu_work_create();
void lu_work_destroy(LuWork*);
// few-RHS solves: sync-free triangular kernels (on) or blocked TRSM + GEMM (off)
void lu_set_trsv(LuWork*, bool on);
bool lu_get_trsv(const LuWork*);
// many right-hand sides: rocBLAS trsm per triangle (on) or the blocked TRSM + GEMM loop (off)
void lu_set_lib_trsm(LuWork*, bool on);
bool lu_get_lib_trsm(const LuWork*);
// polls per wait of the sync-free solve before it gives up and the solve is redone blocked
// (default 2^20, about a second; a real wait lasts microseconds)
void lu_set_spin_limit(LuWork*, int polls);
int lu_get_spin_limit(const LuWork*);
// LU panel: one launch per column (pivot of the next column fused into the update): 1 always,
// 0 never (two launches), 2 for small orders only
void lu_set_fused_panel(LuWork*, int mode);
int lu_get_fused_panel(const LuWork*);
// LU look-ahead: next panel factored on a side stream during the trailing update
void lu_set_look_ahead(LuWork*, int mode);
int lu_get_look_ahead(const LuWork*);
// whole-panel kernel (one launch per panel, grid barrier per column, bounded waits)
void lu_set_panel_lds(LuWork*, int mode);   // 1 within the workgroup budget, 0 never, 2 always
void lu_set_panel_rows32(LuWork*, int);
void lu_set_panel_rows64(LuWork*, int);
void lu_set_tournament(LuWork*, int mode);   // panels by tournament pivoting: 1 always, 0 never
int lu_get_tournament(const LuWork*);
void lu_set_gemm16(LuWork*, int mode);       // FP32 LU rank-512 updates on FP16 matrix cores: 1 split, 2 FP16, 0 off
int lu_get_gemm16(const LuWork*);
void lu_set_look_ahead2(LuWork*, int mode);  // two-level LU: next outer block on the side stream (1) or not (0)
int lu_get_look_ahead2(const LuWork*);
int lu_get_panel_rows32(const LuWork*);
int lu_get_panel_rows64(const LuWork*);
// rocBLAS atomics mode of the context's handle (off: reproducible GEMM results)
void lu_set_blas_atomics(LuWork*, bool on);
// columns per outer block of the two-level LU (64: one level, with the optional look-ahead)
void lu_set_outer_nb(LuWork*, int nb);
// one-kernel triangular solves: FP64 up to max64 right-hand sides; FP32 beyond 8 in chunks (-1: keep)
void lu_set_trsv_limits(LuWork*, int max64, int chunks32, int max32 = -1);
int lu_get_trsv_max64(const LuWork*);
int lu_get_trsv_max32(const LuWork*);
// refinement: also stop only once cond_est x residual <= ferr (0: residual only)
void lu_set_refine_mode(LuWork*, int mode, int restart);   // negative / 0: unchanged
int lu_get_refine_mode(const LuWork*);
int lu_get_refine_restart(const LuWork*);
void lu_set_resid_dd(LuWork*, int mode);
int lu_get_resid_dd(const LuWork*);
// R = B - A X (n x nrhs, complex FP64), sums in FP64 (ZGEMM) or doubled precision (Dot2)
hipError_t residual(hipStream_t s, LuWork* w, int n, const double* A, long long lda, const double* X, long long ldx,
                    const double* B, long long ldb, int nrhs, bool doubled, double* R, long long ldr,
                    const char** msg);
void lu_set_refine_ferr(LuWork*, double ferr);
double lu_get_refine_ferr(const LuWork*);
bool lu_get_trsv_chunks32(const LuWork*);
void lu_set_trsm_lib_panel(LuWork*, int mode);
int lu_get_trsm_lib_panel(const LuWork*);
int lu_get_outer_nb(const LuWork*);
// sync-free triangular solves redone on the blocked path (which 0: a wait expired, 1: ticket fault)
int lu_trsv_retries(const LuWork*, int which);
bool lu_get_blas_atomics(const LuWork*);
int lu_get_panel_lds(const LuWork*);
void lu_set_panel_wait_us(LuWork*, int us);
int lu_get_panel_wait_us(const LuWork*);
int lu_panel_aborts(LuWork*);
int lu_late_admits(const LuWork*);
size_t lu_work_bytes(const LuWork*);
// returns hipSuccess / error; info = 0 or first zero pivot (1-based); msg on library failure
hipError_t lu_factor(hipStream_t s, LuWork* w, int n, double* A, long long lda, int* piv, int* info,
                     const char** msg);
hipError_t lu_solve(hipStream_t s, LuWork* w, int n, const double* A, long long lda, const int* piv, int nrhs,
                    double* B, long long ldb, const char** msg);
// FP32 factors F of the FP64 matrix A (A untouched)
hipError_t lu32_factor(hipStream_t s, LuWork* w, int n, const double* A, long long lda, float* F, long long ldf,
                       int* piv, int* info, const char** msg);
// FP64 iterative refinement on FP32 factors; status 1 = stagnated above tol
// F = U^T A V (random butterflies from seed, diagonals into rbt: 4 (n & ~3) doubles) factored
// without interchanges (piv: identity); A is not modified
hipError_t lu_factor_rbt(hipStream_t s, LuWork* w, int n, const double* A, long long lda, double* F, long long ldf,
                         int* piv, double* rbt, unsigned long long seed, int* info, const char** msg);
// A X = B on those factors, refined against A (classic steps) to tol; status 1: not reached
hipError_t rbt_solve(hipStream_t s, LuWork* w, int n, const double* A, long long lda, const double* F,
                     long long ldf, const int* piv, const double* rbt, int nrhs, double* B, long long ldb, double tol,
                     int max_iter, int* iters, double* resid, int* status, const char** msg);
// FP32 factors of U^T A V (butterflies as lu_factor_rbt), without interchanges
hipError_t lu32_factor_rbt(hipStream_t s, LuWork* w, int n, const double* A, long long lda, float* F, long long ldf,
                           int* piv, double* rbt, unsigned long long seed, int* info, const char** msg);
// rbt: the butterflies' diagonals when F factors U^T A V (lu32_factor_rbt), else null
hipError_t refine_solve(hipStream_t s, LuWork* w, int n, const double* A, long long lda, const float* F,
                        long long ldf, const int* piv, int nrhs, double* B, long long ldb, double tol, int max_iter,
                        int* iters, double* resid, int* status, const char** msg, const double* rbt = nullptr);
// C = alpha A B + beta C, complex FP64 column-major, on the BLAS handle of w
// (matrix cores).  Creates the handle on first use: call once before a
// stream capture.
hipError_t blas_zgemm(hipStream_t s, LuWork* w, long long m, long long n, long long k, double2 alpha,
                      const double* A, long long lda, const double* B, long long ldb, double2 beta, double* C,
                      long long ldc, const char** msg);
// batch GEMMs C_b = alpha A_b B_b + beta C_b with A_b = A + b sa, B_b = B + b sb, C_b = C + b sc
// (strides in complex elements)
hipError_t blas_zgemm_strided(hipStream_t s, LuWork* w, long long m, long long n, long long k, double2 alpha,
                              const double* A, long long lda, long long sa, const double* B, long long ldb,
                              long long sb, double2 beta, double* C, long long ldc, long long sc, long long batch,
                              const char** msg);
// out[:, r] (+)= sum_b part[b][:, r] for an m x nrhs column-major block (ldo; part dense, m rows);
// accumulate = 0 overwrites out
hipError_t launch_sum_partials(hipStream_t s, long long m, int nrhs, int nparts, const double* part, double* out,
                               long long ldo, int accumulate);

}  // namespace bielcc_impl
