/*
 * bielcc_hip.h - C ABI of the MI355X (gfx950) implementation of BIELcc's
 * Stratton-Chu K / R / K0 path.
 *
 * Conventions
 *  - Every function returns an int status (BIELCC_OK = 0) and never throws;
 *    bielcc_last_error() gives the message of the last failure on a context.
 *  - Complex numbers are interleaved (re, im) doubles.  3-vectors are [3].
 *    Cells are [n][nv][3] doubles, nv = 4 (CellQuad) or 3 (CellTriangle).
 *  - Buffers named *_d are DEVICE pointers owned by the caller (hipMalloc,
 *    bielcc_malloc or a framework allocator); everything else is host memory.
 *    Meshes and scratch are owned by the context.
 *  - Calls are ordered on the context's stream (bielcc_stream()); results are
 *    complete after bielcc_sync().  Use one context per host thread.
 *  - The wave number is complex (k_re, k_im): k_im = 0 is the reference's
 *    KType = double, k_im != 0 its KType = std::complex<double> (lossy media).
 *  - Outputs are overwritten, never accumulated (integral_universal.h:53-55).
 */
#ifndef BIELCC_HIP_H
#define BIELCC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BIELCC_OK 0
#define BIELCC_ERR_ARG 1          /* invalid argument */
#define BIELCC_ERR_OOM 2          /* device allocation failed */
#define BIELCC_ERR_HIP 3          /* HIP runtime or library error */
#define BIELCC_ERR_UNSUPPORTED 4  /* valid input outside what is implemented */
#define BIELCC_ERR_SINGULAR 5     /* zero pivot in the LU */
#define BIELCC_ERR_NOTFOUND 6     /* unknown mesh id */

typedef struct bielcc_ctx_s* bielcc_ctx;

/* NumParam (electrodynamics/num_param.h:28-55): the fields the path reads. */
typedef struct {
    double eps, rs, rs_seg, analytic_calc;
    int n_start, n_start_seg, p_max, p_max_seg;
} bielcc_num_param;

/* KernelParam (integrators/kernel_param.h:25-57). */
typedef struct {
    double smoothR, smoothRMin;
    double k_re, k_im;
    double vec_dbl[3];
    double n[3];
    double current[6];
} bielcc_kernel_param;

/* IntegralParam (integrators/integral_param.h:16-45). */
typedef struct {
    int idim, n_start, p_max;
    double eps;
} bielcc_integral_param;

/* Operator ids (electrodynamics/K.h, R.h). */
enum {
    BIELCC_OP_K_NEAR_SMOOTH = 0, /* K_RotRot_Near_Smooth K.h:47-79 */
    BIELCC_OP_K_HS = 1,          /* K_RotRot_HS          K.h:119-145 */
    BIELCC_OP_K_HS_COLLOC = 2,   /* K_RotRot_HS_Colloc   K.h:179-204 */
    BIELCC_OP_K_HS2_COLLOC = 3,  /* K_RotRot_HS2_Colloc  K.h:230-252 */
    BIELCC_OP_K_FAR = 4,         /* K_RotRot_Far         K.h:288-301 */
    BIELCC_OP_R_SMOOTH = 5,      /* R_Rot_Smooth         R.h:46-61 */
    BIELCC_OP_R_COLLOC = 6,      /* R_Rot_Colloc         R.h:89-111 */
    BIELCC_OP_R_NEAR = 7         /* R_Rot_Near           R.h:140-175 */
};

/* Integrand ids (electrodynamics/kernel_lib.h). */
enum {
    BIELCC_F_SIMPLEPOT_L = 0,       /* :62-69 */
    BIELCC_F_GRADSIMPLEPOT_L = 1,   /* :87-98 */
    BIELCC_F_DOUBLEPOT_L = 2,       /* :117-129 */
    BIELCC_F_VECTORPOT_L = 3,       /* :149-177 */
    BIELCC_F_SIMPLEPOT_H = 4,       /* :209-217 */
    BIELCC_NUF_SIMPLEPOT_H = 5,     /* :235-245 */
    BIELCC_F_SIMPLEPOT_HM1 = 6,     /* :262-278 */
    BIELCC_F_GRADSIMPLEPOT_H = 7,   /* :295-313 */
    BIELCC_F_GRADSIMPLEPOT_HM1 = 8, /* :332-356 */
    BIELCC_F_DOUBLEPOT_H = 9,       /* :375-390 */
    BIELCC_NF_DOUBLEPOT_H = 10,     /* :409-419 */
    BIELCC_F_ROTROTSIMPLEPOT_H = 11 /* :453-492 */
};

/* ------------------------------------------------------------ context */
int bielcc_version(void);
/* New context on HIP device `device` with its own default-priority stream. */
int bielcc_ctx_create(int device, bielcc_ctx* out);
int bielcc_ctx_destroy(bielcc_ctx ctx);
/* Blocks until all work queued on the context's stream has finished. */
int bielcc_sync(bielcc_ctx ctx);
/* The context's hipStream_t, as void*. */
void* bielcc_stream(bielcc_ctx ctx);
/* Copies the last error message (NUL-terminated, truncated to len). */
int bielcc_last_error(bielcc_ctx ctx, char* buf, int len);

/* ------------------------------------------------------------ memory */
int bielcc_malloc(bielcc_ctx ctx, size_t bytes, void** dptr);
int bielcc_free(bielcc_ctx ctx, void* dptr);
/* kind: 0 host->device, 1 device->host, 2 device->device; synchronous. */
int bielcc_memcpy(bielcc_ctx ctx, void* dst, const void* src, size_t bytes, int kind);

/* ------------------------------------------------------------ mesh
 * Replaces SurfMesh<CellQuad|CellTriangle> (mesh/surf_mesh.h:47-302) and the
 * per-cell CellFill data (mesh/cell.cpp:8-100): cells are uploaded once and
 * their tables (collocation point, normal, basis, area, diameter, K0 edge
 * normals, quadrature nodes) are built on the device. */
int bielcc_mesh_create(bielcc_ctx ctx, int n_cells, int verts_per_cell, const double* verts, int* mesh_id);
int bielcc_mesh_destroy(bielcc_ctx ctx, int mesh_id);
/* Host copy of the per-cell tables, [n][20]: rkt[3] norm[3] tau1[3] tau2[3]
 * area diam norm_func[3] center[3]  (GetRkt/GetNorm/GetTau1/GetTau2/GetArea,
 * get_diam, norm_func, get_center_mass). */
int bielcc_mesh_tables(bielcc_ctx ctx, int mesh_id, double* out);
/* SurfMesh::GetGridStep (mesh/surf_mesh.h:322-328): max cell diameter. */
int bielcc_mesh_grid_step(bielcc_ctx ctx, int mesh_id, double* grid_step);
int bielcc_mesh_size(bielcc_ctx ctx, int mesh_id, int* n_cells, int* verts_per_cell);

/* ------------------------------------------------------------ single functions, batched
 * Each item i is one call of the reference function on its own inputs. */

/* Operators K_RotRot_* / R_Rot_* (K.h:47-301, R.h:46-175):
 * res_d[i] = Op(j_d[i], x_d[i], cells_d[i], num_param, k).  j, res: [n][3] complex. */
int bielcc_op_batch(bielcc_ctx ctx, int op, int nv, int n, const double* cells_d, const double* x_d,
                    const double* j_d, const bielcc_num_param* np, double k_re, double k_im, double* res_d);

/* Integrands F_* (kernel_lib.h:62-492) at (x_d[i], y_d[i]); res [n][3] complex. */
int bielcc_kernel_batch(bielcc_ctx ctx, int f, int n, const double* x_d, const double* y_d,
                        const bielcc_kernel_param* kp, double* res_d);

/* Averaging functions (electrodynamics/kernel_lib.h:560-634) at (x_d[i], y_d[i]),
 * res [n][3] real: which 0 Ker_AverDiv (:573-583), 1 Ker_AverDiv_NearEdge
 * (:599-609), both with eps_edge = eps (the reference's templates read
 * KernelParam::eps_edge, a member KernelParam does not have, so they cannot be
 * instantiated there); 2 Ker_BasicAverFunc (:626-634) in res[i][0], the other
 * two entries 0. */
int bielcc_aver_batch(bielcc_ctx ctx, int which, int n, const double* x_d, const double* y_d, double eps,
                      double* res_d);

/* IntegralUniversalPnt (integral_universal.h:38-220; nv = 4 or 3, cells_d
 * [n][nv][3]) or IntegralUniversalSegPnt (integral_universal_seg.h:38-98;
 * nv = 2, cells_d = segments [n][2][3]).  kp is the starting KernelParam;
 * smoothR_d[i] (optional) receives its mutated smoothR. */
int bielcc_integral_batch(bielcc_ctx ctx, int f, int nv, int n, const double* cells_d, const double* x_d,
                          const bielcc_kernel_param* kp, const bielcc_integral_param* ip, double* res_d,
                          double* smoothR_d);

/* F(y) integrands of the IntegralUniversal forms. */
enum {
    BIELCC_FY_UNIT = 0,        /* f = 1 (examples/integrators/main0_integrals.cpp:77-81): area / length */
    BIELCC_FY_PLANE_WAVE_E = 1 /* PlaneWaveFieldE (incident_field.h:20-29): e0 e^{i vec_dbl . y} */
};

/* IntegralUniversal, the F(y) forms (integral_universal.h:252-333 quads,
 * :347-431 triangles; nv = 4 or 3, cells_d [n][nv][3]) or the F(y) form of
 * IntegralUniversalSegPnt (integral_universal_seg.h:120-179; nv = 2,
 * cells_d = segments [n][2][3]).  f: BIELCC_FY_*.  kp->vec_dbl is the plane
 * wave's wave vector and e0 (host [3]) its polarisation (KernelParam::e0; read
 * for BIELCC_FY_PLANE_WAVE_E only).  res_d [n][3] complex (entries past idim
 * 0); smoothR_d[i] (optional) receives the mutated smoothR. */
int bielcc_integral_y_batch(bielcc_ctx ctx, int f, int nv, int n, const double* cells_d,
                            const bielcc_kernel_param* kp, const double* e0, const bielcc_integral_param* ip,
                            double* res_d, double* smoothR_d);

/* variant 0: K0 (K0.h:36-86), 1: K0_Common (:112-166), 2: K0_GradSimplePot_H_HS
 * (:197-249, uses analyticR and kp->k only).  ip->idim must be 3. */
int bielcc_k0_batch(bielcc_ctx ctx, int variant, int f, int nv, int n, const double* cells_d, const double* x_d,
                    const double* j_d, const bielcc_kernel_param* kp, const bielcc_integral_param* ip,
                    double analyticR, double* res_d, double* smoothR_d);

/* which 0: integral1Divr (integrals_analytic.cpp:218-235, cells), 1:
 * integralGrad1DivrSeg (:240-257, segments), 2: IntegralSurf_SimplePot_H_HS
 * (integrals_analytic.h:136-164, cells), 3: IntegralSeg_GradSimplePot_H_HS
 * (:193-221, segments).  res [n][3] complex. */
int bielcc_analytic_batch(bielcc_ctx ctx, int which, int nv, int n, const double* cells_d, const double* x_d,
                          const bielcc_integral_param* ip, double analyticR, double k_re, double k_im,
                          double* res_d);

/* PlaneWaveFieldE (incident_field.h:20-29): out [n][3] complex. */
int bielcc_plane_wave_e(bielcc_ctx ctx, int n, const double* pts_d, const double* kvec, const double* e0,
                        double* out_d);

/* PlaneWaveFieldH (incident_field.h:40-59): (kvec x e0) e^{i kvec . x} /
 * (2 pi freq) / mu at n points, out [n][3] complex.  mu is the absolute
 * permeability MaterialParam::GetMu returns: real (mu_im = 0, divided as a
 * double, the reference's double branch) or complex.  In the normalised units
 * of the solver (eps = mu = 1, omega = k) freq = k / (2 pi) and mu = 1. */
int bielcc_plane_wave_h(bielcc_ctx ctx, int n, const double* pts_d, const double* kvec, const double* e0,
                        double freq, double mu_re, double mu_im, double* out_d);

/* Total magnetic field from a scattered one: H[r][p] = scale * H[r][p] +
 * PlaneWaveFieldH(x_p) of incidence r (inc host [nrhs][6] = (kvec, e0), the
 * layout of bielcc_field_batch), for H_d [nrhs][M][3] complex on the device.
 * With currents solved for mult = 1 (j' = multE j_E, multE = i / (omega eps)),
 * the Stratton-Chu H of j_E is R[j'] / multE: scale = 1 / multE. */
int bielcc_field_add_incident_h(bielcc_ctx ctx, int M, int nrhs, const double* pts_d, const double* inc,
                                double freq, double mu_re, double mu_im, double scale_re, double scale_im,
                                double* H_d);

/* ------------------------------------------------------------ RWG topology (host side, no context)
 * RwgSurfMesh (mesh/rwg_surf_mesh.h:300-526) of a cell list [n][nv][3]: unique
 * points (1e-5 tolerance, geometry/tolerance_constants.h:12), unique edges,
 * edge-cell incidence, cell sides with their direction (+1 / -1; a side whose
 * points coincide is a pseudo-edge: edge -5, direction 0), edge mid points and
 * orts, and the cells beside each side (-5: none).  Same orders as the
 * reference, entry for entry. */
typedef struct bielcc_rwg_s* bielcc_rwg;
int bielcc_rwg_create(int n_cells, int verts_per_cell, const double* verts, bielcc_rwg* out);
int bielcc_rwg_destroy(bielcc_rwg r);
/* n_neighbours = total entries of the neighbour lists */
int bielcc_rwg_sizes(bielcc_rwg r, int* n_points, int* n_edges, int* n_neighbours);
/* Host arrays, any may be NULL: points [np][3], cell_points [n][nv],
 * point_num_cells [np], point_num_edges [np], edges [ne][2] (lower, higher
 * point), edge_num_cells [ne], cell_edges [n][nv], cell_edge_dir [n][nv],
 * edge_mid [ne][3], edge_ort [ne][3], neighbour_offsets [n*nv+1] and
 * neighbours [n_neighbours] (the list of side (c, v) is neighbours[off[c*nv+v]
 * .. off[c*nv+v+1])). */
int bielcc_rwg_get(bielcc_rwg r, double* points, int* cell_points, int* point_num_cells, int* point_num_edges,
                   int* edges, int* edge_num_cells, int* cell_edges, int* cell_edge_dir, double* edge_mid,
                   double* edge_ort, int* neighbour_offsets, int* neighbours);
/* RwgSurfMesh::ExportRwgfMeshToVTK (rwg_surf_mesh.h:541-660); cell_normals [n][3]
 * are the cells' CellFill normals (bielcc_mesh_tables). */
int bielcc_rwg_write_vtk(bielcc_rwg r, const double* cell_normals, const char* path);

/* ------------------------------------------------------------ drivers */

/* Collocation matrix (no reference loop; operators named at K.h:156-178, R.h:65-88):
 *   Z[2i+a, 2j+b] = tau_i^a . Op(tau_j^b, rkt_i, cell_j)
 * which 0: Op = K_RotRot_HS_Colloc, 1: Op = R_Rot_Colloc.  Rows row0 ..
 * row0+nrows-1 (both even) of the 2N x 2N matrix go to Z_d rows 0..nrows-1,
 * column-major complex with leading dimension ld (in complex elements).  The
 * list of near pairs (written by the exact per-pair kernel) is built on the
 * first call for a row range and kept with the mesh for the next call with the
 * same rows and analytic_calc. */
int bielcc_assemble(bielcc_ctx ctx, int mesh_id, int which, const bielcc_num_param* np, double k_re, double k_im,
                    int row0, int nrows, double* Z_d, int64_t ld);

/* rhs[2i+a, r] = -tau_i^a . E_inc_r(rkt_i), E_inc_r = e0_r exp(i kvec_r . x)
 * (PlaneWaveFieldE).  kvec, e0: host [nrhs][3].  B_d column-major, ldb >= 2N. */
int bielcc_rhs_planewave(bielcc_ctx ctx, int mesh_id, int nrhs, const double* kvec, const double* e0, double* B_d,
                         int64_t ldb);

/* LU with partial pivoting of the n x n complex column-major A_d (in place);
 * piv_d[n] receives 0-based row interchanges.  Returns BIELCC_ERR_SINGULAR
 * (and *info = first zero pivot, 1-based) for an exactly singular matrix. */
int bielcc_lu_factor(bielcc_ctx ctx, int n, double* A_d, int64_t lda, int* piv_d, int* info);
/* Solves A X = B with the factors of bielcc_lu_factor; B_d (n x nrhs, ldb) is overwritten by X. */
int bielcc_lu_solve(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, const int* piv_d, int nrhs,
                    double* B_d, int64_t ldb);

/* SurfaceCurrents_byBasis (currents.h:22-33) for nrhs coefficient columns:
 * j_d[r][c][3] complex = b[2c, r] tau_c^1 + b[2c+1, r] tau_c^2. */
int bielcc_currents_by_basis(bielcc_ctx ctx, int mesh_id, int nrhs, const double* b_d, int64_t ldb, double* j_d);

/* Stratton-Chu field at M points (pts_d [M][3]):
 *   E = multE sum_j K[jE_j](x) - sum_j R[jM_j](x) (+ E_inc if kvec != NULL)
 *   H = multH sum_j K[jM_j](x) + sum_j R[jE_j](x)
 * (examples/field/main0_field_test.cpp:237-319).  near_mode 0: near iff
 * dist(x, centroid_j) < analytic_calc * diam_j (get_field.h:58), K near =
 * K_RotRot_Near_Smooth; near_mode 1: far iff dist > analytic_calc * grid_step
 * (main0_field_test.cpp:242), K near = K_RotRot_HS.  Far: K_RotRot_Far and
 * R_Rot_Colloc; near R: R_Rot_Smooth.  jE_d, jM_d: [N][3] complex (jM_d may
 * be NULL); E_d, H_d: [M][3] complex (H_d may be NULL). */
int bielcc_field(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im,
                 const double* jE_d, const double* jM_d, double multE_re, double multE_im, double multH_re,
                 double multH_im, const double* kvec, const double* e0, int near_mode, int M, const double* pts_d,
                 double* E_d, double* H_d);

/* bielcc_field for one current set in FP64 that allocates nothing on the
 * device.  No counterpart in the reference: it is for callers that cannot
 * allocate (tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, semantics and results as bielcc_field; its work memory (partial
 * sums, near-pair lists and results) is one file-scope device array of
 * 64 MiB per GPU, held until the process ends.  Limits: the mesh's quadrature
 * nodes for np->n_start must already exist (one bielcc_field or
 * bielcc_assemble call on the mesh builds them); inputs whose work memory
 * exceeds 64 MiB are rejected with BIELCC_ERR_UNSUPPORTED: 108 bytes per
 * point and split of the cell range, plus 104 bytes per near pair (200 with
 * jM) - about 620,000 points without near pairs (beyond 32 MiB of partial
 * sums fewer splits are used than bielcc_field uses, which changes only the
 * summation order).  Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU race.  The caller passes
 * and owns the inputs and outputs. */
int bielcc_field_noalloc(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im,
                         const double* jE_d, const double* jM_d, double multE_re, double multE_im, double multH_re,
                         double multH_im, const double* kvec, const double* e0, int near_mode, int M,
                         const double* pts_d, double* E_d, double* H_d);

/* GetFieldE_Ideal (get_field.h:35-78) with EInc_Func = PlaneWaveFieldE:
 * E = mult sum_j (near ? K_RotRot_Near_Smooth : K_RotRot_Far)[jE_j](x) + E_inc(x). */
int bielcc_get_field_e(bielcc_ctx ctx, int mesh_id, const double* jE_d, double mult_re, double mult_im,
                       const double* kvec, const double* e0, const bielcc_num_param* np, double k_re, double k_im,
                       int M, const double* pts_d, double* E_d);

/* ------------------------------------------------------------ batched and mixed-precision forms
 * No reference counterpart: the reference evaluates one current set per
 * GetFieldE_Ideal call (get_field.h:35-78) and has no solver.
 *
 * Precision of the far-pair surface quadrature.  Near and self pairs, the
 * analytic terms, the edge contours (K0) and every sum across cells stay FP64
 * in both forms. */
#define BIELCC_PREC_FP64 0   /* reference-exact */
#define BIELCC_PREC_MIXED 1  /* FP32 node sums in cell-local coordinates */

/* bielcc_field for nrhs current sets at once.  jE_d, jM_d: [nrhs][N][3]
 * complex (jM_d may be NULL); inc: host [nrhs][6] = (kvec, e0) per set, or
 * NULL for no incident field; E_d, H_d: [nrhs][M][3] complex. */
int bielcc_field_batch(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im,
                       int nrhs, const double* jE_d, const double* jM_d, double multE_re, double multE_im,
                       double multH_re, double multH_im, const double* inc, int near_mode, int precision, int M,
                       const double* pts_d, double* E_d, double* H_d);

/* bielcc_assemble with a choice of far-pair precision. */
int bielcc_assemble_ex(bielcc_ctx ctx, int mesh_id, int which, const bielcc_num_param* np, double k_re,
                       double k_im, int row0, int nrows, int precision, double* Z_d, int64_t ld);

/* FP32 complex LU with partial pivoting on the matrix cores: F_d (n x n
 * complex float, column-major, ldf) receives the factors of the FP64 matrix
 * A_d rounded to FP32; A_d is not modified.  piv_d[n], info as bielcc_lu_factor. */
int bielcc_lu32_factor(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, float* F_d, int64_t ldf, int* piv_d,
                       int* info);
/* LU without interchanges of a randomly transformed matrix (no reference counterpart:
 * the reference has no solver).  F_d (n x n complex FP64, ldf) receives the factors
 * L U = U^T A V, where U and V are random butterflies of depth 2 on the leading
 * 4 floor(n/4) rows / columns (Parker 1995; Baboulin, Dongarra, Herrmann, Tomov, ACM
 * TOMS 2013), drawn from seed; their diagonals go to rbt_d (device, 32 n bytes), piv_d
 * receives the identity.  A_d is not modified (the solve refines against it).  Each
 * panel is one launch in which no workgroup waits for another (the pivoting LU has a
 * device-wide barrier per column).  info as bielcc_lu_factor (zero pivot of F). */
int bielcc_lu_factor_rbt(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, double* F_d, int64_t ldf,
                         int* piv_d, double* rbt_d, uint64_t seed, int* info);
/* A X = B on bielcc_lu_factor_rbt's factors: X0 = V (L U)^{-1} U^T B, then classic
 * refinement steps X += V (L U)^{-1} U^T (B - A X) until max_r ||B_r - A X_r|| / ||B_r||
 * <= tol or max_iter steps.  B_d is overwritten by X.  Returns BIELCC_ERR_UNSUPPORTED when
 * the residual stops shrinking (by less than 10% a step) above tol, or max_iter steps
 * leave it there. */
int bielcc_rbt_solve(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, const double* F_d, int64_t ldf,
                     const int* piv_d, const double* rbt_d, int nrhs, double* B_d, int64_t ldb, double tol,
                     int max_iter, int* iters, double* resid);
/* FP32 factors of U^T A V (butterflies as bielcc_lu_factor_rbt), without interchanges:
 * F_d complex float, piv_d the identity, rbt_d the diagonals; A_d is not modified.  The
 * FP32 counterpart of bielcc_lu32_factor, for bielcc_refine_solve_rbt. */
int bielcc_lu32_factor_rbt(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, float* F_d, int64_t ldf,
                           int* piv_d, double* rbt_d, uint64_t seed, int* info);
/* bielcc_refine_solve on bielcc_lu32_factor_rbt's factors: every FP32 solve applies
 * V F^{-1} U^T.  Arguments and results as bielcc_refine_solve. */
int bielcc_refine_solve_rbt(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, const float* F_d, int64_t ldf,
                            const int* piv_d, const double* rbt_d, int nrhs, double* B_d, int64_t ldb, double tol,
                            int max_iter, int* iters, double* resid);

/* Solves A X = B in FP64 on the FP32 factors, from X0 = F^{-1} B, until
 * max_r ||B_r - A X_r||_2 / ||B_r||_2 <= tol or max_iter corrections.  By
 * default (tunable "refine_mode" = 1) the corrections are the steps of a
 * restarted flexible GMRES in FP64 (basis of "refine_restart" [20] vectors per
 * cycle), right-preconditioned by the FP32 factors; each cycle starts from the
 * FP64 residual.  With "refine_mode" = 0 they are classic refinement steps,
 * X += F^{-1} (B - A X).  B_d is overwritten by X; *iters = corrections applied
 * (FP32 solves after X0), *resid = final relative residual.  Returns
 * BIELCC_ERR_UNSUPPORTED when the residual stops shrinking above tol (classic
 * steps: by less than 10% a step; FGMRES: by less than half a cycle), or when
 * max_iter corrections leave it above tol; X is then the last iterate. */
int bielcc_refine_solve(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, const float* F_d, int64_t ldf,
                        const int* piv_d, int nrhs, double* B_d, int64_t ldb, double tol, int max_iter,
                        int* iters, double* resid);

/* R = B - A X for n x n A and n x nrhs X, B, R (complex FP64, column-major).
 * doubled = 0: sums in FP64 (ZGEMM); 1: every sum carried as a pair s + c
 * (Ogita-Rump-Oishi Dot2), as accurate as sums in twice the FP64 precision,
 * rounded once to FP64 -- the FP64 residual of a solution near the rounding
 * floor is mostly rounding error (about sqrt(n) u64 |A||X|), this one is not.
 * With "refine_ferr_exp" set, the FGMRES refinement switches to doubled
 * residuals once its FP64 residual has reached that floor ("refine_resid_dd"
 * [1]; 0: never, 2: from the first residual), so its solution converges to the
 * rounding of X itself.  No counterpart in the reference (which has no solver). */
int bielcc_residual(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, const double* X_d, int64_t ldx,
                    const double* B_d, int64_t ldb, int nrhs, int doubled, double* R_d, int64_t ldr);

/* Far-field pattern F(d) = lim_{R->inf} R e^{-ikR} (E, H)(R d) of the field of
 * bielcc_field (no reference counterpart: the limit of K_RotRot_Far and
 * R_Rot_Colloc, kernel_lib.h:453-492 / :295-313, over the same cell nodes):
 *   E_inf = multE k^2/(4pi) (I - d d) P[jE] - ik/(4pi) d x P[jM],
 *   H_inf = multH k^2/(4pi) (I - d d) P[jM] + ik/(4pi) d x P[jE],
 *   P[j] = sum_c (sum_q w_q e^{-ik d.y_q}) j_c.
 * dirs_d [ndir][3] unit vectors; jE_d, jM_d [nrhs][N][3] (jM_d may be NULL);
 * Einf_d, Hinf_d [nrhs][ndir][3] complex (Hinf_d may be NULL).  Uses
 * np->n_start only. */
int bielcc_far_pattern(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im,
                       int nrhs, const double* jE_d, const double* jM_d, double multE_re, double multE_im,
                       double multH_re, double multH_im, int precision, int ndir, const double* dirs_d,
                       double* Einf_d, double* Hinf_d);

/* ------------------------------------------------------------ plans: several meshes per launch
 * No reference counterpart (the reference evaluates one mesh per call and has
 * no assembly loop).  A plan fixes its items -- mesh, wave number and the
 * caller's device buffers -- at creation, when it builds the tile list of its
 * persistent kernel, the near-pair lists and its workspace (plan-owned).
 * bielcc_plan_run then enqueues kernels and GEMMs only: no allocation, no
 * host synchronisation, no memset -- except in mixed field plans, whose FP32
 * error control ("mixed_fixup") runs after the launch or replay: per item it
 * reads a count back and sums the points that fail the error model again in
 * FP64 (growing the context's scratch on first use).  When one run recomputes
 * more than "mixed_probe_pct" percent of the plan's points, the plan sums in
 * FP64 from its next run on, bit for bit as an FP64 plan, and a captured plan
 * is captured again at that run (info key "fp64_sums").
 * bielcc_plan_capture runs the plan once
 * and records it into a HIP graph; later bielcc_plan_run calls replay it.
 * Between runs the caller may change the CONTENTS of the item buffers
 * (currents, incidences, points), not their addresses.  Tile shapes come from
 * the "plan_*" tunables at creation. */
typedef struct bielcc_plan_s* bielcc_plan;

typedef struct {
    int mesh_id;
    double k_re, k_im;
    double* Z_d;      /* 2N x 2N complex, column-major, leading dimension ld */
    int64_t ld;
} bielcc_assemble_item;

/* bielcc_assemble_ex (all rows) for every item: the far pairs of all items in
 * one persistent launch, then the exact near pairs per item (K only). */
int bielcc_plan_assemble(bielcc_ctx ctx, int nitems, const bielcc_assemble_item* items, int which,
                         const bielcc_num_param* np, int precision, bielcc_plan* out);

typedef struct {
    int mesh_id;
    double k_re, k_im;
    int nrhs;
    const double* jE_d;      /* [nrhs][N][3] complex electric currents */
    const double* inc_d;     /* DEVICE [nrhs][6] (kvec, e0) plane waves added to E, or NULL */
    double multE_re, multE_im;
    int M;
    const double* pts_d;     /* [M][3] */
    double* E_d;             /* [nrhs][M][3] complex */
    double* H_d;             /* [nrhs][M][3] complex, or NULL */
} bielcc_field_item;

/* bielcc_field_batch with jM = NULL for every item:
 *   E = multE sum_j K[jE_j] (+ E_inc),  H = sum_j R[jE_j].
 * Evaluated through the operator matrices WK, WR ([3M x 3N] per item): one
 * persistent launch writes the far-pair blocks of all items, the exact
 * kernel the near-pair blocks, and E, H are complex GEMMs over all current
 * sets.  Items whose matrices exceed the "plan_ws_mb" workspace are split
 * into point ranges and groups of launches. */
int bielcc_plan_field(bielcc_ctx ctx, int nitems, const bielcc_field_item* items, const bielcc_num_param* np,
                      int near_mode, int precision, bielcc_plan* out);
int bielcc_plan_run(bielcc_ctx ctx, bielcc_plan plan);
int bielcc_plan_capture(bielcc_ctx ctx, bielcc_plan plan);
int bielcc_plan_destroy(bielcc_ctx ctx, bielcc_plan plan);
/* keys: "tiles", "launches", "grid", "pairs", "near_pairs", "segments",
 * "workspace_bytes", "captured", "fp64_sums" */
int bielcc_plan_info(bielcc_ctx ctx, bielcc_plan plan, const char* key, double* value);

/* Launch-shape tunables (the autotuner sets these; defaults in brackets):
 * "field_wg" [2048] and "assemble_wg" [2048] = workgroups a tile launch aims
 * for (cells are split until the grid reaches it); "field_rchunk" [4] =
 * current sets per field pass (1, 2 or 4); "trsv" [1] = small solves use the
 * one-kernel-per-triangle solve (0: as larger ones): FP64 solves with at most
 * "trsv_max64" [1] right-hand sides, FP32 solves with at most "trsv_max32"
 * [16] (0..16), and, with "trsv_chunks32" [0] = 1, FP32 solves with more, in
 * chunks of 16;
 * "trsm_lib" [1] = the other solves use rocBLAS trsm per triangle (0: the
 * library's blocked TRSM + GEMM loop); "refine_ferr_exp" [0] = e > 0: the
 * refinement does not stop at tol alone: classic steps also require an
 * estimated forward error (cond estimated from the residual's contraction,
 * times the residual) below 10^-e, FGMRES requires the residual to have
 * reached the FP64 floor (below 1e-14, or well above what the last cycle's
 * recurrence predicted); "refine_mode" [1] and "refine_restart" [20]: see
 * bielcc_refine_solve; "trsv_spin"
 * [1048576] = polls per wait of the one-kernel solve before it gives up (a
 * solve whose wait gave up is redone on the blocked path; tests lower it);
 * "lu_fused_panel" [2] = one launch per LU panel column (1), two (0), or
 * one up to order 6,144 and two above (2); "lu_outer_nb" [0] = columns
 * per outer block of the two-level LU (64, 128, 256 or 512): the panels'
 * updates stay inside the block and the rest of the matrix gets one update of
 * that rank (64: one level; 0: one level below order 16,000, 512 from there;
 * for FP32 LUs with "lu32_gemm16" on, 512 from order 10,000);
 * "lu32_gemm16" [1] = the FP32 LU's rank-512 updates run on the FP16 matrix
 * cores with split operands (each FP32 operand as an FP16 high part plus a
 * 2^-11-scaled FP16 low part, power-of-two scaled; errors of FP32 size);
 * 2: FP16 high parts only (fast, FP16-sized errors; experiments); 0: the
 * complex FP32 GEMM; "lu_tournament" [0] = 1: LU panels by tournament
 * pivoting (blocks of 128 rows factored in registers, candidates merged up a
 * binary tree; no launch waits for another workgroup, pivots differ from
 * partial pivoting; 0: partial pivoting); "lu_look_ahead2" [0] = 1: in the
 * two-level LU the next outer block is factored on a side stream (tournament
 * panels) while the bulk of the rank-512 update runs;
 * "lu_trsm_lib" [2] = the factorisation's U12 triangles on the library's
 * lane-substitution kernel, one row per lane (1: rocBLAS trsm; 0: the
 * library's LDS sweep); "lu_look_ahead" [2] = with a one-level LU, the
 * next LU panel is factored on a side stream during the trailing update:
 * always (1), never (0), or for FP64 above order 20,000 (2); "lu_panel_lds" [1] =
 * LU panels of at most 32,768 rows (FP64) or 65,536 (FP32) run as one launch
 * that holds the panel in LDS with one grid barrier per column (0: the
 * per-column launches); an LU uses it if its tallest panel's workgroups fit,
 * beside those of the LUs running at the same time, in two per compute unit,
 * else it takes the per-column path (2: every LU uses it, tests only);
 * "lu_panel_wait_us"
 * [250000] = bound of each of its barrier waits: an expired wait aborts the
 * launch before it writes the matrix and one workgroup then factors the
 * panel (slow, correct; tests set 0 to force it); "blas_atomics" [0] =
 * rocBLAS may use kernels that sum partial results with atomics (1); by
 * default it may not, so factors and solves are the same bits on every run.
 * Plans: "plan_field_cells" [64] and "plan_asm_cells" [64] =
 * cells per tile, "plan_wg_per_cu" [4] = persistent workgroups per compute
 * unit, "plan_ws_mb" [24576] = operator-matrix workspace cap in MiB.
 * Mixed fields: "mixed_far32" [0] = the field driver's far sums run in FP32
 * for every mixed call (1); by default a mixed call with a real k, whose node
 * phases use the centre-phase series (|k| times the largest node distance
 * from a cell centre at most 0.5), runs with FP64 sums: there the FP64 node
 * loop takes fewer VALU instructions than the FP32 one, which issues at the
 * same rate, and the result is the FP64 field bit for bit (plans and the
 * far-field pattern keep their FP32 sums); "mixed_fixup" [1] = after the FP32 sums, the points whose
 * modelled FP32 error is not small against their scattered field are
 * recomputed with FP64 sums (0: never); "mixed_probe_pct" [25] = calls of at
 * least 4,096 points first run the FP32 sums and the error model on 1,024
 * evenly strided points, and when more than this percentage of them would be
 * recomputed, the whole call runs with FP64 sums (the FP64 field, bit for
 * bit; 0: no sample).
 * bielcc_get_tuning also reads "mixed_fixup_points" (points recomputed, for
 * the context), "mixed_probe_fp64" (mixed calls the sample sent to FP64
 * sums) and "mixed_auto_fp64" (mixed calls run with FP64 sums by the rule of
 * "mixed_far32"),
 * "n_cu" (compute units) and "lu_panel_aborts" (whole-panel launches that
 * aborted to one workgroup, for the context, since its first LU),
 * "lu_panel_late_admits" (LUs that missed the whole-panel budget at their
 * first panel and were admitted at a later one),
 * "trsv_expired" and "trsv_ticket_faults" (one-kernel triangular solves
 * redone on the blocked path because a wait expired, or because the ticket
 * counter was out of step), and "scratch_kb" (device bytes the context holds
 * in its grow-only work buffers - field partial sums, near lists, operator
 * matrices, LU and solve work, and the near-pair list each mesh keeps from its
 * last assembly for the next one of the same rows - in KiB, rounded up; a
 * caller that budgets device memory charges it).  Returns BIELCC_ERR_ARG for an unknown key or a
 * value out of range. */
int bielcc_set_tuning(bielcc_ctx ctx, const char* key, int value);
int bielcc_get_tuning(bielcc_ctx ctx, const char* key, int* value);

/* ------------------------------------------------------------ timing
 * With timing on, the drivers bracket their dominant kernels with HIP events
 * on the context's stream; bielcc_timing reads the accumulated milliseconds
 * and launch count of one of: "field_far", "assemble_far", "lu_factor",
 * "lu_solve", "near", "assemble_generic", "lu32_factor", "refine_solve",
 * "field_opmat", "plan_assemble", "plan_field" (a whole plan run) and, for
 * eager plan runs, "plan_assemble_tiles", "plan_field_tiles" (the persistent
 * kernels) and "plan_field_gemm" (the GEMMs). */
int bielcc_timing_enable(bielcc_ctx ctx, int on);
int bielcc_timing_reset(bielcc_ctx ctx);
int bielcc_timing(bielcc_ctx ctx, const char* name, double* total_ms, int* count);

#ifdef __cplusplus
}
#endif
#endif /* BIELCC_HIP_H */
