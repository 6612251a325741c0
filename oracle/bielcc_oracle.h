/*
 * CPU restatement of BIELcc's Stratton-Chu K / R / K0 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle: tests/, the smoke
 * check and bench.py's cpu_baseline leg call it as the checker.  No product
 * code links, imports or falls back to it.
 *
 * Conventions: complex numbers are interleaved (re, im) doubles; vectors are
 * [3]; cells are [nv][3] with nv = 4 (quad) or 3 (triangle); the wave number
 * is complex (k_re, k_im) - a real wave number is k_im = 0, which gives the
 * same values as the reference's KType = double instantiation.
 */
#ifndef BIELCC_ORACLE_H
#define BIELCC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors NumParam (electrodynamics/num_param.h:28-55), fields used by the path. */
typedef struct {
    double eps, rs, rs_seg, analytic_calc;
    int n_start, n_start_seg, p_max, p_max_seg;
} orc_num_param;

/* Mirrors KernelParam (integrators/kernel_param.h:25-57). */
typedef struct {
    double smoothR, smoothRMin;
    double k_re, k_im;
    double vec_dbl[3];
    double n[3];
    double current[6];
} orc_kernel_param;

/* Mirrors IntegralParam (integrators/integral_param.h:16-45). */
typedef struct {
    int idim, n_start, p_max;
    double eps;
} orc_integral_param;

/* out[20]: rkt[3] norm[3] tau1[3] tau2[3] area diam norm_func[3] center_mass[3] */
int orc_cellfill(int nv, const double* v, double* out);
int orc_kernel(int f, const double* x, const double* y, const orc_kernel_param* kp, double* res6);
int orc_integral_pnt(int f, int nv, const double* cell, const double* x, orc_kernel_param* kp,
                     const orc_integral_param* ip, double* res6);
int orc_integral_seg(int f, const double* a, const double* b, const double* x, orc_kernel_param* kp,
                     const orc_integral_param* ip, double* res6);
int orc_k0(int variant, int f, int nv, const double* cell, const double* x, const double* j6,
           orc_kernel_param* kp, const orc_integral_param* ip, double* res6);
int orc_k0_hs(int nv, const double* cell, const double* x, const double* j6,
              const orc_integral_param* ip, double analyticR, double kre, double kim, double* res6);
double orc_integral1divr(int nv, const double* cell, const double* x);
void orc_integral_grad1divr_seg(const double* a, const double* b, const double* x, double* res3);
int orc_surf_hs(int nv, const double* cell, const double* x, double analyticR,
                const orc_integral_param* ip, double kre, double kim, double* res2);
int orc_seg_hs(const double* a, const double* b, const double* x, double analyticR,
               const orc_integral_param* ip, double kre, double kim, double* res6);
int orc_op(int op, int nv, const double* cell, const double* x, const double* j6,
           const orc_num_param* np, double kre, double kim, double* res6);
int orc_op_pairs(int op, int nv, int ncell, const double* verts, const double* j6,
                 int npts, const double* pts, const orc_num_param* np, double kre, double kim,
                 double* out);
int orc_get_field_e(int nv, int ncell, const double* verts, const double* jE,
                    double mre, double mim, const double* kvec, const double* e0,
                    double kre, double kim, const orc_num_param* np, int npts,
                    const double* pts, double* out);
int orc_field_eh(int nv, int ncell, const double* verts, const double* jE, const double* jM,
                 double multE_re, double multE_im, double multH_re, double multH_im,
                 int add_incident, const double* kvec, const double* e0,
                 double kre, double kim, const orc_num_param* np, int near_mode,
                 int npts, const double* pts, double* E, double* H);
int orc_assemble(int which, int nv, int ncell, const double* verts, const orc_num_param* np,
                 double kre, double kim, int row0, int nrows, double* Z, long long ld);
int orc_rhs_planewave(int nv, int ncell, const double* verts, const double* kvec,
                      const double* e0, double* rhs);
int orc_currents_by_basis(int nv, int ncell, const double* verts, const double* b, double* j);
int orc_plane_wave_e(int npts, const double* pts, const double* kvec, const double* e0, double* out);
int orc_plane_wave_h(int npts, const double* pts, const double* kvec, const double* e0, double freq,
                     double mu_re, double mu_im, double* out);
int orc_integral_y(int f, int nv, const double* cell, const double* e0, orc_kernel_param* kp,
                   const orc_integral_param* ip, double* res6);
int orc_lu_solve(int n, double* A, long long lda, double* B, int nrhs, long long ldb);
int orc_far_pattern(int nv, int ncell, const double* verts, const double* jE, const double* jM,
                    double multE_re, double multE_im, double multH_re, double multH_im,
                    double kre, double kim, int n_start, int ndir, const double* dirs, double* Einf,
                    double* Hinf);

#ifdef __cplusplus
}
#endif
#endif
