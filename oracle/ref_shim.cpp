// extern "C" wrappers that instantiate the UNMODIFIED reference templates
// (BIELcc, compiled straight from its own sources by oracle/Makefile into
// oracle/_ref/).  This file is test infrastructure: it exists only so that
// tests/golden/make_golden.py can record input/output vectors of the reference
// and so that the CPU baseline can time the reference itself.  No product
// code links it.
//
// Complex numbers cross this boundary as interleaved (re, im) doubles.
// `kt` selects the reference's KType: 0 -> double (real wave number, k_im
// ignored), 1 -> std::complex<double>.

#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <array>
#include <unistd.h>

#include "element_geom.h"
#include "cell.h"
#include "surf_mesh.h"
#include "io_vtk.h"
#include "io_aereco.h"
#include "kernel_param.h"
#include "integral_param.h"
#include "integral_universal.h"
#include "integral_universal_seg.h"
#include "kernel_lib.h"
#include "integrals_analytic.h"
#include "K0.h"
#include "K.h"
#include "R.h"
#include "get_field.h"
#include "currents.h"
#include "incident_field.h"
#include "material.h"
#include "num_param.h"
#include "rwg_surf_mesh.h"

using namespace bielcc;
using cd = std::complex<double>;

extern "C" {
typedef struct {
    double eps, rs, rs_seg, analytic_calc;
    int n_start, n_start_seg, p_max, p_max_seg;
} shim_num_param;

typedef struct {
    double smoothR, smoothRMin;
    double k_re, k_im;
    double vec_dbl[3];
    double n[3];
    double current[6];
} shim_kernel_param;

typedef struct {
    int idim, n_start, p_max;
    double eps;
} shim_integral_param;
}

namespace {

template <typename KType> KType mk_k(double re, double im);
template <> double mk_k<double>(double re, double) { return re; }
template <> cd mk_k<cd>(double re, double im) { return cd(re, im); }

template <typename KType>
void to_kp(const shim_kernel_param* s, KernelParam<KType>& kp) {
    kp.smoothR = s->smoothR;
    kp.smoothRMin = s->smoothRMin;
    kp.k = mk_k<KType>(s->k_re, s->k_im);
    for (int i = 0; i < 3; i++) {
        kp.vec_dbl[i] = s->vec_dbl[i];
        kp.n[i] = s->n[i];
        kp.current[i] = cd(s->current[2 * i], s->current[2 * i + 1]);
        kp.vec_cmplx[i] = cd(0., 0.);
        kp.ort[i] = 0.;
        kp.e0[i] = 0.;
    }
    kp.freq = 0.;
}

NumParam to_np(const shim_num_param* s) {
    return NumParam(s->eps, s->rs, s->rs_seg, s->analytic_calc, s->n_start,
                    s->n_start_seg, s->p_max, s->p_max_seg);
}

// Kernel ids shared with the oracle and the HIP library.
enum {
    F_SIMPLEPOT_L = 0, F_GRADSIMPLEPOT_L, F_DOUBLEPOT_L, F_VECTORPOT_L,
    F_SIMPLEPOT_H, NUF_SIMPLEPOT_H, F_SIMPLEPOT_HM1, F_GRADSIMPLEPOT_H,
    F_GRADSIMPLEPOT_HM1, F_DOUBLEPOT_H, NF_DOUBLEPOT_H, F_ROTROTSIMPLEPOT_H,
    F_COUNT
};

bool is_real_kernel(int f) { return f <= F_VECTORPOT_L; }

template <typename KType>
using RealF = void (*)(const double*, const double*, const KernelParam<KType>&, double*);
template <typename KType>
using CplxF = void (*)(const double*, const double*, const KernelParam<KType>&, cd*);

template <typename KType> RealF<KType> real_f(int f) {
    switch (f) {
    case F_SIMPLEPOT_L: return F_SimplePot_L<KType>;
    case F_GRADSIMPLEPOT_L: return F_GradSimplePot_L<KType>;
    case F_DOUBLEPOT_L: return F_DoublePot_L<KType>;
    case F_VECTORPOT_L: return F_VectorPot_L<KType>;
    }
    return nullptr;
}
template <typename KType> CplxF<KType> cplx_f(int f) {
    switch (f) {
    case F_SIMPLEPOT_H: return F_SimplePot_H<KType>;
    case NUF_SIMPLEPOT_H: return nuF_SimplePot_H<KType>;
    case F_SIMPLEPOT_HM1: return F_SimplePot_Hm1<KType>;
    case F_GRADSIMPLEPOT_H: return F_GradSimplePot_H<KType>;
    case F_GRADSIMPLEPOT_HM1: return F_GradSimplePot_Hm1<KType>;
    case F_DOUBLEPOT_H: return F_DoublePot_H<KType>;
    case NF_DOUBLEPOT_H: return nF_DoublePot_H<KType>;
    case F_ROTROTSIMPLEPOT_H: return F_RotRotSimplePot_H<KType>;
    }
    return nullptr;
}

template <size_t NV> void load_cell(const double* v, double (&c)[NV][3]) {
    for (size_t i = 0; i < NV; i++)
        for (int d = 0; d < 3; d++) c[i][d] = v[3 * i + d];
}

void put(const cd* r, int n, double* out) {
    for (int i = 0; i < n; i++) { out[2 * i] = r[i].real(); out[2 * i + 1] = r[i].imag(); }
}
void put(const double* r, int n, double* out) {
    for (int i = 0; i < n; i++) { out[2 * i] = r[i]; out[2 * i + 1] = 0.; }
}

template <typename KType>
int kernel_eval(int f, const double* x, const double* y, const shim_kernel_param* s, double* res) {
    KernelParam<KType> kp; to_kp(s, kp);
    if (is_real_kernel(f)) {
        double r[3] = {0, 0, 0}; real_f<KType>(f)(x, y, kp, r); put(r, 3, res);
    } else {
        cd r[3] = {}; cplx_f<KType>(f)(x, y, kp, r); put(r, 3, res);
    }
    return 0;
}

template <typename KType, size_t NV>
int integral_pnt(int f, const double* cellv, const double* x, shim_kernel_param* s,
                 const shim_integral_param* ip, double* res) {
    double cell[NV][3]; load_cell(cellv, cell);
    KernelParam<KType> kp; to_kp(s, kp);
    IntegralParam ipr(ip->idim, ip->n_start, ip->p_max, ip->eps);
    if (is_real_kernel(f)) {
        double r[3] = {0, 0, 0};
        IntegralUniversalPnt(x, cell, real_f<KType>(f), kp, ipr, r); put(r, 3, res);
    } else {
        cd r[3] = {};
        IntegralUniversalPnt(x, cell, cplx_f<KType>(f), kp, ipr, r); put(r, 3, res);
    }
    s->smoothR = kp.smoothR;
    return 0;
}

template <typename KType>
int integral_seg(int f, const double* a, const double* b, const double* x, shim_kernel_param* s,
                 const shim_integral_param* ip, double* res) {
    KernelParam<KType> kp; to_kp(s, kp);
    IntegralParam ipr(ip->idim, ip->n_start, ip->p_max, ip->eps);
    if (is_real_kernel(f)) {
        double r[3] = {0, 0, 0};
        IntegralUniversalSegPnt(a, b, x, real_f<KType>(f), kp, ipr, r); put(r, 3, res);
    } else {
        cd r[3] = {};
        IntegralUniversalSegPnt(a, b, x, cplx_f<KType>(f), kp, ipr, r); put(r, 3, res);
    }
    s->smoothR = kp.smoothR;
    return 0;
}

template <typename KType, size_t NV>
int k0(int variant, int f, const double* cellv, const double* x, const double* j,
       shim_kernel_param* s, const shim_integral_param* ip, double* res) {
    double cell[NV][3]; load_cell(cellv, cell);
    KernelParam<KType> kp; to_kp(s, kp);
    IntegralParam ipr(ip->idim, ip->n_start, ip->p_max, ip->eps);
    if (is_real_kernel(f)) {
        double jr[3] = {j[0], j[2], j[4]};
        double r[3] = {0, 0, 0};
        if (variant == 0) K0(jr, x, cell, real_f<KType>(f), kp, ipr, r);
        else K0_Common(x, cell, real_f<KType>(f), kp, ipr, r);
        put(r, 3, res);
    } else {
        cd jc[3] = {cd(j[0], j[1]), cd(j[2], j[3]), cd(j[4], j[5])};
        cd r[3] = {};
        if (variant == 0) K0(jc, x, cell, cplx_f<KType>(f), kp, ipr, r);
        else K0_Common(x, cell, cplx_f<KType>(f), kp, ipr, r);
        put(r, 3, res);
    }
    s->smoothR = kp.smoothR;
    return 0;
}

template <typename KType, size_t NV>
int k0_hs(const double* cellv, const double* x, const double* j, const shim_integral_param* ip,
          double analyticR, double kre, double kim, double* res) {
    double cell[NV][3]; load_cell(cellv, cell);
    IntegralParam ipr(ip->idim, ip->n_start, ip->p_max, ip->eps);
    cd jc[3] = {cd(j[0], j[1]), cd(j[2], j[3]), cd(j[4], j[5])};
    cd r[3] = {};
    K0_GradSimplePot_H_HS(jc, x, cell, ipr, analyticR, mk_k<KType>(kre, kim), r);
    put(r, 3, res);
    return 0;
}

template <typename KType, size_t NV>
int surf_hs(const double* cellv, const double* x, double analyticR, const shim_integral_param* ip,
            double kre, double kim, double* res) {
    double cell[NV][3]; load_cell(cellv, cell);
    IntegralParam ipr(ip->idim, ip->n_start, ip->p_max, ip->eps);
    cd r = IntegralSurf_SimplePot_H_HS(cell, x, analyticR, ipr, mk_k<KType>(kre, kim));
    put(&r, 1, res);
    return 0;
}

template <typename KType>
int seg_hs(const double* a, const double* b, const double* x, double analyticR,
           const shim_integral_param* ip, double kre, double kim, double* res) {
    IntegralParam ipr(ip->idim, ip->n_start, ip->p_max, ip->eps);
    cd r[3] = {};
    IntegralSeg_GradSimplePot_H_HS(a, b, x, analyticR, ipr, mk_k<KType>(kre, kim), r);
    put(r, 3, res);
    return 0;
}

template <typename KType, size_t NV>
int op_eval(int op, const double* cellv, const double* x, const double* j,
            const shim_num_param* snp, double kre, double kim, double* res) {
    double cell[NV][3]; load_cell(cellv, cell);
    NumParam np = to_np(snp);
    KType k = mk_k<KType>(kre, kim);
    cd jc[3] = {cd(j[0], j[1]), cd(j[2], j[3]), cd(j[4], j[5])};
    cd r[3] = {};
    switch (op) {
    case 0: K_RotRot_Near_Smooth(jc, x, cell, np, k, r); break;
    case 1: K_RotRot_HS(jc, x, cell, np, k, r); break;
    case 2: K_RotRot_HS_Colloc(jc, x, cell, np, k, r); break;
    case 3: K_RotRot_HS2_Colloc(jc, x, cell, np, k, r); break;
    case 4: K_RotRot_Far(jc, x, cell, np, k, r); break;
    case 5: R_Rot_Smooth(jc, x, cell, np, k, r); break;
    case 6: R_Rot_Colloc(jc, x, cell, np, k, r); break;
    case 7: R_Rot_Near(jc, x, cell, np, k, r); break;
    default: return -1;
    }
    put(r, 3, res);
    return 0;
}

// Writes cells in the reference's .dat layout (mesh/surf_mesh.cpp:102-175)
// so that the reference's own reader builds the SurfMesh.
bool write_dat(const std::string& path, int nv, int ncell, const double* verts) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "1\n1\n1 %d 0 0\n", ncell);
    for (int c = 0; c < ncell; c++) {
        for (int v = 0; v < nv; v++) {
            const double* p = verts + (size_t)(c * nv + v) * 3;
            std::fprintf(f, "%.17g %.17g %.17g\n", p[0], p[1], p[2]);
        }
    }
    std::fprintf(f, "0\n");
    std::fclose(f);
    return true;
}

template <typename CellT, typename KType>
int get_field_e(int ncell, int nv, const double* verts, const double* jE, double mre, double mim,
                const double* kvec, const double* e0, double kre, double kim,
                const shim_num_param* snp, int npts, const double* pts, double* out,
                const char* tmpdir) {
    std::string path = std::string(tmpdir) + "/ref_mesh_" + std::to_string(getpid()) + ".dat";
    if (!write_dat(path, nv, ncell, verts)) return -2;
    SurfMesh<CellT> mesh(path);
    std::remove(path.c_str());
    if (mesh.GetNFrm() != ncell) return -3;
    std::vector<std::array<cd, 3>> j(ncell);
    for (int c = 0; c < ncell; c++)
        for (int d = 0; d < 3; d++) j[c][d] = cd(jE[6 * c + 2 * d], jE[6 * c + 2 * d + 1]);
    KernelParam<KType> einc;
    einc.k = mk_k<KType>(kre, kim);
    einc.smoothR = 0.; einc.smoothRMin = 0.; einc.freq = 0.;
    for (int d = 0; d < 3; d++) { einc.vec_dbl[d] = kvec[d]; einc.e0[d] = e0[d]; }
    std::vector<std::array<double, 3>> p(npts);
    for (int i = 0; i < npts; i++)
        for (int d = 0; d < 3; d++) p[i][d] = pts[3 * i + d];
    std::vector<std::array<cd, 3>> fe(npts);
    NumParam np = to_np(snp);
    GetFieldE_Ideal(mesh, j, cd(mre, mim), einc, PlaneWaveFieldE<KType>, np, p, fe);
    for (int i = 0; i < npts; i++)
        for (int d = 0; d < 3; d++) { out[6 * i + 2 * d] = fe[i][d].real(); out[6 * i + 2 * d + 1] = fe[i][d].imag(); }
    return 0;
}

template <typename CellT>
int mesh_tables(const std::string& path, int nv, int* ncell_out, double* verts, double* rkt,
                double* norm, double* tau, double* grid_step, int cap) {
    SurfMesh<CellT> mesh(path);
    int n = mesh.GetNFrm();
    *ncell_out = n;
    *grid_step = mesh.GetGridStep();
    if (n > cap) return 1;
    auto r = mesh.FormCollocPoints();
    auto nn = mesh.FormNormalVectors();
    auto t = mesh.FormSurfaceBasis();
    for (int c = 0; c < n; c++) {
        const auto& cc = mesh.GetCell(c).GetCellCoords();
        for (int v = 0; v < nv; v++)
            for (int d = 0; d < 3; d++) verts[(size_t)(c * nv + v) * 3 + d] = cc[v][d];
        for (int d = 0; d < 3; d++) {
            rkt[3 * c + d] = r[c][d];
            norm[3 * c + d] = nn[c][d];
            tau[6 * c + d] = t[c][0][d];
            tau[6 * c + 3 + d] = t[c][1][d];
        }
    }
    return 0;
}

}  // namespace

namespace {
// F(y) integrand of examples/integrators/main0_integrals.cpp:77-81 (the area test):
// restated, since the example defines it in its own translation unit
template <typename KType>
void unit_fy(const double*, const KernelParam<KType>&, double* ff) { ff[0] = 1.; }

template <typename KType>
int integral_y(int f, int nv, const double* cellv, const double* e0, shim_kernel_param* s,
               const shim_integral_param* ip, double* res) {
    KernelParam<KType> kp; to_kp(s, kp);
    for (int d = 0; d < 3; d++) kp.e0[d] = e0[d];
    IntegralParam ipr(ip->idim, ip->n_start, ip->p_max, ip->eps);
    if (f == 0) {
        double r[3] = {0, 0, 0};
        if (nv == 4) { double c[4][3]; load_cell(cellv, c); IntegralUniversal(c, unit_fy<KType>, kp, ipr, r); }
        else if (nv == 3) { double c[3][3]; load_cell(cellv, c); IntegralUniversal(c, unit_fy<KType>, kp, ipr, r); }
        else IntegralUniversalSegPnt(cellv, cellv + 3, unit_fy<KType>, kp, ipr, r);
        put(r, 3, res);
    } else {
        cd r[3] = {};
        if (nv == 4) { double c[4][3]; load_cell(cellv, c); IntegralUniversal(c, PlaneWaveFieldE<KType>, kp, ipr, r); }
        else if (nv == 3) { double c[3][3]; load_cell(cellv, c); IntegralUniversal(c, PlaneWaveFieldE<KType>, kp, ipr, r); }
        else IntegralUniversalSegPnt(cellv, cellv + 3, PlaneWaveFieldE<KType>, kp, ipr, r);
        put(r, 3, res);
    }
    s->smoothR = kp.smoothR;
    return 0;
}
}  // namespace

extern "C" {

int ref_cellfill(int nv, const double* v, double* out) {
    // out: rkt[3] norm[3] tau1[3] tau2[3] area diam norm_func[3] center_mass[3]
    if (nv == 4) {
        double c[4][3]; load_cell(v, c);
        CellQuad q; q.CellFill(c);
        for (int d = 0; d < 3; d++) {
            out[d] = q.GetRkt()[d]; out[3 + d] = q.GetNorm()[d];
            out[6 + d] = q.GetTau1()[d]; out[9 + d] = q.GetTau2()[d];
        }
        out[12] = q.GetArea(); out[13] = get_diam(c);
        norm_func(c, out + 14); get_center_mass(c, out + 17);
    } else if (nv == 3) {
        double c[3][3]; load_cell(v, c);
        CellTriangle t; t.CellFill(c);
        for (int d = 0; d < 3; d++) {
            out[d] = t.GetRkt()[d]; out[3 + d] = t.GetNorm()[d];
            out[6 + d] = t.GetTau1()[d]; out[9 + d] = t.GetTau2()[d];
        }
        out[12] = t.GetArea(); out[13] = get_diam(c);
        norm_func(c, out + 14); get_center_mass(c, out + 17);
    } else return -1;
    return 0;
}

int ref_kernel(int f, int kt, const double* x, const double* y, const shim_kernel_param* s, double* res) {
    if (f < 0 || f >= F_COUNT) return -1;
    return kt ? kernel_eval<cd>(f, x, y, s, res) : kernel_eval<double>(f, x, y, s, res);
}

int ref_integral_pnt(int f, int kt, int nv, const double* cell, const double* x,
                     shim_kernel_param* s, const shim_integral_param* ip, double* res) {
    if (f < 0 || f >= F_COUNT) return -1;
    if (nv == 4) return kt ? integral_pnt<cd, 4>(f, cell, x, s, ip, res) : integral_pnt<double, 4>(f, cell, x, s, ip, res);
    if (nv == 3) return kt ? integral_pnt<cd, 3>(f, cell, x, s, ip, res) : integral_pnt<double, 3>(f, cell, x, s, ip, res);
    return -1;
}

int ref_integral_seg(int f, int kt, const double* a, const double* b, const double* x,
                     shim_kernel_param* s, const shim_integral_param* ip, double* res) {
    if (f < 0 || f >= F_COUNT) return -1;
    return kt ? integral_seg<cd>(f, a, b, x, s, ip, res) : integral_seg<double>(f, a, b, x, s, ip, res);
}

int ref_k0(int variant, int f, int kt, int nv, const double* cell, const double* x, const double* j,
           shim_kernel_param* s, const shim_integral_param* ip, double* res) {
    if (f < 0 || f >= F_COUNT) return -1;
    if (nv == 4) return kt ? k0<cd, 4>(variant, f, cell, x, j, s, ip, res) : k0<double, 4>(variant, f, cell, x, j, s, ip, res);
    if (nv == 3) return kt ? k0<cd, 3>(variant, f, cell, x, j, s, ip, res) : k0<double, 3>(variant, f, cell, x, j, s, ip, res);
    return -1;
}

int ref_k0_hs(int kt, int nv, const double* cell, const double* x, const double* j,
              const shim_integral_param* ip, double analyticR, double kre, double kim, double* res) {
    if (nv == 4) return kt ? k0_hs<cd, 4>(cell, x, j, ip, analyticR, kre, kim, res) : k0_hs<double, 4>(cell, x, j, ip, analyticR, kre, kim, res);
    if (nv == 3) return kt ? k0_hs<cd, 3>(cell, x, j, ip, analyticR, kre, kim, res) : k0_hs<double, 3>(cell, x, j, ip, analyticR, kre, kim, res);
    return -1;
}

double ref_integral1divr(int nv, const double* cellv, const double* x) {
    if (nv == 4) { double c[4][3]; load_cell(cellv, c); return integral1Divr(c, x); }
    double c[3][3]; load_cell(cellv, c); return integral1Divr(c, x);
}

void ref_integral_grad1divr_seg(const double* a, const double* b, const double* x, double* res) {
    integralGrad1DivrSeg(a, b, x, res);
}

int ref_surf_hs(int kt, int nv, const double* cell, const double* x, double analyticR,
                const shim_integral_param* ip, double kre, double kim, double* res) {
    if (nv == 4) return kt ? surf_hs<cd, 4>(cell, x, analyticR, ip, kre, kim, res) : surf_hs<double, 4>(cell, x, analyticR, ip, kre, kim, res);
    if (nv == 3) return kt ? surf_hs<cd, 3>(cell, x, analyticR, ip, kre, kim, res) : surf_hs<double, 3>(cell, x, analyticR, ip, kre, kim, res);
    return -1;
}

int ref_seg_hs(int kt, const double* a, const double* b, const double* x, double analyticR,
               const shim_integral_param* ip, double kre, double kim, double* res) {
    return kt ? seg_hs<cd>(a, b, x, analyticR, ip, kre, kim, res) : seg_hs<double>(a, b, x, analyticR, ip, kre, kim, res);
}

int ref_op(int op, int kt, int nv, const double* cell, const double* x, const double* j,
           const shim_num_param* np, double kre, double kim, double* res) {
    if (nv == 4) return kt ? op_eval<cd, 4>(op, cell, x, j, np, kre, kim, res) : op_eval<double, 4>(op, cell, x, j, np, kre, kim, res);
    if (nv == 3) return kt ? op_eval<cd, 3>(op, cell, x, j, np, kre, kim, res) : op_eval<double, 3>(op, cell, x, j, np, kre, kim, res);
    return -1;
}

// Batched operator evaluation over all (point, cell) pairs with OpenMP over
// points, the reference's own parallel scheme (electrodynamics/get_field.h:49).
// out[(p * ncell + c) * 6 ...] = Op[cell c, j_c](x_p).
int ref_op_pairs(int op, int kt, int nv, int ncell, const double* verts, const double* j,
                 int npts, const double* pts, const shim_num_param* np, double kre, double kim,
                 double* out) {
    int bad = 0;
    #pragma omp parallel for schedule(static) reduction(+:bad)
    for (int p = 0; p < npts; p++)
        for (int c = 0; c < ncell; c++)
            bad += ref_op(op, kt, nv, verts + (size_t)c * nv * 3, pts + 3 * (size_t)p, j + 6 * (size_t)c,
                          np, kre, kim, out + ((size_t)p * ncell + c) * 6) != 0;
    return bad ? -1 : 0;
}

int ref_get_field_e(int kt, int nv, int ncell, const double* verts, const double* jE,
                    double mre, double mim, const double* kvec, const double* e0,
                    double kre, double kim, const shim_num_param* np, int npts,
                    const double* pts, double* out, const char* tmpdir) {
    if (nv == 4) return kt ? get_field_e<CellQuad, cd>(ncell, nv, verts, jE, mre, mim, kvec, e0, kre, kim, np, npts, pts, out, tmpdir)
                           : get_field_e<CellQuad, double>(ncell, nv, verts, jE, mre, mim, kvec, e0, kre, kim, np, npts, pts, out, tmpdir);
    if (nv == 3) return kt ? get_field_e<CellTriangle, cd>(ncell, nv, verts, jE, mre, mim, kvec, e0, kre, kim, np, npts, pts, out, tmpdir)
                           : get_field_e<CellTriangle, double>(ncell, nv, verts, jE, mre, mim, kvec, e0, kre, kim, np, npts, pts, out, tmpdir);
    return -1;
}

int ref_plane_wave_e(int npts, const double* pts, const double* kvec, const double* e0, double* out) {
    KernelParam<double> p;
    p.k = 0.; p.smoothR = 0.; p.smoothRMin = 0.; p.freq = 0.;
    for (int d = 0; d < 3; d++) { p.vec_dbl[d] = kvec[d]; p.e0[d] = e0[d]; }
    for (int i = 0; i < npts; i++) {
        cd e[3];
        PlaneWaveFieldE(pts + 3 * i, p, e);
        put(e, 3, out + 6 * i);
    }
    return 0;
}

// PlaneWaveFieldH through a MaterialParam of relative permeability mu_r at freq
// (MaterialParam::Calc: mu = mu0 mu_r, electrodynamics/material.cpp:59-115)
int ref_plane_wave_h(int npts, const double* pts, const double* kvec, const double* e0, double freq, double mu_r,
                     double* out, double* mu_out) {
    KernelParam<double> p;
    p.k = 0.; p.smoothR = 0.; p.smoothRMin = 0.; p.freq = freq;
    for (int d = 0; d < 3; d++) { p.vec_dbl[d] = kvec[d]; p.e0[d] = e0[d]; }
    MaterialParam mat("shim", ChangeMaterialTypeDir::Isotropic, ChangeMaterialTypeFreq::NonDispersive, 1., mu_r, 0.,
                      0., 0.);
    mat.Calc(freq);
    *mu_out = std::get<double>(mat.GetMu());
    for (int i = 0; i < npts; i++) {
        cd h[3];
        PlaneWaveFieldH(pts + 3 * i, p, mat, h);
        put(h, 3, out + 6 * i);
    }
    return 0;
}

int ref_integral_y(int f, int kt, int nv, const double* cell, const double* e0, shim_kernel_param* s,
                   const shim_integral_param* ip, double* res) {
    if (f < 0 || f > 1 || nv < 2 || nv > 4) return -1;
    return kt ? integral_y<cd>(f, nv, cell, e0, s, ip, res) : integral_y<double>(f, nv, cell, e0, s, ip, res);
}

int ref_currents_by_basis(int ncell, const double* tau, const double* b, double* j) {
    std::vector<std::array<std::array<double, 3>, 2>> t(ncell);
    for (int c = 0; c < ncell; c++)
        for (int d = 0; d < 3; d++) { t[c][0][d] = tau[6 * c + d]; t[c][1][d] = tau[6 * c + 3 + d]; }
    std::vector<cd> bb(2 * (size_t)ncell);
    for (size_t i = 0; i < bb.size(); i++) bb[i] = cd(b[2 * i], b[2 * i + 1]);
    std::vector<std::array<cd, 3>> jv(ncell);
    SurfaceCurrents_byBasis(t, bb.data(), jv);
    for (int c = 0; c < ncell; c++) put(jv[c].data(), 3, j + 6 * c);
    return 0;
}

// Reads a mesh file (.dat or .vtk) with the reference's reader and returns its
// vertices and the per-cell tables it derives.
int ref_read_mesh(const char* path, int nv, int cap, int* ncell, double* verts, double* rkt,
                  double* norm, double* tau, double* grid_step) {
    try {
        if (nv == 4) return mesh_tables<CellQuad>(path, 4, ncell, verts, rkt, norm, tau, grid_step, cap);
        if (nv == 3) return mesh_tables<CellTriangle>(path, 3, ncell, verts, rkt, norm, tau, grid_step, cap);
    } catch (...) {
        return -2;
    }
    return -1;
}


// The reference's writers, for golden output files: points + complex vectors
// (IO_VTK::WritePntsWithField_C, io/io_vtk.cpp:482-551), complex vectors as
// two aereco files (IO_AERECO::WriteVecs_C, io/io_aereco.cpp:98-128) and the
// complex field as three aereco files (IO_AERECO::WriteField_C, :135-173).
static std::vector<std::array<cd, 3>> load_cvecs(int n, const double* v) {
    std::vector<std::array<cd, 3>> out(n);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) out[i][d] = cd(v[6 * i + 2 * d], v[6 * i + 2 * d + 1]);
    return out;
}

int ref_write_vtk_field_c(const char* path, int n, const double* pts, const double* vals) {
    std::vector<std::array<double, 3>> p(n);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) p[i][d] = pts[3 * i + d];
    try {
        IO_VTK io;
        io.WritePntsWithField_C(path, p, load_cvecs(n, vals));
    } catch (...) {
        return -2;
    }
    return 0;
}

int ref_write_aereco_vecs_c(const char* path_re, const char* path_im, int n, const double* vals) {
    try {
        IO_AERECO io;
        io.WriteVecs_C(path_re, path_im, load_cvecs(n, vals));
    } catch (...) {
        return -2;
    }
    return 0;
}

int ref_write_aereco_field_c(const char* path_re, const char* path_im, const char* path_abs, int n,
                             const double* vals) {
    // the reference writes (n/2)^2 rows: only call it where that stays inside the n values
    if ((long long)(n / 2) * (n / 2) > n) return -1;
    try {
        IO_AERECO io;
        io.WriteField_C(path_re, path_im, path_abs, load_cvecs(n, vals));
    } catch (...) {
        return -2;
    }
    return 0;
}

}  // extern "C"

// RwgSurfMesh (mesh/rwg_surf_mesh.h:300-526) of a mesh file: the topology
// arrays, flattened, and the reference's own VTK export of it.  counts[0..2] =
// points, edges, neighbour entries; every array is caller-sized (cap entries).
template <typename CellT>
static int rwg_dump(const char* path, int cap, int* ncell, int* counts, double* points, int* cell_points,
                    int* point_num_cells, int* point_num_edges, int* edges, int* edge_num_cells, int* cell_edges,
                    int* cell_dir, double* edge_mid, double* edge_ort, int* nghb_off, int* nghb, const char* vtk) {
    SurfMesh<CellT> base(path);
    RwgSurfMesh<CellT> m(base);
    const int n = m.GetNFrm(), np_ = m.GetNPoints(), ne = m.GetNEdges();
    constexpr int nv = std::is_same<CellT, CellQuad>::value ? 4 : 3;
    *ncell = n;
    counts[0] = np_;
    counts[1] = ne;
    if (n * nv > cap || np_ > cap || ne > cap) return 1;
    for (int i = 0; i < np_; i++) {
        for (int d = 0; d < 3; d++) points[3 * i + d] = m.GetPoint(i)[d];
        point_num_cells[i] = m.GetPointNumCells(i);
        point_num_edges[i] = (int)m.GetPointEdges(i).size();
    }
    for (int c = 0; c < n; c++)
        for (int v = 0; v < nv; v++) {
            cell_points[c * nv + v] = m.GetCellPoints(c)[v];
            cell_edges[c * nv + v] = m.GetCellEdge(c, v);
            cell_dir[c * nv + v] = m.GetCellEdgeDirIndx(c, v);
        }
    for (int e = 0; e < ne; e++) {
        edges[2 * e] = (int)m.GetEdge(e)[0];
        edges[2 * e + 1] = (int)m.GetEdge(e)[1];
        edge_num_cells[e] = m.GetEdgeNumCells(e);
        for (int d = 0; d < 3; d++) {
            edge_mid[3 * e + d] = m.GetEdgeMid(e)[d];
            edge_ort[3 * e + d] = m.GetEdgeOrt(e)[d];
        }
    }
    int k = 0;
    nghb_off[0] = 0;
    for (int c = 0; c < n; c++)
        for (int v = 0; v < nv; v++) {
            for (int x : m.GetCellNghbsByEdge(c, v)) {
                if (k >= 4 * cap) return 1;
                nghb[k++] = x;
            }
            nghb_off[c * nv + v + 1] = k;
        }
    counts[2] = k;
    if (vtk && vtk[0]) m.ExportRwgfMeshToVTK(vtk);
    return 0;
}

extern "C" {

int ref_rwg(const char* path, int nv, int cap, int* ncell, int* counts, double* points, int* cell_points,
            int* point_num_cells, int* point_num_edges, int* edges, int* edge_num_cells, int* cell_edges,
            int* cell_dir, double* edge_mid, double* edge_ort, int* nghb_off, int* nghb, const char* vtk) {
    try {
        if (nv == 4)
            return rwg_dump<CellQuad>(path, cap, ncell, counts, points, cell_points, point_num_cells, point_num_edges,
                                      edges, edge_num_cells, cell_edges, cell_dir, edge_mid, edge_ort, nghb_off, nghb,
                                      vtk);
        if (nv == 3)
            return rwg_dump<CellTriangle>(path, cap, ncell, counts, points, cell_points, point_num_cells,
                                          point_num_edges, edges, edge_num_cells, cell_edges, cell_dir, edge_mid,
                                          edge_ort, nghb_off, nghb, vtk);
    } catch (...) {
        return -2;
    }
    return -1;
}

// Ker_BasicAverFunc (electrodynamics/kernel_lib.h:626-634)
double ref_basic_aver(const double* x, const double* y, double eps) { return Ker_BasicAverFunc(x, y, eps); }

}  // extern "C"
