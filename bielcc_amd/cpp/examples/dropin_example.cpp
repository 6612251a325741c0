// Reference-style program on the drop-in C++ layer: the same calls a BIELcc
// user makes (examples/ED_Operators/main0_K.cpp, main0_R.cpp,
// examples/integrators/main0_integrals.cpp, electrodynamics/get_field.h),
// now executed on the GPU.  Prints "key values..." lines for the tests.
//
// usage: dropin_example <sphere_quad.dat>
#include "bielcc_hip.hpp"

#include <cstdio>
#include <iostream>

using namespace bielcc;
using cd = std::complex<double>;

static void print(const char* key, const cd* v, int n) {
    std::printf("%s", key);
    for (int i = 0; i < n; i++) std::printf(" %.17g %.17g", v[i].real(), v[i].imag());
    std::printf("\n");
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <mesh.dat>\n", argv[0]);
        return 2;
    }
    const double rut1[4][3] = {{0., 0., 0.}, {2., 0., 0.}, {3., 2., 0.}, {0., 2., 0.}};
    const double rut2[3][3] = {{0., 2., 0.}, {0., 0., 0.}, {2., 0., 0.}};
    const double x[3] = {2., 1.5, 0.};
    const double xf[3] = {10., 10., 10.};
    NumParam np(Calculation_Constants::INTEGRAL_ACCURACY, Calculation_Constants::SMOOTHING_DIST_SURF_INT,
                Calculation_Constants::SMOOTHING_DIST_SEG_INT, Calculation_Constants::ANALYTIC_CALC_DIST,
                Calculation_Constants::START_CELL_SPLIT, Calculation_Constants::START_SEG_SPLIT,
                Calculation_Constants::PMAX_CELL_SPLIT, Calculation_Constants::PMAX_SEG_SPLIT);
    const cd k(1., 0.);
    const cd j[3] = {1., 0., 0.};
    cd res[3];

    K_RotRot_Near_Smooth(j, x, rut1, np, k, res); print("K_RotRot_Near_Smooth", res, 3);
    K_RotRot_HS(j, x, rut1, np, k, res); print("K_RotRot_HS", res, 3);
    K_RotRot_HS_Colloc(j, x, rut1, np, k, res); print("K_RotRot_HS_Colloc", res, 3);
    K_RotRot_Far(j, xf, rut1, np, k, res); print("K_RotRot_Far", res, 3);
    R_Rot_Smooth(j, x, rut2, np, k, res); print("R_Rot_Smooth", res, 3);
    R_Rot_Colloc(j, xf, rut2, np, k, res); print("R_Rot_Colloc", res, 3);

    // integrator with a caller-visible smoothing radius (KernelParam& is mutated)
    KernelParam<double> kp(1.0, 0.5, 1e-6);
    IntegralParam ip(1, 8, 3, 1e-6);
    cd s[1];
    IntegralUniversalPnt(x, rut1, F_SimplePot_H, kp, ip, s);
    print("IntegralUniversalPnt", s, 1);
    std::printf("smoothR_after %.17g\n", kp.smoothR);
    std::printf("integral1Divr %.17g\n", integral1Divr(rut1, x));

    // F(y) forms: the area test of main0_integrals.cpp (rut1 = 5) and a plane wave over a cell
    KernelParam<double> kp0;
    kp0.smoothR = kp0.smoothRMin = 1e-16;
    double area[1];
    IntegralUniversal(rut1, FY_Unit<double>, kp0, IntegralParam(1, 8, 1, 1e-6), area);
    std::printf("area_rut1 %.17g\n", area[0]);
    KernelParam<double> pw;
    pw.smoothR = pw.smoothRMin = 1e-16;
    pw.vec_dbl[0] = 0.3; pw.vec_dbl[1] = -1.2; pw.vec_dbl[2] = 0.5;
    pw.e0[0] = 1.; pw.e0[1] = 0.5; pw.e0[2] = 0.;
    pw.freq = 0.2;
    cd fy[3];
    IntegralUniversal(rut1, PlaneWaveFieldE<double>, pw, IntegralParam(3, 8, 1, 1e-6), fy);
    print("IntegralUniversal_PlaneWaveFieldE", fy, 3);
    MaterialParam vac("vacuum", ChangeMaterialTypeDir::Isotropic, ChangeMaterialTypeFreq::NonDispersive, 1., 1., 0.,
                      0., 0.);
    vac.Calc(pw.freq);
    cd h[3];
    PlaneWaveFieldH(x, pw, vac, h);
    print("PlaneWaveFieldH", h, 3);
    // a reference-style loop of single-point calls: no device allocation per call
    cd e_sum[3] = {0., 0., 0.};
    for (int i = 0; i < 2000; i++) {
        const double xi[3] = {0.001 * i, 0.5, -0.25};
        cd e[3];
        PlaneWaveFieldE(xi, pw, e);
        for (int g = 0; g < 3; g++) e_sum[g] += e[g];
    }
    print("PlaneWaveFieldE_sum2000", e_sum, 3);

    // whole problem: mesh file -> currents -> GetFieldE_Ideal
    SurfMesh<CellQuad> mesh(argv[1]);
    std::printf("cells %d\n", mesh.GetNFrm());
    std::printf("grid_step %.17g\n", mesh.GetGridStep());
    const double kw = 2.0, kvec[3] = {0., 0., kw}, e0[3] = {1., 0., 0.};
    auto b = SolveCurrents_PlaneWave(mesh, np, kw, kvec, e0);
    std::vector<std::array<cd, 3>> jE;
    SurfaceCurrents_byBasis(mesh, b.data(), jE);
    std::vector<std::array<double, 3>> pts;
    for (int i = 0; i < 16; i++) {
        double z = -0.9 + 1.8 * i / 15.0, ph = 2.399963229728653 * i, rr = std::sqrt(1 - z * z);
        pts.push_back({3.0 * rr * std::cos(ph), 3.0 * rr * std::sin(ph), 3.0 * z});
    }
    KernelParam<double> einc;
    einc.k = kw;
    for (int d = 0; d < 3; d++) { einc.vec_dbl[d] = kvec[d]; einc.e0[d] = e0[d]; }
    std::vector<std::array<cd, 3>> E(pts.size());
    GetFieldE_Ideal(mesh, jE, cd(1., 0.), einc, PlaneWaveFieldE<double>, np, pts, E);
    for (size_t i = 0; i < pts.size(); i++) {
        std::printf("point %.17g %.17g %.17g\n", pts[i][0], pts[i][1], pts[i][2]);
        print("E", E[i].data(), 3);
    }
    return 0;
}
