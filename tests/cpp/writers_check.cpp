// Writes the reference-format files with the drop-in header's IO_VTK /
// IO_AERECO from values read as text: n, then n lines "px py pz" and n lines
// "re0 im0 re1 im1 re2 im2" (%.17g).  usage: writers_check <in.txt> <outdir>
#include "bielcc_hip.hpp"

#include <cstdio>

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    std::ifstream in(argv[1]);
    int n = 0;
    in >> n;
    std::vector<std::array<double, 3>> pts(n);
    std::vector<std::array<std::complex<double>, 3>> v(n);
    for (int i = 0; i < n; i++) in >> pts[i][0] >> pts[i][1] >> pts[i][2];
    for (int i = 0; i < n; i++)
        for (int g = 0; g < 3; g++) {
            double re, im;
            in >> re >> im;
            v[i][g] = {re, im};
        }
    const std::string d = argv[2];
    bielcc::IO_VTK().WritePntsWithField_C(d + "/field_c.vtk", pts, v);
    bielcc::IO_AERECO().WriteVecs_C(d + "/vecs_real.gv", d + "/vecs_image.gv", v);
    std::vector<std::array<std::complex<double>, 3>> v4(v.begin(), v.begin() + 4);
    bielcc::IO_AERECO().WriteField_C(d + "/field_real.gv", d + "/field_image.gv", d + "/field_abs.gdr", v4);
    return 0;
}
