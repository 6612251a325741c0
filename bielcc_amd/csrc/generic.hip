// Batched exact evaluation of single reference functions, one item per lane:
// operators (K.h, R.h), integrands (kernel_lib.h), the cell and segment
// integrators (integral_universal.h, integral_universal_seg.h), the contour
// operators (K0.h) and the semi-analytic integrals (integrals_analytic.*).
// These back the per-call drop-in API and the per-function parity tests.
#include "bielcc_internal.h"

using namespace bielcc_dev;

namespace bielcc_impl {

template <int NV>
__device__ inline void load_cell(const double* cells, long long i, double (*cv)[3]) {
    for (int v = 0; v < NV; v++)
        for (int d = 0; d < 3; d++) cv[v][d] = cells[((size_t)i * NV + v) * 3 + d];
}

__device__ inline void store3(const cplx* r, int n, double* o) {
    for (int g = 0; g < n; g++) { o[2 * g] = r[g].re; o[2 * g + 1] = r[g].im; }
}

template <int NV>
__global__ void k_op_batch(int op, int n, const double* __restrict__ cells, const double* __restrict__ x,
                           const double* __restrict__ j, NumP np, double k_re, double k_im,
                           double* __restrict__ res) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double cv[NV][3];
    load_cell<NV>(cells, i, cv);
    cplx jj[3], r[3];
    for (int d = 0; d < 3; d++) jj[d] = mk(j[6 * i + 2 * d], j[6 * i + 2 * d + 1]);
    op_eval<NV>(op, cv, x + 3 * i, jj, np, mk(k_re, k_im), r);
    store3(r, 3, res + 6 * i);
}

hipError_t launch_op_batch(hipStream_t s, int op, int nv, int n, const double* cells, const double* x,
                           const double* j, const NumP& np, double k_re, double k_im, double* res) {
    if (n <= 0) return hipSuccess;
    dim3 g((n + 63) / 64), b(64);
    if (nv == 4) hipLaunchKernelGGL(k_op_batch<4>, g, b, 0, s, op, n, cells, x, j, np, k_re, k_im, res);
    else hipLaunchKernelGGL(k_op_batch<3>, g, b, 0, s, op, n, cells, x, j, np, k_re, k_im, res);
    return hipGetLastError();
}

// Run `body` with F bound to the compile-time kernel id f.
#define BIELCC_KERNEL_SWITCH(f, body)                               \
    switch (f) {                                                    \
    case F_SIMPLEPOT_L: { constexpr int F = F_SIMPLEPOT_L; body; } break;             \
    case F_GRADSIMPLEPOT_L: { constexpr int F = F_GRADSIMPLEPOT_L; body; } break;     \
    case F_DOUBLEPOT_L: { constexpr int F = F_DOUBLEPOT_L; body; } break;             \
    case F_VECTORPOT_L: { constexpr int F = F_VECTORPOT_L; body; } break;             \
    case F_SIMPLEPOT_H: { constexpr int F = F_SIMPLEPOT_H; body; } break;             \
    case NUF_SIMPLEPOT_H: { constexpr int F = NUF_SIMPLEPOT_H; body; } break;         \
    case F_SIMPLEPOT_HM1: { constexpr int F = F_SIMPLEPOT_HM1; body; } break;         \
    case F_GRADSIMPLEPOT_H: { constexpr int F = F_GRADSIMPLEPOT_H; body; } break;     \
    case F_GRADSIMPLEPOT_HM1: { constexpr int F = F_GRADSIMPLEPOT_HM1; body; } break; \
    case F_DOUBLEPOT_H: { constexpr int F = F_DOUBLEPOT_H; body; } break;             \
    case NF_DOUBLEPOT_H: { constexpr int F = NF_DOUBLEPOT_H; body; } break;           \
    case F_ROTROTSIMPLEPOT_H: { constexpr int F = F_ROTROTSIMPLEPOT_H; body; } break; \
    default: break;                                                 \
    }

__global__ void k_kernel_batch(int f, int n, const double* __restrict__ x, const double* __restrict__ y, KP kp,
                               double* __restrict__ res) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cplx r[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    BIELCC_KERNEL_SWITCH(f, (kernel_eval<F>(x + 3 * i, y + 3 * i, kp, r)));
    store3(r, 3, res + 6 * i);
}

hipError_t launch_kernel_batch(hipStream_t s, int f, int n, const double* x, const double* y, const KP& kp,
                               double* res) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kernel_batch, dim3((n + 63) / 64), dim3(64), 0, s, f, n, x, y, kp, res);
    return hipGetLastError();
}

// Averaging functions (electrodynamics/kernel_lib.h:560-634), rounded as the
// reference writes them (no contraction): which 0 Ker_AverDiv, 1
// Ker_AverDiv_NearEdge (eps = eps_edge), 2 Ker_BasicAverFunc (in res[i][0]).
#pragma clang fp contract(off)
__global__ void k_aver_batch(int which, int n, const double* __restrict__ x, const double* __restrict__ y, double eps,
                             double* __restrict__ res) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* a = x + 3 * i;
    const double* b = y + 3 * i;
    double* o = res + 3 * i;
    const double d = sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
    if (which == 2) {
        double v = 0.;
        if (!(d > eps)) {
            const double deg = d * d / eps / eps;
            v = exp(deg) / kPi / eps / eps;
        }
        o[0] = v; o[1] = 0.; o[2] = 0.;
        return;
    }
    const double reps = d / eps;
    const double deg = which == 0
        ? (-6.0 + 2. * reps * reps) * exp(-reps * reps) / kPi / (eps * eps) / (eps * eps)
        : (-20.0 + 8. * reps * reps) * exp(-reps * reps) / kPi / eps / eps / eps / eps;
    for (int g = 0; g < 3; g++) o[g] = (a[g] - b[g]) * deg;
}
#pragma clang fp contract(fast)

hipError_t launch_aver_batch(hipStream_t s, int which, int n, const double* x, const double* y, double eps,
                             double* res) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_aver_batch, dim3((n + 255) / 256), dim3(256), 0, s, which, n, x, y, eps, res);
    return hipGetLastError();
}

// nv = 4 / 3: IntegralUniversalPnt over cells [n][nv][3]; nv = 2: IntegralUniversalSegPnt
// over segments [n][2][3].  r_out (optional) receives the mutated smoothR.
template <int NV>
__global__ void k_integral_batch(int f, int n, const double* __restrict__ cells, const double* __restrict__ x,
                                 KP kp0, int idim, int n_start, int p_max, double eps, double* __restrict__ res,
                                 double* __restrict__ r_out) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    KP kp = kp0;
    cplx r[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    double cv[NV][3];
    load_cell<NV>(cells, i, cv);
    BIELCC_KERNEL_SWITCH(f, (integral_pnt<F, NV>(cv, x + 3 * i, kp, idim, n_start, p_max, eps, r)));
    store3(r, 3, res + 6 * i);
    if (r_out) r_out[i] = kp.R;
}

__global__ void k_seg_batch(int f, int n, const double* __restrict__ segs, const double* __restrict__ x, KP kp0,
                            int idim, int n_start, int p_max, double eps, double* __restrict__ res,
                            double* __restrict__ r_out) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    KP kp = kp0;
    cplx r[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    const double* ab = segs + 6 * i;
    BIELCC_KERNEL_SWITCH(f, (integral_seg<F>(ab, ab + 3, x + 3 * i, kp, idim, n_start, p_max, eps, r)));
    store3(r, 3, res + 6 * i);
    if (r_out) r_out[i] = kp.R;
}

hipError_t launch_integral_batch(hipStream_t s, int f, int nv, int n, const double* cells, const double* x,
                                 const KP& kp, int idim, int n_start, int p_max, double eps, double* res,
                                 double* r_out) {
    if (n <= 0) return hipSuccess;
    dim3 g((n + 63) / 64), b(64);
    if (nv == 4) hipLaunchKernelGGL(k_integral_batch<4>, g, b, 0, s, f, n, cells, x, kp, idim, n_start, p_max, eps, res, r_out);
    else if (nv == 3) hipLaunchKernelGGL(k_integral_batch<3>, g, b, 0, s, f, n, cells, x, kp, idim, n_start, p_max, eps, res, r_out);
    else hipLaunchKernelGGL(k_seg_batch, g, b, 0, s, f, n, cells, x, kp, idim, n_start, p_max, eps, res, r_out);
    return hipGetLastError();
}

// F(y) integrands of the IntegralUniversal forms (integral_universal.h:252-431,
// integral_universal_seg.h:120-179):
//   BIELCC_FY_UNIT  ff[0] = 1, ff[1..2] = 0 (the constant integrand f1 of
//                   examples/integrators/main0_integrals.cpp:77-81: the integral is the area / length);
//   BIELCC_FY_PLANE_WAVE_E  ff = e0 e^{i vec_dbl.y} (PlaneWaveFieldE, incident_field.h:20-29).
struct FyEval {
    int f;
    double e0[3];
    __device__ void operator()(const double* y, const KP& kp, cplx* ff) const {
        if (f == 0) {
            ff[0] = mk(1., 0.);
            ff[1] = mk(0., 0.);
            ff[2] = mk(0., 0.);
            return;
        }
        // scal_prod(param.vec_dbl, x): element_geom.cpp:10-12
        const double ph = kp.vec[0] * y[0] + kp.vec[1] * y[1] + kp.vec[2] * y[2];
        const cplx deg = cexp(cplx{0., 1.} * ph);
        for (int g = 0; g < 3; g++) ff[g] = deg * e0[g];
    }
};

// nv = 4 / 3: IntegralUniversal over cells [n][nv][3]; nv = 2: the F(y) form of
// IntegralUniversalSegPnt over segments [n][2][3].  r_out (optional): the mutated smoothR.
template <int NV>
__global__ void k_integral_y_batch(FyEval ev, int n, const double* __restrict__ cells, KP kp0, int idim,
                                   int n_start, int p_max, double eps, double* __restrict__ res,
                                   double* __restrict__ r_out) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    KP kp = kp0;
    cplx r[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    if constexpr (NV == 2) {
        const double* ab = cells + 6 * i;
        integral_seg_rule(ab, ab + 3, ev, kp, idim, n_start, p_max, eps, r);
    } else {
        double cv[NV][3];
        load_cell<NV>(cells, i, cv);
        integral_rule<NV>(cv, ev, kp, idim, n_start, p_max, eps, r);
    }
    store3(r, 3, res + 6 * i);
    if (r_out) r_out[i] = kp.R;
}

hipError_t launch_integral_y_batch(hipStream_t s, int f, int nv, int n, const double* cells, const KP& kp,
                                   const double* e0, int idim, int n_start, int p_max, double eps, double* res,
                                   double* r_out) {
    if (n <= 0) return hipSuccess;
    FyEval ev{f, {e0[0], e0[1], e0[2]}};
    dim3 g((n + 63) / 64), b(64);
    if (nv == 4) hipLaunchKernelGGL(k_integral_y_batch<4>, g, b, 0, s, ev, n, cells, kp, idim, n_start, p_max, eps, res, r_out);
    else if (nv == 3) hipLaunchKernelGGL(k_integral_y_batch<3>, g, b, 0, s, ev, n, cells, kp, idim, n_start, p_max, eps, res, r_out);
    else hipLaunchKernelGGL(k_integral_y_batch<2>, g, b, 0, s, ev, n, cells, kp, idim, n_start, p_max, eps, res, r_out);
    return hipGetLastError();
}

// variant 0: K0 (K0.h:36-86), 1: K0_Common (:112-166), 2: K0_GradSimplePot_H_HS (:197-249)
template <int NV>
__global__ void k_k0_batch(int variant, int f, int n, const double* __restrict__ cells,
                           const double* __restrict__ x, const double* __restrict__ j, KP kp0, int n_start,
                           int p_max, double eps, double analyticR, double* __restrict__ res,
                           double* __restrict__ r_out) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    KP kp = kp0;
    double cv[NV][3];
    load_cell<NV>(cells, i, cv);
    cplx jj[3], r[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    for (int d = 0; d < 3; d++) jj[d] = mk(j[6 * i + 2 * d], j[6 * i + 2 * d + 1]);
    if (variant == 0) {
        BIELCC_KERNEL_SWITCH(f, (k0<F, NV>(cv, x + 3 * i, jj, kp, n_start, p_max, eps, r)));
    } else if (variant == 1) {
        BIELCC_KERNEL_SWITCH(f, (k0_common<F, NV>(cv, x + 3 * i, kp, n_start, p_max, eps, r)));
    } else {
        k0_hs<NV>(cv, x + 3 * i, jj, n_start, p_max, eps, analyticR, kp.k, r);
    }
    store3(r, 3, res + 6 * i);
    if (r_out) r_out[i] = kp.R;
}

hipError_t launch_k0_batch(hipStream_t s, int variant, int f, int nv, int n, const double* cells,
                           const double* x, const double* j, const KP& kp, int n_start, int p_max, double eps,
                           double analyticR, double* res, double* r_out) {
    if (n <= 0) return hipSuccess;
    dim3 g((n + 63) / 64), b(64);
    if (nv == 4) hipLaunchKernelGGL(k_k0_batch<4>, g, b, 0, s, variant, f, n, cells, x, j, kp, n_start, p_max, eps, analyticR, res, r_out);
    else hipLaunchKernelGGL(k_k0_batch<3>, g, b, 0, s, variant, f, n, cells, x, j, kp, n_start, p_max, eps, analyticR, res, r_out);
    return hipGetLastError();
}

// which 0: integral1Divr (cells), 1: integralGrad1DivrSeg (segments [n][2][3]),
// 2: IntegralSurf_SimplePot_H_HS (cells), 3: IntegralSeg_GradSimplePot_H_HS (segments).
// res: [n][3] complex (unused entries zero).
template <int NV>
__global__ void k_analytic_batch(int which, int n, const double* __restrict__ cells,
                                 const double* __restrict__ x, int n_start, int p_max, double eps,
                                 double analyticR, double k_re, double k_im, double* __restrict__ res) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cplx r[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    const double* xi = x + 3 * i;
    if (which == 1 || which == 3) {
        const double* ab = cells + 6 * i;
        if (which == 1) {
            double g[3];
            grad1divr_seg(ab, ab + 3, xi, g);
            for (int d = 0; d < 3; d++) r[d] = mk(g[d], 0.);
        } else {
            seg_hs(ab, ab + 3, xi, analyticR, n_start, p_max, eps, mk(k_re, k_im), r);
        }
    } else {
        double cv[NV][3];
        load_cell<NV>(cells, i, cv);
        if (which == 0) r[0] = mk(integral1divr<NV>(cv, xi), 0.);
        else r[0] = surf_hs<NV>(cv, xi, analyticR, n_start, p_max, eps, mk(k_re, k_im));
    }
    store3(r, 3, res + 6 * i);
}

hipError_t launch_analytic_batch(hipStream_t s, int which, int nv, int n, const double* cells, const double* x,
                                 int n_start, int p_max, double eps, double analyticR, double k_re, double k_im,
                                 double* res) {
    if (n <= 0) return hipSuccess;
    dim3 g((n + 63) / 64), b(64);
    if (nv == 3) hipLaunchKernelGGL(k_analytic_batch<3>, g, b, 0, s, which, n, cells, x, n_start, p_max, eps, analyticR, k_re, k_im, res);
    else hipLaunchKernelGGL(k_analytic_batch<4>, g, b, 0, s, which, n, cells, x, n_start, p_max, eps, analyticR, k_re, k_im, res);
    return hipGetLastError();
}

}  // namespace bielcc_impl
