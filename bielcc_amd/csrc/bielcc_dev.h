// Device-side building blocks of the K / R / K0 path for gfx950.
//
// Two tiers live here:
//  * exact per-pair evaluation (`gen::` namespace): the reference's
//    integrands, adaptive integrators (with the caller-visible smoothing-radius
//    halving), contour operators, semi-analytic integrals and operators, one
//    (point, cell) pair per lane.  Used for near/self pairs, for p_max > 1 and
//    for the batched single-operator entry points.
//  * node-table helpers for the tiled far-pair kernels (bielcc_kernels.hip),
//    which evaluate the same quadrature rules from precomputed x-independent
//    sub-cell centroids and weights.
//
// Complex values are `cplx` (re, im) in registers; in memory they are
// interleaved doubles.
//
// Everything in this header is compiled WITHOUT floating-point contraction:
// the reference (x86-64, no FMA) rounds every product and sum separately, and
// its branch tests (r < smoothR, |kr| < 1e-7, dist < analytic_calc * diam) and
// its 0/0 at a node that coincides with x (kernel_lib.h:216) depend on those
// exact roundings.  The tiled far-pair loops in field.hip / assemble.hip keep
// contraction; they only call the near/far test from here, so that every
// kernel classifies a pair identically.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

#pragma clang fp contract(off)

namespace bielcc_dev {

constexpr double kMachineZero = DBL_EPSILON;   // geometry/tolerance_constants.h:11
constexpr int kPMaxSmooth = 8;                 // integrators/integral_universal.h:14
constexpr double kPi = 3.14159265358979323846;
constexpr double kFourPi = 4. * kPi;

struct cplx {
    double re, im;
};

__host__ __device__ inline cplx mk(double re, double im) { return cplx{re, im}; }
__host__ __device__ inline cplx operator+(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cplx operator-(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cplx operator-(cplx a) { return cplx{-a.re, -a.im}; }
__host__ __device__ inline cplx operator*(cplx a, cplx b) {
    return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline cplx operator*(cplx a, double d) { return cplx{a.re * d, a.im * d}; }
__host__ __device__ inline cplx operator*(double d, cplx a) { return cplx{d * a.re, d * a.im}; }
__host__ __device__ inline cplx operator/(cplx a, double d) { return cplx{a.re / d, a.im / d}; }
__host__ __device__ inline cplx& operator+=(cplx& a, cplx b) { a.re += b.re; a.im += b.im; return a; }

__device__ inline double cabs(cplx a) { return hypot(a.re, a.im); }
// e^{a} for complex a: e^{re} (cos im, sin im)
__device__ inline cplx cexp(cplx a) {
    double s, c;
    sincos(a.im, &s, &c);
    double e = exp(a.re);
    return cplx{e * c, e * s};
}

// ------------------------------------------------------------------ geometry
// geometry/element_geom.{h,cpp}
__device__ inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline void cross3(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double len3(const double* a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
__device__ inline double dist3(const double* a, const double* b) {
    double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}
__device__ inline cplx cdot_cr(const cplx* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline void ccross(const cplx* a, const cplx* b, cplx* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

// norm_func: diagonal cross product (element_geom.cpp:130-156)
template <int NV>
__device__ inline void norm_func(const double (*c)[3], double* nr) {
    double u[3], w[3], v[3];
    if (NV == 4) {
        for (int i = 0; i < 3; i++) { u[i] = c[2][i] - c[0][i]; w[i] = c[3][i] - c[1][i]; }
        cross3(u, w, v);
    } else {
        for (int i = 0; i < 3; i++) { u[i] = c[2][i] - c[0][i]; w[i] = c[1][i] - c[0][i]; }
        cross3(w, u, v);
    }
    double l = len3(v);
    nr[0] = v[0] / l; nr[1] = v[1] / l; nr[2] = v[2] / l;
}

// get_diam (element_geom.cpp:275-317)
template <int NV>
__device__ inline double get_diam(const double (*c)[3]) {
    if (NV == 3) {
        double l1 = dist3(c[0], c[1]), l2 = dist3(c[1], c[2]), l3 = dist3(c[2], c[0]);
        double m = (l1 < l2) ? l2 : l1;
        return (m < l3) ? l3 : m;
    }
    double d1[3], d2[3];
    for (int g = 0; g < 3; g++) { d1[g] = c[2][g] - c[0][g]; d2[g] = c[3][g] - c[1][g]; }
    double res = len3(d1) > len3(d2) ? len3(d1) : len3(d2);
    for (int s = 0; s < 4; s++) {
        double e[3];
        int t = (s + 1) & 3;
        for (int g = 0; g < 3; g++) e[g] = c[t][g] - c[s][g];
        double l = len3(e);
        if (l > res) res = l;
    }
    return res;
}

template <int NV>
__device__ inline void center_mass(const double (*c)[3], double* r) {
    for (int i = 0; i < 3; i++) {
        if (NV == 4) r[i] = (c[0][i] + c[1][i] + c[2][i] + c[3][i]) * 0.25;
        else r[i] = (c[0][i] + c[1][i] + c[2][i]) / 3.;
    }
}

// K0 edge normal: nu = (B - A) x n, normalised unless |nu| <= DBL_EPSILON (K0.h:69-79)
__device__ inline void edge_nu(const double* A, const double* B, const double* nrm, double* nu) {
    double diff[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    cross3(diff, nrm, nu);
    double l = len3(nu);
    if (l > kMachineZero) { nu[0] /= l; nu[1] /= l; nu[2] /= l; }
}

// ------------------------------------------------------------------ kernels
// electrodynamics/kernel_lib.h.  Ids match the C-ABI's bielcc_kernel enum.
enum KernelId {
    F_SIMPLEPOT_L = 0, F_GRADSIMPLEPOT_L, F_DOUBLEPOT_L, F_VECTORPOT_L,
    F_SIMPLEPOT_H, NUF_SIMPLEPOT_H, F_SIMPLEPOT_HM1, F_GRADSIMPLEPOT_H,
    F_GRADSIMPLEPOT_HM1, F_DOUBLEPOT_H, NF_DOUBLEPOT_H, F_ROTROTSIMPLEPOT_H,
    F_COUNT
};

struct KP {                 // KernelParam (integrators/kernel_param.h:25-57)
    double R, Rmin;         // smoothR, smoothRMin (R is mutated by the integrators)
    cplx k;
    double vec[3];
    double n[3];
    cplx cur[3];
};

__device__ inline KP kp_make(double R, double Rmin, cplx k) {
    KP p;
    p.R = R; p.Rmin = Rmin; p.k = k;
    for (int i = 0; i < 3; i++) { p.vec[i] = 0.; p.n[i] = 0.; p.cur[i] = mk(0., 0.); }
    return p;
}

// std::complex<double>(0., r) * k
__device__ inline cplx ir_k(double r, cplx k) { return cplx{0., r} * k; }

// Smoothing factors kept as boolean products (kernel_lib.h:216,276,310,388):
// with r = 0 they give the reference's NaN.
__device__ inline double sm_pot(double r, double R) {
    double t = r / R;
    return (double)(t < 1) * (t * (3. - 2. * t) / R) + (double)(t >= 1) / r;
}
__device__ inline double sm_grad(double r, double R) {
    double t = r / R;
    return (double)(t < 1) * (3. - 2. * t) / (r * (R * R)) + (double)(t >= 1) / ((r * r) * r);
}

template <int F>
__device__ inline void kernel_eval(const double* x, const double* y, const KP& kp, cplx* res) {
    double r = dist3(x, y);
    if (F == F_SIMPLEPOT_L) {
        res[0] = mk(sm_pot(r, kp.R), 0.);
    } else if (F == F_GRADSIMPLEPOT_L) {
        double s = sm_grad(r, kp.R);
        for (int i = 0; i < 3; i++) res[i] = mk((y[i] - x[i]) * s, 0.);
    } else if (F == F_DOUBLEPOT_L) {
        double diff[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
        res[0] = mk(dot3(kp.n, diff) * sm_grad(r, kp.R), 0.);
    } else if (F == F_VECTORPOT_L) {
        if (r < kMachineZero) {
            for (int i = 0; i < 3; i++) res[i] = mk(0., 0.);
        } else {
            double f = 1. / r / r / r, d[3], o[3];
            for (int i = 0; i < 3; i++) d[i] = (y[i] - x[i]) * f;
            cross3(d, kp.vec, o);
            if (r >= kMachineZero && r < kp.R) {
                double t = r / kp.R;
                for (int i = 0; i < 3; i++) o[i] = o[i] * (3 * t * t - 2 * t * t * t);
            }
            for (int i = 0; i < 3; i++) res[i] = mk(o[i], 0.);
        }
    } else if (F == F_SIMPLEPOT_H || F == NUF_SIMPLEPOT_H) {
        cplx e = cexp(ir_k(r, kp.k)) * sm_pot(r, kp.R);
        if (F == F_SIMPLEPOT_H) res[0] = e;
        else for (int i = 0; i < 3; i++) res[i] = e * kp.vec[i];
    } else if (F == F_SIMPLEPOT_HM1) {
        cplx ikr = ir_k(r, kp.k);
        if (cabs(ikr) < 1e-7) {
            res[0] = cplx{0., 1.} * kp.k - (kp.k * kp.k) * r / 2.;
        } else {
            res[0] = (cexp(ikr) - mk(1., 0.)) * sm_pot(r, kp.R);
        }
    } else if (F == F_GRADSIMPLEPOT_H) {
        double diff[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
        cplx ikr = ir_k(r, kp.k);
        cplx e = cexp(ikr);
        cplx m = ikr - mk(1., 0.);
        double s = sm_grad(r, kp.R);
        for (int i = 0; i < 3; i++) res[i] = ((e * diff[i]) * m) * s;
    } else if (F == F_GRADSIMPLEPOT_HM1) {
        cplx ikr = ir_k(r, kp.k);
        double a = cabs(ikr);
        if (a < 1e-6) {
            for (int i = 0; i < 3; i++) res[i] = mk(0., 0.);
        } else if (a < 1e-4) {
            for (int i = 0; i < 3; i++) res[i] = ((kp.k * (y[i] - x[i])) * kp.k) / (2. * r);
        } else {
            cplx q = mk(1., 0.) + (ikr - mk(1., 0.)) * cexp(ikr);
            for (int i = 0; i < 3; i++) res[i] = (q * (x[i] - y[i])) / (r * r * r);
        }
    } else if (F == F_DOUBLEPOT_H || F == NF_DOUBLEPOT_H) {
        double diff[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
        cplx ikr = ir_k(r, kp.k);
        cplx v = ((cexp(ikr) * dot3(kp.n, diff)) * (ikr - mk(1., 0.))) * sm_grad(r, kp.R);
        if (F == F_DOUBLEPOT_H) res[0] = v;
        else for (int i = 0; i < 3; i++) res[i] = v * kp.n[i];
    } else if (F == F_ROTROTSIMPLEPOT_H) {
        if (r < kMachineZero) {
            for (int i = 0; i < 3; i++) res[i] = mk(0., 0.);
            return;
        }
        cplx ik = cplx{0., 1.} * kp.k;
        cplx k2 = kp.k * kp.k;
        double d[3];
        for (int i = 0; i < 3; i++) d[i] = (x[i] - y[i]) / r;
        double r2 = r * r;
        cplx f1 = mk(-1.0 / r2 / r, 0.) + ik / r2 + k2 / r;
        cplx f2 = mk(3.0 / r2 / r, 0.) - (ik * 3.0) / r2 - k2 / r;
        cplx e = cexp(ik * r);
        cplx dj = kp.cur[0] * d[0] + kp.cur[1] * d[1] + kp.cur[2] * d[2];
        f2 = f2 * dj;
        for (int i = 0; i < 3; i++) res[i] = (f2 * d[i] + kp.cur[i] * f1) * e;
        if (r < kp.R) {
            double t = r / kp.R;
            for (int i = 0; i < 3; i++) res[i] = res[i] * (t * t * (3 - 2 * t));
        }
    }
}

// ------------------------------------------------------------------ integrators
// Tail of every adaptive pass (integral_universal.h:96-112,
// integral_universal_seg.h:74-90).  Returns true when the loop stops.
__device__ inline bool pass_end(int p_n, int idim, double eps, cplx* res, cplx* prev, KP& kp) {
    double delta = 0.;
    for (int g = 0; g < idim; g++) {
        double a = cabs(res[g] - prev[g]);
        delta += a * a;
    }
    if (delta < eps * eps && p_n != 0) return true;
    for (int g = 0; g < idim; g++) {
        prev[g] = res[g];
        res[g] = mk(0., 0.);
        if (p_n < kPMaxSmooth && kp.R > kp.Rmin) kp.R /= 2.;
    }
    return false;
}

// The cell rules with the integrand as a functor ev(y, kp, ff).
// F(x, y) forms, IntegralUniversalPnt: quad integral_universal.h:38-120, triangle :136-220;
// F(y) forms, IntegralUniversal: quad :252-333, triangle :347-431 (the same rules).
template <int NV, class Ev>
__device__ void integral_rule(const double (*c)[3], Ev ev, KP& kp, int idim, int n_start, int p_max,
                              double eps, cplx* res) {
    int n = n_start, p_n;
    cplx ff[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, prev[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    for (int g = 0; g < idim; g++) res[g] = mk(0., 0.);
    double s_tri = 0.;
    if (NV == 3) {
        double ab[3], ac[3], nn[3];
        for (int i = 0; i < 3; i++) { ab[i] = c[1][i] - c[0][i]; ac[i] = c[2][i] - c[0][i]; }
        cross3(ab, ac, nn);
        s_tri = sqrt(dot3(nn, nn)) / 2.;
    }
    for (p_n = 0; p_n < p_max; p_n++) {
        if (NV == 4) {
            for (int i = 0; i < n; i++) {
                for (int j = 0; j < n; j++) {
                    double p = (double)i / (double)n, q = (double)j / (double)n;
                    double p1 = ((double)i + 1.) / (double)n, q1 = ((double)j + 1.) / (double)n;
                    double a1[3], a2[3], a3[3], a4[3], rc[3], m1[3], m2[3], rn[3];
                    for (int k = 0; k < 3; k++) {
                        double a = q * c[1][k] + (1 - q) * c[0][k];
                        double b = q * c[2][k] + (1 - q) * c[3][k];
                        a1[k] = p * b + (1 - p) * a;
                        a4[k] = p1 * b + (1 - p1) * a;
                        a = q1 * c[1][k] + (1 - q1) * c[0][k];
                        b = q1 * c[2][k] + (1 - q1) * c[3][k];
                        a2[k] = p * b + (1 - p) * a;
                        a3[k] = p1 * b + (1 - p1) * a;
                    }
                    for (int k = 0; k < 3; k++) {
                        rc[k] = (a1[k] + a2[k] + a3[k] + a4[k]) / 4.0;
                        m1[k] = ((a2[k] + a3[k]) - (a1[k] + a4[k])) / 2.0;
                        m2[k] = ((a3[k] + a4[k]) - (a1[k] + a2[k])) / 2.0;
                    }
                    cross3(m1, m2, rn);
                    double s = len3(rn);
                    ev(rc, kp, ff);
                    for (int g = 0; g < idim; g++) res[g] += ff[g] * s;
                }
            }
        } else {
            double pv[3], qv[3], rv[3];
            for (int k = 0; k < 3; k++) {
                pv[k] = (c[1][k] - c[0][k]) / n;
                qv[k] = (c[2][k] - c[0][k]) / n;
                rv[k] = (c[2][k] - c[1][k]) / n;
            }
            for (int i = 0; i < n; i++) {
                for (int j = 0; j < n - i; j++) {
                    double B[3], rc[3];
                    for (int k = 0; k < 3; k++) {
                        B[k] = c[0][k] + (i + 1) * pv[k] + j * qv[k];
                        double AM = B[k] - pv[k];
                        double CM = AM + qv[k];
                        rc[k] = (B[k] + AM + CM) / 3.0;
                    }
                    ev(rc, kp, ff);
                    for (int g = 0; g < idim; g++) res[g] += ff[g];
                    if (j < n - 1 - i) {
                        for (int k = 0; k < 3; k++) {
                            double AP = B[k] + rv[k];
                            double CP = B[k] + qv[k];
                            rc[k] = (B[k] + AP + CP) / 3.0;
                        }
                        ev(rc, kp, ff);
                        for (int g = 0; g < idim; g++) res[g] += ff[g];
                    }
                }
            }
            for (int g = 0; g < idim; g++) res[g] = ((res[g] * s_tri) / (double)n) / (double)n;
        }
        if (pass_end(p_n, idim, eps, res, prev, kp)) break;
        n = n * 2;
    }
    if (p_n == p_max)
        for (int g = 0; g < idim; g++) res[g] = prev[g];
}

template <int F, int NV>
__device__ void integral_pnt(const double (*c)[3], const double* x, KP& kp, int idim, int n_start,
                             int p_max, double eps, cplx* res) {
    integral_rule<NV>(c, [x](const double* y, const KP& k, cplx* ff) { kernel_eval<F>(x, y, k, ff); }, kp, idim,
                      n_start, p_max, eps, res);
}

// The segment rule with the integrand as a functor ev(y, kp, ff): IntegralUniversalSegPnt,
// F(x, y) form integral_universal_seg.h:38-98, F(y) form :120-179 (the same rule).
template <class Ev>
__device__ void integral_seg_rule(const double* a, const double* b, Ev ev, KP& kp, int idim, int n_start,
                                  int p_max, double eps, cplx* res) {
    int n = n_start, p_n;
    cplx ff[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, prev[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    for (int g = 0; g < idim; g++) res[g] = mk(0., 0.);
    for (p_n = 0; p_n < p_max; p_n++) {
        double d[3], y[3];
        for (int j = 0; j < 3; j++) d[j] = (b[j] - a[j]) / n;
        double dl = len3(d);
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < 3; j++) y[j] = a[j] + (i + 0.5) * d[j];
            ev(y, kp, ff);
            for (int g = 0; g < idim; g++) res[g] += ff[g] * dl;
        }
        if (pass_end(p_n, idim, eps, res, prev, kp)) break;
        n = n * 2;
    }
    if (p_n == p_max)
        for (int g = 0; g < idim; g++) res[g] = prev[g];
}

template <int F>
__device__ void integral_seg(const double* a, const double* b, const double* x, KP& kp, int idim,
                             int n_start, int p_max, double eps, cplx* res) {
    integral_seg_rule(a, b, [x](const double* y, const KP& k, cplx* ff) { kernel_eval<F>(x, y, k, ff); }, kp,
                      idim, n_start, p_max, eps, res);
}

// ------------------------------------------------------------------ analytic
// integrators/integrals_analytic.cpp
__device__ inline void lartgp(double x, double y, double& cs, double& sn, double& r) {   // :15-26
    double tmp = sqrt(x * x + y * y);
    if (tmp == 0) { r = 0; cs = 1; sn = 0; }
    else { cs = x / tmp; sn = y / tmp; r = tmp; }
}
__device__ inline void rot(double& x, double& y, double cs, double sn) {   // :29-34
    double tmp = x * cs + y * sn;
    y = y * cs - x * sn;
    x = tmp;
}

// Exact integral of 1/|x - y| over a flat triangle (_Integral1Divr_inv :37-200):
// Givens-rotation QR of the edge matrix, then per-edge log terms and an
// atan2 solid-angle term.  t[vertex][coord].
__device__ inline double int1divr_tri(const double* t0, const double* t1, const double* t2, const double* d) {
    double e[3][3], Q[3][2], D[7][3], cs[3], sn[3], n[3];
    for (int i = 0; i < 3; ++i) {
        e[i][0] = t1[i] - t0[i];
        e[i][1] = t2[i] - t1[i];
        e[i][2] = t0[i] - t2[i];
    }
    for (int i = 0; i < 3; ++i) {
        const double* v = (i == 0) ? t0 : (i == 1) ? t1 : t2;
        for (int j = 0; j < 3; ++j) D[j][i] = v[j] - d[j];
        D[3][i] = sqrt(D[0][i] * D[0][i] + D[1][i] * D[1][i] + D[2][i] * D[2][i]);
        Q[i][0] = e[i][0];
        Q[i][1] = e[i][1];
        n[i] = 0;
    }
    D[4][0] = D[3][1]; D[4][1] = D[3][2]; D[4][2] = D[3][0];
    lartgp(Q[0][0], Q[1][0], cs[0], sn[0], Q[0][0]);
    lartgp(Q[0][0], Q[2][0], cs[1], sn[1], Q[0][0]);
    rot(Q[0][1], Q[1][1], cs[0], sn[0]);
    rot(Q[0][1], Q[2][1], cs[1], sn[1]);
    lartgp(Q[1][1], Q[2][1], cs[2], sn[2], Q[1][1]);
    n[2] = 1.0;
    rot(n[1], n[2], cs[2], -sn[2]);
    rot(n[0], n[2], cs[1], -sn[1]);
    rot(n[0], n[1], cs[0], -sn[0]);
    lartgp(n[0], n[1], cs[0], sn[0], n[0]);
    for (int i = 0; i < 3; ++i) { rot(e[0][i], e[1][i], cs[0], sn[0]); rot(D[0][i], D[1][i], cs[0], sn[0]); }
    lartgp(n[0], n[2], cs[0], sn[0], n[0]);
    for (int i = 0; i < 3; ++i) { rot(e[0][i], e[2][i], cs[0], sn[0]); rot(D[0][i], D[2][i], cs[0], sn[0]); }
    lartgp(e[1][0], e[2][0], cs[0], sn[0], e[1][0]);
    lartgp(e[1][1], e[2][1], cs[1], sn[1], e[1][1]);
    lartgp(e[1][2], e[2][2], cs[2], sn[2], e[1][2]);
    D[5][0] = cs[0] * D[1][1] + sn[0] * D[2][1];
    D[5][1] = cs[1] * D[1][2] + sn[1] * D[2][2];
    D[5][2] = cs[2] * D[1][0] + sn[2] * D[2][0];
    D[6][0] = -sn[0] * D[1][1] + cs[0] * D[2][1];
    D[6][1] = -sn[1] * D[1][2] + cs[1] * D[2][2];
    D[6][2] = -sn[2] * D[1][0] + cs[2] * D[2][0];
    rot(D[1][0], D[2][0], cs[0], sn[0]);
    rot(D[1][1], D[2][1], cs[1], sn[1]);
    rot(D[1][2], D[2][2], cs[2], sn[2]);
    double int_1r = 0, proj = 0;
    for (int ed = 0; ed < 3; ++ed) {
        double ad = 0;
        double r = D[3][ed];
        if (fabs(r) > 0 && fabs(D[4][ed]) > 0) {
            double p0 = -D[2][ed] / r;
            double l = D[1][ed] / r;
            double ratio = D[4][ed] / r;
            double l_pl = D[5][ed] / D[4][ed];
            if (l > -1 && l_pl > -1) ad = p0 * log(ratio * (1 + l_pl) / (1 + l));
            int emn = (ed + 2) % 3;
            proj = proj + (D[0][ed] * D[0][ed] + D[1][ed] * D[5][ed] + D[2][ed] * D[6][ed]) * D[3][emn];
            int_1r = int_1r + r * ad;
        }
    }
    double c0 = fabs(D[0][0]) * Q[0][0] * Q[1][1];
    double c1 = D[3][0] * D[3][1] * D[3][2] + proj;
    if (c0 > 0 || fabs(c1) > 0) int_1r = int_1r - fabs(D[0][0]) * 2 * atan2(c0, c1);
    return int_1r;
}

// integral1Divr: quad = tri(0,1,2) + tri(2,3,0) (:218-235)
template <int NV>
__device__ inline double integral1divr(const double (*c)[3], const double* x) {
    if (NV == 3) return int1divr_tri(c[0], c[1], c[2], x);
    return int1divr_tri(c[0], c[1], c[2], x) + int1divr_tri(c[2], c[3], c[0], x);
}

// integralGrad1DivrSeg (:240-257)
__device__ inline void grad1divr_seg(const double* A, const double* B, const double* x, double* res) {
    double xmA[3] = {x[0] - A[0], x[1] - A[1], x[2] - A[2]};
    double xmB[3] = {x[0] - B[0], x[1] - B[1], x[2] - B[2]};
    double xA = dist3(x, A), xB = dist3(x, B), AB = dist3(A, B);
    double den = dot3(xmA, xmB) + xA * xB;
    for (int i = 0; i < 3; i++) res[i] = -AB * (xmA[i] / xA + xmB[i] / xB) / den;
}

// IntegralSurf_SimplePot_H_HS integrals_analytic.h:136-164
template <int NV>
__device__ inline cplx surf_hs(const double (*c)[3], const double* x, double analyticR, int n_start,
                               int p_max, double eps, cplx k) {
    double y[3];
    center_mass<NV>(c, y);
    KP kp = kp_make(kMachineZero, kMachineZero, k);
    double calc_dist = analyticR * get_diam<NV>(c);
    cplx tmp[3];
    if (dist3(x, y) < calc_dist) {
        cplx ff = mk(integral1divr<NV>(c, x), 0.);
        integral_pnt<F_SIMPLEPOT_HM1, NV>(c, x, kp, 1, n_start, p_max, eps, tmp);
        return ff + tmp[0];
    }
    integral_pnt<F_SIMPLEPOT_H, NV>(c, x, kp, 1, n_start, p_max, eps, tmp);
    return tmp[0];
}

// IntegralSeg_GradSimplePot_H_HS integrals_analytic.h:193-221 (its smoothRMin
// is uninitialised in the reference; 0 here)
__device__ inline void seg_hs(const double* A, const double* B, const double* x, double analyticR,
                              int n_start, int p_max, double eps, cplx k, cplx* res) {
    KP kp = kp_make(kMachineZero, 0., k);
    double y[3] = {(B[0] + A[0]) / 2., (B[1] + A[1]) / 2., (B[2] + A[2]) / 2.};
    if (dist3(x, y) < analyticR * dist3(A, B)) {
        double tmp[3];
        grad1divr_seg(A, B, x, tmp);
        integral_seg<F_GRADSIMPLEPOT_HM1>(A, B, x, kp, 3, n_start, p_max, eps, res);
        for (int i = 0; i < 3; i++) res[i] += mk(tmp[i], 0.);
    } else {
        integral_seg<F_GRADSIMPLEPOT_H>(A, B, x, kp, 3, n_start, p_max, eps, res);
    }
}

// ------------------------------------------------------------------ contour operators
// K0 (K0.h:36-86): -sum_e (j . nu_e) int_e F, one KernelParam shared by all
// edges (so its smoothing radius keeps shrinking from edge to edge).
template <int F, int NV>
__device__ void k0(const double (*c)[3], const double* x, const cplx* j, KP& kp, int n_start, int p_max,
                   double eps, cplx* res) {
    double nrm[3], nu[3];
    cplx ks[3];
    for (int g = 0; g < 3; g++) res[g] = mk(0., 0.);
    norm_func<NV>(c, nrm);
    for (int s = 0; s < NV; s++) {
        int A = s, B = (s != NV - 1) ? s + 1 : 0;
        integral_seg<F>(c[A], c[B], x, kp, 3, n_start, p_max, eps, ks);
        edge_nu(c[A], c[B], nrm, nu);
        cplx jn = -cdot_cr(j, nu);
        for (int g = 0; g < 3; g++) res[g] += ks[g] * jn;
    }
}

// K0_Common (K0.h:112-166): sum_e int_e F with kp.vec = nu_e
template <int F, int NV>
__device__ void k0_common(const double (*c)[3], const double* x, KP& kp, int n_start, int p_max,
                          double eps, cplx* res) {
    double nrm[3], nu[3];
    cplx ks[3];
    for (int g = 0; g < 3; g++) res[g] = mk(0., 0.);
    norm_func<NV>(c, nrm);
    for (int s = 0; s < NV; s++) {
        int A = s, B = (s != NV - 1) ? s + 1 : 0;
        edge_nu(c[A], c[B], nrm, nu);
        for (int g = 0; g < 3; g++) kp.vec[g] = nu[g];
        integral_seg<F>(c[A], c[B], x, kp, 3, n_start, p_max, eps, ks);
        for (int g = 0; g < 3; g++) res[g] += ks[g];
    }
}

// K0_GradSimplePot_H_HS (K0.h:197-249)
template <int NV>
__device__ void k0_hs(const double (*c)[3], const double* x, const cplx* j, int n_start, int p_max,
                      double eps, double analyticR, cplx k, cplx* res) {
    double nrm[3], nu[3];
    cplx ks[3];
    for (int g = 0; g < 3; g++) res[g] = mk(0., 0.);
    norm_func<NV>(c, nrm);
    for (int s = 0; s < NV; s++) {
        int A = s, B = (s != NV - 1) ? s + 1 : 0;
        seg_hs(c[A], c[B], x, analyticR, n_start, p_max, eps, k, ks);
        edge_nu(c[A], c[B], nrm, nu);
        cplx jn = -cdot_cr(j, nu);
        for (int g = 0; g < 3; g++) res[g] += ks[g] * jn;
    }
}

// ------------------------------------------------------------------ operators
struct NumP {              // NumParam (electrodynamics/num_param.h:28-55)
    double eps, rs, rs_seg, analytic_calc;
    int n_start, n_start_seg, p_max, p_max_seg;
};

enum OpId {
    OP_K_NEAR_SMOOTH = 0,  // K_RotRot_Near_Smooth K.h:47-79
    OP_K_HS = 1,           // K_RotRot_HS K.h:119-145
    OP_K_HS_COLLOC = 2,    // K_RotRot_HS_Colloc K.h:179-204
    OP_K_HS2_COLLOC = 3,   // K_RotRot_HS2_Colloc K.h:230-252
    OP_K_FAR = 4,          // K_RotRot_Far K.h:288-301
    OP_R_SMOOTH = 5,       // R_Rot_Smooth R.h:46-61
    OP_R_COLLOC = 6,       // R_Rot_Colloc R.h:89-111
    OP_R_NEAR = 7,         // R_Rot_Near R.h:140-175
    OP_COUNT
};

template <int NV>
__device__ void op_eval(int op, const double (*c)[3], const double* x, const cplx* j, const NumP& np,
                        cplx k, cplx* res) {
    double diam = get_diam<NV>(c);
    cplx t1[3], t3[3];
    switch (op) {
    case OP_K_NEAR_SMOOTH: {
        KP kf = kp_make(np.rs * diam / np.n_start, np.eps, k);
        integral_pnt<F_SIMPLEPOT_H, NV>(c, x, kf, 1, np.n_start, np.p_max, np.eps, t1);
        cplx surf = ((t1[0] * k) * k) / (4. * kPi);
        KP kg = kp_make(np.rs_seg * diam / np.n_start_seg, np.eps, k);
        k0<F_GRADSIMPLEPOT_H, NV>(c, x, j, kg, np.n_start_seg, np.p_max_seg, np.eps, t3);
        for (int i = 0; i < 3; i++) res[i] = surf * j[i] + t3[i] / (4. * kPi);
        break;
    }
    case OP_K_HS:
    case OP_K_HS_COLLOC:
    case OP_K_HS2_COLLOC: {
        cplx s = surf_hs<NV>(c, x, np.analytic_calc, np.n_start, np.p_max, np.eps, k);
        for (int i = 0; i < 3; i++) res[i] = (((s * k) * k) * j[i]) / (4. * kPi);
        if (op == OP_K_HS2_COLLOC) {
            k0_hs<NV>(c, x, j, np.n_start_seg, np.p_max_seg, np.eps, np.analytic_calc, k, t3);
        } else {
            KP kg = (op == OP_K_HS) ? kp_make(np.rs_seg * diam / np.n_start_seg, np.eps, k)
                                    : kp_make(kMachineZero, kMachineZero, k);
            k0<F_GRADSIMPLEPOT_H, NV>(c, x, j, kg, np.n_start_seg, np.p_max_seg, np.eps, t3);
        }
        for (int i = 0; i < 3; i++) res[i] += t3[i] / (4. * kPi);
        break;
    }
    case OP_K_FAR: {
        KP kf = kp_make(kMachineZero, kMachineZero, k);
        for (int i = 0; i < 3; i++) kf.cur[i] = j[i];
        integral_pnt<F_ROTROTSIMPLEPOT_H, NV>(c, x, kf, 3, np.n_start, np.p_max, np.eps, res);
        for (int i = 0; i < 3; i++) res[i] = res[i] / (4. * kPi);
        break;
    }
    case OP_R_SMOOTH:
    case OP_R_COLLOC: {
        KP kf;
        if (op == OP_R_COLLOC) {
            double y[3];
            center_mass<NV>(c, y);
            if (dist3(x, y) < kMachineZero) {
                for (int i = 0; i < 3; i++) res[i] = mk(0., 0.);
                break;
            }
            kf = kp_make(kMachineZero, kMachineZero, k);
        } else {
            kf = kp_make(np.rs * diam / np.n_start, diam / 100., k);
        }
        integral_pnt<F_GRADSIMPLEPOT_H, NV>(c, x, kf, 3, np.n_start, np.p_max, np.eps, t3);
        ccross(t3, j, res);
        for (int i = 0; i < 3; i++) res[i] = res[i] / (4. * kPi);
        break;
    }
    case OP_R_NEAR: {
        double nrm[3];
        norm_func<NV>(c, nrm);
        KP kf = kp_make(np.rs * diam / np.n_start, diam / 100., k);
        for (int i = 0; i < 3; i++) kf.n[i] = nrm[i];
        integral_pnt<NF_DOUBLEPOT_H, NV>(c, x, kf, 3, np.n_start, np.p_max, np.eps, t1);
        KP ks = kp_make(kMachineZero, kMachineZero, k);
        k0_common<NUF_SIMPLEPOT_H, NV>(c, x, ks, np.n_start_seg, np.p_max_seg, np.eps, t3);
        cplx tmp[3];
        for (int i = 0; i < 3; i++) tmp[i] = (t1[i] + t3[i]) / (4. * kPi);
        ccross(j, tmp, res);
        break;
    }
    default:
        for (int i = 0; i < 3; i++) res[i] = mk(__builtin_nan(""), __builtin_nan(""));
    }
}

// ------------------------------------------------------------------ near/far rule
// The one near test used by every kernel.  dist as element_geom.h:91-98.
// mode 0 (get_field.h:58, integrals_analytic.h:151-153): near iff dist < ac * diam;
// mode 1 (main0_field_test.cpp:242): near iff !(dist > thresh), thresh = ac * grid step.
__device__ inline double center_dist(double x0, double x1, double x2, const double* c) {
    double d0 = x0 - c[0], d1 = x1 - c[1], d2 = x2 - c[2];
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}
__device__ inline bool near_rule(double dist, double diam, int mode, double ac, double thresh) {
    if (mode == 0) return dist < ac * diam;
    return !(dist > thresh);
}

// 1/sqrt(x) for finite x > 0 (far pairs: x is at least the square of a node
// distance well away from zero).  v_rsq_f64 is correct to about 2^-23 relative;
// one third-order step y + y e (1/2 + 3/8 e), e = 1 - x y^2, brings that below
// rounding.  The library rsqrt does the same step and adds a class test and two
// selects for 0, inf and denormals, which these arguments never are.
__device__ __forceinline__ double rsqrt_pos(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

}  // namespace bielcc_dev

#pragma clang fp contract(fast)
