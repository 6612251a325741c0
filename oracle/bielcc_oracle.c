/*
 * CPU restatement of BIELcc's Stratton-Chu K / R / K0 path - the parity
 * oracle.  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's
 * cpu_baseline leg use it as the checker; the product never links it.
 *
 * Every function restates a reference function step by step (same node
 * order, same branch tests, same adaptive loop and the same mutation of the
 * smoothing radius) and cites the file:line it follows.  Own code: complex
 * arithmetic is written out on (re, im) pairs.
 *
 * Known departures (documented, harmless at the configurations used):
 *  - IntegralSeg_GradSimplePot_H_HS leaves KernelParam::smoothRMin
 *    uninitialised (integrators/integrals_analytic.h:199-201); here it is 0.
 *    Neither F_GradSimplePot_Hm1 nor F_GradSimplePot_H with smoothR ~ 2^-52
 *    changes value with it unless a node sits within 1e-16 of x.
 *  - K0 / K0_Common / K0_GradSimplePot_H_HS are restated for idim = 3 (the
 *    only value any operator passes); the reference reads uninitialised
 *    entries for idim < 3 (electrodynamics/K0.h:44-79).
 */
#include "bielcc_oracle.h"

#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>

#define MACHINE_ZERO DBL_EPSILON           /* geometry/tolerance_constants.h:11 */
#define PMAX_SMOOTH 8                      /* integrators/integral_universal.h:14 */
#define FOUR_PI (4. * M_PI)

/* ---------------------------------------------------------------- complex */
typedef struct { double re, im; } cplx;

static inline cplx C(double re, double im) { cplx z = {re, im}; return z; }
static inline cplx cadd(cplx a, cplx b) { return C(a.re + b.re, a.im + b.im); }
static inline cplx csub(cplx a, cplx b) { return C(a.re - b.re, a.im - b.im); }
static inline cplx cmul(cplx a, cplx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cplx cmuld(cplx a, double d) { return C(a.re * d, a.im * d); }
static inline cplx cdivd(cplx a, double d) { return C(a.re / d, a.im / d); }
static inline cplx cneg(cplx a) { return C(-a.re, -a.im); }
static inline double cabs_(cplx a) { return hypot(a.re, a.im); }
/* std::exp(complex): e^re (cos im, sin im) */
static inline cplx cexp_(cplx a) {
    double e = exp(a.re);
    return C(e * cos(a.im), e * sin(a.im));
}

/* ---------------------------------------------------------------- geometry
 * geometry/element_geom.{h,cpp} */
static inline double scal(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void vprod(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double vlen(const double* a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static inline double vdist(const double* a, const double* b) {   /* element_geom.h:91-98 */
    return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}
/* complex-vector x real-vector dot: scal_prod(complex*, double*) element_geom.cpp:14-16 */
static inline cplx cscal(const cplx* a, const double* b) {
    return cadd(cadd(cmuld(a[0], b[0]), cmuld(a[1], b[1])), cmuld(a[2], b[2]));
}
/* vec_prod(complex*, complex*) element_geom.cpp:47-51 */
static inline void cvprod(const cplx* a, const cplx* b, cplx* r) {
    r[0] = csub(cmul(a[1], b[2]), cmul(a[2], b[1]));
    r[1] = csub(cmul(a[2], b[0]), cmul(a[0], b[2]));
    r[2] = csub(cmul(a[0], b[1]), cmul(a[1], b[0]));
}

static double tr_square(const double* p1, const double* p2, const double* p3) {   /* element_geom.cpp:55-66 */
    double ab[3], ac[3], n[3];
    for (int i = 0; i < 3; i++) { ab[i] = p2[i] - p1[i]; ac[i] = p3[i] - p1[i]; }
    vprod(ab, ac, n);
    return sqrt(scal(n, n)) / 2.;
}

static void norm_func(int nv, const double (*c)[3], double* nr) {   /* element_geom.cpp:130-156 */
    double u[3], w[3], v[3];
    if (nv == 4) {
        for (int i = 0; i < 3; i++) { u[i] = c[2][i] - c[0][i]; w[i] = c[3][i] - c[1][i]; }
        vprod(u, w, v);
    } else {
        for (int i = 0; i < 3; i++) { u[i] = c[2][i] - c[0][i]; w[i] = c[1][i] - c[0][i]; }
        vprod(w, u, v);
    }
    double l = vlen(v);
    nr[0] = v[0] / l; nr[1] = v[1] / l; nr[2] = v[2] / l;
}

static double get_diam(int nv, const double (*c)[3]) {   /* element_geom.cpp:275-317 */
    if (nv == 3) {
        double l1 = vdist(c[0], c[1]), l2 = vdist(c[1], c[2]), l3 = vdist(c[2], c[0]);
        double m = l1 > l2 ? l1 : l2;        /* std::max(a, b) returns a unless a < b */
        if (m < l2) m = l2;
        return (m < l3) ? l3 : m;
    }
    double d1[3], d2[3], e[4][3];
    for (int g = 0; g < 3; g++) {
        d1[g] = c[2][g] - c[0][g]; d2[g] = c[3][g] - c[1][g];
        e[0][g] = c[1][g] - c[0][g]; e[1][g] = c[2][g] - c[1][g];
        e[2][g] = c[3][g] - c[2][g]; e[3][g] = c[0][g] - c[3][g];
    }
    double res = vlen(d1) > vlen(d2) ? vlen(d1) : vlen(d2);
    for (int s = 0; s < 4; s++) if (vlen(e[s]) > res) res = vlen(e[s]);
    return res;
}

static void center_mass(int nv, const double (*c)[3], double* r) {   /* element_geom.cpp:332-342 */
    for (int i = 0; i < 3; i++) {
        if (nv == 4) r[i] = (c[0][i] + c[1][i] + c[2][i] + c[3][i]) * 0.25;
        else r[i] = (c[0][i] + c[1][i] + c[2][i]) / 3.;
    }
}

int orc_cellfill(int nv, const double* v, double* out) {
    /* CellQuad::CellFill mesh/cell.cpp:8-51, CellTriangle::CellFill :61-100 */
    const double (*c)[3] = (const double (*)[3])v;
    double nrm[3], t1[3], t2[3], len, area;
    if (nv == 4) {
        double a[3], b[3], cc[3], d[3], ac[3], bd[3];
        for (int i = 0; i < 3; i++)
            out[i] = (c[0][i] + c[1][i] + c[2][i] + c[3][i]) * 0.25;
        for (int i = 0; i < 3; i++) {
            a[i] = (c[0][i] + c[1][i]) / 2.0; b[i] = (c[1][i] + c[2][i]) / 2.0;
            cc[i] = (c[2][i] + c[3][i]) / 2.0; d[i] = (c[3][i] + c[0][i]) / 2.0;
            ac[i] = cc[i] - a[i]; bd[i] = d[i] - b[i];
        }
        vprod(ac, bd, nrm);
        len = vlen(nrm);
        for (int k = 0; k < 3; k++) nrm[k] /= len;
        area = len;
        len = vlen(ac);
        for (int k = 0; k < 3; k++) t1[k] = ac[k] / len;
    } else if (nv == 3) {
        double ab[3], ac[3];
        for (int i = 0; i < 3; i++) out[i] = (c[0][i] + c[1][i] + c[2][i]) / 3.;
        for (int k = 0; k < 3; k++) { ab[k] = c[1][k] - c[0][k]; ac[k] = c[2][k] - c[0][k]; }
        vprod(ab, ac, nrm);
        len = vlen(nrm);
        for (int k = 0; k < 3; k++) nrm[k] /= len;
        area = len / 2;
        len = vlen(ab);
        for (int k = 0; k < 3; k++) t1[k] = ab[k] / len;
    } else {
        return -1;
    }
    vprod(nrm, t1, t2);
    len = vlen(t2);
    for (int k = 0; k < 3; k++) t2[k] /= len;
    for (int k = 0; k < 3; k++) { out[3 + k] = nrm[k]; out[6 + k] = t1[k]; out[9 + k] = t2[k]; }
    out[12] = area;
    out[13] = get_diam(nv, c);
    norm_func(nv, c, out + 14);
    center_mass(nv, c, out + 17);
    return 0;
}

/* ---------------------------------------------------------------- kernels
 * electrodynamics/kernel_lib.h.  Output: up to 3 complex values (real kernels
 * return im = 0).  `smooth` factors keep the reference's boolean products so
 * that r = 0 gives the same NaN. */
enum {
    F_SIMPLEPOT_L = 0, F_GRADSIMPLEPOT_L, F_DOUBLEPOT_L, F_VECTORPOT_L,
    F_SIMPLEPOT_H, NUF_SIMPLEPOT_H, F_SIMPLEPOT_HM1, F_GRADSIMPLEPOT_H,
    F_GRADSIMPLEPOT_HM1, F_DOUBLEPOT_H, NF_DOUBLEPOT_H, F_ROTROTSIMPLEPOT_H,
    F_COUNT
};

static inline cplx kval(const orc_kernel_param* kp) { return C(kp->k_re, kp->k_im); }
/* std::complex<double>(0., r) * k */
static inline cplx ir_times_k(double r, cplx k) { return cmul(C(0., r), k); }

static void kernel(int f, const double* x, const double* y, const orc_kernel_param* kp, cplx* res) {
    double r = vdist(x, y);
    double R = kp->smoothR;
    double t = r / R;
    cplx k = kval(kp);
    switch (f) {
    case F_SIMPLEPOT_L: {            /* kernel_lib.h:62-69 */
        res[0] = C((t < 1) * (t * (3. - 2. * t) / R) + (t >= 1) / r, 0.);
        break;
    }
    case F_GRADSIMPLEPOT_L: {        /* :87-98 */
        for (int i = 0; i < 3; i++) {
            double v = (y[i] - x[i]);
            v *= (t < 1) * (3. - 2. * t) / (r * (R * R)) + (t >= 1) / ((r * r) * r);
            res[i] = C(v, 0.);
        }
        break;
    }
    case F_DOUBLEPOT_L: {            /* :117-129 */
        double diff[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
        double v = scal(kp->n, diff);
        v *= (t < 1) * (3. - 2. * t) / (r * (R * R)) + (t >= 1) / ((r * r) * r);
        res[0] = C(v, 0.);
        break;
    }
    case F_VECTORPOT_L: {            /* :149-177 */
        if (r < MACHINE_ZERO) {
            for (int i = 0; i < 3; i++) res[i] = C(0., 0.);
        } else {
            double ff = 1. / r / r / r, diff[3], o[3];
            diff[0] = (y[0] - x[0]) * ff; diff[1] = (y[1] - x[1]) * ff; diff[2] = (y[2] - x[2]) * ff;
            vprod(diff, kp->vec_dbl, o);
            if (r >= MACHINE_ZERO && r < R) {
                double tt = r / R;
                for (int i = 0; i < 3; i++) o[i] = o[i] * (3 * tt * tt - 2 * tt * tt * tt);
            }
            for (int i = 0; i < 3; i++) res[i] = C(o[i], 0.);
        }
        break;
    }
    case F_SIMPLEPOT_H:              /* :209-217 */
    case NUF_SIMPLEPOT_H: {          /* :235-245 */
        cplx e = cexp_(ir_times_k(r, k));
        e = cmuld(e, (t < 1) * (t * (3. - 2. * t) / R) + (t >= 1) / r);
        if (f == F_SIMPLEPOT_H) res[0] = e;
        else for (int i = 0; i < 3; i++) res[i] = cmuld(e, kp->vec_dbl[i]);
        break;
    }
    case F_SIMPLEPOT_HM1: {          /* :262-278 */
        cplx ikr = ir_times_k(r, k);
        if (cabs_(ikr) < 1e-7) {
            res[0] = csub(cmul(C(0., 1.), k), cdivd(cmuld(cmul(k, k), r), 2.));
        } else {
            cplx e = csub(cexp_(ikr), C(1., 0.));
            res[0] = cmuld(e, (t < 1) * (t * (3. - 2. * t) / R) + (t >= 1) / r);
        }
        break;
    }
    case F_GRADSIMPLEPOT_H: {        /* :295-313 */
        double diff[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
        cplx ikr = ir_times_k(r, k);
        cplx e = cexp_(ikr);
        cplx m = csub(ikr, C(1., 0.));
        double s = (t < 1) * (3. - 2. * t) / (r * (R * R)) + (t >= 1) / ((r * r) * r);
        for (int i = 0; i < 3; i++) res[i] = cmuld(cmul(cmuld(e, diff[i]), m), s);
        break;
    }
    case F_GRADSIMPLEPOT_HM1: {      /* :332-356 */
        double ikr_abs = cabs_(ir_times_k(r, k));
        if (ikr_abs < 1e-6) {
            for (int i = 0; i < 3; i++) res[i] = C(0., 0.);
        } else if (ikr_abs < 1e-4) {
            for (int i = 0; i < 3; i++)
                res[i] = cdivd(cmul(cmuld(k, y[i] - x[i]), k), 2. * r);
        } else {
            cplx ikr = ir_times_k(r, k);
            cplx q = cadd(C(1., 0.), cmul(csub(ikr, C(1., 0.)), cexp_(ikr)));
            for (int i = 0; i < 3; i++) res[i] = cdivd(cmuld(q, x[i] - y[i]), r * r * r);
        }
        break;
    }
    case F_DOUBLEPOT_H:              /* :375-390 */
    case NF_DOUBLEPOT_H: {           /* :409-419 */
        double diff[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
        cplx ikr = ir_times_k(r, k);
        cplx v = cmul(cmuld(cexp_(ikr), scal(kp->n, diff)), csub(ikr, C(1., 0.)));
        v = cmuld(v, (t < 1) * (3 - 2 * t) / (r * (R * R)) + (t >= 1) / ((r * r) * r));
        if (f == F_DOUBLEPOT_H) res[0] = v;
        else for (int i = 0; i < 3; i++) res[i] = cmuld(v, kp->n[i]);
        break;
    }
    case F_ROTROTSIMPLEPOT_H: {      /* :453-492 */
        if (r < MACHINE_ZERO) {
            for (int i = 0; i < 3; i++) res[i] = C(0., 0.);
            break;
        }
        cplx ik = cmul(C(0., 1.), k);
        cplx k2 = cmul(k, k);
        double d[3];
        for (int i = 0; i < 3; i++) d[i] = (x[i] - y[i]) / r;
        double r2 = r * r;
        cplx f1 = cadd(cadd(C(-1.0 / r2 / r, 0.), cdivd(ik, r2)), cdivd(k2, r));
        cplx f2 = csub(csub(C(3.0 / r2 / r, 0.), cdivd(cmuld(ik, 3.0), r2)), cdivd(k2, r));
        cplx e = cexp_(cmuld(ik, r));
        const cplx cur[3] = {C(kp->current[0], kp->current[1]), C(kp->current[2], kp->current[3]),
                             C(kp->current[4], kp->current[5])};
        /* scal_prod(double*, complex*) */
        cplx dj = cadd(cadd(cmuld(cur[0], d[0]), cmuld(cur[1], d[1])), cmuld(cur[2], d[2]));
        f2 = cmul(f2, dj);
        for (int i = 0; i < 3; i++) res[i] = cmul(cadd(cmuld(f2, d[i]), cmul(cur[i], f1)), e);
        if (r < R) {
            double tt = r / R;
            for (int i = 0; i < 3; i++) res[i] = cmuld(res[i], tt * tt * (3 - 2 * tt));
        }
        break;
    }
    default:
        for (int i = 0; i < 3; i++) res[i] = C(NAN, NAN);
    }
}

int orc_kernel(int f, const double* x, const double* y, const orc_kernel_param* kp, double* res6) {
    if (f < 0 || f >= F_COUNT) return -1;
    cplx r[3] = {{0, 0}, {0, 0}, {0, 0}};
    kernel(f, x, y, kp, r);
    for (int i = 0; i < 3; i++) { res6[2 * i] = r[i].re; res6[2 * i + 1] = r[i].im; }
    return 0;
}

/* ---------------------------------------------------------------- quadrature */
/* Shared tail of every adaptive pass (integral_universal.h:96-112 and
 * integral_universal_seg.h:74-90).  Returns 1 when the loop must stop. */
static int pass_end(int p_n, int idim, double eps, cplx* res, cplx* prev, orc_kernel_param* kp) {
    double delta = 0.;
    for (int g = 0; g < idim; g++) {
        double a = cabs_(csub(res[g], prev[g]));
        delta += a * a;
    }
    if (delta < eps * eps && p_n != 0) return 1;
    for (int g = 0; g < idim; g++) {
        prev[g] = res[g];
        res[g] = C(0., 0.);
        if (p_n < PMAX_SMOOTH && kp->smoothR > kp->smoothRMin) kp->smoothR /= 2.;
    }
    return 0;
}

/* The integrand of a rule: ev(env, y, kp, ff). */
typedef void (*eval_fn)(const void* env, const double* y, const orc_kernel_param* kp, cplx* ff);

/* Cell rules: IntegralUniversalPnt (F(x, y) forms), quad integral_universal.h:38-120,
 * triangle :136-220, and IntegralUniversal (F(y) forms), quad :252-333, triangle
 * :347-431 - the same rules with the integrand's x fixed or absent. */
static void integral_rule(eval_fn ev, const void* env, int nv, const double (*c)[3], orc_kernel_param* kp,
                          const orc_integral_param* ip, cplx* res) {
    int n = ip->n_start, idim = ip->idim, p_n;
    cplx ff[3] = {{0, 0}, {0, 0}, {0, 0}}, prev[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (int g = 0; g < idim; g++) res[g] = C(0., 0.);
    double s_tri = (nv == 3) ? tr_square(c[0], c[1], c[2]) : 0.;
    for (p_n = 0; p_n < ip->p_max; p_n++) {
        if (nv == 4) {
            for (int i = 0; i < n; i++) {
                for (int j = 0; j < n; j++) {
                    double p = (double)i / (double)n, q = (double)j / (double)n;
                    double p1 = ((double)i + 1.) / (double)n, q1 = ((double)j + 1.) / (double)n;
                    double a[3], b[3], a1[3], a2[3], a3[3], a4[3], rc[3], m1[3], m2[3], rn[3];
                    for (int k = 0; k < 3; k++) {
                        a[k] = q * c[1][k] + (1 - q) * c[0][k];
                        b[k] = q * c[2][k] + (1 - q) * c[3][k];
                        a1[k] = p * b[k] + (1 - p) * a[k];
                        a4[k] = p1 * b[k] + (1 - p1) * a[k];
                    }
                    for (int k = 0; k < 3; k++) {
                        a[k] = q1 * c[1][k] + (1 - q1) * c[0][k];
                        b[k] = q1 * c[2][k] + (1 - q1) * c[3][k];
                        a2[k] = p * b[k] + (1 - p) * a[k];
                        a3[k] = p1 * b[k] + (1 - p1) * a[k];
                    }
                    for (int k = 0; k < 3; k++) {
                        rc[k] = (a1[k] + a2[k] + a3[k] + a4[k]) / 4.0;
                        m1[k] = ((a2[k] + a3[k]) - (a1[k] + a4[k])) / 2.0;
                        m2[k] = ((a3[k] + a4[k]) - (a1[k] + a2[k])) / 2.0;
                    }
                    vprod(m1, m2, rn);
                    double s = vlen(rn);
                    ev(env, rc, kp, ff);
                    for (int g = 0; g < idim; g++) res[g] = cadd(res[g], cmuld(ff[g], s));
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
                    double B[3], AM[3], CM[3], AP[3], CP[3], rc[3];
                    for (int k = 0; k < 3; k++) {
                        B[k] = c[0][k] + (i + 1) * pv[k] + j * qv[k];
                        AM[k] = B[k] - pv[k];
                        CM[k] = AM[k] + qv[k];
                        rc[k] = (B[k] + AM[k] + CM[k]) / 3.0;
                    }
                    ev(env, rc, kp, ff);
                    for (int g = 0; g < idim; g++) res[g] = cadd(res[g], ff[g]);
                    if (j < n - 1 - i) {
                        for (int k = 0; k < 3; k++) {
                            AP[k] = B[k] + rv[k];
                            CP[k] = B[k] + qv[k];
                            rc[k] = (B[k] + AP[k] + CP[k]) / 3.0;
                        }
                        ev(env, rc, kp, ff);
                        for (int g = 0; g < idim; g++) res[g] = cadd(res[g], ff[g]);
                    }
                }
            }
            for (int g = 0; g < idim; g++) res[g] = cdivd(cdivd(cmuld(res[g], s_tri), (double)n), (double)n);
        }
        if (pass_end(p_n, idim, ip->eps, res, prev, kp)) break;
        n = n * 2;
    }
    if (p_n == ip->p_max)
        for (int g = 0; g < idim; g++) res[g] = prev[g];
}

/* F(x, y) integrand with x fixed */
typedef struct {
    int f;
    const double* x;
} fxy_env;
static void eval_fxy(const void* env, const double* y, const orc_kernel_param* kp, cplx* ff) {
    const fxy_env* e = (const fxy_env*)env;
    kernel(e->f, e->x, y, kp, ff);
}

/* IntegralUniversalPnt, quad integral_universal.h:38-120, triangle :136-220 */
static void integral_pnt(int f, int nv, const double (*c)[3], const double* x, orc_kernel_param* kp,
                         const orc_integral_param* ip, cplx* res) {
    fxy_env e = {f, x};
    integral_rule(eval_fxy, &e, nv, c, kp, ip, res);
}

/* Segment rule: IntegralUniversalSegPnt, F(x, y) form integral_universal_seg.h:38-98,
 * F(y) form :120-179 */
static void integral_seg_rule(eval_fn ev, const void* env, const double* a, const double* b,
                              orc_kernel_param* kp, const orc_integral_param* ip, cplx* res) {
    int n = ip->n_start, idim = ip->idim, p_n;
    cplx ff[3] = {{0, 0}, {0, 0}, {0, 0}}, prev[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (int g = 0; g < idim; g++) res[g] = C(0., 0.);
    for (p_n = 0; p_n < ip->p_max; p_n++) {
        double d[3], y[3];
        for (int j = 0; j < 3; j++) d[j] = (b[j] - a[j]) / n;
        double dl = vlen(d);
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < 3; j++) y[j] = a[j] + (i + 0.5) * d[j];
            ev(env, y, kp, ff);
            for (int g = 0; g < idim; g++) res[g] = cadd(res[g], cmuld(ff[g], dl));
        }
        if (pass_end(p_n, idim, ip->eps, res, prev, kp)) break;
        n = n * 2;
    }
    if (p_n == ip->p_max)
        for (int g = 0; g < idim; g++) res[g] = prev[g];
}

static void integral_seg(int f, const double* a, const double* b, const double* x, orc_kernel_param* kp,
                         const orc_integral_param* ip, cplx* res) {
    fxy_env e = {f, x};
    integral_seg_rule(eval_fxy, &e, a, b, kp, ip, res);
}

static void put3(const cplx* r, int n, double* out) {
    for (int i = 0; i < n; i++) { out[2 * i] = r[i].re; out[2 * i + 1] = r[i].im; }
}

int orc_integral_pnt(int f, int nv, const double* cell, const double* x, orc_kernel_param* kp,
                     const orc_integral_param* ip, double* res6) {
    if (f < 0 || f >= F_COUNT || (nv != 3 && nv != 4) || ip->idim < 1 || ip->idim > 3) return -1;
    cplx r[3] = {{0, 0}, {0, 0}, {0, 0}};
    integral_pnt(f, nv, (const double (*)[3])cell, x, kp, ip, r);
    put3(r, 3, res6);
    return 0;
}

int orc_integral_seg(int f, const double* a, const double* b, const double* x, orc_kernel_param* kp,
                     const orc_integral_param* ip, double* res6) {
    if (f < 0 || f >= F_COUNT || ip->idim < 1 || ip->idim > 3) return -1;
    cplx r[3] = {{0, 0}, {0, 0}, {0, 0}};
    integral_seg(f, a, b, x, kp, ip, r);
    put3(r, 3, res6);
    return 0;
}

/* ---------------------------------------------------------------- analytic
 * integrators/integrals_analytic.cpp */
static void lartgp(double x, double y, double* cs, double* sn, double* r) {   /* :15-26 */
    double tmp = sqrt(x * x + y * y);
    if (tmp == 0) { *r = 0; *cs = 1; *sn = 0; }
    else { *cs = x / tmp; *sn = y / tmp; *r = tmp; }
}
static void rot(double* x, double* y, double cs, double sn) {   /* :29-34 */
    double tmp = *x * cs + *y * sn;
    *y = *y * cs - *x * sn;
    *x = tmp;
}

/* _Integral1Divr_inv :37-200.  abc[coord][vertex], d = observation point. */
static double int1divr_tri(const double abc[3][3], const double* d) {
    double p0, l, l_pl, ratio, r, ad, proj;
    double n[3], e[3][3], Q[3][2], D[7][3], cs[3], sn[3];
    for (int i = 0; i < 3; ++i) {
        e[i][0] = abc[i][1] - abc[i][0];
        e[i][1] = abc[i][2] - abc[i][1];
        e[i][2] = abc[i][0] - abc[i][2];
    }
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) D[j][i] = abc[j][i] - d[j];
        D[3][i] = sqrt(D[0][i] * D[0][i] + D[1][i] * D[1][i] + D[2][i] * D[2][i]);
        Q[i][0] = e[i][0];
        Q[i][1] = e[i][1];
        n[i] = 0;
    }
    D[4][0] = D[3][1]; D[4][1] = D[3][2]; D[4][2] = D[3][0];
    lartgp(Q[0][0], Q[1][0], &cs[0], &sn[0], &Q[0][0]);
    lartgp(Q[0][0], Q[2][0], &cs[1], &sn[1], &Q[0][0]);
    rot(&Q[0][1], &Q[1][1], cs[0], sn[0]);
    rot(&Q[0][1], &Q[2][1], cs[1], sn[1]);
    lartgp(Q[1][1], Q[2][1], &cs[2], &sn[2], &Q[1][1]);
    n[2] = 1.0;
    rot(&n[1], &n[2], cs[2], -sn[2]);
    rot(&n[0], &n[2], cs[1], -sn[1]);
    rot(&n[0], &n[1], cs[0], -sn[0]);
    lartgp(n[0], n[1], &cs[0], &sn[0], &n[0]);
    for (int i = 0; i < 3; ++i) {
        rot(&e[0][i], &e[1][i], cs[0], sn[0]);
        rot(&D[0][i], &D[1][i], cs[0], sn[0]);
    }
    lartgp(n[0], n[2], &cs[0], &sn[0], &n[0]);
    for (int i = 0; i < 3; ++i) {
        rot(&e[0][i], &e[2][i], cs[0], sn[0]);
        rot(&D[0][i], &D[2][i], cs[0], sn[0]);
    }
    lartgp(e[1][0], e[2][0], &cs[0], &sn[0], &e[1][0]);
    lartgp(e[1][1], e[2][1], &cs[1], &sn[1], &e[1][1]);
    lartgp(e[1][2], e[2][2], &cs[2], &sn[2], &e[1][2]);
    D[5][0] = cs[0] * D[1][1] + sn[0] * D[2][1];
    D[5][1] = cs[1] * D[1][2] + sn[1] * D[2][2];
    D[5][2] = cs[2] * D[1][0] + sn[2] * D[2][0];
    D[6][0] = -sn[0] * D[1][1] + cs[0] * D[2][1];
    D[6][1] = -sn[1] * D[1][2] + cs[1] * D[2][2];
    D[6][2] = -sn[2] * D[1][0] + cs[2] * D[2][0];
    rot(&D[1][0], &D[2][0], cs[0], sn[0]);
    rot(&D[1][1], &D[2][1], cs[1], sn[1]);
    rot(&D[1][2], &D[2][2], cs[2], sn[2]);
    double int_1r = 0;
    proj = 0;
    for (int ed = 0; ed < 3; ++ed) {
        ad = 0;
        r = D[3][ed];
        if (fabs(r) > 0 && fabs(D[4][ed]) > 0) {
            p0 = -D[2][ed] / r;
            l = D[1][ed] / r;
            ratio = D[4][ed] / r;
            l_pl = D[5][ed] / D[4][ed];
            if (l > -1 && l_pl > -1) ad = p0 * log(ratio * (1 + l_pl) / (1 + l));
            int emn = (ed + 2) % 3;
            proj = proj + (D[0][ed] * D[0][ed] + D[1][ed] * D[5][ed] + D[2][ed] * D[6][ed]) * D[3][emn];
            int_1r = int_1r + r * ad;
        }
    }
    cs[0] = fabs(D[0][0]) * Q[0][0] * Q[1][1];
    cs[1] = D[3][0] * D[3][1] * D[3][2] + proj;
    if (cs[0] > 0 || fabs(cs[1]) > 0) int_1r = int_1r - fabs(D[0][0]) * 2 * atan2(cs[0], cs[1]);
    return int_1r;
}

static double int1divr_t(const double (*t)[3], const double* x) {   /* :206-213 */
    double abc[3][3];
    for (int i = 0; i < 3; i++)
        for (int v = 0; v < 3; v++) abc[i][v] = t[v][i];
    return int1divr_tri(abc, x);
}

static double integral1divr(int nv, const double (*c)[3], const double* x) {   /* :218-235 */
    if (nv == 3) return int1divr_t(c, x);
    double t1[3][3], t2[3][3];
    for (int i = 0; i < 3; i++) {
        t1[0][i] = c[0][i]; t1[1][i] = c[1][i]; t1[2][i] = c[2][i];
        t2[0][i] = c[2][i]; t2[1][i] = c[3][i]; t2[2][i] = c[0][i];
    }
    return int1divr_t((const double (*)[3])t1, x) + int1divr_t((const double (*)[3])t2, x);
}

double orc_integral1divr(int nv, const double* cell, const double* x) {
    return integral1divr(nv, (const double (*)[3])cell, x);
}

void orc_integral_grad1divr_seg(const double* A, const double* B, const double* x, double* res) {
    /* integralGrad1DivrSeg :240-257 */
    double xmA[3] = {x[0] - A[0], x[1] - A[1], x[2] - A[2]};
    double xmB[3] = {x[0] - B[0], x[1] - B[1], x[2] - B[2]};
    double xA = vdist(x, A), xB = vdist(x, B), AB = vdist(A, B);
    for (int i = 0; i < 3; i++)
        res[i] = -AB * (xmA[i] / xA + xmB[i] / xB) / (scal(xmA, xmB) + xA * xB);
}

/* IntegralSurf_SimplePot_H_HS integrals_analytic.h:136-164 */
static cplx surf_hs(int nv, const double (*c)[3], const double* x, double analyticR,
                    const orc_integral_param* ip, cplx k) {
    double y[3];
    center_mass(nv, c, y);
    orc_kernel_param kp;
    memset(&kp, 0, sizeof kp);
    kp.k_re = k.re; kp.k_im = k.im;
    kp.smoothR = MACHINE_ZERO; kp.smoothRMin = MACHINE_ZERO;
    double calc_dist = analyticR * get_diam(nv, c);
    double r = vdist(x, y);
    cplx tmp[3], ff;
    if (r < calc_dist) {
        ff = C(integral1divr(nv, c, x), 0.);
        integral_pnt(F_SIMPLEPOT_HM1, nv, c, x, &kp, ip, tmp);
        ff = cadd(ff, tmp[0]);
    } else {
        integral_pnt(F_SIMPLEPOT_H, nv, c, x, &kp, ip, tmp);
        ff = tmp[0];
    }
    return ff;
}

int orc_surf_hs(int nv, const double* cell, const double* x, double analyticR,
                const orc_integral_param* ip, double kre, double kim, double* res2) {
    if (nv != 3 && nv != 4) return -1;
    cplx r = surf_hs(nv, (const double (*)[3])cell, x, analyticR, ip, C(kre, kim));
    res2[0] = r.re; res2[1] = r.im;
    return 0;
}

/* IntegralSeg_GradSimplePot_H_HS integrals_analytic.h:193-221 */
static void seg_hs(const double* A, const double* B, const double* x, double analyticR,
                   const orc_integral_param* ip, cplx k, cplx* res) {
    orc_kernel_param kp;
    memset(&kp, 0, sizeof kp);
    kp.k_re = k.re; kp.k_im = k.im;
    kp.smoothR = MACHINE_ZERO;
    kp.smoothRMin = 0.;   /* uninitialised in the reference */
    double y[3] = {(B[0] + A[0]) / 2., (B[1] + A[1]) / 2., (B[2] + A[2]) / 2.};
    double calc_dist = analyticR * vdist(A, B);
    double r = vdist(x, y);
    if (r < calc_dist) {
        double tmp[3] = {0, 0, 0};
        orc_integral_grad1divr_seg(A, B, x, tmp);
        integral_seg(F_GRADSIMPLEPOT_HM1, A, B, x, &kp, ip, res);
        for (int i = 0; i < 3; i++) res[i] = cadd(res[i], C(tmp[i], 0.));
    } else {
        integral_seg(F_GRADSIMPLEPOT_H, A, B, x, &kp, ip, res);
    }
}

int orc_seg_hs(const double* a, const double* b, const double* x, double analyticR,
               const orc_integral_param* ip, double kre, double kim, double* res6) {
    cplx r[3] = {{0, 0}, {0, 0}, {0, 0}};
    seg_hs(a, b, x, analyticR, ip, C(kre, kim), r);
    put3(r, 3, res6);
    return 0;
}

/* ---------------------------------------------------------------- K0
 * electrodynamics/K0.h.  variant 0: K0 (:36-86), 1: K0_Common (:112-166),
 * 2: K0_GradSimplePot_H_HS (:197-249). */
static void edge_nu(const double* A, const double* B, const double* nrm, double* nu) {
    double diff[3];
    for (int g = 0; g < 3; g++) diff[g] = B[g] - A[g];
    vprod(diff, nrm, nu);
    double len = vlen(nu);
    if (len > MACHINE_ZERO)
        for (int g = 0; g < 3; g++) nu[g] /= len;
}

static void k0(int variant, int f, int nv, const double (*c)[3], const double* x, const cplx* j,
               orc_kernel_param* kp, const orc_integral_param* ip, double analyticR, cplx k,
               cplx* res) {
    double nrm[3], nu[3];
    cplx ks[3];
    for (int g = 0; g < 3; g++) res[g] = C(0., 0.);
    norm_func(nv, c, nrm);
    for (int s = 0; s < nv; s++) {
        int A = s, B = (s != nv - 1) ? s + 1 : 0;
        for (int g = 0; g < 3; g++) ks[g] = C(0., 0.);
        if (variant == 1) {
            edge_nu(c[A], c[B], nrm, nu);
            for (int g = 0; g < 3; g++) kp->vec_dbl[g] = nu[g];
            integral_seg(f, c[A], c[B], x, kp, ip, ks);
            for (int g = 0; g < 3; g++) res[g] = cadd(res[g], ks[g]);
            continue;
        }
        if (variant == 0) integral_seg(f, c[A], c[B], x, kp, ip, ks);
        else seg_hs(c[A], c[B], x, analyticR, ip, k, ks);
        edge_nu(c[A], c[B], nrm, nu);
        cplx jn = cneg(cscal(j, nu));
        for (int g = 0; g < 3; g++) {
            ks[g] = cmul(ks[g], jn);
            res[g] = cadd(res[g], ks[g]);
        }
    }
}

static void load_j(const double* j6, cplx* j) {
    for (int i = 0; i < 3; i++) j[i] = C(j6[2 * i], j6[2 * i + 1]);
}

int orc_k0(int variant, int f, int nv, const double* cell, const double* x, const double* j6,
           orc_kernel_param* kp, const orc_integral_param* ip, double* res6) {
    if (f < 0 || f >= F_COUNT || (nv != 3 && nv != 4) || ip->idim != 3 || variant < 0 || variant > 1)
        return -1;
    cplx j[3], r[3];
    load_j(j6, j);
    k0(variant, f, nv, (const double (*)[3])cell, x, j, kp, ip, 0., C(0, 0), r);
    put3(r, 3, res6);
    return 0;
}

int orc_k0_hs(int nv, const double* cell, const double* x, const double* j6,
              const orc_integral_param* ip, double analyticR, double kre, double kim, double* res6) {
    if ((nv != 3 && nv != 4) || ip->idim != 3) return -1;
    cplx j[3], r[3];
    load_j(j6, j);
    k0(2, 0, nv, (const double (*)[3])cell, x, j, NULL, ip, analyticR, C(kre, kim), r);
    put3(r, 3, res6);
    return 0;
}

/* ---------------------------------------------------------------- operators
 * electrodynamics/K.h and R.h.  op ids:
 * 0 K_RotRot_Near_Smooth K.h:47-79     1 K_RotRot_HS K.h:119-145
 * 2 K_RotRot_HS_Colloc K.h:179-204     3 K_RotRot_HS2_Colloc K.h:230-252
 * 4 K_RotRot_Far K.h:288-301           5 R_Rot_Smooth R.h:46-61
 * 6 R_Rot_Colloc R.h:89-111            7 R_Rot_Near R.h:140-175 */
static void op_eval(int op, int nv, const double (*c)[3], const double* x, const cplx* j,
                    const orc_num_param* np, cplx k, cplx* res) {
    double diam = get_diam(nv, c);
    orc_kernel_param kp;
    memset(&kp, 0, sizeof kp);
    kp.k_re = k.re; kp.k_im = k.im;
    orc_integral_param ipF = {1, np->n_start, np->p_max, np->eps};
    orc_integral_param ipG = {3, np->n_start_seg, np->p_max_seg, np->eps};
    orc_integral_param ipS3 = {3, np->n_start, np->p_max, np->eps};
    cplx t1[3], t3[3];
    switch (op) {
    case 0: {
        kp.smoothR = np->rs * diam / np->n_start;
        kp.smoothRMin = np->eps;
        integral_pnt(F_SIMPLEPOT_H, nv, c, x, &kp, &ipF, t1);
        cplx surf = cdivd(cmul(cmul(t1[0], k), k), 4. * M_PI);
        orc_kernel_param kg;
        memset(&kg, 0, sizeof kg);
        kg.k_re = k.re; kg.k_im = k.im;
        kg.smoothR = np->rs_seg * diam / np->n_start_seg;
        kg.smoothRMin = np->eps;
        k0(0, F_GRADSIMPLEPOT_H, nv, c, x, j, &kg, &ipG, 0., k, t3);
        for (int i = 0; i < 3; i++) res[i] = cadd(cmul(surf, j[i]), cdivd(t3[i], 4. * M_PI));
        break;
    }
    case 1:
    case 2:
    case 3: {
        cplx s = surf_hs(nv, c, x, np->analytic_calc, &ipF, k);
        for (int i = 0; i < 3; i++) res[i] = cdivd(cmul(cmul(cmul(s, k), k), j[i]), 4. * M_PI);
        if (op == 3) {
            k0(2, 0, nv, c, x, j, NULL, &ipG, np->analytic_calc, k, t3);
        } else {
            if (op == 1) {
                kp.smoothR = np->rs_seg * diam / np->n_start_seg;
                kp.smoothRMin = np->eps;
            } else {
                kp.smoothR = MACHINE_ZERO;
                kp.smoothRMin = MACHINE_ZERO;
            }
            k0(0, F_GRADSIMPLEPOT_H, nv, c, x, j, &kp, &ipG, 0., k, t3);
        }
        for (int i = 0; i < 3; i++) res[i] = cadd(res[i], cdivd(t3[i], 4. * M_PI));
        break;
    }
    case 4: {
        kp.smoothR = MACHINE_ZERO;
        kp.smoothRMin = MACHINE_ZERO;
        for (int i = 0; i < 3; i++) { kp.current[2 * i] = j[i].re; kp.current[2 * i + 1] = j[i].im; }
        integral_pnt(F_ROTROTSIMPLEPOT_H, nv, c, x, &kp, &ipS3, res);
        for (int i = 0; i < 3; i++) res[i] = cdivd(res[i], 4. * M_PI);
        break;
    }
    case 5:
    case 6: {
        if (op == 6) {
            double y[3];
            center_mass(nv, c, y);
            if (vdist(x, y) < MACHINE_ZERO) {
                for (int i = 0; i < 3; i++) res[i] = C(0., 0.);
                break;
            }
            kp.smoothR = MACHINE_ZERO;
            kp.smoothRMin = MACHINE_ZERO;
        } else {
            kp.smoothR = np->rs * diam / np->n_start;
            kp.smoothRMin = diam / 100.;
        }
        integral_pnt(F_GRADSIMPLEPOT_H, nv, c, x, &kp, &ipS3, t3);
        cvprod(t3, j, res);
        for (int i = 0; i < 3; i++) res[i] = cdivd(res[i], 4. * M_PI);
        break;
    }
    case 7: {
        double nrm[3];
        norm_func(nv, c, nrm);
        kp.smoothR = np->rs * diam / np->n_start;
        kp.smoothRMin = diam / 100.;
        for (int i = 0; i < 3; i++) kp.n[i] = nrm[i];
        cplx rs_[3], rc[3], tmp[3];
        integral_pnt(NF_DOUBLEPOT_H, nv, c, x, &kp, &ipS3, rs_);
        orc_kernel_param ks;
        memset(&ks, 0, sizeof ks);
        ks.k_re = k.re; ks.k_im = k.im;
        ks.smoothR = MACHINE_ZERO; ks.smoothRMin = MACHINE_ZERO;
        k0(1, NUF_SIMPLEPOT_H, nv, c, x, j, &ks, &ipG, 0., k, rc);
        for (int i = 0; i < 3; i++) tmp[i] = cdivd(cadd(rs_[i], rc[i]), 4. * M_PI);
        cvprod(j, tmp, res);
        break;
    }
    default:
        for (int i = 0; i < 3; i++) res[i] = C(NAN, NAN);
    }
}

int orc_op(int op, int nv, const double* cell, const double* x, const double* j6,
           const orc_num_param* np, double kre, double kim, double* res6) {
    if (op < 0 || op > 7 || (nv != 3 && nv != 4)) return -1;
    cplx j[3], r[3];
    load_j(j6, j);
    op_eval(op, nv, (const double (*)[3])cell, x, j, np, C(kre, kim), r);
    put3(r, 3, res6);
    return 0;
}

int orc_op_pairs(int op, int nv, int ncell, const double* verts, const double* j6,
                 int npts, const double* pts, const orc_num_param* np, double kre, double kim,
                 double* out) {
    if (op < 0 || op > 7 || (nv != 3 && nv != 4)) return -1;
    #pragma omp parallel for schedule(static)
    for (int p = 0; p < npts; p++)
        for (int c = 0; c < ncell; c++)
            orc_op(op, nv, verts + (size_t)c * nv * 3, pts + 3 * (size_t)p, j6 + 6 * (size_t)c, np,
                   kre, kim, out + ((size_t)p * ncell + c) * 6);
    return 0;
}

/* ---------------------------------------------------------------- drivers */
/* PlaneWaveFieldE electrodynamics/incident_field.h:20-29 */
static void plane_wave_e(const double* x, const double* kvec, const double* e0, cplx* e) {
    cplx deg = cexp_(cmuld(C(0., 1.), scal(kvec, x)));
    for (int i = 0; i < 3; i++) e[i] = cmuld(deg, e0[i]);
}

int orc_plane_wave_e(int npts, const double* pts, const double* kvec, const double* e0, double* out) {
    for (int i = 0; i < npts; i++) {
        cplx e[3];
        plane_wave_e(pts + 3 * i, kvec, e0, e);
        put3(e, 3, out + 6 * i);
    }
    return 0;
}

/* PlaneWaveFieldH electrodynamics/incident_field.h:40-59:
 * (vec_dbl x e0) e^{i vec_dbl.x} / (2 pi freq) / mu, mu real (mu_im = 0) or complex */
static void plane_wave_h(const double* x, const double* kvec, const double* e0, double freq, double mu_re,
                         double mu_im, cplx* h) {
    double vp[3];
    vprod(kvec, e0, vp);
    cplx deg = cexp_(cmuld(C(0., 1.), scal(kvec, x)));
    deg = cdivd(deg, 2. * M_PI * freq);
    if (mu_im != 0.) {
        double den = mu_re * mu_re + mu_im * mu_im;
        deg = C((deg.re * mu_re + deg.im * mu_im) / den, (deg.im * mu_re - deg.re * mu_im) / den);
    } else {
        deg = cdivd(deg, mu_re);
    }
    for (int i = 0; i < 3; i++) h[i] = cmuld(deg, vp[i]);
}

int orc_plane_wave_h(int npts, const double* pts, const double* kvec, const double* e0, double freq,
                     double mu_re, double mu_im, double* out) {
    for (int i = 0; i < npts; i++) {
        cplx h[3];
        plane_wave_h(pts + 3 * i, kvec, e0, freq, mu_re, mu_im, h);
        put3(h, 3, out + 6 * i);
    }
    return 0;
}

/* F(y) integrands: 0 the constant 1 (examples/integrators/main0_integrals.cpp:77-81),
 * 1 PlaneWaveFieldE with kp->vec_dbl as the wave vector */
typedef struct {
    int f;
    const double* e0;
} fy_env;
static void eval_fy(const void* env, const double* y, const orc_kernel_param* kp, cplx* ff) {
    const fy_env* e = (const fy_env*)env;
    if (e->f == 0) {
        ff[0] = C(1., 0.);
        ff[1] = ff[2] = C(0., 0.);
    } else {
        plane_wave_e(y, kp->vec_dbl, e->e0, ff);
    }
}

/* IntegralUniversal (F(y) forms) over a cell (nv = 4, 3) or the F(y) form of
 * IntegralUniversalSegPnt over a segment (nv = 2, cell = a then b). */
int orc_integral_y(int f, int nv, const double* cell, const double* e0, orc_kernel_param* kp,
                   const orc_integral_param* ip, double* res6) {
    if (f < 0 || f > 1 || nv < 2 || nv > 4 || ip->idim < 1 || ip->idim > 3) return -1;
    static const double zero[3] = {0., 0., 0.};
    fy_env e = {f, e0 ? e0 : zero};
    cplx r[3] = {{0, 0}, {0, 0}, {0, 0}};
    if (nv == 2) integral_seg_rule(eval_fy, &e, cell, cell + 3, kp, ip, r);
    else integral_rule(eval_fy, &e, nv, (const double (*)[3])cell, kp, ip, r);
    put3(r, 3, res6);
    return 0;
}

/* GetFieldE_Ideal electrodynamics/get_field.h:35-78 */
int orc_get_field_e(int nv, int ncell, const double* verts, const double* jE,
                    double mre, double mim, const double* kvec, const double* e0,
                    double kre, double kim, const orc_num_param* np, int npts,
                    const double* pts, double* out) {
    if (nv != 3 && nv != 4) return -1;
    cplx k = C(kre, kim), mult = C(mre, mim);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < npts; i++) {
        const double* x = pts + 3 * (size_t)i;
        cplx fe[3] = {{0, 0}, {0, 0}, {0, 0}}, tmp[3], j[3];
        for (int c = 0; c < ncell; c++) {
            const double (*cell)[3] = (const double (*)[3])(verts + (size_t)c * nv * 3);
            double y[3];
            center_mass(nv, cell, y);
            load_j(jE + 6 * (size_t)c, j);
            int near = vdist(x, y) < np->analytic_calc * get_diam(nv, cell);
            op_eval(near ? 0 : 4, nv, cell, x, j, np, k, tmp);
            for (int g = 0; g < 3; g++) fe[g] = cadd(fe[g], tmp[g]);
        }
        for (int g = 0; g < 3; g++) fe[g] = cmul(fe[g], mult);
        cplx ei[3];
        plane_wave_e(x, kvec, e0, ei);
        for (int g = 0; g < 3; g++) fe[g] = cadd(fe[g], ei[g]);
        put3(fe, 3, out + 6 * (size_t)i);
    }
    return 0;
}

/* Fused Stratton-Chu field driver of this library (no single reference
 * function): for every point x and cell j,
 *   near_mode 0 (electrodynamics/get_field.h:58-64): near iff
 *     dist(x, centroid_j) < analytic_calc * diam_j; near K = K_RotRot_Near_Smooth,
 *     near R = R_Rot_Smooth; far K = K_RotRot_Far, far R = R_Rot_Colloc.
 *   near_mode 1 (examples/field/main0_field_test.cpp:237-311): far iff
 *     dist(x, centroid_j) > analytic_calc * grid_step (max cell diameter);
 *     near K = K_RotRot_HS, near R = R_Rot_Smooth.
 *   E = multE * sum K[jE] - sum R[jM] (+ E_inc),  H = multH * sum K[jM] + sum R[jE]
 *   (main0_field_test.cpp:313-319).  jM may be NULL (perfect conductor). */
int orc_field_eh(int nv, int ncell, const double* verts, const double* jE, const double* jM,
                 double multE_re, double multE_im, double multH_re, double multH_im,
                 int add_incident, const double* kvec, const double* e0,
                 double kre, double kim, const orc_num_param* np, int near_mode,
                 int npts, const double* pts, double* E, double* H) {
    if (nv != 3 && nv != 4) return -1;
    cplx k = C(kre, kim), mE = C(multE_re, multE_im), mH = C(multH_re, multH_im);
    double grid_step = 0.;
    for (int c = 0; c < ncell; c++) {
        double d = get_diam(nv, (const double (*)[3])(verts + (size_t)c * nv * 3));
        if (d > grid_step) grid_step = d;
    }
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < npts; i++) {
        const double* x = pts + 3 * (size_t)i;
        cplx sKE[3] = {{0, 0}, {0, 0}, {0, 0}}, sRE[3] = {{0, 0}, {0, 0}, {0, 0}};
        cplx sKH[3] = {{0, 0}, {0, 0}, {0, 0}}, sRH[3] = {{0, 0}, {0, 0}, {0, 0}};
        cplx tmp[3], je[3], jm[3];
        for (int c = 0; c < ncell; c++) {
            const double (*cell)[3] = (const double (*)[3])(verts + (size_t)c * nv * 3);
            double y[3];
            center_mass(nv, cell, y);
            int near;
            if (near_mode == 0) near = vdist(x, y) < np->analytic_calc * get_diam(nv, cell);
            else near = !(vdist(x, y) > np->analytic_calc * grid_step);
            int opK = near ? (near_mode == 0 ? 0 : 1) : 4;
            int opR = near ? 5 : 6;
            load_j(jE + 6 * (size_t)c, je);
            op_eval(opK, nv, cell, x, je, np, k, tmp);
            for (int g = 0; g < 3; g++) sKE[g] = cadd(sKE[g], tmp[g]);
            op_eval(opR, nv, cell, x, je, np, k, tmp);
            for (int g = 0; g < 3; g++) sRH[g] = cadd(sRH[g], tmp[g]);
            if (jM) {
                load_j(jM + 6 * (size_t)c, jm);
                op_eval(opR, nv, cell, x, jm, np, k, tmp);
                for (int g = 0; g < 3; g++) sRE[g] = cadd(sRE[g], tmp[g]);
                op_eval(opK, nv, cell, x, jm, np, k, tmp);
                for (int g = 0; g < 3; g++) sKH[g] = cadd(sKH[g], tmp[g]);
            }
        }
        cplx e[3], h[3], ei[3];
        for (int g = 0; g < 3; g++) {
            e[g] = csub(cmul(mE, sKE[g]), sRE[g]);
            h[g] = cadd(cmul(mH, sKH[g]), sRH[g]);
        }
        if (add_incident) {
            plane_wave_e(x, kvec, e0, ei);
            for (int g = 0; g < 3; g++) e[g] = cadd(e[g], ei[g]);
        }
        put3(e, 3, E + 6 * (size_t)i);
        put3(h, 3, H + 6 * (size_t)i);
    }
    return 0;
}

/* Quadrature nodes and weights of one cell at partition n, in the order and
 * with the weights of IntegralUniversalPnt's first pass
 * (integral_universal.h:38-120 quads: sub-quad centroid, |m1 x m2|;
 * :136-220 triangles: sub-triangle centroids, S / n^2). */
static int cell_nodes(int nv, const double (*c)[3], int n, double* y, double* w) {
    int cnt = 0;
    if (nv == 4) {
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n; j++) {
                double p = (double)i / (double)n, q = (double)j / (double)n;
                double p1 = ((double)i + 1.) / (double)n, q1 = ((double)j + 1.) / (double)n;
                double a[3], b[3], a1[3], a2[3], a3[3], a4[3], m1[3], m2[3], rn[3];
                for (int k = 0; k < 3; k++) {
                    a[k] = q * c[1][k] + (1 - q) * c[0][k];
                    b[k] = q * c[2][k] + (1 - q) * c[3][k];
                    a1[k] = p * b[k] + (1 - p) * a[k];
                    a4[k] = p1 * b[k] + (1 - p1) * a[k];
                    a[k] = q1 * c[1][k] + (1 - q1) * c[0][k];
                    b[k] = q1 * c[2][k] + (1 - q1) * c[3][k];
                    a2[k] = p * b[k] + (1 - p) * a[k];
                    a3[k] = p1 * b[k] + (1 - p1) * a[k];
                }
                for (int k = 0; k < 3; k++) {
                    y[3 * cnt + k] = (a1[k] + a2[k] + a3[k] + a4[k]) / 4.0;
                    m1[k] = ((a2[k] + a3[k]) - (a1[k] + a4[k])) / 2.0;
                    m2[k] = ((a3[k] + a4[k]) - (a1[k] + a2[k])) / 2.0;
                }
                vprod(m1, m2, rn);
                w[cnt++] = vlen(rn);
            }
        }
        return cnt;
    }
    double pv[3], qv[3], rv[3];
    const double wt = tr_square(c[0], c[1], c[2]) / n / n;
    for (int k = 0; k < 3; k++) {
        pv[k] = (c[1][k] - c[0][k]) / n;
        qv[k] = (c[2][k] - c[0][k]) / n;
        rv[k] = (c[2][k] - c[1][k]) / n;
    }
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n - i; j++) {
            double B[3], AM[3], CM[3];
            for (int k = 0; k < 3; k++) {
                B[k] = c[0][k] + (i + 1) * pv[k] + j * qv[k];
                AM[k] = B[k] - pv[k];
                CM[k] = AM[k] + qv[k];
                y[3 * cnt + k] = (B[k] + AM[k] + CM[k]) / 3.0;
            }
            w[cnt++] = wt;
            if (j < n - 1 - i) {
                for (int k = 0; k < 3; k++) y[3 * cnt + k] = (B[k] + (B[k] + rv[k]) + (B[k] + qv[k])) / 3.0;
                w[cnt++] = wt;
            }
        }
    }
    return cnt;
}

/* Far-field pattern of the Stratton-Chu field.  No reference counterpart
 * (the reference evaluates the field at finite points only, get_field.h:35-78);
 * this is the R -> infinity limit of its far operators with the same cell
 * quadrature:  F(d) = lim R e^{-ikR} E(R d), d a unit vector,
 *   K_RotRot_Far  (kernel_lib.h:453-492): f1 -> k^2/r, f2 -> -k^2/r, so
 *                 K[j] -> k^2/(4 pi) int e^{-ik d.y} (j - d (d.j)),
 *   R_Rot_Colloc  (kernel_lib.h:295-313): (ikr-1) e^{ikr} (x-y)/r^3 -> ik d e^{ikr}/r,
 *                 so R[j] -> ik/(4 pi) d x int e^{-ik d.y} j,
 * with e^{ik|x-y|}/|x-y| = e^{ikR}/R e^{-ik d.y} (1 + O(1/R)).  Then
 *   E_inf = multE K_inf[jE] - R_inf[jM],  H_inf = multH K_inf[jM] + R_inf[jE]
 * (main0_field_test.cpp:313-319).  Einf, Hinf: [ndir][3] complex. */
int orc_far_pattern(int nv, int ncell, const double* verts, const double* jE, const double* jM,
                    double multE_re, double multE_im, double multH_re, double multH_im,
                    double kre, double kim, int n_start, int ndir, const double* dirs, double* Einf,
                    double* Hinf) {
    if ((nv != 3 && nv != 4) || n_start < 1) return -1;
    const int nn = n_start * n_start;
    double* y = (double*)malloc(sizeof(double) * 3 * (size_t)nn * ncell);
    double* w = (double*)malloc(sizeof(double) * (size_t)nn * ncell);
    if (!y || !w) { free(y); free(w); return -2; }
    for (int c = 0; c < ncell; c++)
        cell_nodes(nv, (const double (*)[3])(verts + (size_t)c * nv * 3), n_start, y + 3 * (size_t)c * nn,
                   w + (size_t)c * nn);
    const cplx k = C(kre, kim), mE = C(multE_re, multE_im), mH = C(multH_re, multH_im);
    const cplx ik = C(-kim, kre), k2 = cmul(k, k);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < ndir; i++) {
        const double* d = dirs + 3 * (size_t)i;
        cplx Pe[3] = {{0, 0}, {0, 0}, {0, 0}}, Pm[3] = {{0, 0}, {0, 0}, {0, 0}}, je[3], jm[3];
        for (int c = 0; c < ncell; c++) {
            cplx S = C(0., 0.);
            for (int q = 0; q < nn; q++) {
                const double* yq = y + 3 * ((size_t)c * nn + q);
                const double dy = d[0] * yq[0] + d[1] * yq[1] + d[2] * yq[2];
                /* e^{-ik d.y} = e^{ki d.y} (cos(kr d.y) - i sin(kr d.y)) */
                const double a = exp(kim * dy) * w[(size_t)c * nn + q];
                S = cadd(S, C(a * cos(kre * dy), -a * sin(kre * dy)));
            }
            load_j(jE + 6 * (size_t)c, je);
            for (int g = 0; g < 3; g++) Pe[g] = cadd(Pe[g], cmul(S, je[g]));
            if (jM) {
                load_j(jM + 6 * (size_t)c, jm);
                for (int g = 0; g < 3; g++) Pm[g] = cadd(Pm[g], cmul(S, jm[g]));
            }
        }
        cplx e[3], h[3], te[3], tm[3], xe[3], xm[3];
        const cplx de = cadd(cadd(cmuld(Pe[0], d[0]), cmuld(Pe[1], d[1])), cmuld(Pe[2], d[2]));
        const cplx dm = cadd(cadd(cmuld(Pm[0], d[0]), cmuld(Pm[1], d[1])), cmuld(Pm[2], d[2]));
        for (int g = 0; g < 3; g++) {
            te[g] = csub(Pe[g], cmuld(de, d[g]));   /* (I - d d) Pe */
            tm[g] = csub(Pm[g], cmuld(dm, d[g]));
        }
        /* d x P */
        xe[0] = csub(cmuld(Pe[2], d[1]), cmuld(Pe[1], d[2]));
        xe[1] = csub(cmuld(Pe[0], d[2]), cmuld(Pe[2], d[0]));
        xe[2] = csub(cmuld(Pe[1], d[0]), cmuld(Pe[0], d[1]));
        xm[0] = csub(cmuld(Pm[2], d[1]), cmuld(Pm[1], d[2]));
        xm[1] = csub(cmuld(Pm[0], d[2]), cmuld(Pm[2], d[0]));
        xm[2] = csub(cmuld(Pm[1], d[0]), cmuld(Pm[0], d[1]));
        for (int g = 0; g < 3; g++) {
            e[g] = cdivd(csub(cmul(mE, cmul(k2, te[g])), cmul(ik, xm[g])), FOUR_PI);
            h[g] = cdivd(cadd(cmul(mH, cmul(k2, tm[g])), cmul(ik, xe[g])), FOUR_PI);
        }
        put3(e, 3, Einf + 6 * (size_t)i);
        if (Hinf) put3(h, 3, Hinf + 6 * (size_t)i);
    }
    free(y);
    free(w);
    return 0;
}

/* Collocation matrix of this library (the reference names the operators but
 * has no assembly loop, electrodynamics/K.h:156-178, R.h:65-88):
 *   Z[2i+a, 2j+b] = tau_i^a . Op(tau_j^b, rkt_i, cell_j)
 * with Op = K_RotRot_HS_Colloc (which = 0) or R_Rot_Colloc (which = 1),
 * rkt and tau from CellFill, unknown order cell-major / basis-minor
 * (electrodynamics/currents.h:22-33).  Z is column-major with leading
 * dimension ld; rows row0 .. row0+nrows-1 of the full 2N x 2N matrix are
 * written to Z rows 0 .. nrows-1. */
int orc_assemble(int which, int nv, int ncell, const double* verts, const orc_num_param* np,
                 double kre, double kim, int row0, int nrows, double* Z, long long ld) {
    if ((nv != 3 && nv != 4) || which < 0 || which > 1) return -1;
    double* tab = (double*)malloc(sizeof(double) * 20 * (size_t)ncell);
    if (!tab) return -2;
    for (int c = 0; c < ncell; c++) orc_cellfill(nv, verts + (size_t)c * nv * 3, tab + 20 * (size_t)c);
    cplx k = C(kre, kim);
    #pragma omp parallel for schedule(static)
    for (int rr = 0; rr < nrows; rr++) {
        int row = row0 + rr;
        int i = row / 2, a = row % 2;
        const double* ti = tab + 20 * (size_t)i;
        const double* x = ti;              /* rkt */
        const double* tau_a = ti + 6 + 3 * a;
        for (int c = 0; c < ncell; c++) {
            const double* tc = tab + 20 * (size_t)c;
            const double (*cell)[3] = (const double (*)[3])(verts + (size_t)c * nv * 3);
            for (int b = 0; b < 2; b++) {
                const double* tb = tc + 6 + 3 * b;
                cplx j[3] = {C(tb[0], 0.), C(tb[1], 0.), C(tb[2], 0.)}, res[3];
                op_eval(which == 0 ? 2 : 6, nv, cell, x, j, np, k, res);
                cplx z = cadd(cadd(cmuld(res[0], tau_a[0]), cmuld(res[1], tau_a[1])), cmuld(res[2], tau_a[2]));
                size_t col = 2 * (size_t)c + b;
                Z[2 * ((size_t)rr + col * (size_t)ld)] = z.re;
                Z[2 * ((size_t)rr + col * (size_t)ld) + 1] = z.im;
            }
        }
    }
    free(tab);
    return 0;
}

/* rhs[2i+a] = -tau_i^a . E_inc(rkt_i)  (PEC: tangential total E vanishes) */
int orc_rhs_planewave(int nv, int ncell, const double* verts, const double* kvec,
                      const double* e0, double* rhs) {
    if (nv != 3 && nv != 4) return -1;
    for (int c = 0; c < ncell; c++) {
        double t[20];
        orc_cellfill(nv, verts + (size_t)c * nv * 3, t);
        cplx e[3];
        plane_wave_e(t, kvec, e0, e);
        for (int a = 0; a < 2; a++) {
            const double* ta = t + 6 + 3 * a;
            cplx z = cneg(cadd(cadd(cmuld(e[0], ta[0]), cmuld(e[1], ta[1])), cmuld(e[2], ta[2])));
            rhs[2 * (2 * (size_t)c + a)] = z.re;
            rhs[2 * (2 * (size_t)c + a) + 1] = z.im;
        }
    }
    return 0;
}

/* SurfaceCurrents_byBasis electrodynamics/currents.h:22-33 */
int orc_currents_by_basis(int nv, int ncell, const double* verts, const double* b, double* j) {
    if (nv != 3 && nv != 4) return -1;
    for (int c = 0; c < ncell; c++) {
        double t[20];
        orc_cellfill(nv, verts + (size_t)c * nv * 3, t);
        cplx b0 = C(b[4 * (size_t)c], b[4 * (size_t)c + 1]), b1 = C(b[4 * (size_t)c + 2], b[4 * (size_t)c + 3]);
        for (int d = 0; d < 3; d++) {
            cplx v = cadd(cmuld(b0, t[6 + d]), cmuld(b1, t[9 + d]));
            j[6 * (size_t)c + 2 * d] = v.re;
            j[6 * (size_t)c + 2 * d + 1] = v.im;
        }
    }
    return 0;
}

/* Dense complex LU with partial pivoting (pivot = max |re| + |im|, as LAPACK's
 * izamax) followed by forward/back substitution; column-major, in place:
 * A is overwritten by its factors, B by the solution.  The reference links
 * LAPACK but never calls it (CMakeLists.txt:18-19,56). */
int orc_lu_solve(int n, double* A, long long lda, double* B, int nrhs, long long ldb) {
    cplx* a = (cplx*)A;
    cplx* bm = (cplx*)B;
    #define AE(i, j) a[(size_t)(i) + (size_t)(j) * (size_t)lda]
    #define BE(i, j) bm[(size_t)(i) + (size_t)(j) * (size_t)ldb]
    for (int kk = 0; kk < n; kk++) {
        int p = kk;
        double best = -1.;
        for (int i = kk; i < n; i++) {
            double v = fabs(AE(i, kk).re) + fabs(AE(i, kk).im);
            if (v > best) { best = v; p = i; }
        }
        if (best == 0.) return kk + 1;
        if (p != kk) {
            for (int j = 0; j < n; j++) { cplx t = AE(kk, j); AE(kk, j) = AE(p, j); AE(p, j) = t; }
            for (int j = 0; j < nrhs; j++) { cplx t = BE(kk, j); BE(kk, j) = BE(p, j); BE(p, j) = t; }
        }
        cplx piv = AE(kk, kk);
        double den = piv.re * piv.re + piv.im * piv.im;
        cplx inv = C(piv.re / den, -piv.im / den);
        for (int i = kk + 1; i < n; i++) AE(i, kk) = cmul(AE(i, kk), inv);
        #pragma omp parallel for schedule(static)
        for (int j = kk + 1; j < n; j++) {
            cplx u = AE(kk, j);
            if (u.re == 0. && u.im == 0.) continue;
            for (int i = kk + 1; i < n; i++) AE(i, j) = csub(AE(i, j), cmul(AE(i, kk), u));
        }
    }
    for (int r = 0; r < nrhs; r++) {
        for (int i = 0; i < n; i++)
            for (int kk = 0; kk < i; kk++) BE(i, r) = csub(BE(i, r), cmul(AE(i, kk), BE(kk, r)));
        for (int i = n - 1; i >= 0; i--) {
            for (int kk = i + 1; kk < n; kk++) BE(i, r) = csub(BE(i, r), cmul(AE(i, kk), BE(kk, r)));
            cplx piv = AE(i, i);
            double den = piv.re * piv.re + piv.im * piv.im;
            BE(i, r) = cmul(BE(i, r), C(piv.re / den, -piv.im / den));
        }
    }
    #undef AE
    #undef BE
    return 0;
}
