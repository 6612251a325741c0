// e^{i k r} for the quadrature nodes of one cell, seen from one point.
//
// The tile kernels evaluate cos/sin of phi = k r at every node.  All nodes of
// a cell lie within diam of its centroid, so phi stays within k*diam of the
// centroid phase phi_c.  Per (point, cell) the kernels pick n = rint(phi_c*2/pi)
// once; per node they reduce t = phi - n*pi/2 with a three-piece Cody-Waite
// split of pi/2 (each piece has at most 33 significant bits, so n*piece is
// exact), and evaluate sin t and cos t by their Taylor series (coefficients
// (-1)^i/(2i+1)!, (-1)^i/(2i)!, rounded to double) on |t| <= 2.  The factor
// e^{i n pi/2} = i^n is the same for every node of the cell; since all
// accumulated sums are linear in e^{ikr}, it is applied once per cell (rot_q).
// Accuracy on |t| <= 2: |error| <= 3.5e-16 for sin and cos (checked against
// long double).  The host uses this path only when pi/4 + |k| * max diameter
// <= kPhaseRange, else the kernels call the general sincos.
#pragma once

#include "bielcc_dev.h"

namespace bielcc_impl {

constexpr double kPio2_1 = 0x1.921fb544p+0;          // pi/2, first 33 bits
constexpr double kPio2_2 = 0x1.0b4611a6p-34;         // next 33 bits
constexpr double kPio2_3 = 0x1.3198a2e037073p-69;    // remainder
constexpr double kTwoOverPi = 0.63661977236758138;
constexpr double kPhaseRange = 1.95;

struct PhaseRef {
    double o1, o2, o3;
    int q;
};

// (re, im) * i^q
__device__ inline void rot_q(double& re, double& im, int q) {
    double a = re, b = im;
    if (q == 1) { re = -b; im = a; }
    else if (q == 2) { re = -a; im = -b; }
    else if (q == 3) { re = b; im = -a; }
}

__device__ inline PhaseRef phase_ref(double phi_c) {
    const double n = rint(phi_c * kTwoOverPi);
    return PhaseRef{n * kPio2_1, n * kPio2_2, n * kPio2_3, static_cast<int>(n) & 3};
}

// sin and cos of (kr * r - n*pi/2)
__device__ inline void sincos_near(double kr, double r, const PhaseRef& p, double& s, double& c) {
    double t = fma(kr, r, -p.o1);
    t -= p.o2;
    t -= p.o3;
    const double t2 = t * t;
    double ps = -3.8681701706306841e-23;
    ps = fma(ps, t2, 1.9572941063391263e-20);
    ps = fma(ps, t2, -8.2206352466243295e-18);
    ps = fma(ps, t2, 2.8114572543455206e-15);
    ps = fma(ps, t2, -7.6471637318198164e-13);
    ps = fma(ps, t2, 1.6059043836821613e-10);
    ps = fma(ps, t2, -2.505210838544172e-08);
    ps = fma(ps, t2, 2.7557319223985893e-06);
    ps = fma(ps, t2, -0.00019841269841269841);
    ps = fma(ps, t2, 0.0083333333333333332);
    ps = fma(ps, t2, -0.16666666666666666);
    ps = ps * t2;
    s = fma(t, ps, t);
    double pc = -8.8967913924505741e-22;
    pc = fma(pc, t2, 4.1103176233121648e-19);
    pc = fma(pc, t2, -1.5619206968586225e-16);
    pc = fma(pc, t2, 4.7794773323873853e-14);
    pc = fma(pc, t2, -1.1470745597729725e-11);
    pc = fma(pc, t2, 2.08767569878681e-09);
    pc = fma(pc, t2, -2.7557319223985888e-07);
    pc = fma(pc, t2, 2.4801587301587302e-05);
    pc = fma(pc, t2, -0.0013888888888888889);
    pc = fma(pc, t2, 0.041666666666666664);
    pc = fma(pc, t2, -0.5);
    c = fma(pc, t2, 1.0);
}

// Centre-phase reference (phase mode 2).  Per (point, cell) pair the reference
// phase is h = fl(k dcen), and e^{ih} is evaluated once (quadrant reduction + the
// series above, reducing h itself); per node the phase is t = k r - h = fma(k, r, -h),
// one rounding of the exact difference, so |t| <= |k| |r - dcen| + |k dcen - h|
// <= |k| max |y - c| + 2^-53 |k dcen| (every node y lies within max|y - c| of the
// centre c).  The host picks this mode when |k| max|y - c| <= kShortRange = 0.5, where the
// Taylor series of sin through t^15 and of cos through t^14 truncate below 2e-20
// and 8e-19 (sin: 7 terms instead of 11, cos: 7 instead of 11); the accumulated
// sums, linear in e^{ikr}, are multiplied by e^{ih} once per pair (rot_c).
constexpr double kShortRange = 0.5;

struct PhaseC {
    double hi;       // h = fl(k dcen)
    double c, s;     // e^{ih}
};

__device__ inline PhaseC phase_centre(double kr, double dcen) {
    PhaseC p;
    p.hi = kr * dcen;
    const PhaseRef q = phase_ref(p.hi);
    double s, c;
    sincos_near(p.hi, 1., q, s, c);   // sin, cos of h - n pi/2
    rot_q(c, s, q.q);
    p.c = c;
    p.s = s;
    return p;
}

// sin and cos of (kr * r - h) on |t| <= kShortRange
__device__ inline void sincos_short(double kr, double r, const PhaseC& p, double& s, double& c) {
    const double t = fma(kr, r, -p.hi);
    const double t2 = t * t;
    double ps = -7.6471637318198164e-13;        // -1/15!
    ps = fma(ps, t2, 1.6059043836821613e-10);    // 1/13!
    ps = fma(ps, t2, -2.505210838544172e-08);    // -1/11!
    ps = fma(ps, t2, 2.7557319223985893e-06);    // 1/9!
    ps = fma(ps, t2, -0.00019841269841269841);   // -1/7!
    ps = fma(ps, t2, 0.0083333333333333332);     // 1/5!
    ps = fma(ps, t2, -0.16666666666666666);      // -1/3!
    ps = ps * t2;
    s = fma(t, ps, t);
    double pc = -1.1470745597729725e-11;         // -1/14!
    pc = fma(pc, t2, 2.08767569878681e-09);      // 1/12!
    pc = fma(pc, t2, -2.7557319223985888e-07);   // -1/10!
    pc = fma(pc, t2, 2.4801587301587302e-05);    // 1/8!
    pc = fma(pc, t2, -0.0013888888888888889);    // -1/6!
    pc = fma(pc, t2, 0.041666666666666664);      // 1/4!
    pc = fma(pc, t2, -0.5);
    c = fma(pc, t2, 1.0);
}

// Phase mode 3, the centre-phase reference on a shorter range, |t| <= kVShortRange = 0.25:
// sin through t^11 (Taylor; first omitted term t^13/13! < 3e-18) and cos as a degree-5
// polynomial in t^2, the Chebyshev economisation on t^2 in [0, 1/16] of the Taylor series
// through t^16 (truncation below 7e-20; every coefficient rounded to double, the constant
// kept at 1).  Evaluated in double on 24,000 points of the range the results stay within
// 1.0 ulp (sin) and 0.55 ulp (cos) of the exact values, the same as the Taylor series of
// cos through t^12, with one multiply-add fewer.
constexpr double kVShortRange = 0.25;

__device__ inline void sincos_vshort(double kr, double r, const PhaseC& p, double& s, double& c) {
    const double t = fma(kr, r, -p.hi);
    const double t2 = t * t;
    double ps = -2.505210838544172e-08;          // -1/11!
    ps = fma(ps, t2, 2.7557319223985893e-06);    // 1/9!
    ps = fma(ps, t2, -0.00019841269841269841);   // -1/7!
    ps = fma(ps, t2, 0.0083333333333333332);     // 1/5!
    ps = fma(ps, t2, -0.16666666666666666);      // -1/3!
    ps = ps * t2;
    s = fma(t, ps, t);
    double pc = -2.751820077874607e-07;
    pc = fma(pc, t2, 2.480155980239979e-05);
    pc = fma(pc, t2, -0.0013888888879978083);
    pc = fma(pc, t2, 0.04166666666665361);
    pc = fma(pc, t2, -0.49999999999999994);
    c = fma(pc, t2, 1.0);
}

// (re, im) * e^{ih}
__device__ inline void rot_c(double& re, double& im, const PhaseC& p) {
    const double a = re, b = im;
    re = a * p.c - b * p.s;
    im = a * p.s + b * p.c;
}

// e^{ikr} at one node by phase mode PH (0: general sincos, 1: pr, 2 and 3: pc)
template <int PH>
__device__ inline void phase_eval(double kr, double r, const PhaseRef& pr, const PhaseC& pc, double& s, double& c) {
    if (PH == 1) sincos_near(kr, r, pr, s, c);
    else if (PH == 2) sincos_short(kr, r, pc, s, c);
    else if (PH == 3) sincos_vshort(kr, r, pc, s, c);
    else sincos(kr * r, &s, &c);
}

// sin t and cos t on |t| <= kPhaseRange in FP32.  One more quadrant step in
// FP32 (m = rint(2t/pi) in {-1, 0, 1}; t - m pi/2 with a two-piece pi/2, the
// first subtraction exact by Sterbenz) leaves |t'| <= pi/4, where Taylor
// series through t'^9 and t'^10 truncate below 2e-9; the result is rotated
// back by m quarter turns.  Max error 6.2e-8 (sin) / 5.2e-8 (cos) over the
// range, checked against FP64 sin / cos with every operation rounded to FP32.
constexpr float kPio2Hi32 = 1.57079637e+00f;    // nearest float to pi/2
constexpr float kPio2Lo32 = -4.37113883e-08f;   // pi/2 - kPio2Hi32

__device__ inline void sincos_near32(float t, float& s, float& c) {
    const float m = rintf(t * 0.636619772f);
    float tr = fmaf(-m, kPio2Hi32, t);
    tr = fmaf(-m, kPio2Lo32, tr);
    const float t2 = tr * tr;
    float ps = 2.7557319e-06f;
    ps = fmaf(ps, t2, -1.9841270e-04f);
    ps = fmaf(ps, t2, 8.3333333e-03f);
    ps = fmaf(ps, t2, -1.6666667e-01f);
    const float s1 = fmaf(tr * t2, ps, tr);
    float pc = -2.7557319e-07f;
    pc = fmaf(pc, t2, 2.4801587e-05f);
    pc = fmaf(pc, t2, -1.3888889e-03f);
    pc = fmaf(pc, t2, 4.1666667e-02f);
    pc = fmaf(pc, t2, -0.5f);
    const float c1 = fmaf(pc, t2, 1.0f);
    const bool rot = m != 0.f;   // m = +-1: (sin, cos)(t' + m pi/2) = (m cos t', -m sin t')
    s = rot ? m * c1 : s1;
    c = rot ? -m * s1 : c1;
}


}  // namespace bielcc_impl
