// Far-field pattern F(d) = lim R e^{-ikR} E(R d) of the Stratton-Chu field
// (the R -> infinity limit of K_RotRot_Far, kernel_lib.h:453-492, and of
// R_Rot_Colloc, kernel_lib.h:295-313, over the same cell quadrature):
//   K_inf[j] = k^2/(4 pi) (I - d d) P[j],   R_inf[j] = ik/(4 pi) d x P[j],
//   P[j] = sum_c S_c(d) j_c,   S_c(d) = sum_q w_q e^{-ik d.y_q},
//   E_inf = multE K_inf[jE] - R_inf[jM],  H_inf = multH K_inf[jM] + R_inf[jE].
// The reference evaluates fields at finite points only (get_field.h:35-78);
// the pattern is new.
//
// Tile: 256 directions (one per lane) x a range of cells staged in LDS.  Per
// (direction, cell) the phase splits into the centre part -k d.c (FP64, one
// sincos per cell) and the node offsets -k d.(y - c), |k d.(y - c)| <= k diam,
// evaluated by the short polynomials (FP64: phase.h's Taylor series on |t| <= 2;
// FP32 form: cell-local FP32 offsets and sincos_near32, flushed to FP64 every
// kGroup32 nodes).  The per-set sums P_E, P_M are kept unscaled per split and
// projected once per direction by the finalize kernel.
#include "bielcc_internal.h"
#include "phase.h"

using namespace bielcc_dev;

namespace bielcc_impl {

constexpr int kPT = 256;
constexpr int kPNodes = 512;
constexpr int kPCells = 16;
constexpr int kPRec = 16;      // centre 3, jE 6, jM 6
constexpr int kPGroup32 = 8;

// accumulator layout per (split, set): Pe (6 doubles), Pm (6 doubles)
template <bool CK, bool JM, bool FP, bool P32>
__global__ __launch_bounds__(kPT) void k_far_pattern(PatternArgs a, int cells_per_split, int r0) {
    __shared__ double4 snode[kPNodes];   // FP64 nodes, or (P32) FP32 cell-local nodes in the same space
    __shared__ double scell[kPCells][kPRec];
    const int tid = threadIdx.x;
    const int di = blockIdx.x * kPT + tid;
    const bool active = di < a.ndir;
    double d0 = 0., d1 = 0., d2 = 1.;
    if (active) { d0 = a.dirs[3 * (size_t)di]; d1 = a.dirs[3 * (size_t)di + 1]; d2 = a.dirs[3 * (size_t)di + 2]; }
    const int c_begin = blockIdx.y * cells_per_split;
    const int c_end = min(a.N, c_begin + cells_per_split);
    const int nn = a.nn;
    const int cb = max(1, min(kPCells, kPNodes / nn));
    const double kr_ = a.k_re, ki_ = a.k_im;
    const float kf = (float)kr_, kif = (float)ki_;
    const float df0 = (float)d0, df1 = (float)d1, df2 = (float)d2;
    cplx Pe[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, Pm[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    for (int cs = c_begin; cs < c_end; cs += cb) {
        const int ncs = min(cb, c_end - cs);
        __syncthreads();
        if (P32) {
            float4* s32 = reinterpret_cast<float4*>(snode);
            const float4* g4 = a.nodes32 + (size_t)cs * nn;
            for (int idx = tid; idx < ncs * nn; idx += kPT) s32[idx] = g4[idx];
        } else {
            const double4* gn = reinterpret_cast<const double4*>(a.nodes) + (size_t)cs * nn;
            for (int idx = tid; idx < ncs * nn; idx += kPT) snode[idx] = gn[idx];
        }
        for (int idx = tid; idx < ncs * kPRec; idx += kPT) {
            const int cc = idx / kPRec, f = idx % kPRec;
            const size_t c = (size_t)(cs + cc);
            double v = 0.;
            if (f < 3) v = a.ctab[c * CT + CT_CENTER + f];
            else if (f < 9) v = a.jE[((size_t)r0 * a.N + c) * 6 + (f - 3)];
            else if (f < 15 && JM) v = a.jM[((size_t)r0 * a.N + c) * 6 + (f - 9)];
            scell[cc][f] = v;
        }
        __syncthreads();
        if (!active) continue;
        for (int cc = 0; cc < ncs; cc++) {
            const double* rec = scell[cc];
            const double dc = d0 * rec[0] + d1 * rec[1] + d2 * rec[2];
            double Sr = 0., Si = 0.;
            if (P32) {
                const float4* nd = reinterpret_cast<const float4*>(snode) + cc * nn;
                for (int q0 = 0; q0 < nn; q0 += kPGroup32) {
                    float sr = 0.f, si = 0.f;
                    const int q1 = min(nn, q0 + kPGroup32);
                    for (int q = q0; q < q1; q++) {
                        const float4 y = nd[q];
                        const float dy = df0 * y.x + df1 * y.y + df2 * y.z;   // d.(y - c)
                        float s, c;
                        if (FP) sincos_near32(-kf * dy, s, c);
                        else sincosf(-kf * dy, &s, &c);
                        float w = y.w;
                        if (CK) w *= __expf(kif * dy);
                        sr += w * c;
                        si += w * s;
                    }
                    Sr += sr;
                    Si += si;
                }
            } else {
                const double4* nd = snode + cc * nn;
                const PhaseRef zero{0., 0., 0., 0};
                for (int q = 0; q < nn; q++) {
                    const double4 y = nd[q];
                    const double dy = (d0 * y.x + d1 * y.y + d2 * y.z) - dc;   // d.(y - c)
                    double s, c;
                    if (FP) sincos_near(-kr_, dy, zero, s, c);
                    else sincos(-kr_ * dy, &s, &c);
                    double w = y.w;
                    if (CK) w *= exp(ki_ * dy);
                    Sr += w * c;
                    Si += w * s;
                }
            }
            // centre phase e^{-ik d.c} (FP64)
            double sc, cc_;
            sincos(-kr_ * dc, &sc, &cc_);
            if (CK) { const double amp = exp(ki_ * dc); sc *= amp; cc_ *= amp; }
            const cplx S = mk(Sr, Si) * mk(cc_, sc);
            for (int g = 0; g < 3; g++) Pe[g] += S * mk(rec[3 + 2 * g], rec[4 + 2 * g]);
            if (JM)
                for (int g = 0; g < 3; g++) Pm[g] += S * mk(rec[9 + 2 * g], rec[10 + 2 * g]);
        }
    }
    if (!active) return;
    double* out = a.part + ((size_t)blockIdx.y * a.nrhs + r0) * 12 * (size_t)a.ndir + di;
    const size_t st = a.ndir;
    for (int g = 0; g < 3; g++) {
        out[(2 * g) * st] = Pe[g].re; out[(2 * g + 1) * st] = Pe[g].im;
        out[(6 + 2 * g) * st] = Pm[g].re; out[(6 + 2 * g + 1) * st] = Pm[g].im;
    }
}

// One lane per (direction, set): sum the splits, project, scale.
__global__ __launch_bounds__(256) void k_pattern_finalize(PatternArgs a, double mEr, double mEi, double mHr,
                                                          double mHi) {
    const int di = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (di >= a.ndir) return;
    double acc[12];
    for (int t = 0; t < 12; t++) acc[t] = 0.;
    for (int sp = 0; sp < a.nsplit; sp++) {
        const double* src = a.part + ((size_t)sp * a.nrhs + r) * 12 * (size_t)a.ndir + di;
        for (int t = 0; t < 12; t++) acc[t] += src[(size_t)t * a.ndir];
    }
    const double* d = a.dirs + 3 * (size_t)di;
    cplx Pe[3], Pm[3];
    for (int g = 0; g < 3; g++) { Pe[g] = mk(acc[2 * g], acc[2 * g + 1]); Pm[g] = mk(acc[6 + 2 * g], acc[7 + 2 * g]); }
    const cplx k = mk(a.k_re, a.k_im);
    const cplx k2 = k * k, ik = mk(-a.k_im, a.k_re);
    const cplx de = Pe[0] * d[0] + Pe[1] * d[1] + Pe[2] * d[2];
    const cplx dm = Pm[0] * d[0] + Pm[1] * d[1] + Pm[2] * d[2];
    cplx te[3], tm[3], xe[3], xm[3];
    for (int g = 0; g < 3; g++) { te[g] = Pe[g] - de * d[g]; tm[g] = Pm[g] - dm * d[g]; }
    xe[0] = Pe[2] * d[1] - Pe[1] * d[2]; xe[1] = Pe[0] * d[2] - Pe[2] * d[0]; xe[2] = Pe[1] * d[0] - Pe[0] * d[1];
    xm[0] = Pm[2] * d[1] - Pm[1] * d[2]; xm[1] = Pm[0] * d[2] - Pm[2] * d[0]; xm[2] = Pm[1] * d[0] - Pm[0] * d[1];
    const cplx mE = mk(mEr, mEi), mH = mk(mHr, mHi);
    double* Eo = a.Einf + ((size_t)r * a.ndir + di) * 6;
    double* Ho = a.Hinf ? a.Hinf + ((size_t)r * a.ndir + di) * 6 : nullptr;
    for (int g = 0; g < 3; g++) {
        const cplx e = (mE * (k2 * te[g]) - ik * xm[g]) / (4. * kPi);
        Eo[2 * g] = e.re; Eo[2 * g + 1] = e.im;
        if (Ho) {
            const cplx h = (mH * (k2 * tm[g]) + ik * xe[g]) / (4. * kPi);
            Ho[2 * g] = h.re; Ho[2 * g + 1] = h.im;
        }
    }
}

template <bool CK, bool JM, bool FP, bool P32>
static void pattern_t(hipStream_t s, const PatternArgs& a) {
    const int cps = (a.N + a.nsplit - 1) / a.nsplit;
    dim3 g((a.ndir + kPT - 1) / kPT, a.nsplit), b(kPT);
    for (int r = 0; r < a.nrhs; r++) hipLaunchKernelGGL((k_far_pattern<CK, JM, FP, P32>), g, b, 0, s, a, cps, r);
}

template <bool CK, bool JM>
static void pattern_fp(hipStream_t s, const PatternArgs& a) {
    if (a.prec) { if (a.fast_phase) pattern_t<CK, JM, true, true>(s, a); else pattern_t<CK, JM, false, true>(s, a); }
    else { if (a.fast_phase) pattern_t<CK, JM, true, false>(s, a); else pattern_t<CK, JM, false, false>(s, a); }
}

hipError_t launch_far_pattern(hipStream_t s, const PatternArgs& a, double mEr, double mEi, double mHr, double mHi) {
    if (a.ndir <= 0 || a.nrhs <= 0) return hipSuccess;
    if (a.prec && !a.nodes32) return hipErrorInvalidValue;
    const bool ck = a.k_im != 0., jm = a.jM != nullptr;
    if (ck) { if (jm) pattern_fp<true, true>(s, a); else pattern_fp<true, false>(s, a); }
    else { if (jm) pattern_fp<false, true>(s, a); else pattern_fp<false, false>(s, a); }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pattern_finalize, dim3((a.ndir + 255) / 256, a.nrhs), dim3(256), 0, s, a, mEr, mEi, mHr, mHi);
    return hipGetLastError();
}

}  // namespace bielcc_impl
