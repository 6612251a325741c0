// Stratton-Chu field at M observation points over N cells.
//
// Far pairs (the near rule of electrodynamics/get_field.h:58 is false) go
// through one fused tile kernel that evaluates, per quadrature node, the
// K_RotRot_Far integrand (kernel_lib.h:453-492, K.h:288-301) and the
// F_GradSimplePot_H integrand of R_Rot_Colloc (kernel_lib.h:295-313,
// R.h:89-111), sharing r, 1/r and e^{ikr}.  Near pairs are listed and
// evaluated exactly (K_RotRot_Near_Smooth / R_Rot_Smooth or K_RotRot_HS) by
// the per-pair kernel, then everything is reduced per point in a fixed order.
//
// Tile: one workgroup = 256 points (one per lane) x a range of cells; the
// cells' node tables (sub-cell centroid + weight) are staged in LDS and read
// as wave-wide broadcasts.  Per lane and far cell the kernel accumulates
//   A = sum w e f1,  T_ab = sum w e f2 dh_a dh_b,  G = sum w (ikr-1) e d / r^3
// so that K_far[j] = (A j + T j)/4pi and R_colloc[j] = (G x j)/4pi hold for the
// electric and the magnetic current alike.
#include "bielcc_internal.h"
#include "phase.h"

using namespace bielcc_dev;

namespace bielcc_impl {

constexpr int kFT = 256;          // points per workgroup
constexpr int kFNodes = 512;      // node slots staged in LDS (16 KB)
constexpr int kFCells = 16;       // max cells per LDS stage
constexpr int kCellRec = 16;      // staged per-cell record: center 3, diam, jE 6, jM 6

// Accumulator layout of the partial sums, [NACC][M] per split:
//  0-5  sum (A j_E + T j_E)         (E from K[j_E], unscaled)
//  6-11 sum G x j_E                  (H from R[j_E], unscaled)
//  12-17 sum (A j_M + T j_M)        (H from K[j_M], unscaled)   [with j_M]
//  18-23 sum G x j_M                 (E from R[j_M], unscaled)   [with j_M]
template <bool CK, bool JM, bool WH, bool FP>
__global__ __launch_bounds__(kFT) void k_field_far(FieldArgs a, int cells_per_split) {
    __shared__ double4 snode[kFNodes];
    __shared__ double scell[kFCells][kCellRec];
    const int tid = threadIdx.x;
    const int pt = blockIdx.x * kFT + tid;
    const bool active = pt < a.M;
    double x0 = 0., x1 = 0., x2 = 0.;
    if (active) { x0 = a.pts[3 * (size_t)pt]; x1 = a.pts[3 * (size_t)pt + 1]; x2 = a.pts[3 * (size_t)pt + 2]; }
    const int c_begin = blockIdx.y * cells_per_split;
    const int c_end = min(a.N, c_begin + cells_per_split);
    const int nn = a.nn;
    const int cb = max(1, min(kFCells, kFNodes / nn));
    const double kr_ = a.k_re, ki_ = a.k_im;
    const double k2r = kr_ * kr_ - ki_ * ki_, k2i = 2. * kr_ * ki_;

    cplx E[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, H[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    cplx KH[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, RE[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};

    for (int cs = c_begin; cs < c_end; cs += cb) {
        const int ncs = min(cb, c_end - cs);
        __syncthreads();
        const double4* gn = reinterpret_cast<const double4*>(a.nodes) + (size_t)cs * nn;
        for (int idx = tid; idx < ncs * nn; idx += kFT) snode[idx] = gn[idx];
        for (int idx = tid; idx < ncs * kCellRec; idx += kFT) {
            int cc = idx / kCellRec, f = idx % kCellRec;
            const int c = cs + cc;
            double v;
            if (f < 3) v = a.ctab[(size_t)c * CT + CT_CENTER + f];
            else if (f == 3) v = a.ctab[(size_t)c * CT + CT_DIAM];
            else if (f < 10) v = a.jE[6 * (size_t)c + (f - 4)];
            else v = JM ? a.jM[6 * (size_t)c + (f - 10)] : 0.;
            scell[cc][f] = v;
        }
        __syncthreads();
        if (!active) continue;
        for (int cc = 0; cc < ncs; cc++) {
            const double* rec = scell[cc];
            const double dcen = center_dist(x0, x1, x2, rec);
            if (near_rule(dcen, rec[3], a.mode, a.ac, a.thresh)) continue;
            PhaseRef pr{0., 0., 0., 0};
            if (FP) pr = phase_ref(kr_ * dcen);
            // A = sum w e f1 ; T (xx,yy,zz,xy,xz,yz) = sum w e f2 dh dh ; G = sum w (ikr-1) e d / r^3
            double Ar = 0., Ai = 0.;
            double Tr[6] = {0, 0, 0, 0, 0, 0}, Ti[6] = {0, 0, 0, 0, 0, 0};
            double Gr[3] = {0, 0, 0}, Gi[3] = {0, 0, 0};
            const double4* nd = snode + cc * nn;
            for (int q = 0; q < nn; q++) {
                const double4 y = nd[q];
                const double d0 = x0 - y.x, d1 = x1 - y.y, d2 = x2 - y.z;
                const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                const double ri = rsqrt(r2);
                const double r = r2 * ri;
                const double ri2 = ri * ri, ri3 = ri2 * ri;
                double s, c;
                if (FP) sincos_near(kr_, r, pr, s, c);     // e^{i(kr - n pi/2)}
                else sincos(kr_ * r, &s, &c);
                double w = y.w;
                if (CK) w *= exp(-ki_ * r);
                const double wc = w * c, ws = w * s;
                // f1 = -1/r^3 + ik/r^2 + k^2/r ; f2 = 3/r^3 - 3ik/r^2 - k^2/r  (ik = -ki + i kr)
                double f1r, f1i, f2r, f2i;
                if (CK) {
                    f1r = -ri3 - ki_ * ri2 + k2r * ri; f1i = kr_ * ri2 + k2i * ri;
                    f2r = 3. * ri3 + 3. * ki_ * ri2 - k2r * ri; f2i = -3. * kr_ * ri2 - k2i * ri;
                } else {
                    f1r = k2r * ri - ri3; f1i = kr_ * ri2;
                    f2r = 3. * ri3 - k2r * ri; f2i = -3. * f1i;
                }
                Ar += wc * f1r - ws * f1i;
                Ai += wc * f1i + ws * f1r;
                const double e2r = wc * f2r - ws * f2i, e2i = wc * f2i + ws * f2r;
                const double h0 = d0 * ri, h1 = d1 * ri, h2 = d2 * ri;
                const double p00 = h0 * h0, p11 = h1 * h1, p22 = h2 * h2, p01 = h0 * h1, p02 = h0 * h2, p12 = h1 * h2;
                Tr[0] += e2r * p00; Ti[0] += e2i * p00;
                Tr[1] += e2r * p11; Ti[1] += e2i * p11;
                Tr[2] += e2r * p22; Ti[2] += e2i * p22;
                Tr[3] += e2r * p01; Ti[3] += e2i * p01;
                Tr[4] += e2r * p02; Ti[4] += e2i * p02;
                Tr[5] += e2r * p12; Ti[5] += e2i * p12;
                if (WH || JM) {
                    // (ikr - 1) e = (-(1 + ki r) + i kr r) (c + i s)
                    const double mr = CK ? -(1. + ki_ * r) : -1., mi = kr_ * r;
                    const double wr3 = ri3;
                    const double gr = (mr * wc - mi * ws) * wr3, gi = (mr * ws + mi * wc) * wr3;
                    Gr[0] += gr * d0; Gi[0] += gi * d0;
                    Gr[1] += gr * d1; Gi[1] += gi * d1;
                    Gr[2] += gr * d2; Gi[2] += gi * d2;
                }
            }
            if (FP) {   // restore the common factor i^n of the cell's e^{ikr}
                rot_q(Ar, Ai, pr.q);
                for (int t = 0; t < 6; t++) rot_q(Tr[t], Ti[t], pr.q);
                for (int g = 0; g < 3; g++) rot_q(Gr[g], Gi[g], pr.q);
            }
            // apply to the cell's currents
            const cplx A = mk(Ar, Ai);
            const cplx T00 = mk(Tr[0], Ti[0]), T11 = mk(Tr[1], Ti[1]), T22 = mk(Tr[2], Ti[2]);
            const cplx T01 = mk(Tr[3], Ti[3]), T02 = mk(Tr[4], Ti[4]), T12 = mk(Tr[5], Ti[5]);
            const cplx je[3] = {mk(rec[4], rec[5]), mk(rec[6], rec[7]), mk(rec[8], rec[9])};
            E[0] += A * je[0] + T00 * je[0] + T01 * je[1] + T02 * je[2];
            E[1] += A * je[1] + T01 * je[0] + T11 * je[1] + T12 * je[2];
            E[2] += A * je[2] + T02 * je[0] + T12 * je[1] + T22 * je[2];
            const cplx G[3] = {mk(Gr[0], Gi[0]), mk(Gr[1], Gi[1]), mk(Gr[2], Gi[2])};
            // R_Rot_Colloc is exactly zero when x is the cell centroid (R.h:94-97)
            const bool self = dcen < kMachineZero;
            if (WH && !self) {
                cplx t[3];
                ccross(G, je, t);
                H[0] += t[0]; H[1] += t[1]; H[2] += t[2];
            }
            if (JM) {
                const cplx jm[3] = {mk(rec[10], rec[11]), mk(rec[12], rec[13]), mk(rec[14], rec[15])};
                if (WH) {
                    KH[0] += A * jm[0] + T00 * jm[0] + T01 * jm[1] + T02 * jm[2];
                    KH[1] += A * jm[1] + T01 * jm[0] + T11 * jm[1] + T12 * jm[2];
                    KH[2] += A * jm[2] + T02 * jm[0] + T12 * jm[1] + T22 * jm[2];
                }
                if (!self) {
                    cplx t[3];
                    ccross(G, jm, t);
                    RE[0] += t[0]; RE[1] += t[1]; RE[2] += t[2];
                }
            }
        }
    }
    if (!active) return;
    const int nacc = JM ? 24 : 12;
    double* out = a.part + (size_t)blockIdx.y * nacc * a.M + pt;
    const size_t st = a.M;
    for (int g = 0; g < 3; g++) {
        out[(2 * g) * st] = E[g].re; out[(2 * g + 1) * st] = E[g].im;
        out[(6 + 2 * g) * st] = H[g].re; out[(6 + 2 * g + 1) * st] = H[g].im;
        if (JM) {
            out[(12 + 2 * g) * st] = KH[g].re; out[(12 + 2 * g + 1) * st] = KH[g].im;
            out[(18 + 2 * g) * st] = RE[g].re; out[(18 + 2 * g + 1) * st] = RE[g].im;
        }
    }
}

template <bool CK, bool JM, bool WH, bool FP>
static void field_far_t(hipStream_t s, const FieldArgs& a) {
    int cps = (a.N + a.nsplit - 1) / a.nsplit;
    dim3 g((a.M + kFT - 1) / kFT, a.nsplit), b(kFT);
    hipLaunchKernelGGL((k_field_far<CK, JM, WH, FP>), g, b, 0, s, a, cps);
}

template <bool CK, bool JM>
static void field_far_fp(hipStream_t s, const FieldArgs& a) {
    if (a.fast_phase) field_far_t<CK, JM, true, true>(s, a);
    else field_far_t<CK, JM, true, false>(s, a);
}

hipError_t launch_field_far(hipStream_t s, const FieldArgs& a) {
    const bool ck = a.k_im != 0.;
    const bool jm = a.jM != nullptr;
    if (ck) { if (jm) field_far_fp<true, true>(s, a); else field_far_fp<true, false>(s, a); }
    else { if (jm) field_far_fp<false, true>(s, a); else field_far_fp<false, false>(s, a); }
    return hipGetLastError();
}

// ---------------------------------------------------------------- near pairs
// One lane per listed (point, cell) pair, exact reference semantics.
// mode 0: K = K_RotRot_Near_Smooth (get_field.h:60), R = R_Rot_Smooth;
// mode 1: K = K_RotRot_HS, R = R_Rot_Smooth (main0_field_test.cpp:283-310).
// out[pair]: KE(6) RH(6) [RE(6) KH(6)], each already divided by 4pi.
template <int NV>
__global__ __launch_bounds__(64) void k_field_near(long long np_, const int* __restrict__ lq,
                                                   const int* __restrict__ lc, const double* __restrict__ pts,
                                                   const double* __restrict__ verts, const double* __restrict__ jE,
                                                   const double* __restrict__ jM, NumP np, double k_re,
                                                   double k_im, int mode, double* __restrict__ out) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np_) return;
    const int q = lq[p], c = lc[p];
    double cv[NV][3], x[3];
    for (int v = 0; v < NV; v++)
        for (int d = 0; d < 3; d++) cv[v][d] = verts[((size_t)c * NV + v) * 3 + d];
    for (int d = 0; d < 3; d++) x[d] = pts[3 * (size_t)q + d];
    const cplx k = mk(k_re, k_im);
    const int opK = (mode == 0) ? OP_K_NEAR_SMOOTH : OP_K_HS;
    const int nacc = jM ? 24 : 12;
    double* o = out + (size_t)p * nacc;
    cplx j[3], r[3];
    for (int d = 0; d < 3; d++) j[d] = mk(jE[6 * (size_t)c + 2 * d], jE[6 * (size_t)c + 2 * d + 1]);
    op_eval<NV>(opK, cv, x, j, np, k, r);
    for (int d = 0; d < 3; d++) { o[2 * d] = r[d].re; o[2 * d + 1] = r[d].im; }
    op_eval<NV>(OP_R_SMOOTH, cv, x, j, np, k, r);
    for (int d = 0; d < 3; d++) { o[6 + 2 * d] = r[d].re; o[6 + 2 * d + 1] = r[d].im; }
    if (jM) {
        for (int d = 0; d < 3; d++) j[d] = mk(jM[6 * (size_t)c + 2 * d], jM[6 * (size_t)c + 2 * d + 1]);
        op_eval<NV>(OP_R_SMOOTH, cv, x, j, np, k, r);
        for (int d = 0; d < 3; d++) { o[12 + 2 * d] = r[d].re; o[12 + 2 * d + 1] = r[d].im; }
        op_eval<NV>(opK, cv, x, j, np, k, r);
        for (int d = 0; d < 3; d++) { o[18 + 2 * d] = r[d].re; o[18 + 2 * d + 1] = r[d].im; }
    }
}

hipError_t launch_field_near(hipStream_t s, int nv, long long npairs, const int* lq, const int* lc,
                             const double* pts, const double* verts, const double* jE, const double* jM,
                             const NumP& np, double k_re, double k_im, int mode, double* out) {
    if (npairs <= 0) return hipSuccess;
    dim3 g((unsigned)((npairs + 63) / 64)), b(64);
    if (nv == 4) hipLaunchKernelGGL(k_field_near<4>, g, b, 0, s, npairs, lq, lc, pts, verts, jE, jM, np, k_re, k_im, mode, out);
    else hipLaunchKernelGGL(k_field_near<3>, g, b, 0, s, npairs, lq, lc, pts, verts, jE, jM, np, k_re, k_im, mode, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------- finalize
// E = multE (sum K[jE]) - sum R[jM] (+ E_inc),  H = multH (sum K[jM]) + sum R[jE]
// (main0_field_test.cpp:313-319; get_field.h:69-76 for the E-only driver).
__global__ __launch_bounds__(256) void k_field_finalize(int M, int nsplit, const double* __restrict__ part,
                                                        int has_jm, const long long* __restrict__ offs,
                                                        const double* __restrict__ near_out, double mEr,
                                                        double mEi, double mHr, double mHi, int add_inc,
                                                        double kv0, double kv1, double kv2, double e00,
                                                        double e01, double e02, const double* __restrict__ pts,
                                                        double* __restrict__ E, double* __restrict__ H) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= M) return;
    const int nacc = has_jm ? 24 : 12;
    double acc[24];
    for (int a = 0; a < nacc; a++) acc[a] = 0.;
    for (int sp = 0; sp < nsplit; sp++) {
        const double* src = part + (size_t)sp * nacc * M + p;
        for (int a = 0; a < nacc; a++) acc[a] += src[(size_t)a * M];
    }
    // far sums are unscaled integrals: divide by 4pi as K.h:300 / R.h:109 do per cell
    for (int a = 0; a < nacc; a++) acc[a] = acc[a] / (4. * kPi);
    cplx sKE[3], sRH[3], sKH[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, sRE[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    for (int g = 0; g < 3; g++) {
        sKE[g] = mk(acc[2 * g], acc[2 * g + 1]);
        sRH[g] = mk(acc[6 + 2 * g], acc[6 + 2 * g + 1]);
        if (has_jm) {
            sKH[g] = mk(acc[12 + 2 * g], acc[12 + 2 * g + 1]);
            sRE[g] = mk(acc[18 + 2 * g], acc[18 + 2 * g + 1]);
        }
    }
    if (offs) {
        for (long long q = offs[p]; q < offs[p + 1]; q++) {
            const double* o = near_out + (size_t)q * nacc;
            for (int g = 0; g < 3; g++) {
                sKE[g] += mk(o[2 * g], o[2 * g + 1]);
                sRH[g] += mk(o[6 + 2 * g], o[6 + 2 * g + 1]);
                if (has_jm) {
                    sRE[g] += mk(o[12 + 2 * g], o[12 + 2 * g + 1]);
                    sKH[g] += mk(o[18 + 2 * g], o[18 + 2 * g + 1]);
                }
            }
        }
    }
    const cplx mE = mk(mEr, mEi), mH = mk(mHr, mHi);
    cplx e[3], h[3];
    for (int g = 0; g < 3; g++) {
        e[g] = mE * sKE[g] - sRE[g];
        h[g] = mH * sKH[g] + sRH[g];
    }
    if (add_inc) {   // PlaneWaveFieldE incident_field.h:20-29
        const double* x = pts + 3 * (size_t)p;
        double ph = kv0 * x[0] + kv1 * x[1] + kv2 * x[2];
        cplx deg = cexp(cplx{0., 1.} * ph);
        e[0] += deg * e00; e[1] += deg * e01; e[2] += deg * e02;
    }
    for (int g = 0; g < 3; g++) {
        E[6 * (size_t)p + 2 * g] = e[g].re; E[6 * (size_t)p + 2 * g + 1] = e[g].im;
        if (H) { H[6 * (size_t)p + 2 * g] = h[g].re; H[6 * (size_t)p + 2 * g + 1] = h[g].im; }
    }
}

hipError_t launch_field_finalize(hipStream_t s, int M, int nsplit, const double* part, int has_jm,
                                 const long long* near_offs, const double* near_out, double mE_re, double mE_im,
                                 double mH_re, double mH_im, int add_inc, const double* kvec_e0,
                                 const double* pts, double* E, double* H) {
    hipLaunchKernelGGL(k_field_finalize, dim3((M + 255) / 256), dim3(256), 0, s, M, nsplit, part, has_jm, near_offs,
                       near_out, mE_re, mE_im, mH_re, mH_im, add_inc, kvec_e0[0], kvec_e0[1], kvec_e0[2],
                       kvec_e0[3], kvec_e0[4], kvec_e0[5], pts, E, H);
    return hipGetLastError();
}

// PlaneWaveFieldE at n points (incident_field.h:20-29)
__global__ void k_plane_wave(int n, const double* __restrict__ pts, double kv0, double kv1, double kv2,
                             double e00, double e01, double e02, double* __restrict__ out) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double* x = pts + 3 * (size_t)p;
    double ph = kv0 * x[0] + kv1 * x[1] + kv2 * x[2];
    cplx deg = cexp(cplx{0., 1.} * ph);
    cplx e[3] = {deg * e00, deg * e01, deg * e02};
    for (int g = 0; g < 3; g++) { out[6 * (size_t)p + 2 * g] = e[g].re; out[6 * (size_t)p + 2 * g + 1] = e[g].im; }
}

hipError_t launch_plane_wave(hipStream_t s, int n, const double* pts, const double* kvec_e0, double* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_plane_wave, dim3((n + 255) / 256), dim3(256), 0, s, n, pts, kvec_e0[0], kvec_e0[1],
                       kvec_e0[2], kvec_e0[3], kvec_e0[4], kvec_e0[5], out);
    return hipGetLastError();
}

}  // namespace bielcc_impl
