// Stratton-Chu field at M observation points over N cells, for one or more
// sets of surface currents (right-hand sides).
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
// electric and the magnetic current alike.  A, T and G do not depend on the
// currents, so one pass serves R current sets (R = 1, 2 or 4) for the price of
// R small contractions per cell.
//
// Mixed precision (k_field_far32): the node sums run in FP32 in cell-local
// coordinates.  Per (point, cell) the centre distance dcen, the near/far
// decision and the centre phase k*dcen - n*pi/2 stay FP64; per node
//   r - dcen = (|y_c|^2 - 2 x_c.y_c) / (r + dcen)      (x_c = x - c, y_c = y - c)
// is formed from small cell-local quantities, so the node phase carries no
// cancellation error, and the cell's sums are rotated by i^n, scaled and
// contracted with the currents in FP64.
#include "bielcc_internal.h"
#include "phase.h"

using namespace bielcc_dev;

namespace bielcc_impl {

constexpr int kFT = 256;          // points per workgroup
constexpr int kFNodes = 512;      // FP64 node slots staged in LDS (16 KB)
constexpr int kFNodes32 = 1024;   // FP32 node slots (16 KB + 4 KB)
constexpr int kFCells = 16;       // max cells per LDS stage
constexpr int kCellRec = 32;      // staged per-cell record: centre 3, diam, jE[R][6], jM[6]
constexpr int kGroup32 = 8;       // FP32 node sums are flushed to FP64 every kGroup32 nodes

// Stage the per-cell records of cells [cs, cs+ncs): centre, diameter, the R
// electric current sets r0..r0+R-1 and (R = 1 only) the magnetic current.
template <bool JM, int R>
__device__ inline void stage_cells(const FieldArgs& a, int cs, int ncs, int r0, double (*scell)[kCellRec]) {
    constexpr int nf = 4 + 6 * R + (JM ? 6 : 0);
    for (int idx = threadIdx.x; idx < ncs * nf; idx += kFT) {
        const int cc = idx / nf, f = idx % nf;
        const size_t c = (size_t)(cs + cc);
        double v;
        if (f < 3) v = a.ctab[c * CT + CT_CENTER + f];
        else if (f == 3) v = a.ctab[c * CT + CT_DIAM];
        else if (f < 4 + 6 * R) v = a.jE[((size_t)(r0 + (f - 4) / 6) * a.N + c) * 6 + (f - 4) % 6];
        else v = a.jM[((size_t)r0 * a.N + c) * 6 + (f - 4 - 6 * R)];
        scell[cc][f] = v;
    }
}

// One far cell's sums applied to R current sets: K[j] += A j + T j and
// R[j] += G x j; R_Rot_Colloc is exactly zero when x is the cell centroid
// (R.h:94-97).  T = (xx, yy, zz, xy, xz, yz).  FAC: A, T and G still lack the
// pair's common factor f (e^{ikr} = f x the node phase); the contraction is
// linear, so f multiplies the currents instead (3 complex products per current
// set instead of 10 on the sums).
template <bool JM, int R, bool FAC = false>
__device__ inline void far_contract(const double* rec, const cplx& A, const cplx (&T)[6], const cplx (&G)[3],
                                    bool self, cplx (&E)[R][3], cplx (&H)[R][3], cplx (&KH)[3], cplx (&RE)[3],
                                    const cplx f = cplx{1., 0.}) {
    const cplx T00 = A + T[0], T11 = A + T[1], T22 = A + T[2], T01 = T[3], T02 = T[4], T12 = T[5];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const double* j = rec + 4 + 6 * r;
        cplx je[3] = {mk(j[0], j[1]), mk(j[2], j[3]), mk(j[4], j[5])};
        if (FAC)
            for (int g = 0; g < 3; g++) je[g] = f * je[g];
        E[r][0] += T00 * je[0] + T01 * je[1] + T02 * je[2];
        E[r][1] += T01 * je[0] + T11 * je[1] + T12 * je[2];
        E[r][2] += T02 * je[0] + T12 * je[1] + T22 * je[2];
        if (!self) {
            cplx t[3];
            ccross(G, je, t);
            H[r][0] += t[0]; H[r][1] += t[1]; H[r][2] += t[2];
        }
    }
    if (JM) {
        const double* j = rec + 4 + 6 * R;
        cplx jm[3] = {mk(j[0], j[1]), mk(j[2], j[3]), mk(j[4], j[5])};
        if (FAC)
            for (int g = 0; g < 3; g++) jm[g] = f * jm[g];
        KH[0] += T00 * jm[0] + T01 * jm[1] + T02 * jm[2];
        KH[1] += T01 * jm[0] + T11 * jm[1] + T12 * jm[2];
        KH[2] += T02 * jm[0] + T12 * jm[1] + T22 * jm[2];
        if (!self) {
            cplx t[3];
            ccross(G, jm, t);
            RE[0] += t[0]; RE[1] += t[1]; RE[2] += t[2];
        }
    }
}

// Accumulator layout of the partial sums, [nsplit][nrhs][NACC][M]:
//  0-5  sum (A j_E + T j_E)         (E from K[j_E], unscaled)
//  6-11 sum G x j_E                  (H from R[j_E], unscaled)
//  12-17 sum (A j_M + T j_M)        (H from K[j_M], unscaled)   [with j_M]
//  18-23 sum G x j_M                 (E from R[j_M], unscaled)   [with j_M]
template <bool JM, int R>
__device__ inline void write_partials(const FieldArgs& a, int r0, int pt, const cplx (&E)[R][3],
                                      const cplx (&H)[R][3], const cplx (&KH)[3], const cplx (&RE)[3]) {
    const int nacc = JM ? 24 : 12;
    const size_t st = a.M;
#pragma unroll
    for (int r = 0; r < R; r++) {
        double* out = a.part + ((size_t)blockIdx.y * a.nrhs + r0 + r) * nacc * st + pt;
        for (int g = 0; g < 3; g++) {
            out[(2 * g) * st] = E[r][g].re; out[(2 * g + 1) * st] = E[r][g].im;
            out[(6 + 2 * g) * st] = H[r][g].re; out[(6 + 2 * g + 1) * st] = H[r][g].im;
            if (JM) {
                out[(12 + 2 * g) * st] = KH[g].re; out[(12 + 2 * g + 1) * st] = KH[g].im;
                out[(18 + 2 * g) * st] = RE[g].re; out[(18 + 2 * g + 1) * st] = RE[g].im;
            }
        }
    }
}

// Node sums of one far (point, cell) pair in FP64: A = sum w e f1,
// T (xx,yy,zz,xy,xz,yz) = sum w e f2 dh dh, G = sum w (ikr-1) e d / r^3
// (e = e^{ikr}; unscaled, without 1/4pi).  nd: the cell's nn nodes.
// PH: node phase evaluation, 0 general sincos, 1 quadrant reduction, 2 centre-phase
// reduction with the short series (phase.h)
// ROT false: the sums are returned without the pair's common factor, which goes to *fac.
template <bool CK, int PH, bool ROT = true>
__device__ __forceinline__ void pair_sums64(double x0, double x1, double x2, const double4* nd, int nn, double dcen,
                                            double kr_, double ki_, double k2r, double k2i, cplx& A, cplx (&T)[6],
                                            cplx (&G)[3], cplx* fac = nullptr) {
    PhaseRef pr{0., 0., 0., 0};
    PhaseC pc{0., 1., 0.};
    if (PH == 1) pr = phase_ref(kr_ * dcen);
    if (PH >= 2) pc = phase_centre(kr_, dcen);
    // A = sum w e f1 ; T (xx,yy,zz,xy,xz,yz) = sum w e f2 dh dh ; G = sum w (ikr-1) e d / r^3
    double Ar = 0., Ai = 0.;
    double Tr[6] = {0, 0, 0, 0, 0, 0}, Ti[6] = {0, 0, 0, 0, 0, 0};
    double Gr[3] = {0, 0, 0}, Gi[3] = {0, 0, 0};
    if (CK) {
        for (int q = 0; q < nn; q++) {
            const double4 y = nd[q];
            const double d0 = x0 - y.x, d1 = x1 - y.y, d2 = x2 - y.z;
            const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
            const double ri = rsqrt(r2);
            const double r = r2 * ri;
            const double ri2 = ri * ri, ri3 = ri2 * ri;
            double s, c;
            phase_eval<PH>(kr_, r, pr, pc, s, c);     // e^{ikr}, up to the pair's common factor
            const double w = y.w * exp(-ki_ * r);
            const double wc = w * c, ws = w * s;
            // f1 = -1/r^3 + ik/r^2 + k^2/r ; f2 = 3/r^3 - 3ik/r^2 - k^2/r  (ik = -ki + i kr)
            const double f1r = -ri3 - ki_ * ri2 + k2r * ri, f1i = kr_ * ri2 + k2i * ri;
            const double f2r = 3. * ri3 + 3. * ki_ * ri2 - k2r * ri, f2i = -3. * kr_ * ri2 - k2i * ri;
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
            // (ikr - 1) e = (-(1 + ki r) + i kr r) (c + i s)
            const double mr = -(1. + ki_ * r), mi = kr_ * r;
            const double gr = (mr * wc - mi * ws) * ri3, gi = (mr * ws + mi * wc) * ri3;
            Gr[0] += gr * d0; Gi[0] += gi * d0;
            Gr[1] += gr * d1; Gi[1] += gi * d1;
            Gr[2] += gr * d2; Gi[2] += gi * d2;
        }
    } else {
        // Real k, with we = w e^{ikr} per node:
        //   g = w (ikr - 1) e / r^3 = -(w/r^3) (c + k r s) + i (w/r^3) (k r c - s)
        //   w e f2 dh dh = (-3 g/r^2 - k^2 we/r^3) d d
        //   w e f1 = g + k^2 we/r, and since the trace of d d is r^2,
        //   A = sum w e f1 = -(T_xx + T_yy + T_zz) - 2 sum g
        // (65 VALU instructions per node on gfx950, against 83 for the complex-k form).
        for (int q = 0; q < nn; q++) {
            const double4 y = nd[q];
            const double d0 = x0 - y.x, d1 = x1 - y.y, d2 = x2 - y.z;
            const double p00 = d0 * d0, p11 = d1 * d1, p22 = d2 * d2;
            const double r2 = p00 + p11 + p22;
            const double ri = rsqrt_pos(r2);
            const double r = r2 * ri, ri2 = ri * ri;
            double s, c;
            phase_eval<PH>(kr_, r, pr, pc, s, c);     // e^{ikr}, up to the pair's common factor
            const double wr2 = y.w * ri2, wr3 = wr2 * ri, wk = wr2 * kr_;
            const double u = wr3 * c, v = wr3 * s;     // we / r^3
            const double gr = -fma(wk, s, u), gi = fma(wk, c, -v);
            Ar += gr; Ai += gi;
            const double m3 = -3. * ri2;
            const double e2r = fma(gr, m3, -k2r * u), e2i = fma(gi, m3, -k2r * v);
            const double p01 = d0 * d1, p02 = d0 * d2, p12 = d1 * d2;
            Tr[0] += e2r * p00; Ti[0] += e2i * p00;
            Tr[1] += e2r * p11; Ti[1] += e2i * p11;
            Tr[2] += e2r * p22; Ti[2] += e2i * p22;
            Tr[3] += e2r * p01; Ti[3] += e2i * p01;
            Tr[4] += e2r * p02; Ti[4] += e2i * p02;
            Tr[5] += e2r * p12; Ti[5] += e2i * p12;
            Gr[0] += gr * d0; Gi[0] += gi * d0;
            Gr[1] += gr * d1; Gi[1] += gi * d1;
            Gr[2] += gr * d2; Gi[2] += gi * d2;
        }
        Ar = -(Tr[0] + Tr[1] + Tr[2]) - 2. * Ar;
        Ai = -(Ti[0] + Ti[1] + Ti[2]) - 2. * Ai;
    }
    if (!ROT) {
        double fr = 1., fi = 0.;
        if (PH == 1) rot_q(fr, fi, pr.q);
        else if (PH >= 2) { fr = pc.c; fi = pc.s; }
        *fac = mk(fr, fi);
    } else if (PH == 1) {   // restore the common factor i^n of the cell's e^{ikr}
        rot_q(Ar, Ai, pr.q);
        for (int t = 0; t < 6; t++) rot_q(Tr[t], Ti[t], pr.q);
        for (int g = 0; g < 3; g++) rot_q(Gr[g], Gi[g], pr.q);
    } else if (PH >= 2) {   // ... or e^{ih}
        rot_c(Ar, Ai, pc);
        for (int t = 0; t < 6; t++) rot_c(Tr[t], Ti[t], pc);
        for (int g = 0; g < 3; g++) rot_c(Gr[g], Gi[g], pc);
    }
    A = mk(Ar, Ai);
    for (int t = 0; t < 6; t++) T[t] = mk(Tr[t], Ti[t]);
    for (int g = 0; g < 3; g++) G[g] = mk(Gr[g], Gi[g]);
}

// The same sums with FP32 node arithmetic in cell-local coordinates
// (rec = cell centre; nd, ny = the cell's FP32 nodes and |y_c|^2), flushed
// into FP64 every kGroup32 nodes; the cell phase and i^n stay FP64.
template <bool CK, bool FP>
__device__ __forceinline__ void pair_sums32(double x0, double x1, double x2, const double* rec, const float4* nd,
                                            const float* ny, int nn, double dcen, double kr_, double ki_, cplx& A,
                                            cplx (&T)[6], cplx (&G)[3]) {
    const float kf = (float)kr_, kif = (float)ki_;
    const float k2rf = (float)(kr_ * kr_ - ki_ * ki_), k2if = (float)(2. * kr_ * ki_);
    const PhaseRef pr = phase_ref(kr_ * dcen);
    const float t0 = (float)(fma(kr_, dcen, -pr.o1) - pr.o2 - pr.o3);
    const float xf0 = (float)(x0 - rec[0]), xf1 = (float)(x1 - rec[1]), xf2 = (float)(x2 - rec[2]);
    const float m0 = -2.f * xf0, m1 = -2.f * xf1, m2 = -2.f * xf2;
    const float dcf = (float)dcen;
    // FP32 sums over groups of kGroup32 nodes, flushed into FP64 (the
    // recursive FP32 sum of 64 same-signed terms would lose ~sqrt(64) ulp)
    double ArD = 0., AiD = 0., TrD[6] = {0, 0, 0, 0, 0, 0}, TiD[6] = {0, 0, 0, 0, 0, 0};
    double GrD[3] = {0, 0, 0}, GiD[3] = {0, 0, 0}, PrD = 0., PiD = 0.;
    for (int q0 = 0; q0 < nn; q0 += kGroup32) {
        float Ar = 0.f, Ai = 0.f;
        float Tr[6] = {0, 0, 0, 0, 0, 0}, Ti[6] = {0, 0, 0, 0, 0, 0};
        float Gr[3] = {0, 0, 0}, Gi[3] = {0, 0, 0}, Pr = 0.f, Pi = 0.f;
        const int q1 = min(nn, q0 + kGroup32);
        for (int q = q0; q < q1; q++) {
            const float4 y = nd[q];
            const float d0 = xf0 - y.x, d1 = xf1 - y.y, d2 = xf2 - y.z;
            const float r2 = d0 * d0 + d1 * d1 + d2 * d2;
            const float ri = rsqrtf(r2);
            const float r = r2 * ri;
            // r - dcen without cancellation: (|y_c|^2 - 2 x_c.y_c) / (r + dcen)
            const float del = fmaf(m0, y.x, fmaf(m1, y.y, fmaf(m2, y.z, ny[q])));
            const float dr = del * __builtin_amdgcn_rcpf(r + dcf);   // v_rcp_f32 (1 ulp): 2 instructions, not a full division
            const float t = fmaf(kf, dr, t0);
            float s, c;
            if (FP) sincos_near32(t, s, c);
            else sincosf(t, &s, &c);
            const float ri2 = ri * ri, ri3 = ri2 * ri;
            if (CK) {
                const float w = y.w * __expf(-kif * dr);   // e^{-ki dcen} is applied per cell in FP64
                const float wc = w * c, ws = w * s;
                const float f1r = -ri3 - kif * ri2 + k2rf * ri, f1i = kf * ri2 + k2if * ri;
                const float f2r = 3.f * ri3 + 3.f * kif * ri2 - k2rf * ri, f2i = -3.f * kf * ri2 - k2if * ri;
                Ar += wc * f1r - ws * f1i;
                Ai += wc * f1i + ws * f1r;
                const float e2r = wc * f2r - ws * f2i, e2i = wc * f2i + ws * f2r;
                const float h0 = d0 * ri, h1 = d1 * ri, h2 = d2 * ri;
                const float p00 = h0 * h0, p11 = h1 * h1, p22 = h2 * h2, p01 = h0 * h1, p02 = h0 * h2, p12 = h1 * h2;
                Tr[0] += e2r * p00; Ti[0] += e2i * p00;
                Tr[1] += e2r * p11; Ti[1] += e2i * p11;
                Tr[2] += e2r * p22; Ti[2] += e2i * p22;
                Tr[3] += e2r * p01; Ti[3] += e2i * p01;
                Tr[4] += e2r * p02; Ti[4] += e2i * p02;
                Tr[5] += e2r * p12; Ti[5] += e2i * p12;
                const float mr = -(1.f + kif * r), mi = kf * r;
                const float gr = (mr * wc - mi * ws) * ri3, gi = (mr * ws + mi * wc) * ri3;
                Gr[0] += gr * d0; Gi[0] += gi * d0;
                Gr[1] += gr * d1; Gi[1] += gi * d1;
                Gr[2] += gr * d2; Gi[2] += gi * d2;
            } else {   // real k: g and w e f2 dh dh as in pair_sums64; A = sum g + k^2 sum we/r
                       // (summed: the trace form would lose FP32 digits where k r << 1)
                const float wc = y.w * c, ws = y.w * s;
                const float f1i = kf * ri2, u = wc * ri3, v = ws * ri3;
                const float gr = -fmaf(ws, f1i, u), gi = fmaf(wc, f1i, -v);
                Ar += gr; Ai += gi;
                Pr = fmaf(wc, ri, Pr); Pi = fmaf(ws, ri, Pi);
                const float m3 = -3.f * ri2;
                const float e2r = fmaf(gr, m3, -k2rf * u), e2i = fmaf(gi, m3, -k2rf * v);
                const float p00 = d0 * d0, p11 = d1 * d1, p22 = d2 * d2, p01 = d0 * d1, p02 = d0 * d2, p12 = d1 * d2;
                Tr[0] += e2r * p00; Ti[0] += e2i * p00;
                Tr[1] += e2r * p11; Ti[1] += e2i * p11;
                Tr[2] += e2r * p22; Ti[2] += e2i * p22;
                Tr[3] += e2r * p01; Ti[3] += e2i * p01;
                Tr[4] += e2r * p02; Ti[4] += e2i * p02;
                Tr[5] += e2r * p12; Ti[5] += e2i * p12;
                Gr[0] += gr * d0; Gi[0] += gi * d0;
                Gr[1] += gr * d1; Gi[1] += gi * d1;
                Gr[2] += gr * d2; Gi[2] += gi * d2;
            }
        }
        ArD += Ar; AiD += Ai;
        PrD += Pr; PiD += Pi;
        for (int t = 0; t < 6; t++) { TrD[t] += Tr[t]; TiD[t] += Ti[t]; }
        for (int g = 0; g < 3; g++) { GrD[g] += Gr[g]; GiD[g] += Gi[g]; }
    }
    if (!CK) { ArD = fma(kr_ * kr_, PrD, ArD); AiD = fma(kr_ * kr_, PiD, AiD); }   // + k^2 sum we/r
    rot_q(ArD, AiD, pr.q);
    for (int t = 0; t < 6; t++) rot_q(TrD[t], TiD[t], pr.q);
    for (int g = 0; g < 3; g++) rot_q(GrD[g], GiD[g], pr.q);
    if (CK) {
        const double amp = exp(-ki_ * dcen);
        ArD *= amp; AiD *= amp;
        for (int t = 0; t < 6; t++) { TrD[t] *= amp; TiD[t] *= amp; }
        for (int g = 0; g < 3; g++) { GrD[g] *= amp; GiD[g] *= amp; }
    }
    A = mk(ArD, AiD);
    for (int t = 0; t < 6; t++) T[t] = mk(TrD[t], TiD[t]);
    for (int g = 0; g < 3; g++) G[g] = mk(GrD[g], GiD[g]);
}

template <bool CK, bool JM, int PH, int R>
__global__ __launch_bounds__(kFT) void k_field_far(FieldArgs a, int cells_per_split, int r0) {
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

    cplx E[R][3], H[R][3];
    for (int r = 0; r < R; r++)
        for (int g = 0; g < 3; g++) { E[r][g] = mk(0, 0); H[r][g] = mk(0, 0); }
    cplx KH[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, RE[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};

    for (int cs = c_begin; cs < c_end; cs += cb) {
        const int ncs = min(cb, c_end - cs);
        __syncthreads();
        const double4* gn = reinterpret_cast<const double4*>(a.nodes) + (size_t)cs * nn;
        for (int idx = tid; idx < ncs * nn; idx += kFT) snode[idx] = gn[idx];
        stage_cells<JM, R>(a, cs, ncs, r0, scell);
        __syncthreads();
        if (!active) continue;
        for (int cc = 0; cc < ncs; cc++) {
            const double* rec = scell[cc];
            const double dcen = center_dist(x0, x1, x2, rec);
            if (near_rule(dcen, rec[3], a.mode, a.ac, a.thresh)) continue;
            cplx A, T[6], G[3];
            cplx f;
            pair_sums64<CK, PH, false>(x0, x1, x2, snode + cc * nn, nn, dcen, kr_, ki_, k2r, k2i, A, T, G, &f);
            far_contract<JM, R, PH != 0>(rec, A, T, G, dcen < kMachineZero, E, H, KH, RE, f);
        }
    }
    if (!active) return;
    write_partials<JM, R>(a, r0, pt, E, H, KH, RE);
}

template <bool CK, bool JM, bool FP, int R>
__global__ __launch_bounds__(kFT) void k_field_far32(FieldArgs a, int cells_per_split, int r0) {
    __shared__ float4 snode[kFNodes32];
    __shared__ float syy[kFNodes32];
    __shared__ double scell[kFCells][kCellRec];
    const int tid = threadIdx.x;
    const int pt = blockIdx.x * kFT + tid;
    const bool active = pt < a.M;
    double x0 = 0., x1 = 0., x2 = 0.;
    if (active) { x0 = a.pts[3 * (size_t)pt]; x1 = a.pts[3 * (size_t)pt + 1]; x2 = a.pts[3 * (size_t)pt + 2]; }
    const int c_begin = blockIdx.y * cells_per_split;
    const int c_end = min(a.N, c_begin + cells_per_split);
    const int nn = a.nn;
    const int cb = max(1, min(kFCells, kFNodes32 / nn));
    const double kr_ = a.k_re, ki_ = a.k_im;

    cplx E[R][3], H[R][3];
    for (int r = 0; r < R; r++)
        for (int g = 0; g < 3; g++) { E[r][g] = mk(0, 0); H[r][g] = mk(0, 0); }
    cplx KH[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, RE[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};

    for (int cs = c_begin; cs < c_end; cs += cb) {
        const int ncs = min(cb, c_end - cs);
        __syncthreads();
        const float4* gn = a.nodes32 + (size_t)cs * nn;
        const float* gy = a.nodes32_yy + (size_t)cs * nn;
        for (int idx = tid; idx < ncs * nn; idx += kFT) { snode[idx] = gn[idx]; syy[idx] = gy[idx]; }
        stage_cells<JM, R>(a, cs, ncs, r0, scell);
        __syncthreads();
        if (!active) continue;
        for (int cc = 0; cc < ncs; cc++) {
            const double* rec = scell[cc];
            const double dcen = center_dist(x0, x1, x2, rec);
            if (near_rule(dcen, rec[3], a.mode, a.ac, a.thresh)) continue;
            cplx A, T[6], G[3];
            pair_sums32<CK, FP>(x0, x1, x2, rec, snode + cc * nn, syy + cc * nn, nn, dcen, kr_, ki_, A, T, G);
            far_contract<JM, R>(rec, A, T, G, dcen < kMachineZero, E, H, KH, RE);
        }
    }
    if (!active) return;
    write_partials<JM, R>(a, r0, pt, E, H, KH, RE);
}

// FP32 node sums take the quadrant reduction in phase mode 2 as well (phase.h: mode 2
// implies the quadrant range)
template <bool CK, bool JM, int PH, int R>
static void field_far_t(hipStream_t s, const FieldArgs& a, int r0) {
    int cps = (a.N + a.nsplit - 1) / a.nsplit;
    dim3 g((a.M + kFT - 1) / kFT, a.nsplit), b(kFT);
    if (a.prec) hipLaunchKernelGGL((k_field_far32<CK, JM, PH != 0, R>), g, b, 0, s, a, cps, r0);
    else hipLaunchKernelGGL((k_field_far<CK, JM, PH, R>), g, b, 0, s, a, cps, r0);
}

// current sets in passes of up to rchunk (4, 2 or 1; 1 with a magnetic current)
template <bool CK, bool JM, int PH>
static void field_far_r(hipStream_t s, const FieldArgs& a) {
    int r0 = 0;
    if (!JM) {
        if (a.rchunk >= 4)
            for (; r0 + 4 <= a.nrhs; r0 += 4) field_far_t<CK, JM, PH, 4>(s, a, r0);
        if (a.rchunk >= 2)
            for (; r0 + 2 <= a.nrhs; r0 += 2) field_far_t<CK, JM, PH, 2>(s, a, r0);
    }
    for (; r0 < a.nrhs; r0++) field_far_t<CK, JM, PH, 1>(s, a, r0);
}

template <bool CK, bool JM>
static void field_far_fp(hipStream_t s, const FieldArgs& a) {
    if (a.fast_phase == 3) field_far_r<CK, JM, 3>(s, a);
    else if (a.fast_phase == 2) field_far_r<CK, JM, 2>(s, a);
    else if (a.fast_phase) field_far_r<CK, JM, 1>(s, a);
    else field_far_r<CK, JM, 0>(s, a);
}

int resident_slots(const void* kernel, int block) {
    int dev = 0, ncu = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) != hipSuccess) return 0;
    return ncu * nb;
}

template <bool CK, bool JM, int PH, int R>
static const void* field_far_kernel(bool p32) {
    return p32 ? reinterpret_cast<const void*>(&k_field_far32<CK, JM, PH != 0, R>)
               : reinterpret_cast<const void*>(&k_field_far<CK, JM, PH, R>);
}

template <bool CK, bool JM, int PH>
static const void* field_far_kernel_r(const FieldArgs& a) {
    // the instance of the first pass (field_far_r)
    if (!JM && a.rchunk >= 4 && a.nrhs >= 4) return field_far_kernel<CK, JM, PH, 4>(a.prec);
    if (!JM && a.rchunk >= 2 && a.nrhs >= 2) return field_far_kernel<CK, JM, PH, 2>(a.prec);
    return field_far_kernel<CK, JM, PH, 1>(a.prec);
}

template <bool CK, bool JM>
static const void* field_far_kernel_ph(const FieldArgs& a) {
    if (a.fast_phase == 3) return field_far_kernel_r<CK, JM, 3>(a);
    if (a.fast_phase == 2) return field_far_kernel_r<CK, JM, 2>(a);
    if (a.fast_phase) return field_far_kernel_r<CK, JM, 1>(a);
    return field_far_kernel_r<CK, JM, 0>(a);
}

int field_far_splits(const FieldArgs& a, int s_min) {
    const bool ck = a.k_im != 0., jm = a.jM != nullptr;
    const void* k = ck ? (jm ? field_far_kernel_ph<true, true>(a) : field_far_kernel_ph<true, false>(a))
                       : (jm ? field_far_kernel_ph<false, true>(a) : field_far_kernel_ph<false, false>(a));
    const int bx = (a.M + kFT - 1) / kFT;
    return tail_aware_splits(bx, s_min, (a.N + 63) / 64, resident_slots(k, kFT));
}

hipError_t launch_field_far(hipStream_t s, const FieldArgs& a) {
    if (a.prec && (!a.nodes32 || !a.nodes32_yy)) return hipErrorInvalidValue;
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

// Several current sets: the near operators are linear in j, so each pair's
// 3x3 matrices are formed once from the unit vectors (column b = Op(e_b)).
// out[pair]: Kmat (9 complex, column-major) then Rmat (9 complex).
template <int NV>
__global__ __launch_bounds__(64) void k_field_near_mat(long long np_, const int* __restrict__ lq,
                                                       const int* __restrict__ lc, const double* __restrict__ pts,
                                                       const double* __restrict__ verts, NumP np, double k_re,
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
    double* o = out + (size_t)p * 36;
    for (int b = 0; b < 3; b++) {
        cplx j[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, r[3];
        j[b] = mk(1., 0.);
        op_eval<NV>(opK, cv, x, j, np, k, r);
        for (int d = 0; d < 3; d++) { o[2 * (3 * b + d)] = r[d].re; o[2 * (3 * b + d) + 1] = r[d].im; }
        op_eval<NV>(OP_R_SMOOTH, cv, x, j, np, k, r);
        for (int d = 0; d < 3; d++) { o[18 + 2 * (3 * b + d)] = r[d].re; o[18 + 2 * (3 * b + d) + 1] = r[d].im; }
    }
}

hipError_t launch_field_near(hipStream_t s, int nv, long long npairs, const int* lq, const int* lc,
                             const double* pts, const double* verts, const double* jE, const double* jM,
                             const NumP& np, double k_re, double k_im, int mode, int as_matrix, double* out) {
    if (npairs <= 0) return hipSuccess;
    dim3 g((unsigned)((npairs + 63) / 64)), b(64);
    if (as_matrix) {
        if (nv == 4) hipLaunchKernelGGL(k_field_near_mat<4>, g, b, 0, s, npairs, lq, lc, pts, verts, np, k_re, k_im, mode, out);
        else hipLaunchKernelGGL(k_field_near_mat<3>, g, b, 0, s, npairs, lq, lc, pts, verts, np, k_re, k_im, mode, out);
    } else {
        if (nv == 4) hipLaunchKernelGGL(k_field_near<4>, g, b, 0, s, npairs, lq, lc, pts, verts, jE, jM, np, k_re, k_im, mode, out);
        else hipLaunchKernelGGL(k_field_near<3>, g, b, 0, s, npairs, lq, lc, pts, verts, jE, jM, np, k_re, k_im, mode, out);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- finalize
// E = multE (sum K[jE]) - sum R[jM] (+ E_inc),  H = multH (sum K[jM]) + sum R[jE]
// (main0_field_test.cpp:313-319; get_field.h:69-76 for the E-only driver).
// One lane per (point, current set); E, H are [nrhs][M][3] complex.
__device__ inline cplx mat3_row(const double* m, int d, const cplx* j) {
    // row d of a column-major 3x3 complex matrix times j
    cplx s = mk(m[2 * d], m[2 * d + 1]) * j[0];
    s += mk(m[2 * (3 + d)], m[2 * (3 + d) + 1]) * j[1];
    s += mk(m[2 * (6 + d)], m[2 * (6 + d) + 1]) * j[2];
    return s;
}

__global__ __launch_bounds__(256) void k_field_finalize(FinalizeArgs f) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (p >= f.M) return;
    const int has_jm = f.jM != nullptr;
    const int nacc = has_jm ? 24 : 12;
    double acc[24];
    for (int a = 0; a < nacc; a++) acc[a] = 0.;
    for (int sp = 0; sp < f.nsplit; sp++) {
        const double* src = f.part + ((size_t)sp * f.nrhs + r) * nacc * f.M + p;
        for (int a = 0; a < nacc; a++) acc[a] += src[(size_t)a * f.M];
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
    if (f.offs && !f.near_mat) {
        for (long long q = f.offs[p]; q < f.offs[p + 1]; q++) {
            const double* o = f.near_out + (size_t)q * nacc;
            for (int g = 0; g < 3; g++) {
                sKE[g] += mk(o[2 * g], o[2 * g + 1]);
                sRH[g] += mk(o[6 + 2 * g], o[6 + 2 * g + 1]);
                if (has_jm) {
                    sRE[g] += mk(o[12 + 2 * g], o[12 + 2 * g + 1]);
                    sKH[g] += mk(o[18 + 2 * g], o[18 + 2 * g + 1]);
                }
            }
        }
    } else if (f.offs) {
        for (long long q = f.offs[p]; q < f.offs[p + 1]; q++) {
            const double* Km = f.near_out + (size_t)q * 36;
            const double* Rm = Km + 18;
            const size_t c = (size_t)f.lc[q];
            const double* je = f.jE + ((size_t)r * f.N + c) * 6;
            const cplx j[3] = {mk(je[0], je[1]), mk(je[2], je[3]), mk(je[4], je[5])};
            for (int g = 0; g < 3; g++) {
                sKE[g] += mat3_row(Km, g, j);
                sRH[g] += mat3_row(Rm, g, j);
            }
            if (has_jm) {
                const double* jm = f.jM + ((size_t)r * f.N + c) * 6;
                const cplx m[3] = {mk(jm[0], jm[1]), mk(jm[2], jm[3]), mk(jm[4], jm[5])};
                for (int g = 0; g < 3; g++) {
                    sRE[g] += mat3_row(Rm, g, m);
                    sKH[g] += mat3_row(Km, g, m);
                }
            }
        }
    }
    const cplx mE = mk(f.mEr, f.mEi), mH = mk(f.mHr, f.mHi);
    cplx e[3], h[3];
    for (int g = 0; g < 3; g++) {
        e[g] = mE * sKE[g] - sRE[g];
        h[g] = mH * sKH[g] + sRH[g];
    }
    if (f.inc) {   // PlaneWaveFieldE incident_field.h:20-29
        const double* ke = f.inc + 6 * r;
        const double* x = f.pts + 3 * (size_t)p;
        double ph = ke[0] * x[0] + ke[1] * x[1] + ke[2] * x[2];
        cplx deg = cexp(cplx{0., 1.} * ph);
        e[0] += deg * ke[3]; e[1] += deg * ke[4]; e[2] += deg * ke[5];
    }
    double* Eo = f.E + ((size_t)r * f.M + p) * 6;
    double* Ho = f.H ? f.H + ((size_t)r * f.M + p) * 6 : nullptr;
    for (int g = 0; g < 3; g++) {
        Eo[2 * g] = e[g].re; Eo[2 * g + 1] = e[g].im;
        if (Ho) { Ho[2 * g] = h[g].re; Ho[2 * g + 1] = h[g].im; }
    }
}

hipError_t launch_field_finalize(hipStream_t s, const FinalizeArgs& f) {
    if (f.M <= 0 || f.nrhs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_field_finalize, dim3((f.M + 255) / 256, f.nrhs), dim3(256), 0, s, f);
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


// PlaneWaveFieldH (incident_field.h:40-59): (vec_dbl x e0) e^{i vec_dbl.x} / (2 pi freq) / mu,
// mu the absolute permeability (MaterialParam::GetMu, real or complex).
__device__ inline void plane_wave_h(const double* x, const double* ke, double freq, cplx mu, bool mu_cplx,
                                    cplx* h) {
    double vp[3];
    cross3(ke, ke + 3, vp);   // vec_prod(param.vec_dbl, param.e0)
    const double ph = ke[0] * x[0] + ke[1] * x[1] + ke[2] * x[2];
    cplx deg = cexp(cplx{0., 1.} * ph);
    deg = deg / (2. * kPi * freq);
    if (mu_cplx) {   // std::complex division
        const double den = mu.re * mu.re + mu.im * mu.im;
        deg = cplx{(deg.re * mu.re + deg.im * mu.im) / den, (deg.im * mu.re - deg.re * mu.im) / den};
    } else {
        deg = deg / mu.re;
    }
    for (int g = 0; g < 3; g++) h[g] = vp[g] * deg;
}

__global__ void k_plane_wave_h(int n, const double* __restrict__ pts, PlaneWaveH pw, double* __restrict__ out) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    cplx h[3];
    plane_wave_h(pts + 3 * (size_t)p, pw.ke, pw.freq, mk(pw.mu_re, pw.mu_im), pw.mu_im != 0., h);
    for (int g = 0; g < 3; g++) { out[6 * (size_t)p + 2 * g] = h[g].re; out[6 * (size_t)p + 2 * g + 1] = h[g].im; }
}

hipError_t launch_plane_wave_h(hipStream_t s, int n, const double* pts, const PlaneWaveH& pw, double* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_plane_wave_h, dim3((n + 255) / 256), dim3(256), 0, s, n, pts, pw, out);
    return hipGetLastError();
}

// H[r][p] = scale H[r][p] + PlaneWaveFieldH(x_p) of incidence r, inc[r] = (kvec, e0) on the
// device; H is [nrhs][M][3] complex.
__global__ __launch_bounds__(256) void k_add_incident_h(int M, const double* __restrict__ pts,
                                                        const double* __restrict__ inc, double freq, double mu_re,
                                                        double mu_im, double s_re, double s_im,
                                                        double* __restrict__ H) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (p >= M) return;
    cplx h[3];
    plane_wave_h(pts + 3 * (size_t)p, inc + 6 * r, freq, mk(mu_re, mu_im), mu_im != 0., h);
    double* o = H + ((size_t)r * M + p) * 6;
    const cplx sc = mk(s_re, s_im);
    for (int g = 0; g < 3; g++) {
        const cplx v = sc * mk(o[2 * g], o[2 * g + 1]) + h[g];
        o[2 * g] = v.re;
        o[2 * g + 1] = v.im;
    }
}

hipError_t launch_add_incident_h(hipStream_t s, int M, int nrhs, const double* pts, const double* inc, double freq,
                                 double mu_re, double mu_im, double s_re, double s_im, double* H) {
    if (M <= 0 || nrhs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_add_incident_h, dim3((M + 255) / 256, nrhs), dim3(256), 0, s, M, pts, inc, freq, mu_re,
                       mu_im, s_re, s_im, H);
    return hipGetLastError();
}


// ---------------------------------------------------------------- FP32 error control
// The FP32 node sums of the mixed form carry a rounding error per far pair of
// about u = 2^-24 of that pair's contribution, nearly independent from pair to
// pair, so at a point the error of the sum over cells is about
//   u sqrt(S),  S = sum_c |contribution_c|^2,
// while the field itself can be far smaller than sqrt(S): near a null of the
// pattern, or when the currents carry a large non-radiating part (near an
// interior resonance, where |j| is 100x what the field needs).  A point is
// recomputed in FP64 when kFixKappa u sqrt(S) exceeds kFixTau times
// max(|E_p|, 1e-3 max_p |E_p|) (scattered field, per current set; the same for
// H).  S is modelled from cell centres and areas: the far integrands' magnitudes
//   |f1| + |f2| for K (the E of j_E, the H of j_M) and |ikr - 1| / r^2 for R,
// times area_c |j_c|.  On a 7,688-cell sphere with resonant currents the FP32
// error stayed below 1.43 u sqrt(S) of this model at every point, on an 11,880-cell
// triangulated sphere below 8.4 u sqrt(S) (where the rest is a relative error of
// the whole field, harmless); kFixKappa = 4 and kFixTau = 2e-7 leave a factor 5
// to the 1e-6 target.
namespace {
constexpr double kFixKappa = 4.0;
constexpr double kFixTau = 2e-7;
constexpr double kU32 = 5.9604644775390625e-08;   // 2^-24
constexpr int kFixCells = 256;

// positive doubles order like their bit patterns
__device__ inline void atomic_max_pos(double* a, double v) {
    atomicMax(reinterpret_cast<unsigned long long*>(a), (unsigned long long)__double_as_longlong(v));
}
}  // namespace

// max over points of |E - E_inc| and |H| per current set (mx zeroed before)
__global__ __launch_bounds__(256) void k_scat_max(int M, const double* __restrict__ pts, const double* __restrict__ inc,
                                                  const double* __restrict__ E, const double* __restrict__ H,
                                                  double* __restrict__ mx) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    double e = 0., h = 0.;
    if (p < M) {
        const double* o = E + ((size_t)r * M + p) * 6;
        cplx v[3] = {mk(o[0], o[1]), mk(o[2], o[3]), mk(o[4], o[5])};
        if (inc) {
            const double* ke = inc + 6 * r;
            const double* x = pts + 3 * (size_t)p;
            const cplx deg = cexp(cplx{0., 1.} * (ke[0] * x[0] + ke[1] * x[1] + ke[2] * x[2]));
            for (int g = 0; g < 3; g++) v[g] = v[g] - deg * ke[3 + g];
        }
        for (int g = 0; g < 3; g++) e += v[g].re * v[g].re + v[g].im * v[g].im;
        if (H) {
            const double* q = H + ((size_t)r * M + p) * 6;
            for (int a = 0; a < 6; a++) h += q[a] * q[a];
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        e = fmax(e, __shfl_xor(e, off, 64));
        h = fmax(h, __shfl_xor(h, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomic_max_pos(mx + 2 * r, sqrt(e));
        atomic_max_pos(mx + 2 * r + 1, sqrt(h));
    }
}

// Per (point, current set): the error model above; flag[r][p] = 1 where the point
// needs FP64 sums.
__global__ __launch_bounds__(256) void k_fix_flags(FixArgs f) {
    __shared__ double sc[kFixCells][6];   // centre (3), area |jE|, area |jM|, unused
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    double x0 = 0., x1 = 0., x2 = 0.;
    if (p < f.M) { x0 = f.pts[3 * (size_t)p]; x1 = f.pts[3 * (size_t)p + 1]; x2 = f.pts[3 * (size_t)p + 2]; }
    const double k = f.k_re, k2 = k * k;
    double sE = 0., sH = 0.;
    for (int c0 = 0; c0 < f.N; c0 += kFixCells) {
        __syncthreads();
        for (int cc = threadIdx.x; cc < kFixCells && c0 + cc < f.N; cc += blockDim.x) {
            const size_t c = (size_t)(c0 + cc);
            const double* t = f.ctab + c * CT;
            sc[cc][0] = t[CT_CENTER]; sc[cc][1] = t[CT_CENTER + 1]; sc[cc][2] = t[CT_CENTER + 2];
            const double* je = f.jE + ((size_t)r * f.N + c) * 6;
            double a = 0.;
            for (int q = 0; q < 6; q++) a += je[q] * je[q];
            sc[cc][3] = t[CT_AREA] * sqrt(a);
            double b = 0.;
            if (f.jM) {
                const double* jm = f.jM + ((size_t)r * f.N + c) * 6;
                for (int q = 0; q < 6; q++) b += jm[q] * jm[q];
            }
            sc[cc][4] = t[CT_AREA] * sqrt(b);
        }
        __syncthreads();
        const int m = min(kFixCells, f.N - c0);
        for (int cc = 0; cc < m; cc++) {
            const double d0 = x0 - sc[cc][0], d1 = x1 - sc[cc][1], d2 = x2 - sc[cc][2];
            const double R2 = d0 * d0 + d1 * d1 + d2 * d2;
            const double ri = rsqrt(R2), ri2 = ri * ri, ri3 = ri2 * ri;
            const double a1 = k2 * ri - ri3, a2 = k2 * ri - 3. * ri3;
            const double fK = sqrt(a1 * a1 + k2 * ri2 * ri2) + sqrt(a2 * a2 + 9. * k2 * ri2 * ri2);
            const double fR = sqrt(1. + k2 * R2) * ri2;
            const double ke = sc[cc][3], km = sc[cc][4];
            sE += f.mE2 * (ke * fK) * (ke * fK) + (km * fR) * (km * fR);
            sH += (ke * fR) * (ke * fR) + f.mH2 * (km * fK) * (km * fK);
        }
    }
    if (p >= f.M) return;
    const double* o = f.E + ((size_t)r * f.M + p) * 6;
    cplx v[3] = {mk(o[0], o[1]), mk(o[2], o[3]), mk(o[4], o[5])};
    if (f.inc) {
        const double* kv = f.inc + 6 * r;
        const cplx deg = cexp(cplx{0., 1.} * (kv[0] * x0 + kv[1] * x1 + kv[2] * x2));
        for (int g = 0; g < 3; g++) v[g] = v[g] - deg * kv[3 + g];
    }
    double e = 0.;
    for (int g = 0; g < 3; g++) e += v[g].re * v[g].re + v[g].im * v[g].im;
    e = sqrt(e);
    const double estE = kFixKappa * kU32 * sqrt(sE) / (4. * kPi);
    bool flag = estE > kFixTau * fmax(e, 1e-3 * f.mx[2 * r]);
    if (f.H) {
        const double* q = f.H + ((size_t)r * f.M + p) * 6;
        double h = 0.;
        for (int a = 0; a < 6; a++) h += q[a] * q[a];
        const double estH = kFixKappa * kU32 * sqrt(sH) / (4. * kPi);
        flag = flag || estH > kFixTau * fmax(sqrt(h), 1e-3 * f.mx[2 * r + 1]);
    }
    f.flags[(size_t)r * f.M + p] = flag ? 1 : 0;
}

// list of the points flagged for any current set (order arbitrary; each point's
// FP64 result does not depend on the others)
__global__ __launch_bounds__(256) void k_fix_compact(int M, int nrhs, const int* __restrict__ flags,
                                                     int* __restrict__ count, int* __restrict__ list) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= M) return;
    int any = 0;
    for (int r = 0; r < nrhs; r++) any |= flags[(size_t)r * M + p];
    if (any) list[atomicAdd(count, 1)] = p;
}

__global__ void k_fix_zero(int n, double* a, int* count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.;
    if (i == 0) *count = 0;
}

__global__ __launch_bounds__(256) void k_fix_gather(int m, const int* __restrict__ list, const double* __restrict__ pts,
                                                    double* __restrict__ sub) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int p = list[i];
    for (int d = 0; d < 3; d++) sub[3 * (size_t)i + d] = pts[3 * (size_t)p + d];
}

// X[r][list[i]] = Xs[r][i] for both fields ([nrhs][M][3] complex / [nrhs][m][3])
__global__ __launch_bounds__(256) void k_fix_scatter(int m, int M, const int* __restrict__ list,
                                                     const double* __restrict__ Es, const double* __restrict__ Hs,
                                                     double* __restrict__ E, double* __restrict__ H) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= m) return;
    const size_t dst = ((size_t)r * M + list[i]) * 6, src = ((size_t)r * m + i) * 6;
    for (int a = 0; a < 6; a++) E[dst + a] = Es[src + a];
    if (H)
        for (int a = 0; a < 6; a++) H[dst + a] = Hs[src + a];
}

hipError_t launch_fix_flags(hipStream_t s, const FixArgs& f, int* count, int* list) {
    if (f.M <= 0 || f.nrhs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fix_zero, dim3((2 * f.nrhs + 255) / 256), dim3(256), 0, s, 2 * f.nrhs, f.mx, count);
    hipLaunchKernelGGL(k_scat_max, dim3((f.M + 255) / 256, f.nrhs), dim3(256), 0, s, f.M, f.pts, f.inc, f.E, f.H, f.mx);
    hipLaunchKernelGGL(k_fix_flags, dim3((f.M + 255) / 256, f.nrhs), dim3(256), 0, s, f);
    hipLaunchKernelGGL(k_fix_compact, dim3((f.M + 255) / 256), dim3(256), 0, s, f.M, f.nrhs, f.flags, count, list);
    return hipGetLastError();
}

hipError_t launch_fix_gather(hipStream_t s, int m, const int* list, const double* pts, double* sub) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fix_gather, dim3((m + 255) / 256), dim3(256), 0, s, m, list, pts, sub);
    return hipGetLastError();
}

hipError_t launch_fix_scatter(hipStream_t s, int m, int M, int nrhs, const int* list, const double* Es,
                              const double* Hs, double* E, double* H) {
    if (m <= 0 || nrhs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fix_scatter, dim3((m + 255) / 256, nrhs), dim3(256), 0, s, m, M, list, Es, Hs, E, H);
    return hipGetLastError();
}


// ---------------------------------------------------------------- operator matrix
// Field of many current sets: per segment (one mesh, a range of points) the
// kernels write the operator matrices
//   WK[(3p+g), (3c+h)] = K_RotRot_Far[e_h](x_p, cell_c)_g    (K.h:288-301)
//   WR[(3p+g), (3c+h)] = R_Rot_Colloc[e_h](x_p, cell_c)_g    (R.h:89-111)
// column-major with leading dimension 3M (complex), so that E = WK J and
// H = WR J are complex GEMMs with J = the current sets as a [3N x nrhs]
// matrix.  Far pairs come from the persistent tile kernel (node sums shared
// with k_field_far: K = (A I + T)/4pi, R = [G]x/4pi); near pairs from the
// exact per-pair kernel below.  Each entry is written by exactly one of them.
namespace {
constexpr double kInv4Pi = 0.079577471545947673;   // 1/(4 pi)

__device__ inline void put(double* W, size_t idx, double re, double im) {
    reinterpret_cast<double2*>(W)[idx] = make_double2(re, im);
}
}  // namespace

// Persistent: workgroup b takes tiles b, b + gridDim.x, ...; a tile is
// (segment, first point, cell range) = 256 points x up to plan cells.
template <bool CK, bool P32>
__global__ __launch_bounds__(kFT) void k_field_opmat(const OpSeg* __restrict__ segs, const int4* __restrict__ tiles,
                                                     int ntiles) {
    __shared__ double4 snode[P32 ? 1 : kFNodes];
    __shared__ float4 snode32[P32 ? kFNodes32 : 1];
    __shared__ float syy[P32 ? kFNodes32 : 1];
    __shared__ double scell[kFCells][4];
    const int tid = threadIdx.x;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tiles[t];
        const OpSeg sg = segs[tl.x];
        const int pt = tl.y + tid;
        const bool active = pt < sg.M;
        double x0 = 0., x1 = 0., x2 = 0.;
        if (active) { x0 = sg.pts[3 * (size_t)pt]; x1 = sg.pts[3 * (size_t)pt + 1]; x2 = sg.pts[3 * (size_t)pt + 2]; }
        const int nn = sg.nn;
        const int cb = max(1, min(kFCells, (P32 ? kFNodes32 : kFNodes) / nn));
        const double kr_ = sg.k_re, ki_ = sg.k_im;
        const double k2r = kr_ * kr_ - ki_ * ki_, k2i = 2. * kr_ * ki_;
        const size_t ld = 3 * (size_t)sg.M;
        for (int cs = tl.z; cs < tl.w; cs += cb) {
            const int ncs = min(cb, tl.w - cs);
            __syncthreads();
            if (P32) {
                const float4* gn = sg.nodes32 + (size_t)cs * nn;
                const float* gy = sg.nodes32_yy + (size_t)cs * nn;
                for (int idx = tid; idx < ncs * nn; idx += kFT) { snode32[idx] = gn[idx]; syy[idx] = gy[idx]; }
            } else {
                const double4* gn = reinterpret_cast<const double4*>(sg.nodes) + (size_t)cs * nn;
                for (int idx = tid; idx < ncs * nn; idx += kFT) snode[idx] = gn[idx];
            }
            for (int idx = tid; idx < ncs * 4; idx += kFT) {
                const int cc = idx / 4, f = idx % 4;
                scell[cc][f] = sg.ctab[(size_t)(cs + cc) * CT + (f < 3 ? CT_CENTER + f : CT_DIAM)];
            }
            __syncthreads();
            if (!active) continue;
            for (int cc = 0; cc < ncs; cc++) {
                const double* rec = scell[cc];
                const double dcen = center_dist(x0, x1, x2, rec);
                if (near_rule(dcen, rec[3], sg.mode, sg.ac, sg.thresh)) continue;
                cplx A, T[6], G[3];
                if (P32) {
                    if (sg.fast_phase)
                        pair_sums32<CK, true>(x0, x1, x2, rec, snode32 + cc * nn, syy + cc * nn, nn, dcen, kr_, ki_, A, T, G);
                    else
                        pair_sums32<CK, false>(x0, x1, x2, rec, snode32 + cc * nn, syy + cc * nn, nn, dcen, kr_, ki_, A, T, G);
                } else {
                    if (sg.fast_phase)
                        pair_sums64<CK, 1>(x0, x1, x2, snode + cc * nn, nn, dcen, kr_, ki_, k2r, k2i, A, T, G);
                    else
                        pair_sums64<CK, 0>(x0, x1, x2, snode + cc * nn, nn, dcen, kr_, ki_, k2r, k2i, A, T, G);
                }
                const int c = cs + cc;
                // K = (A I + T)/4pi, T = (xx, yy, zz, xy, xz, yz)
                const int tix[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
                for (int h = 0; h < 3; h++) {
                    const size_t col = (size_t)(3 * c + h) * ld + 3 * (size_t)pt;
                    for (int g = 0; g < 3; g++) {
                        cplx v = T[tix[g][h]];
                        if (g == h) v += A;
                        put(sg.WK, col + g, v.re * kInv4Pi, v.im * kInv4Pi);
                    }
                }
                if (sg.WR) {
                    // R[j] = (G x j)/4pi; exactly zero when x is the centroid (R.h:94-97)
                    const double z = dcen < kMachineZero ? 0. : kInv4Pi;
                    const cplx Gs[3] = {G[0] * z, G[1] * z, G[2] * z};
                    const cplx zero = mk(0., 0.);
                    const cplx Rm[3][3] = {{zero, -Gs[2], Gs[1]}, {Gs[2], zero, -Gs[0]}, {-Gs[1], Gs[0], zero}};
                    for (int h = 0; h < 3; h++) {
                        const size_t col = (size_t)(3 * c + h) * ld + 3 * (size_t)pt;
                        for (int g = 0; g < 3; g++) put(sg.WR, col + g, Rm[g][h].re, Rm[g][h].im);
                    }
                }
            }
        }
    }
}

hipError_t launch_field_opmat(hipStream_t s, const OpSeg* segs_d, const int4* tiles_d, int ntiles, int grid, bool ck,
                              bool p32) {
    if (ntiles <= 0) return hipSuccess;
    dim3 g(grid), b(kFT);
    if (ck) {
        if (p32) hipLaunchKernelGGL((k_field_opmat<true, true>), g, b, 0, s, segs_d, tiles_d, ntiles);
        else hipLaunchKernelGGL((k_field_opmat<true, false>), g, b, 0, s, segs_d, tiles_d, ntiles);
    } else {
        if (p32) hipLaunchKernelGGL((k_field_opmat<false, true>), g, b, 0, s, segs_d, tiles_d, ntiles);
        else hipLaunchKernelGGL((k_field_opmat<false, false>), g, b, 0, s, segs_d, tiles_d, ntiles);
    }
    return hipGetLastError();
}

// Near pairs of one segment, exact (as k_field_near_mat), written into WK/WR:
// column h of a pair's block = Op(e_h).  lq = point index within the segment.
template <int NV>
__global__ __launch_bounds__(64) void k_opmat_near(long long np_, const int* __restrict__ lq,
                                                   const int* __restrict__ lc, OpSeg sg, NumP np, int mode) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np_) return;
    const int q = lq[p], c = lc[p];
    double cv[NV][3], x[3];
    for (int v = 0; v < NV; v++)
        for (int d = 0; d < 3; d++) cv[v][d] = sg.verts[((size_t)c * NV + v) * 3 + d];
    for (int d = 0; d < 3; d++) x[d] = sg.pts[3 * (size_t)q + d];
    const cplx k = mk(sg.k_re, sg.k_im);
    const int opK = (mode == 0) ? OP_K_NEAR_SMOOTH : OP_K_HS;
    const size_t ld = 3 * (size_t)sg.M;
    for (int h = 0; h < 3; h++) {
        cplx j[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, r[3];
        j[h] = mk(1., 0.);
        const size_t col = (size_t)(3 * c + h) * ld + 3 * (size_t)q;
        op_eval<NV>(opK, cv, x, j, np, k, r);
        for (int g = 0; g < 3; g++) put(sg.WK, col + g, r[g].re, r[g].im);
        if (sg.WR) {
            op_eval<NV>(OP_R_SMOOTH, cv, x, j, np, k, r);
            for (int g = 0; g < 3; g++) put(sg.WR, col + g, r[g].re, r[g].im);
        }
    }
}

hipError_t launch_opmat_near(hipStream_t s, long long npairs, const int* lq, const int* lc, const OpSeg& sg,
                             const NumP& np, int mode) {
    if (npairs <= 0) return hipSuccess;
    dim3 g((unsigned)((npairs + 63) / 64)), b(64);
    if (sg.nv == 4) hipLaunchKernelGGL(k_opmat_near<4>, g, b, 0, s, npairs, lq, lc, sg, np, mode);
    else hipLaunchKernelGGL(k_opmat_near<3>, g, b, 0, s, npairs, lq, lc, sg, np, mode);
    return hipGetLastError();
}

// E[r][p] = e0_r exp(i kvec_r . x_p) (PlaneWaveFieldE, incident_field.h:20-29)
// from inc[r] = (kvec, e0), or 0 without inc; E is [nrhs][M][3] complex.
__global__ __launch_bounds__(256) void k_init_incident(int M, const double* __restrict__ pts,
                                                       const double* __restrict__ inc, double* __restrict__ E) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (p >= M) return;
    double* o = E + ((size_t)r * M + p) * 6;
    if (!inc) {
        for (int a = 0; a < 6; a++) o[a] = 0.;
        return;
    }
    const double* ke = inc + 6 * r;
    const double* x = pts + 3 * (size_t)p;
    const double ph = ke[0] * x[0] + ke[1] * x[1] + ke[2] * x[2];
    const cplx deg = cexp(cplx{0., 1.} * ph);
    for (int g = 0; g < 3; g++) {
        const cplx e = deg * ke[3 + g];
        o[2 * g] = e.re;
        o[2 * g + 1] = e.im;
    }
}

hipError_t launch_init_incident(hipStream_t s, int M, int nrhs, const double* pts, const double* inc, double* E) {
    if (M <= 0 || nrhs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_init_incident, dim3((M + 255) / 256, nrhs), dim3(256), 0, s, M, pts, inc, E);
    return hipGetLastError();
}

// out[:, r] (+)= sum_b part[b][:, r]: the split-K partial products of the
// operator-matrix GEMMs (each part m x nrhs, dense), in a fixed order.
__global__ __launch_bounds__(256) void k_sum_partials(long long m, int nrhs, int nparts,
                                                      const double2* __restrict__ part, double2* __restrict__ out,
                                                      long long ldo, int accumulate) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= m) return;
    double2 acc = accumulate ? out[(size_t)r * ldo + i] : make_double2(0., 0.);
    const size_t stride = (size_t)m * nrhs;
    for (int b = 0; b < nparts; b++) {
        const double2 v = part[b * stride + (size_t)r * m + i];
        acc.x += v.x;
        acc.y += v.y;
    }
    out[(size_t)r * ldo + i] = acc;
}

hipError_t launch_sum_partials(hipStream_t s, long long m, int nrhs, int nparts, const double* part, double* out,
                               long long ldo, int accumulate) {
    if (m <= 0 || nrhs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)((m + 255) / 256), nrhs), dim3(256), 0, s, m, nrhs, nparts,
                       reinterpret_cast<const double2*>(part), reinterpret_cast<double2*>(out), ldo, accumulate);
    return hipGetLastError();
}

}  // namespace bielcc_impl
