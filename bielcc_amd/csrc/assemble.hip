// Collocation matrix, right-hand sides and currents.
//
// The reference documents K_RotRot_HS_Colloc (K.h:156-204) and R_Rot_Colloc
// (R.h:65-111) as the operators for "assembling the matrix at collocation
// points" but has no assembly loop.  This library defines it as
//   Z[2i+a, 2j+b] = tau_i^a . Op(tau_j^b, rkt_i, cell_j)
// (unknowns cell-major / basis-minor as SurfaceCurrents_byBasis reads them,
// electrodynamics/currents.h:22-33), stored column-major, complex FP64.
//
// Far pairs: one tile kernel; 256 collocation points per workgroup (one per
// lane) x a range of cells whose surface and edge node tables sit in LDS.
// For K the per-pair work is S = sum w e^{ikr}/r (64 nodes) and, per edge e,
// g_e = sum dl (ikr-1) e^{ikr} (x-y)/r^3 (10 nodes), shared by both basis
// vectors b: K0_b = -sum_e (tau_j^b . nu_e) g_e.  Near pairs
// (IntegralSurf_SimplePot_H_HS's analytic branch, integrals_analytic.h:151-153)
// are skipped here and written by the exact per-pair kernel.
//
// Mixed precision (P32): the surface sums (S for K, G for R) run in FP32 in
// cell-local coordinates as in the field kernel (node phase from
// r - dcen = (|y_c|^2 - 2 x_c.y_c)/(r + dcen), centre phase and quadrant in
// FP64); the edge contour of K0, whose four edge sums cancel to O(diam/r) of
// each other, stays FP64.
#include "bielcc_internal.h"
#include "phase.h"

using namespace bielcc_dev;

namespace bielcc_impl {

constexpr int kAT = 256;
constexpr int kASurf = 512;     // surface node slots in LDS
constexpr int kASeg = 256;      // edge node slots in LDS
constexpr int kACells = 8;
constexpr int kARec = 24;       // center 3, diam, tau1 3, tau2 3, nu 12, pad 2

// FP32 surface sums of one (point, cell) pair in cell-local coordinates.
// WHICH 0: S = sum w e^{ikr}/r; WHICH 1: G = sum w (ikr-1) e^{ikr} (x-y)/r^3.
// Returned rotated by i^n and scaled by e^{-ki dcen} (FP64).
template <bool CK, int WHICH, bool FP>
__device__ inline void surf_sums32(const float4* nd, const float* ny, int nn, const double* x, const double* cen,
                                   double dcen, double kr_, double ki_, double* out_re, double* out_im) {
    const PhaseRef pr = phase_ref(kr_ * dcen);
    const float t0 = (float)(fma(kr_, dcen, -pr.o1) - pr.o2 - pr.o3);
    const float xf0 = (float)(x[0] - cen[0]), xf1 = (float)(x[1] - cen[1]), xf2 = (float)(x[2] - cen[2]);
    const float m0 = -2.f * xf0, m1 = -2.f * xf1, m2 = -2.f * xf2;
    const float dcf = (float)dcen, kf = (float)kr_, kif = (float)ki_;
    float Sr = 0.f, Si = 0.f;
    float Gr[3] = {0, 0, 0}, Gi[3] = {0, 0, 0};
    for (int q = 0; q < nn; q++) {
        const float4 y = nd[q];
        const float d0 = xf0 - y.x, d1 = xf1 - y.y, d2 = xf2 - y.z;
        const float r2 = d0 * d0 + d1 * d1 + d2 * d2;
        const float ri = rsqrtf(r2);
        const float r = r2 * ri;
        const float del = fmaf(m0, y.x, fmaf(m1, y.y, fmaf(m2, y.z, ny[q])));
        const float dr = del * __builtin_amdgcn_rcpf(r + dcf);   // v_rcp_f32 (1 ulp): 2 instructions, not a full division
        const float t = fmaf(kf, dr, t0);
        float s, c;
        if (FP) sincos_near32(t, s, c);
        else sincosf(t, &s, &c);
        if (WHICH == 0) {
            float w = y.w * ri;
            if (CK) w *= __expf(-kif * dr);
            Sr += w * c;
            Si += w * s;
        } else {
            float w = y.w * ri * ri * ri;
            if (CK) w *= __expf(-kif * dr);
            const float mr = CK ? -(1.f + kif * r) : -1.f, mi = kf * r;
            const float gr = (mr * c - mi * s) * w, gi = (mr * s + mi * c) * w;
            Gr[0] += gr * d0; Gi[0] += gi * d0;
            Gr[1] += gr * d1; Gi[1] += gi * d1;
            Gr[2] += gr * d2; Gi[2] += gi * d2;
        }
    }
    const double amp = CK ? exp(-ki_ * dcen) : 1.;
    const int ncomp = WHICH == 0 ? 1 : 3;
    for (int g = 0; g < ncomp; g++) {
        double re = WHICH == 0 ? (double)Sr : (double)Gr[g], im = WHICH == 0 ? (double)Si : (double)Gi[g];
        rot_q(re, im, pr.q);
        out_re[g] = re * amp;
        out_im[g] = im * amp;
    }
}

// One tile of far pairs: collocation points bx*256 .. +255 (one per lane) x
// cells [c_begin, c_end), staged through the LDS arrays handed in.
// PH: node phase evaluation, 0 general sincos, 1 quadrant reduction (phase.h),
// 2 centre-phase reduction with the short series (phase.h)
template <bool CK, int WHICH, int PH, bool P32>
__device__ __forceinline__ void assemble_tile(const AssembleArgs& a, int bx, int c_begin, int c_end, double4* ssurf,
                                              double4* sseg, double (*scell)[kARec]) {
    const int tid = threadIdx.x;
    const int li = bx * kAT + tid;          // local point index
    const bool active = li < a.ni;
    const int i = a.i0 + li;
    double x[3] = {0., 0., 0.}, ta[2][3] = {{0, 0, 0}, {0, 0, 0}};
    if (active) {
        const double* t = a.ctab + (size_t)i * CT;
        for (int d = 0; d < 3; d++) { x[d] = t[CT_RKT + d]; ta[0][d] = t[CT_TAU1 + d]; ta[1][d] = t[CT_TAU2 + d]; }
    }
    const int nn = a.nn, nv = a.nv, ns = a.ns, nseg = nv * ns;
    int cb = kACells;
    if (WHICH == 0) cb = min(cb, min(kASurf / nn, kASeg / nseg));
    else cb = min(cb, kASurf / nn);
    cb = max(cb, 1);
    const double kr_ = a.k_re, ki_ = a.k_im;
    const cplx k = mk(kr_, ki_);

    for (int cs = c_begin; cs < c_end; cs += cb) {
        const int ncs = min(cb, c_end - cs);
        __syncthreads();
        if (P32) {   // FP32 nodes (16 B) and |y_c|^2 (4 B) in the space of the FP64 table
            float4* s32 = reinterpret_cast<float4*>(ssurf);
            float* sy32 = reinterpret_cast<float*>(s32 + kASurf);
            const float4* g4 = a.nodes32 + (size_t)cs * nn;
            const float* gy = a.nodes32_yy + (size_t)cs * nn;
            for (int idx = tid; idx < ncs * nn; idx += kAT) { s32[idx] = g4[idx]; sy32[idx] = gy[idx]; }
        } else {
            const double4* gs = reinterpret_cast<const double4*>(a.nodes) + (size_t)cs * nn;
            for (int idx = tid; idx < ncs * nn; idx += kAT) ssurf[idx] = gs[idx];
        }
        if (WHICH == 0) {
            const double4* gg = reinterpret_cast<const double4*>(a.segs) + (size_t)cs * nseg;
            for (int idx = tid; idx < ncs * nseg; idx += kAT) sseg[idx] = gg[idx];
        }
        for (int idx = tid; idx < ncs * kARec; idx += kAT) {
            int cc = idx / kARec, f = idx % kARec;
            const double* t = a.ctab + (size_t)(cs + cc) * CT;
            double v = 0.;
            if (f < 3) v = t[CT_CENTER + f];
            else if (f == 3) v = t[CT_DIAM];
            else if (f < 7) v = t[CT_TAU1 + f - 4];
            else if (f < 10) v = t[CT_TAU2 + f - 7];
            else if (f < 22) v = t[CT_NU + f - 10];
            scell[cc][f] = v;
        }
        __syncthreads();
        if (!active) continue;
        for (int cc = 0; cc < ncs; cc++) {
            const int c = cs + cc;
            const double* rec = scell[cc];
            const double dcen = center_dist(x[0], x[1], x[2], rec);
            constexpr bool FP = PH != 0;   // FP32 node sums: quadrant reduction in either mode
            PhaseRef pr{0., 0., 0., 0};
            PhaseC pc{0., 1., 0.};
            if (PH == 1) pr = phase_ref(kr_ * dcen);
            cplx z[2][2];
            if (WHICH == 0) {
                if (near_rule(dcen, rec[3], 0, a.ac, 0.)) continue;   // near pair: exact kernel writes it
                if (PH >= 2) pc = phase_centre(kr_, dcen);
                // S = int e^{ikr}/r  (F_SimplePot_H, smoothR = DBL_EPSILON: factor 1/r)
                double Sr = 0., Si = 0.;
                const double4* nd = ssurf + cc * nn;
                if (P32) {
                    const float4* s32 = reinterpret_cast<const float4*>(ssurf);
                    const float* sy32 = reinterpret_cast<const float*>(s32 + kASurf);
                    surf_sums32<CK, 0, FP>(s32 + cc * nn, sy32 + cc * nn, nn, x, rec, dcen, kr_, ki_, &Sr, &Si);
                }
                for (int q = 0; q < (P32 ? 0 : nn); q++) {
                    const double4 y = nd[q];
                    const double d0 = x[0] - y.x, d1 = x[1] - y.y, d2 = x[2] - y.z;
                    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                    const double ri = rsqrt_pos(r2);
                    const double r = r2 * ri;
                    double s, co;
                    phase_eval<PH>(kr_, r, pr, pc, s, co);
                    double w = y.w * ri;
                    if (CK) w *= exp(-ki_ * r);
                    Sr += w * co;
                    Si += w * s;
                }
                // K0 with F_GradSimplePot_H, smoothR = smoothRMin = DBL_EPSILON (K.h:189-200)
                cplx k0b[2][3] = {{mk(0, 0), mk(0, 0), mk(0, 0)}, {mk(0, 0), mk(0, 0), mk(0, 0)}};
                const double4* sg = sseg + cc * nseg;
                for (int e = 0; e < nv; e++) {
                    double gr0 = 0., gi0 = 0., gr1 = 0., gi1 = 0., gr2 = 0., gi2 = 0.;
                    for (int q = 0; q < ns; q++) {
                        const double4 y = sg[e * ns + q];
                        const double d0 = x[0] - y.x, d1 = x[1] - y.y, d2 = x[2] - y.z;
                        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                        const double ri = rsqrt_pos(r2);
                        const double r = r2 * ri;
                        double s, co;
                        phase_eval<PH>(kr_, r, pr, pc, s, co);
                        double w = y.w * ri * ri * ri;
                        if (CK) w *= exp(-ki_ * r);
                        const double mr = CK ? -(1. + ki_ * r) : -1., mi = kr_ * r;
                        const double gr = (mr * co - mi * s) * w, gi = (mr * s + mi * co) * w;
                        gr0 += gr * d0; gi0 += gi * d0;
                        gr1 += gr * d1; gi1 += gi * d1;
                        gr2 += gr * d2; gi2 += gi * d2;
                    }
                    const double* nu = rec + 10 + 3 * e;
                    for (int b = 0; b < 2; b++) {
                        const double* tb = rec + 4 + 3 * b;
                        const double jn = -(tb[0] * nu[0] + tb[1] * nu[1] + tb[2] * nu[2]);
                        k0b[b][0] += mk(gr0, gi0) * jn;
                        k0b[b][1] += mk(gr1, gi1) * jn;
                        k0b[b][2] += mk(gr2, gi2) * jn;
                    }
                }
                if (PH == 1) {   // restore the common factor i^n of the cell's e^{ikr}
                    if (!P32) rot_q(Sr, Si, pr.q);
                    for (int b = 0; b < 2; b++)
                        for (int g = 0; g < 3; g++) rot_q(k0b[b][g].re, k0b[b][g].im, pr.q);
                } else if (PH >= 2) {   // ... or e^{i k dcen}
                    if (!P32) rot_c(Sr, Si, pc);
                    for (int b = 0; b < 2; b++)
                        for (int g = 0; g < 3; g++) rot_c(k0b[b][g].re, k0b[b][g].im, pc);
                }
                const cplx Sk2 = (mk(Sr, Si) * k) * k;
                for (int b = 0; b < 2; b++) {
                    const double* tb = rec + 4 + 3 * b;
                    cplx res[3];
                    for (int g = 0; g < 3; g++) res[g] = (Sk2 * tb[g]) / (4. * kPi) + k0b[b][g] / (4. * kPi);
                    for (int aa = 0; aa < 2; aa++)
                        z[aa][b] = res[0] * ta[aa][0] + res[1] * ta[aa][1] + res[2] * ta[aa][2];
                }
            } else {
                if (dcen < kMachineZero) {
                    z[0][0] = z[0][1] = z[1][0] = z[1][1] = mk(0., 0.);
                } else {
                    double Gr[3] = {0, 0, 0}, Gi[3] = {0, 0, 0};
                    // near pairs keep FP64 sums: for almost coplanar neighbours the
                    // tau_i . (G x tau_j) projection cancels to O(diam/R) of |G|.
                    // In mixed mode LDS holds the FP32 table, so they read the FP64
                    // nodes from global memory (few pairs, L2-resident).
                    const bool use32 = P32 && !near_rule(dcen, rec[3], 0, a.ac, 0.);
                    if (PH >= 2 && !use32) pc = phase_centre(kr_, dcen);
                    const double4* nd = P32 ? reinterpret_cast<const double4*>(a.nodes) + (size_t)c * nn
                                            : ssurf + cc * nn;
                    if (use32) {
                        const float4* s32 = reinterpret_cast<const float4*>(ssurf);
                        const float* sy32 = reinterpret_cast<const float*>(s32 + kASurf);
                        surf_sums32<CK, 1, FP>(s32 + cc * nn, sy32 + cc * nn, nn, x, rec, dcen, kr_, ki_, Gr, Gi);
                    }
                    for (int q = 0; q < (use32 ? 0 : nn); q++) {
                        const double4 y = nd[q];
                        const double d0 = x[0] - y.x, d1 = x[1] - y.y, d2 = x[2] - y.z;
                        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                        const double ri = rsqrt_pos(r2);
                        const double r = r2 * ri;
                        double s, co;
                        phase_eval<PH>(kr_, r, pr, pc, s, co);
                        double w = y.w * ri * ri * ri;
                        if (CK) w *= exp(-ki_ * r);
                        const double mr = CK ? -(1. + ki_ * r) : -1., mi = kr_ * r;
                        const double gr = (mr * co - mi * s) * w, gi = (mr * s + mi * co) * w;
                        Gr[0] += gr * d0; Gi[0] += gi * d0;
                        Gr[1] += gr * d1; Gi[1] += gi * d1;
                        Gr[2] += gr * d2; Gi[2] += gi * d2;
                    }
                    if (PH == 1 && !use32)
                        for (int g = 0; g < 3; g++) rot_q(Gr[g], Gi[g], pr.q);
                    if (PH >= 2 && !use32)
                        for (int g = 0; g < 3; g++) rot_c(Gr[g], Gi[g], pc);
                    const cplx G[3] = {mk(Gr[0], Gi[0]), mk(Gr[1], Gi[1]), mk(Gr[2], Gi[2])};
                    for (int b = 0; b < 2; b++) {
                        const double* tb = rec + 4 + 3 * b;
                        cplx res[3] = {(G[1] * tb[2] - G[2] * tb[1]) / (4. * kPi),
                                       (G[2] * tb[0] - G[0] * tb[2]) / (4. * kPi),
                                       (G[0] * tb[1] - G[1] * tb[0]) / (4. * kPi)};
                        for (int aa = 0; aa < 2; aa++)
                            z[aa][b] = res[0] * ta[aa][0] + res[1] * ta[aa][1] + res[2] * ta[aa][2];
                    }
                }
            }
            for (int b = 0; b < 2; b++) {
                double2* col = reinterpret_cast<double2*>(a.Z) + (size_t)(2 * c + b) * a.ld + 2 * (size_t)li;
                col[0] = make_double2(z[0][b].re, z[0][b].im);
                col[1] = make_double2(z[1][b].re, z[1][b].im);
            }
        }
    }
}

template <bool CK, int WHICH, int PH, bool P32>
__global__ __launch_bounds__(kAT) void k_assemble_far(AssembleArgs a, int cells_per_split) {
    __shared__ double4 ssurf[kASurf];
    __shared__ double4 sseg[kASeg];
    __shared__ double scell[kACells][kARec];
    const int c_begin = blockIdx.y * cells_per_split;
    assemble_tile<CK, WHICH, PH, P32>(a, blockIdx.x, c_begin, min(a.N, c_begin + cells_per_split), ssurf, sseg,
                                      scell);
}

// Persistent form over several meshes: workgroup b takes tiles b, b + gridDim.x,
// ...; a tile is (item, point block, cell begin, cell end).  The reduced-phase
// choice is per item.
template <bool CK, int WHICH, bool P32>
__global__ __launch_bounds__(kAT) void k_assemble_multi(const AssembleArgs* __restrict__ items,
                                                        const int4* __restrict__ tiles, int ntiles) {
    __shared__ double4 ssurf[kASurf];
    __shared__ double4 sseg[kASeg];
    __shared__ double scell[kACells][kARec];
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tiles[t];
        const AssembleArgs a = items[tl.x];
        if (a.fast_phase) assemble_tile<CK, WHICH, 1, P32>(a, tl.y, tl.z, tl.w, ssurf, sseg, scell);
        else assemble_tile<CK, WHICH, 0, P32>(a, tl.y, tl.z, tl.w, ssurf, sseg, scell);
    }
}

template <bool CK, int W, bool P32>
static void assemble_multi_t(hipStream_t s, const AssembleArgs* items_d, const int4* tiles_d, int ntiles, int grid) {
    hipLaunchKernelGGL((k_assemble_multi<CK, W, P32>), dim3(grid), dim3(kAT), 0, s, items_d, tiles_d, ntiles);
}

hipError_t launch_assemble_multi(hipStream_t s, const AssembleArgs* items_d, const int4* tiles_d, int ntiles,
                                 int grid, int which, bool ck, bool p32) {
    if (ntiles <= 0) return hipSuccess;
    if (which == 0) {
        if (ck) { if (p32) assemble_multi_t<true, 0, true>(s, items_d, tiles_d, ntiles, grid);
                  else assemble_multi_t<true, 0, false>(s, items_d, tiles_d, ntiles, grid); }
        else { if (p32) assemble_multi_t<false, 0, true>(s, items_d, tiles_d, ntiles, grid);
               else assemble_multi_t<false, 0, false>(s, items_d, tiles_d, ntiles, grid); }
    } else {
        if (ck) { if (p32) assemble_multi_t<true, 1, true>(s, items_d, tiles_d, ntiles, grid);
                  else assemble_multi_t<true, 1, false>(s, items_d, tiles_d, ntiles, grid); }
        else { if (p32) assemble_multi_t<false, 1, true>(s, items_d, tiles_d, ntiles, grid);
               else assemble_multi_t<false, 1, false>(s, items_d, tiles_d, ntiles, grid); }
    }
    return hipGetLastError();
}

template <bool CK, int W, int PH, bool P32>
static void assemble_far_ph(hipStream_t s, const AssembleArgs& a, int s_min) {
    const int bx = (a.ni + kAT - 1) / kAT;
    const void* kern = reinterpret_cast<const void*>(&k_assemble_far<CK, W, PH, P32>);
    const int nsplit = tail_aware_splits(bx, s_min, (a.N + 63) / 64, resident_slots(kern, kAT));
    int cps = (a.N + nsplit - 1) / nsplit;
    dim3 g(bx, nsplit), b(kAT);
    hipLaunchKernelGGL((k_assemble_far<CK, W, PH, P32>), g, b, 0, s, a, cps);
}

template <bool CK, int W, bool P32>
static void assemble_far_p(hipStream_t s, const AssembleArgs& a, int s_min) {
    if (a.fast_phase == 3) assemble_far_ph<CK, W, 3, P32>(s, a, s_min);
    else if (a.fast_phase == 2) assemble_far_ph<CK, W, 2, P32>(s, a, s_min);
    else if (a.fast_phase) assemble_far_ph<CK, W, 1, P32>(s, a, s_min);
    else assemble_far_ph<CK, W, 0, P32>(s, a, s_min);
}

template <bool CK, int W>
static void assemble_far_t(hipStream_t s, const AssembleArgs& a, int nsplit) {
    if (a.prec) assemble_far_p<CK, W, true>(s, a, nsplit);
    else assemble_far_p<CK, W, false>(s, a, nsplit);
}

bool fast_phase_ok(double k_re, double grid_step) {
    return 0.7853981633974483 + fabs(k_re) * grid_step <= kPhaseRange;
}

int phase_mode(double k_re, double grid_step, double max_rad) {
    if (fabs(k_re) * max_rad <= kVShortRange) return 3;
    if (fabs(k_re) * max_rad <= kShortRange) return 2;
    return fast_phase_ok(k_re, grid_step) ? 1 : 0;
}

hipError_t launch_assemble_far(hipStream_t s, const AssembleArgs& a) {
    if (a.prec && (!a.nodes32 || !a.nodes32_yy)) return hipErrorInvalidValue;
    // enough workgroups to fill 256 CUs several times over
    int bx = (a.ni + kAT - 1) / kAT;
    const int wg = a.wg_target > 0 ? a.wg_target : 2048;
    int nsplit = (wg + bx - 1) / bx;
    nsplit = max(1, min(nsplit, (a.N + 63) / 64));
    bool ck = a.k_im != 0.;
    if (a.which == 0) { if (ck) assemble_far_t<true, 0>(s, a, nsplit); else assemble_far_t<false, 0>(s, a, nsplit); }
    else { if (ck) assemble_far_t<true, 1>(s, a, nsplit); else assemble_far_t<false, 1>(s, a, nsplit); }
    return hipGetLastError();
}

// Exact per-pair blocks: listed pairs (lq = point index relative to i0, lc =
// cell) or, with lq == nullptr, every pair of points i0..i0+ni-1 x all cells.  Z_ab = tau_i^a . Op(tau_j^b, rkt_i, cell_j).
template <int NV>
__global__ __launch_bounds__(64) void k_assemble_generic(int which, long long npairs, const int* __restrict__ lq,
                                                         const int* __restrict__ lc, int N, int i0,
                                                         const double* __restrict__ verts,
                                                         const double* __restrict__ ctab, NumP np, double k_re,
                                                         double k_im, double* __restrict__ Z, long long ld) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    int i, c;
    if (lq) { i = i0 + lq[p]; c = lc[p]; }
    else { i = i0 + (int)(p / N); c = (int)(p % N); }
    double cv[NV][3], x[3], ta[2][3];
    for (int v = 0; v < NV; v++)
        for (int d = 0; d < 3; d++) cv[v][d] = verts[((size_t)c * NV + v) * 3 + d];
    const double* ti = ctab + (size_t)i * CT;
    const double* tc = ctab + (size_t)c * CT;
    for (int d = 0; d < 3; d++) { x[d] = ti[CT_RKT + d]; ta[0][d] = ti[CT_TAU1 + d]; ta[1][d] = ti[CT_TAU2 + d]; }
    const cplx k = mk(k_re, k_im);
    for (int b = 0; b < 2; b++) {
        const double* tb = tc + CT_TAU1 + 3 * b;
        cplx j[3] = {mk(tb[0], 0.), mk(tb[1], 0.), mk(tb[2], 0.)}, res[3];
        op_eval<NV>(which == 0 ? OP_K_HS_COLLOC : OP_R_COLLOC, cv, x, j, np, k, res);
        double2* col = reinterpret_cast<double2*>(Z) + (size_t)(2 * c + b) * ld + 2 * (size_t)(i - i0);
        for (int aa = 0; aa < 2; aa++) {
            cplx z = res[0] * ta[aa][0] + res[1] * ta[aa][1] + res[2] * ta[aa][2];
            col[aa] = make_double2(z.re, z.im);
        }
    }
}

hipError_t launch_assemble_generic(hipStream_t s, int nv, int which, long long npairs, const int* lq,
                                   const int* lc, int N, int i0, int ni, const double* verts, const double* ctab,
                                   const NumP& np, double k_re, double k_im, double* Z, long long ld) {
    (void)ni;
    if (npairs <= 0) return hipSuccess;
    dim3 g((unsigned)((npairs + 63) / 64)), b(64);
    if (nv == 4) hipLaunchKernelGGL(k_assemble_generic<4>, g, b, 0, s, which, npairs, lq, lc, N, i0, verts, ctab, np, k_re, k_im, Z, ld);
    else hipLaunchKernelGGL(k_assemble_generic<3>, g, b, 0, s, which, npairs, lq, lc, N, i0, verts, ctab, np, k_re, k_im, Z, ld);
    return hipGetLastError();
}

// rhs[2i+a, r] = -tau_i^a . E_inc_r(rkt_i), E_inc_r = e0_r exp(i kvec_r . x)
// (PlaneWaveFieldE incident_field.h:20-29; perfect conductor: the tangential
// total field vanishes).  kvec_e0: [nrhs][6] device array (kvec, e0).
__global__ void k_rhs_planewave(int N, const double* __restrict__ ctab, int nrhs,
                                const double* __restrict__ kvec_e0, double* __restrict__ B, long long ldb) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)N * nrhs) return;
    int c = (int)(idx % N), r = (int)(idx / N);
    const double* t = ctab + (size_t)c * CT;
    const double* ke = kvec_e0 + 6 * (size_t)r;
    double ph = ke[0] * t[CT_RKT] + ke[1] * t[CT_RKT + 1] + ke[2] * t[CT_RKT + 2];
    cplx deg = cexp(cplx{0., 1.} * ph);
    cplx e[3] = {deg * ke[3], deg * ke[4], deg * ke[5]};
    double2* out = reinterpret_cast<double2*>(B) + (size_t)r * ldb + 2 * (size_t)c;
    for (int aa = 0; aa < 2; aa++) {
        const double* ta = t + CT_TAU1 + 3 * aa;
        cplx z = -(e[0] * ta[0] + e[1] * ta[1] + e[2] * ta[2]);
        out[aa] = make_double2(z.re, z.im);
    }
}

hipError_t launch_rhs_planewave(hipStream_t s, int N, const double* ctab, int nrhs, const double* kvec_e0,
                                double* B, long long ldb) {
    long long tot = (long long)N * nrhs;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rhs_planewave, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, N, ctab, nrhs,
                       kvec_e0, B, ldb);
    return hipGetLastError();
}

// SurfaceCurrents_byBasis (currents.h:22-33): j_c = b_{2c} tau_c^1 + b_{2c+1} tau_c^2
__global__ void k_currents(int N, const double* __restrict__ ctab, int nrhs, const double* __restrict__ b,
                           long long ldb, double* __restrict__ j) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)N * nrhs) return;
    int c = (int)(idx % N), r = (int)(idx / N);
    const double* t = ctab + (size_t)c * CT;
    const double2* bb = reinterpret_cast<const double2*>(b) + (size_t)r * ldb + 2 * (size_t)c;
    cplx b0 = mk(bb[0].x, bb[0].y), b1 = mk(bb[1].x, bb[1].y);
    double* o = j + ((size_t)r * N + c) * 6;
    for (int d = 0; d < 3; d++) {
        cplx v = b0 * t[CT_TAU1 + d] + b1 * t[CT_TAU2 + d];
        o[2 * d] = v.re; o[2 * d + 1] = v.im;
    }
}

hipError_t launch_currents(hipStream_t s, int N, const double* ctab, int nrhs, const double* b, long long ldb,
                           double* j) {
    long long tot = (long long)N * nrhs;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_currents, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, N, ctab, nrhs, b, ldb, j);
    return hipGetLastError();
}

}  // namespace bielcc_impl
