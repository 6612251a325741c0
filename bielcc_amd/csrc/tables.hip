// Per-cell tables and near-pair lists.
//
// The reference recomputes every x-independent quantity (CellFill data,
// diameters, sub-cell centroids and weights, edge nodes) for each (point,
// cell) pair (integrators/integral_universal.h:59-94, electrodynamics/K0.h:52-79).
// Here they are computed once per mesh, in FP64 with the reference's formulas
// and node order, and stored in HBM for the tiled kernels.
#include "bielcc_internal.h"

// tables and near lists round like the reference (no contraction): node
// positions equal the reference's bit for bit
#pragma clang fp contract(off)

using namespace bielcc_dev;

namespace bielcc_impl {

template <int NV>
__global__ void k_cell_tables(int n, const double* __restrict__ verts, double* __restrict__ ctab) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double cv[NV][3];
    for (int v = 0; v < NV; v++)
        for (int d = 0; d < 3; d++) cv[v][d] = verts[((size_t)c * NV + v) * 3 + d];
    double* t = ctab + (size_t)c * CT;
    double nrm[3], t1[3], t2[3], len, area;
    if (NV == 4) {   // CellQuad::CellFill mesh/cell.cpp:8-51
        double a[3], b[3], cc[3], d[3], ac[3], bd[3];
        for (int i = 0; i < 3; i++) t[CT_RKT + i] = (cv[0][i] + cv[1][i] + cv[2][i] + cv[3][i]) * 0.25;
        for (int i = 0; i < 3; i++) {
            a[i] = (cv[0][i] + cv[1][i]) / 2.0; b[i] = (cv[1][i] + cv[2][i]) / 2.0;
            cc[i] = (cv[2][i] + cv[3][i]) / 2.0; d[i] = (cv[3][i] + cv[0][i]) / 2.0;
            ac[i] = cc[i] - a[i]; bd[i] = d[i] - b[i];
        }
        cross3(ac, bd, nrm);
        len = len3(nrm);
        for (int k = 0; k < 3; k++) nrm[k] /= len;
        area = len;
        len = len3(ac);
        for (int k = 0; k < 3; k++) t1[k] = ac[k] / len;
    } else {         // CellTriangle::CellFill mesh/cell.cpp:61-100
        double ab[3], ac[3];
        for (int i = 0; i < 3; i++) t[CT_RKT + i] = (cv[0][i] + cv[1][i] + cv[2][i]) / 3.;
        for (int k = 0; k < 3; k++) { ab[k] = cv[1][k] - cv[0][k]; ac[k] = cv[2][k] - cv[0][k]; }
        cross3(ab, ac, nrm);
        len = len3(nrm);
        for (int k = 0; k < 3; k++) nrm[k] /= len;
        area = len / 2;
        len = len3(ab);
        for (int k = 0; k < 3; k++) t1[k] = ab[k] / len;
    }
    cross3(nrm, t1, t2);
    len = len3(t2);
    for (int k = 0; k < 3; k++) t2[k] /= len;
    for (int k = 0; k < 3; k++) { t[CT_NORM + k] = nrm[k]; t[CT_TAU1 + k] = t1[k]; t[CT_TAU2 + k] = t2[k]; }
    t[CT_AREA] = area;
    t[CT_DIAM] = get_diam<NV>(cv);
    double nf[3], cm[3];
    norm_func<NV>(cv, nf);
    center_mass<NV>(cv, cm);
    for (int k = 0; k < 3; k++) { t[CT_NFUNC + k] = nf[k]; t[CT_CENTER + k] = cm[k]; }
    for (int s = 0; s < 4; s++) {
        double nu[3] = {0., 0., 0.};
        if (s < NV) edge_nu(cv[s], cv[(s + 1) % NV], nf, nu);
        for (int k = 0; k < 3; k++) t[CT_NU + 3 * s + k] = nu[k];
    }
    for (int k = CT_NU + 12; k < CT; k++) t[k] = 0.;
}

// Sub-cell centroid rc and weight of each quadrature node, in the reference's
// order: quads integral_universal.h:59-94 (weight |m1 x m2|), triangles
// :166-193 (weight S/n^2 - the reference scales the sum by S/n/n, :197).
template <int NV>
__global__ void k_node_tables(int n, const double* __restrict__ verts, int ns, double* __restrict__ nodes) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double cv[NV][3];
    for (int v = 0; v < NV; v++)
        for (int d = 0; d < 3; d++) cv[v][d] = verts[((size_t)c * NV + v) * 3 + d];
    double* out = nodes + (size_t)c * ns * ns * 4;
    int idx = 0;
    if (NV == 4) {
        for (int i = 0; i < ns; i++) {
            for (int j = 0; j < ns; j++) {
                double p = (double)i / (double)ns, q = (double)j / (double)ns;
                double p1 = ((double)i + 1.) / (double)ns, q1 = ((double)j + 1.) / (double)ns;
                double a1[3], a2[3], a3[3], a4[3], m1[3], m2[3], rn[3];
                for (int k = 0; k < 3; k++) {
                    double a = q * cv[1][k] + (1 - q) * cv[0][k];
                    double b = q * cv[2][k] + (1 - q) * cv[3][k];
                    a1[k] = p * b + (1 - p) * a;
                    a4[k] = p1 * b + (1 - p1) * a;
                    a = q1 * cv[1][k] + (1 - q1) * cv[0][k];
                    b = q1 * cv[2][k] + (1 - q1) * cv[3][k];
                    a2[k] = p * b + (1 - p) * a;
                    a3[k] = p1 * b + (1 - p1) * a;
                }
                for (int k = 0; k < 3; k++) {
                    out[4 * idx + k] = (a1[k] + a2[k] + a3[k] + a4[k]) / 4.0;
                    m1[k] = ((a2[k] + a3[k]) - (a1[k] + a4[k])) / 2.0;
                    m2[k] = ((a3[k] + a4[k]) - (a1[k] + a2[k])) / 2.0;
                }
                cross3(m1, m2, rn);
                out[4 * idx + 3] = len3(rn);
                idx++;
            }
        }
    } else {
        double ab[3], ac[3], nn[3], pv[3], qv[3], rv[3];
        for (int k = 0; k < 3; k++) { ab[k] = cv[1][k] - cv[0][k]; ac[k] = cv[2][k] - cv[0][k]; }
        cross3(ab, ac, nn);
        double w = ((sqrt(dot3(nn, nn)) / 2.) / (double)ns) / (double)ns;
        for (int k = 0; k < 3; k++) {
            pv[k] = (cv[1][k] - cv[0][k]) / ns;
            qv[k] = (cv[2][k] - cv[0][k]) / ns;
            rv[k] = (cv[2][k] - cv[1][k]) / ns;
        }
        for (int i = 0; i < ns; i++) {
            for (int j = 0; j < ns - i; j++) {
                double B[3];
                for (int k = 0; k < 3; k++) {
                    B[k] = cv[0][k] + (i + 1) * pv[k] + j * qv[k];
                    double AM = B[k] - pv[k];
                    double CM = AM + qv[k];
                    out[4 * idx + k] = (B[k] + AM + CM) / 3.0;
                }
                out[4 * idx + 3] = w;
                idx++;
                if (j < ns - 1 - i) {
                    for (int k = 0; k < 3; k++) out[4 * idx + k] = (B[k] + (B[k] + rv[k]) + (B[k] + qv[k])) / 3.0;
                    out[4 * idx + 3] = w;
                    idx++;
                }
            }
        }
    }
}

// Edge midpoint nodes of K0's segment rule (integral_universal_seg.h:58-72),
// edges (0,1), (1,2), ..., (nv-1, 0) as K0 walks them (K0.h:55-65).
template <int NV>
__global__ void k_seg_tables(int n, const double* __restrict__ verts, int ns, double* __restrict__ segs) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double cv[NV][3];
    for (int v = 0; v < NV; v++)
        for (int d = 0; d < 3; d++) cv[v][d] = verts[((size_t)c * NV + v) * 3 + d];
    double* out = segs + (size_t)c * NV * ns * 4;
    for (int s = 0; s < NV; s++) {
        const double* a = cv[s];
        const double* b = cv[(s + 1) % NV];
        double d[3];
        for (int j = 0; j < 3; j++) d[j] = (b[j] - a[j]) / ns;
        double dl = len3(d);
        for (int i = 0; i < ns; i++) {
            double* o = out + 4 * (s * ns + i);
            for (int j = 0; j < 3; j++) o[j] = a[j] + (i + 0.5) * d[j];
            o[3] = dl;
        }
    }
}

hipError_t launch_cell_tables(hipStream_t s, int nv, int n, const double* verts, double* ctab) {
    dim3 g((n + 127) / 128), b(128);
    if (nv == 4) hipLaunchKernelGGL(k_cell_tables<4>, g, b, 0, s, n, verts, ctab);
    else hipLaunchKernelGGL(k_cell_tables<3>, g, b, 0, s, n, verts, ctab);
    return hipGetLastError();
}

hipError_t launch_node_tables(hipStream_t s, int nv, int n, const double* verts, int n_start, double* nodes) {
    dim3 g((n + 63) / 64), b(64);
    if (nv == 4) hipLaunchKernelGGL(k_node_tables<4>, g, b, 0, s, n, verts, n_start, nodes);
    else hipLaunchKernelGGL(k_node_tables<3>, g, b, 0, s, n, verts, n_start, nodes);
    return hipGetLastError();
}

hipError_t launch_seg_tables(hipStream_t s, int nv, int n, const double* verts, int n_seg, double* segs) {
    dim3 g((n + 63) / 64), b(64);
    if (nv == 4) hipLaunchKernelGGL(k_seg_tables<4>, g, b, 0, s, n, verts, n_seg, segs);
    else hipLaunchKernelGGL(k_seg_tables<3>, g, b, 0, s, n, verts, n_seg, segs);
    return hipGetLastError();
}

// Cell-local FP32 copy of a node table for the mixed-precision tile kernels:
// loc[c][q] = (y - centre_c, weight), yy[c][q] = |y - centre_c|^2, with the
// offsets and their squared length formed in FP64 and rounded once.
__global__ void k_local32(int n, int cnt, const double* __restrict__ nodes, const double* __restrict__ ctab,
                          float4* __restrict__ loc, float* __restrict__ yy) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * cnt) return;
    const int c = (int)(idx / cnt);
    const double* y = nodes + 4 * (size_t)idx;
    const double* cen = ctab + (size_t)c * CT + CT_CENTER;
    const double d0 = y[0] - cen[0], d1 = y[1] - cen[1], d2 = y[2] - cen[2];
    loc[idx] = make_float4((float)d0, (float)d1, (float)d2, (float)y[3]);
    yy[idx] = (float)(d0 * d0 + d1 * d1 + d2 * d2);
}

hipError_t launch_local32(hipStream_t s, int n, int cnt, const double* nodes, const double* ctab, float4* loc,
                          float* yy) {
    const long long tot = (long long)n * cnt;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_local32, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, cnt, nodes, ctab, loc, yy);
    return hipGetLastError();
}

// ---------------------------------------------------------------- near lists
constexpr int kNearTile = 256;

// The point is farther from every cell centre than the largest near distance: beyond the
// sphere around the centres (fr[0..2], radius fr[3]) by more than the reach fr[4], with a
// relative margin for rounding.  NaN coordinates are never far (the scan decides).
struct FarArgs {
    double c0, c1, c2, r, reach;
    int on;
};
__device__ inline bool point_far(const double* x, const FarArgs& f) {
    if (!f.on) return false;
    const double d0 = x[0] - f.c0, d1 = x[1] - f.c1, d2 = x[2] - f.c2;
    const double lim = (f.r + f.reach) * (1. + 1e-12);
    return d0 * d0 + d1 * d1 + d2 * d2 > lim * lim;
}

// One thread per query point; cell centres and diameters staged through LDS.
__global__ __launch_bounds__(256) void k_near_count(int nq, const double* __restrict__ q, int q_stride, int n,
                                                    const double* __restrict__ ctab, int mode, double ac,
                                                    double thresh, int* __restrict__ counts, FarArgs fa) {
    __shared__ double sc[kNearTile][4];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double x[3] = {0., 0., 0.};
    if (i < nq) for (int d = 0; d < 3; d++) x[d] = q[(size_t)i * q_stride + d];
    if (__syncthreads_and(i >= nq || point_far(x, fa))) {   // no near pair in this workgroup
        if (i < nq) counts[i] = 0;
        return;
    }
    int cnt = 0;
    for (int c0 = 0; c0 < n; c0 += kNearTile) {
        int c = c0 + threadIdx.x;
        __syncthreads();
        if (c < n) {
            const double* t = ctab + (size_t)c * CT;
            sc[threadIdx.x][0] = t[CT_CENTER]; sc[threadIdx.x][1] = t[CT_CENTER + 1];
            sc[threadIdx.x][2] = t[CT_CENTER + 2]; sc[threadIdx.x][3] = t[CT_DIAM];
        }
        __syncthreads();
        int m = min(kNearTile, n - c0);
        for (int cc = 0; cc < m; cc++) {
            double cm[3] = {sc[cc][0], sc[cc][1], sc[cc][2]};
            cnt += near_rule(center_dist(x[0], x[1], x[2], cm), sc[cc][3], mode, ac, thresh);
        }
    }
    if (i < nq) counts[i] = cnt;
}

__global__ __launch_bounds__(256) void k_near_fill(int nq, const double* __restrict__ q, int q_stride, int n,
                                                   const double* __restrict__ ctab, int mode, double ac,
                                                   double thresh, const long long* __restrict__ offs,
                                                   int* __restrict__ lq, int* __restrict__ lc, FarArgs fa) {
    __shared__ double sc[kNearTile][4];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double x[3] = {0., 0., 0.};
    if (i < nq) for (int d = 0; d < 3; d++) x[d] = q[(size_t)i * q_stride + d];
    if (__syncthreads_and(i >= nq || point_far(x, fa))) return;   // no near pair in this workgroup
    long long pos = (i < nq) ? offs[i] : 0;
    for (int c0 = 0; c0 < n; c0 += kNearTile) {
        int c = c0 + threadIdx.x;
        __syncthreads();
        if (c < n) {
            const double* t = ctab + (size_t)c * CT;
            sc[threadIdx.x][0] = t[CT_CENTER]; sc[threadIdx.x][1] = t[CT_CENTER + 1];
            sc[threadIdx.x][2] = t[CT_CENTER + 2]; sc[threadIdx.x][3] = t[CT_DIAM];
        }
        __syncthreads();
        int m = min(kNearTile, n - c0);
        for (int cc = 0; cc < m; cc++) {
            double cm[3] = {sc[cc][0], sc[cc][1], sc[cc][2]};
            if (i < nq && near_rule(center_dist(x[0], x[1], x[2], cm), sc[cc][3], mode, ac, thresh)) {
                lq[pos] = i;
                lc[pos] = c0 + cc;
                pos++;
            }
        }
    }
}

static FarArgs far_args(const double* far) {
    FarArgs f{0., 0., 0., 0., 0., 0};
    if (far) f = FarArgs{far[0], far[1], far[2], far[3], far[4], 1};
    return f;
}

hipError_t launch_near_count(hipStream_t s, int nq, const double* q, int q_stride, int n, const double* ctab,
                             int mode, double ac, double thresh, int* counts, const double* far) {
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_near_count, dim3((nq + 255) / 256), dim3(256), 0, s, nq, q, q_stride, n, ctab, mode,
                       ac, thresh, counts, far_args(far));
    return hipGetLastError();
}

hipError_t launch_near_fill(hipStream_t s, int nq, const double* q, int q_stride, int n, const double* ctab,
                            int mode, double ac, double thresh, const long long* offs, int* lq, int* lc,
                            const double* far) {
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_near_fill, dim3((nq + 255) / 256), dim3(256), 0, s, nq, q, q_stride, n, ctab, mode, ac,
                       thresh, offs, lq, lc, far_args(far));
    return hipGetLastError();
}

// Exclusive prefix sum of n ints in one workgroup of 1024 lanes: each lane
// sums a contiguous chunk, the 1024 chunk sums are scanned in LDS, then each
// lane writes its chunk's prefixes.  out has n + 1 entries (out[n] = total).
__global__ __launch_bounds__(1024) void k_scan(int n, const int* __restrict__ in, long long* __restrict__ out,
                                               long long* __restrict__ total) {
    __shared__ long long sums[1024];
    int t = threadIdx.x;
    int chunk = (n + 1023) / 1024;
    int b = t * chunk, e = min(n, b + chunk);
    long long s = 0;
    for (int i = b; i < e; i++) s += in[i];
    sums[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        long long v = (t >= off) ? sums[t - off] : 0;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    long long run = sums[t] - s;
    for (int i = b; i < e; i++) {
        out[i] = run;
        run += in[i];
    }
    if (t == 1023) {
        *total = sums[1023];
        out[n] = sums[1023];
    }
}

hipError_t launch_scan(hipStream_t s, int n, const int* in, long long* out_excl, long long* total) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, n, in, out_excl, total);
    return hipGetLastError();
}

}  // namespace bielcc_impl
