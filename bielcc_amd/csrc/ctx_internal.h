// Library context and the host helpers shared by the C-ABI translation units
// (capi.hip, plan.hip).  Internal; the public surface is include/bielcc_hip.h.
#pragma once

#include "bielcc_hip.h"
#include "bielcc_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

using bielcc_dev::NumP;
using bielcc_dev::KP;

struct bielcc_plan_s;

struct bielcc_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<int, bielcc_impl::DevMesh> meshes;
    int next_id = 1;
    bielcc_impl::LuWork* lu = nullptr;
    void* scratch = nullptr;
    size_t scratch_cap = 0;
    void* lists = nullptr;
    size_t lists_cap = 0;
    bool timing = false;
    bool no_fast_phase = false;   // BIELCC_GENERAL_SINCOS=1: tile kernels use the general sincos
    void* near_buf = nullptr;     // per-pair near results of the field driver
    size_t near_cap = 0;
    void* opmat_ws = nullptr;     // operator matrices of bielcc_field_batch's many-set path (grow-only)
    size_t opmat_ws_cap = 0;
    void* opmat_meta = nullptr;   // its segment descriptors and tiles (grow-only)
    size_t opmat_meta_cap = 0;
    void* opmat_part = nullptr;   // its split-K GEMM partial products (grow-only)
    size_t opmat_part_cap = 0;
    void* small = nullptr;        // staging of small host arrays (incidences) (grow-only)
    size_t small_cap = 0;
    void* fix = nullptr;          // FP32 error control of the mixed field: flags, lists, FP64 results (grow-only)
    size_t fix_cap = 0;
    bool mixed_fixup = true;      // mixed field: points the FP32 error model does not clear are recomputed in FP64
    long long fixup_points = 0;   // points so recomputed, cumulative
    int mixed_probe_pct = 25;     // mixed field: if more than this % of a strided sample of the points would go
                                  // back to FP64 sums, the whole field runs in FP64 (0: no sample)
    bool in_probe = false;        // the field driver is computing that sample
    void* probe = nullptr;        // the sample's points, fields, flags (grow-only)
    size_t probe_cap = 0;
    long long probe_fp64 = 0;     // mixed field calls run in FP64 after the sample, cumulative
    bool mixed_far32 = false;     // mixed field: FP32 far sums always (false: FP64 far sums where the node
                                  // phase uses the centre-phase series and k is real, where the FP64 node
                                  // loop is the shorter of the two)
    long long auto_fp64 = 0;      // mixed field calls so run with FP64 far sums, cumulative
    int field_wg = 2048, assemble_wg = 2048, field_rchunk = 4;   // launch-shape tunables
    // plan (multi-mesh) tunables: cells per field tile and per assembly tile,
    // persistent workgroups per CU, operator-matrix workspace cap (MiB), and the
    // number of current sets from which bielcc_field_batch uses the operator matrix
    int plan_field_cells = 64, plan_asm_cells = 64, plan_wg_per_cu = 4, plan_ws_mb = 24576, opmat_min_rhs = 9;
    int n_cu = 256;               // compute units of the device (hipDeviceProp_t::multiProcessorCount)
    bool capturing = false;       // a HIP graph capture is in progress on the stream: no host syncs
    std::map<std::string, std::pair<double, int>> times;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace bielcc_capi {

using namespace bielcc_impl;


inline int fail(bielcc_ctx ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

inline int hip_fail(bielcc_ctx ctx, hipError_t e, const char* where) {
    std::string m = std::string(where) + ": " + hipGetErrorString(e);
    return fail(ctx, e == hipErrorOutOfMemory ? BIELCC_ERR_OOM : BIELCC_ERR_HIP, m);
}

#define HIPCHK(expr)                                              \
    do {                                                          \
        hipError_t _e = (expr);                                   \
        if (_e != hipSuccess) return hip_fail(ctx, _e, #expr);   \
    } while (0)

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// grow-only device buffer
inline hipError_t ensure(void*& buf, size_t& cap, size_t need) {
    if (cap >= need) return hipSuccess;
    size_t n = std::max(need, cap * 3 / 2);
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&buf, n);
    if (e == hipSuccess) cap = n;
    return e;
}

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

inline NumP to_np(const bielcc_num_param* p) {
    NumP n;
    n.eps = p->eps; n.rs = p->rs; n.rs_seg = p->rs_seg; n.analytic_calc = p->analytic_calc;
    n.n_start = p->n_start; n.n_start_seg = p->n_start_seg; n.p_max = p->p_max; n.p_max_seg = p->p_max_seg;
    return n;
}

inline KP to_kp(const bielcc_kernel_param* p) {
    KP k;
    k.R = p->smoothR; k.Rmin = p->smoothRMin; k.k = bielcc_dev::mk(p->k_re, p->k_im);
    for (int i = 0; i < 3; i++) {
        k.vec[i] = p->vec_dbl[i];
        k.n[i] = p->n[i];
        k.cur[i] = bielcc_dev::mk(p->current[2 * i], p->current[2 * i + 1]);
    }
    return k;
}

inline int check_np(bielcc_ctx ctx, const bielcc_num_param* np) {
    if (!np) return fail(ctx, BIELCC_ERR_ARG, "num_param is NULL");
    if (np->n_start < 1 || np->n_start_seg < 1 || np->p_max < 1 || np->p_max_seg < 1)
        return fail(ctx, BIELCC_ERR_ARG, "n_start, n_start_seg, p_max and p_max_seg must be >= 1");
    if ((long long)np->n_start << (np->p_max - 1) > 4096 || (long long)np->n_start_seg << (np->p_max_seg - 1) > (1 << 20))
        return fail(ctx, BIELCC_ERR_UNSUPPORTED, "quadrature partition beyond 4096 (cell) / 2^20 (segment)");
    return BIELCC_OK;
}

inline DevMesh* find_mesh(bielcc_ctx ctx, int id) {
    auto it = ctx->meshes.find(id);
    return it == ctx->meshes.end() ? nullptr : &it->second;
}

inline int ensure_tables(bielcc_ctx ctx, DevMesh& m, int n_start, int n_seg) {
    if (m.nodes_n != n_start) {
        if (m.nodes) (void)hipFree(m.nodes);
        m.nodes = nullptr;
        m.nodes_n = 0;
        HIPCHK(hipMalloc(&m.nodes, (size_t)m.n * n_start * n_start * 4 * sizeof(double)));
        HIPCHK(launch_node_tables(ctx->stream, m.nv, m.n, m.verts, n_start, m.nodes));
        m.nodes_n = n_start;
    }
    if (n_seg > 0 && m.segs_n != n_seg) {
        if (m.segs) (void)hipFree(m.segs);
        m.segs = nullptr;
        m.segs_n = 0;
        HIPCHK(hipMalloc(&m.segs, (size_t)m.n * m.nv * n_seg * 4 * sizeof(double)));
        HIPCHK(launch_seg_tables(ctx->stream, m.nv, m.n, m.verts, n_seg, m.segs));
        m.segs_n = n_seg;
    }
    return BIELCC_OK;
}

inline void free_mesh(DevMesh& m) {
    (void)hipFree(m.verts); (void)hipFree(m.ctab); (void)hipFree(m.nodes); (void)hipFree(m.segs);
    (void)hipFree(m.nodes32); (void)hipFree(m.nodes32_yy);
    (void)hipFree(m.near_lq); (void)hipFree(m.near_lc);
}

// cell-local FP32 surface nodes for the mixed-precision kernels
inline int ensure_tables32(bielcc_ctx ctx, DevMesh& m, int n_start) {
    if (int rc = ensure_tables(ctx, m, n_start, 0)) return rc;
    if (m.nodes32_n == n_start) return BIELCC_OK;
    (void)hipFree(m.nodes32);
    (void)hipFree(m.nodes32_yy);
    m.nodes32 = nullptr;
    m.nodes32_yy = nullptr;
    m.nodes32_n = 0;
    const size_t cnt = (size_t)m.n * n_start * n_start;
    HIPCHK(hipMalloc(&m.nodes32, cnt * sizeof(float4)));
    HIPCHK(hipMalloc(&m.nodes32_yy, cnt * sizeof(float)));
    HIPCHK(launch_local32(ctx->stream, m.n, n_start * n_start, m.nodes, m.ctab, m.nodes32, m.nodes32_yy));
    m.nodes32_n = n_start;
    return BIELCC_OK;
}

struct Timer {
    bielcc_ctx ctx;
    const char* name;
    explicit Timer(bielcc_ctx c, const char* n) : ctx(c), name(n) {
        if (ctx->timing && !ctx->capturing) (void)hipEventRecord(ctx->ev0, ctx->stream);
    }
    void stop() {
        if (!ctx->timing || ctx->capturing) return;
        (void)hipEventRecord(ctx->ev1, ctx->stream);
        (void)hipEventSynchronize(ctx->ev1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
        auto& t = ctx->times[name];
        t.first += ms;
        t.second += 1;
    }
};

// Near-pair list for nq query points: counts -> exclusive scan -> fill.
// Layout in ctx->scratch region handed in: counts [nq], offs [nq+1], total [1].
// The lists themselves (lq, lc) live in ctx->lists.
inline int build_near_list(bielcc_ctx ctx, int nq, const double* q, int q_stride, const DevMesh& m, int mode, double ac,
                    double thresh, int* counts, long long* offs, long long* total_d, long long* total_h,
                    int** lq, int** lc) {
    const NearFar nf = near_far(m, mode, ac, thresh);
    HIPCHK(launch_near_count(ctx->stream, nq, q, q_stride, m.n, m.ctab, mode, ac, thresh, counts, nf.v));
    HIPCHK(launch_scan(ctx->stream, nq, counts, offs, total_d));
    HIPCHK(hipMemcpyAsync(total_h, total_d, sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const long long T = *total_h;
    *lq = *lc = nullptr;
    if (T > 0) {
        HIPCHK(ensure(ctx->lists, ctx->lists_cap, align256((size_t)T * sizeof(int)) * 2));
        *lq = reinterpret_cast<int*>(ctx->lists);
        *lc = reinterpret_cast<int*>(reinterpret_cast<char*>(ctx->lists) + align256((size_t)T * sizeof(int)));
        HIPCHK(launch_near_fill(ctx->stream, nq, q, q_stride, m.n, m.ctab, mode, ac, thresh, offs, *lq, *lc, nf.v));
    }
    return BIELCC_OK;
}

// Split-K parts for the operator-matrix GEMMs of a segment of Ms points and N
// cells: enough that the (3Ms / 128-row tiles) x parts fill the device twice.
inline int opmat_parts(bielcc_ctx ctx, int Ms, int N) {
    const long long tiles_m = (3LL * Ms + 127) / 128;
    long long s = (2LL * ctx->n_cu + tiles_m - 1) / tiles_m;
    s = std::min<long long>(s, std::max(1, N / 32));   // at least 32 cells per part
    return (int)std::max<long long>(1, std::min<long long>(s, 64));
}

// E = multE WK J + E, H = WR J for one segment (E, H at the segment's first
// point, leading dimension ldo); with nparts > 1 the cells are split into
// parts whose products go to part (nparts x 3Ms x nrhs complex) and are summed.
int opmat_gemms(bielcc_ctx ctx, const bielcc_impl::OpSeg& sg, int nrhs, const double* jE, double2 mE, double* E,
                double* H, long long ldo, double* part, int nparts);

// bielcc_field_batch for many current sets (electric only) through the
// operator matrices in context-owned grow-only buffers (plan.hip).  inc_d:
// device [nrhs][6] or null.  counts/offs/tot_d: near-list scratch for M points.
// FP32 error control of a mixed field evaluated for electric currents only (capi.hip): points
// the error model does not clear are recomputed with FP64 sums.  inc: host [nrhs][6] or null,
// inc_d: its device copy.  Synchronises the stream.
int field_mixed_fixup(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im, int nrhs,
                      const double* jE_d, double mEr, double mEi, const double* inc, const double* inc_d, int near_mode,
                      int M, const double* pts_d, double* E_d, double* H_d);

// bielcc_field_noalloc's per-device scratch address (noalloc.hip), taken at context creation.
int noalloc_setup(bielcc_ctx ctx);

int field_opmat(bielcc_ctx ctx, DevMesh& m, const bielcc_num_param* np, double k_re, double k_im, int nrhs,
                const double* jE_d, double mEr, double mEi, const double* inc_d, int near_mode, int precision, int M,
                const double* pts_d, double* E_d, double* H_d, int* counts, long long* offs, long long* tot_d);

}  // namespace bielcc_capi
