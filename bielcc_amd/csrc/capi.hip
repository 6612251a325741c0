// Implementation of the C ABI declared in include/bielcc_hip.h.
#include "ctx_internal.h"
#include <climits>

using namespace bielcc_impl;
using namespace bielcc_capi;

extern "C" {

int bielcc_version(void) { return 100; }

int bielcc_ctx_create(int device, bielcc_ctx* out) {
    if (!out) return BIELCC_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BIELCC_ERR_HIP;
    if (device < 0 || device >= ndev) return BIELCC_ERR_ARG;
    bielcc_ctx ctx = new bielcc_ctx_s();
    ctx->device = device;
    DeviceGuard g(device);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        delete ctx;
        return BIELCC_ERR_HIP;
    }
    ctx->lu = lu_work_create();
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->n_cu = prop.multiProcessorCount;
    const char* gs = std::getenv("BIELCC_GENERAL_SINCOS");
    ctx->no_fast_phase = gs && gs[0] == '1';
    if (noalloc_setup(ctx) != BIELCC_OK) {   // per-device address of bielcc_field_noalloc's scratch
        (void)hipEventDestroy(ctx->ev0);
        (void)hipEventDestroy(ctx->ev1);
        (void)hipStreamDestroy(ctx->stream);
        lu_work_destroy(ctx->lu);
        delete ctx;
        return BIELCC_ERR_HIP;
    }
    *out = ctx;
    return BIELCC_OK;
}

int bielcc_ctx_destroy(bielcc_ctx ctx) {
    if (!ctx) return BIELCC_ERR_ARG;
    {
        DeviceGuard g(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto& kv : ctx->meshes) free_mesh(kv.second);
        if (ctx->scratch) (void)hipFree(ctx->scratch);
        if (ctx->lists) (void)hipFree(ctx->lists);
        if (ctx->near_buf) (void)hipFree(ctx->near_buf);
        if (ctx->opmat_ws) (void)hipFree(ctx->opmat_ws);
        if (ctx->opmat_meta) (void)hipFree(ctx->opmat_meta);
        if (ctx->opmat_part) (void)hipFree(ctx->opmat_part);
        if (ctx->small) (void)hipFree(ctx->small);
        if (ctx->fix) (void)hipFree(ctx->fix);
        if (ctx->probe) (void)hipFree(ctx->probe);
        lu_work_destroy(ctx->lu);
        (void)hipEventDestroy(ctx->ev0);
        (void)hipEventDestroy(ctx->ev1);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return BIELCC_OK;
}

int bielcc_sync(bielcc_ctx ctx) {
    if (!ctx) return BIELCC_ERR_ARG;
    DeviceGuard g(ctx->device);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return BIELCC_OK;
}

void* bielcc_stream(bielcc_ctx ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int bielcc_last_error(bielcc_ctx ctx, char* buf, int len) {
    if (!ctx || !buf || len <= 0) return BIELCC_ERR_ARG;
    std::snprintf(buf, (size_t)len, "%s", ctx->err.c_str());
    return BIELCC_OK;
}

int bielcc_malloc(bielcc_ctx ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return BIELCC_ERR_ARG;
    DeviceGuard g(ctx->device);
    *dptr = nullptr;
    HIPCHK(hipMalloc(dptr, bytes ? bytes : 1));
    return BIELCC_OK;
}

int bielcc_free(bielcc_ctx ctx, void* dptr) {
    if (!ctx) return BIELCC_ERR_ARG;
    DeviceGuard g(ctx->device);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipFree(dptr));
    return BIELCC_OK;
}

int bielcc_memcpy(bielcc_ctx ctx, void* dst, const void* src, size_t bytes, int kind) {
    if (!ctx || kind < 0 || kind > 2) return fail(ctx, BIELCC_ERR_ARG, "bad memcpy kind");
    if (bytes == 0) return BIELCC_OK;
    if (!dst || !src) return fail(ctx, BIELCC_ERR_ARG, "NULL memcpy pointer");
    DeviceGuard g(ctx->device);
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return BIELCC_OK;
}

int bielcc_mesh_create(bielcc_ctx ctx, int n_cells, int nv, const double* verts, int* mesh_id) {
    if (!ctx || !verts || !mesh_id) return fail(ctx, BIELCC_ERR_ARG, "NULL argument");
    if (n_cells <= 0 || (nv != 3 && nv != 4)) return fail(ctx, BIELCC_ERR_ARG, "n_cells must be > 0 and nv 3 or 4");
    for (size_t i = 0; i < (size_t)n_cells * nv * 3; i++)
        if (!std::isfinite(verts[i])) return fail(ctx, BIELCC_ERR_ARG, "non-finite vertex coordinate");
    DeviceGuard g(ctx->device);
    DevMesh m;
    m.n = n_cells;
    m.nv = nv;
    size_t vb = (size_t)n_cells * nv * 3 * sizeof(double);
    HIPCHK(hipMalloc(&m.verts, vb));
    hipError_t e = hipMalloc(&m.ctab, (size_t)n_cells * CT * sizeof(double));
    if (e != hipSuccess) { (void)hipFree(m.verts); return hip_fail(ctx, e, "hipMalloc(ctab)"); }
    HIPCHK(hipMemcpyAsync(m.verts, verts, vb, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(launch_cell_tables(ctx->stream, nv, n_cells, m.verts, m.ctab));
    // grid step = max diameter (host reduction of the diameter column)
    std::vector<double> tab((size_t)n_cells * CT);
    HIPCHK(hipMemcpyAsync(tab.data(), m.ctab, tab.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double gs = 0., mr = 0.;
    for (int c = 0; c < n_cells; c++) {
        gs = std::max(gs, tab[(size_t)c * CT + CT_DIAM]);
        for (int v = 0; v < nv; v++) {   // quadrature nodes lie in the cell: within max_rad of its centre
            double d2 = 0.;
            for (int d = 0; d < 3; d++) {
                const double x = verts[((size_t)c * nv + v) * 3 + d] - tab[(size_t)c * CT + CT_CENTER + d];
                d2 += x * x;
            }
            mr = std::max(mr, std::sqrt(d2));
        }
    }
    m.grid_step = gs;
    m.max_rad = mr;
    // a sphere around every cell centre: the centre of their bounding box, the largest distance
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int c = 0; c < n_cells; c++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], tab[(size_t)c * CT + CT_CENTER + d]);
            hi[d] = std::max(hi[d], tab[(size_t)c * CT + CT_CENTER + d]);
        }
    for (int d = 0; d < 3; d++) m.bc[d] = 0.5 * (lo[d] + hi[d]);
    double br2 = 0.;
    for (int c = 0; c < n_cells; c++) {
        double d2 = 0.;
        for (int d = 0; d < 3; d++) {
            const double x = tab[(size_t)c * CT + CT_CENTER + d] - m.bc[d];
            d2 += x * x;
        }
        br2 = std::max(br2, d2);
    }
    m.br = std::sqrt(br2);
    int id = ctx->next_id++;
    ctx->meshes[id] = m;
    *mesh_id = id;
    return BIELCC_OK;
}

int bielcc_mesh_destroy(bielcc_ctx ctx, int mesh_id) {
    if (!ctx) return BIELCC_ERR_ARG;
    DevMesh* m = find_mesh(ctx, mesh_id);
    if (!m) return fail(ctx, BIELCC_ERR_NOTFOUND, "unknown mesh id");
    DeviceGuard g(ctx->device);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    free_mesh(*m);
    ctx->meshes.erase(mesh_id);
    return BIELCC_OK;
}

int bielcc_mesh_tables(bielcc_ctx ctx, int mesh_id, double* out) {
    if (!ctx || !out) return fail(ctx, BIELCC_ERR_ARG, "NULL argument");
    DevMesh* m = find_mesh(ctx, mesh_id);
    if (!m) return fail(ctx, BIELCC_ERR_NOTFOUND, "unknown mesh id");
    DeviceGuard g(ctx->device);
    std::vector<double> tab((size_t)m->n * CT);
    HIPCHK(hipMemcpyAsync(tab.data(), m->ctab, tab.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < m->n; c++) std::memcpy(out + 20 * (size_t)c, tab.data() + (size_t)c * CT, 20 * sizeof(double));
    return BIELCC_OK;
}

int bielcc_mesh_grid_step(bielcc_ctx ctx, int mesh_id, double* grid_step) {
    if (!ctx || !grid_step) return fail(ctx, BIELCC_ERR_ARG, "NULL argument");
    DevMesh* m = find_mesh(ctx, mesh_id);
    if (!m) return fail(ctx, BIELCC_ERR_NOTFOUND, "unknown mesh id");
    *grid_step = m->grid_step;
    return BIELCC_OK;
}

int bielcc_mesh_size(bielcc_ctx ctx, int mesh_id, int* n_cells, int* nv) {
    if (!ctx || !n_cells || !nv) return fail(ctx, BIELCC_ERR_ARG, "NULL argument");
    DevMesh* m = find_mesh(ctx, mesh_id);
    if (!m) return fail(ctx, BIELCC_ERR_NOTFOUND, "unknown mesh id");
    *n_cells = m->n;
    *nv = m->nv;
    return BIELCC_OK;
}

int bielcc_op_batch(bielcc_ctx ctx, int op, int nv, int n, const double* cells_d, const double* x_d,
                    const double* j_d, const bielcc_num_param* np, double k_re, double k_im, double* res_d) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (op < 0 || op > 7 || (nv != 3 && nv != 4) || n < 0) return fail(ctx, BIELCC_ERR_ARG, "bad op, nv or n");
    if (int rc = check_np(ctx, np)) return rc;
    if (n > 0 && (!cells_d || !x_d || !j_d || !res_d)) return fail(ctx, BIELCC_ERR_ARG, "NULL buffer");
    DeviceGuard g(ctx->device);
    HIPCHK(launch_op_batch(ctx->stream, op, nv, n, cells_d, x_d, j_d, to_np(np), k_re, k_im, res_d));
    return BIELCC_OK;
}

int bielcc_kernel_batch(bielcc_ctx ctx, int f, int n, const double* x_d, const double* y_d,
                        const bielcc_kernel_param* kp, double* res_d) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (f < 0 || f > 11 || n < 0 || !kp) return fail(ctx, BIELCC_ERR_ARG, "bad kernel id, n or kp");
    if (n > 0 && (!x_d || !y_d || !res_d)) return fail(ctx, BIELCC_ERR_ARG, "NULL buffer");
    DeviceGuard g(ctx->device);
    HIPCHK(launch_kernel_batch(ctx->stream, f, n, x_d, y_d, to_kp(kp), res_d));
    return BIELCC_OK;
}

int bielcc_aver_batch(bielcc_ctx ctx, int which, int n, const double* x_d, const double* y_d, double eps,
                      double* res_d) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (which < 0 || which > 2 || n < 0) return fail(ctx, BIELCC_ERR_ARG, "which must be 0, 1 or 2 and n >= 0");
    if (!(eps > 0.) || !std::isfinite(eps)) return fail(ctx, BIELCC_ERR_ARG, "eps must be positive and finite");
    if (n > 0 && (!x_d || !y_d || !res_d)) return fail(ctx, BIELCC_ERR_ARG, "NULL buffer");
    DeviceGuard g(ctx->device);
    HIPCHK(launch_aver_batch(ctx->stream, which, n, x_d, y_d, eps, res_d));
    return BIELCC_OK;
}

int bielcc_integral_batch(bielcc_ctx ctx, int f, int nv, int n, const double* cells_d, const double* x_d,
                          const bielcc_kernel_param* kp, const bielcc_integral_param* ip, double* res_d,
                          double* smoothR_d) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (f < 0 || f > 11 || n < 0 || !kp || !ip || (nv < 2 || nv > 4)) return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    if (ip->idim < 1 || ip->idim > 3 || ip->n_start < 1 || ip->p_max < 1)
        return fail(ctx, BIELCC_ERR_ARG, "idim must be 1..3, n_start and p_max >= 1");
    if (n > 0 && (!cells_d || !x_d || !res_d)) return fail(ctx, BIELCC_ERR_ARG, "NULL buffer");
    DeviceGuard g(ctx->device);
    HIPCHK(launch_integral_batch(ctx->stream, f, nv, n, cells_d, x_d, to_kp(kp), ip->idim, ip->n_start, ip->p_max,
                                 ip->eps, res_d, smoothR_d));
    return BIELCC_OK;
}

int bielcc_k0_batch(bielcc_ctx ctx, int variant, int f, int nv, int n, const double* cells_d, const double* x_d,
                    const double* j_d, const bielcc_kernel_param* kp, const bielcc_integral_param* ip,
                    double analyticR, double* res_d, double* smoothR_d) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (variant < 0 || variant > 2 || f < 0 || f > 11 || n < 0 || !kp || !ip || (nv != 3 && nv != 4))
        return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    if (ip->idim != 3) return fail(ctx, BIELCC_ERR_UNSUPPORTED, "K0 is defined for idim = 3 only");
    if (n > 0 && (!cells_d || !x_d || !res_d || (variant != 1 && !j_d))) return fail(ctx, BIELCC_ERR_ARG, "NULL buffer");
    DeviceGuard g(ctx->device);
    HIPCHK(launch_k0_batch(ctx->stream, variant, f, nv, n, cells_d, x_d, variant == 1 ? x_d : j_d, to_kp(kp),
                           ip->n_start, ip->p_max, ip->eps, analyticR, res_d, smoothR_d));
    return BIELCC_OK;
}

int bielcc_analytic_batch(bielcc_ctx ctx, int which, int nv, int n, const double* cells_d, const double* x_d,
                          const bielcc_integral_param* ip, double analyticR, double k_re, double k_im,
                          double* res_d) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (which < 0 || which > 3 || n < 0 || ((which == 0 || which == 2) && nv != 3 && nv != 4))
        return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    if ((which >= 2) && (!ip || ip->n_start < 1 || ip->p_max < 1)) return fail(ctx, BIELCC_ERR_ARG, "bad integral param");
    if (n > 0 && (!cells_d || !x_d || !res_d)) return fail(ctx, BIELCC_ERR_ARG, "NULL buffer");
    DeviceGuard g(ctx->device);
    int ns = ip ? ip->n_start : 1, pm = ip ? ip->p_max : 1;
    double eps = ip ? ip->eps : 0.;
    HIPCHK(launch_analytic_batch(ctx->stream, which, nv, n, cells_d, x_d, ns, pm, eps, analyticR, k_re, k_im, res_d));
    return BIELCC_OK;
}

int bielcc_plane_wave_e(bielcc_ctx ctx, int n, const double* pts_d, const double* kvec, const double* e0,
                        double* out_d) {
    if (!ctx || !kvec || !e0 || n < 0) return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    DeviceGuard g(ctx->device);
    double ke[6] = {kvec[0], kvec[1], kvec[2], e0[0], e0[1], e0[2]};
    HIPCHK(launch_plane_wave(ctx->stream, n, pts_d, ke, out_d));
    return BIELCC_OK;
}

int bielcc_plane_wave_h(bielcc_ctx ctx, int n, const double* pts_d, const double* kvec, const double* e0,
                        double freq, double mu_re, double mu_im, double* out_d) {
    if (!ctx || !kvec || !e0 || n < 0) return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    if (n > 0 && (!pts_d || !out_d)) return fail(ctx, BIELCC_ERR_ARG, "NULL buffer");
    DeviceGuard g(ctx->device);
    PlaneWaveH pw{{kvec[0], kvec[1], kvec[2], e0[0], e0[1], e0[2]}, freq, mu_re, mu_im};
    HIPCHK(launch_plane_wave_h(ctx->stream, n, pts_d, pw, out_d));
    return BIELCC_OK;
}

int bielcc_field_add_incident_h(bielcc_ctx ctx, int M, int nrhs, const double* pts_d, const double* inc,
                                double freq, double mu_re, double mu_im, double scale_re, double scale_im,
                                double* H_d) {
    if (!ctx || M < 0 || nrhs < 0) return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    if (M == 0 || nrhs == 0) return BIELCC_OK;
    if (!pts_d || !inc || !H_d) return fail(ctx, BIELCC_ERR_ARG, "NULL buffer");
    DeviceGuard g(ctx->device);
    const size_t b = (size_t)nrhs * 6 * sizeof(double);
    HIPCHK(ensure(ctx->small, ctx->small_cap, b));
    HIPCHK(hipMemcpyAsync(ctx->small, inc, b, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(launch_add_incident_h(ctx->stream, M, nrhs, pts_d, static_cast<const double*>(ctx->small), freq, mu_re,
                                 mu_im, scale_re, scale_im, H_d));
    return BIELCC_OK;
}

int bielcc_integral_y_batch(bielcc_ctx ctx, int f, int nv, int n, const double* cells_d,
                            const bielcc_kernel_param* kp, const double* e0, const bielcc_integral_param* ip,
                            double* res_d, double* smoothR_d) {
    if (!ctx) return BIELCC_ERR_ARG;
    if ((f != BIELCC_FY_UNIT && f != BIELCC_FY_PLANE_WAVE_E) || n < 0 || !kp || !ip || nv < 2 || nv > 4)
        return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    if (f == BIELCC_FY_PLANE_WAVE_E && !e0) return fail(ctx, BIELCC_ERR_ARG, "PlaneWaveFieldE needs e0");
    if (ip->idim < 1 || ip->idim > 3 || ip->n_start < 1 || ip->p_max < 1)
        return fail(ctx, BIELCC_ERR_ARG, "idim must be 1..3, n_start and p_max >= 1");
    if ((long long)ip->n_start << (ip->p_max - 1) > (nv == 2 ? (1 << 20) : 4096))
        return fail(ctx, BIELCC_ERR_UNSUPPORTED, "quadrature partition beyond 4096 (cell) / 2^20 (segment)");
    if (n > 0 && (!cells_d || !res_d)) return fail(ctx, BIELCC_ERR_ARG, "NULL buffer");
    DeviceGuard g(ctx->device);
    const double zero[3] = {0., 0., 0.};
    HIPCHK(launch_integral_y_batch(ctx->stream, f, nv, n, cells_d, to_kp(kp), e0 ? e0 : zero, ip->idim,
                                   ip->n_start, ip->p_max, ip->eps, res_d, smoothR_d));
    return BIELCC_OK;
}

static int assemble_impl(bielcc_ctx ctx, int mesh_id, int which, const bielcc_num_param* np, double k_re,
                         double k_im, int row0, int nrows, int precision, double* Z_d, int64_t ld) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (int rc = check_np(ctx, np)) return rc;
    if (precision != BIELCC_PREC_FP64 && precision != BIELCC_PREC_MIXED)
        return fail(ctx, BIELCC_ERR_ARG, "precision must be BIELCC_PREC_FP64 or BIELCC_PREC_MIXED");
    DevMesh* m = find_mesh(ctx, mesh_id);
    if (!m) return fail(ctx, BIELCC_ERR_NOTFOUND, "unknown mesh id");
    if (which != 0 && which != 1) return fail(ctx, BIELCC_ERR_ARG, "which must be 0 (K) or 1 (R)");
    if (row0 < 0 || nrows < 0 || (row0 & 1) || (nrows & 1) || row0 + nrows > 2 * m->n)
        return fail(ctx, BIELCC_ERR_ARG, "row range must be even and inside [0, 2N)");
    if (nrows == 0) return BIELCC_OK;
    if (!Z_d || ld < nrows) return fail(ctx, BIELCC_ERR_ARG, "Z is NULL or ld < nrows");
    if (!std::isfinite(k_re) || !std::isfinite(k_im)) return fail(ctx, BIELCC_ERR_ARG, "non-finite wave number");
    DeviceGuard g(ctx->device);
    const NumP npp = to_np(np);
    const int i0 = row0 / 2, ni = nrows / 2;
    const int nn = np->n_start * np->n_start, nseg = m->nv * np->n_start_seg;
    const bool fast = np->p_max == 1 && np->p_max_seg == 1 && nn <= 512 && nseg <= 256 && np->analytic_calc > 1.;
    if (!fast) {
        Timer t(ctx, "assemble_generic");
        HIPCHK(launch_assemble_generic(ctx->stream, m->nv, which, (long long)ni * m->n, nullptr, nullptr, m->n, i0,
                                       ni, m->verts, m->ctab, npp, k_re, k_im, Z_d, ld));
        t.stop();
        return BIELCC_OK;
    }
    if (int rc = ensure_tables(ctx, *m, np->n_start, which == 0 ? np->n_start_seg : 0)) return rc;
    if (precision == BIELCC_PREC_MIXED)
        if (int rc = ensure_tables32(ctx, *m, np->n_start)) return rc;
    AssembleArgs a;
    a.nv = m->nv; a.N = m->n; a.nn = nn; a.ns = np->n_start_seg; a.which = which;
    a.i0 = i0; a.ni = ni; a.ctab = m->ctab; a.nodes = m->nodes; a.segs = m->segs;
    a.k_re = k_re; a.k_im = k_im; a.ac = np->analytic_calc; a.Z = Z_d; a.ld = ld;
    a.fast_phase = ctx->no_fast_phase ? 0 : phase_mode(k_re, m->grid_step, m->max_rad);
    a.prec = precision;
    a.nodes32 = m->nodes32;
    a.nodes32_yy = m->nodes32_yy;
    a.wg_target = ctx->assemble_wg;
    {
        Timer t(ctx, "assemble_far");
        HIPCHK(launch_assemble_far(ctx->stream, a));
        t.stop();
    }
    if (which == 0) {
        // near pairs (analytic branch of IntegralSurf_SimplePot_H_HS): exact per-pair kernel.
        // The list depends on the mesh, the rows and the rule only: the mesh keeps the last one.
        if (!(m->near_tot >= 0 && m->near_i0 == i0 && m->near_ni == ni && m->near_ac == np->analytic_calc)) {
            size_t need = align256((size_t)ni * sizeof(int)) + align256((size_t)(ni + 1) * sizeof(long long)) + 256;
            HIPCHK(ensure(ctx->scratch, ctx->scratch_cap, need));
            char* base = reinterpret_cast<char*>(ctx->scratch);
            int* counts = reinterpret_cast<int*>(base);
            long long* offs = reinterpret_cast<long long*>(base + align256((size_t)ni * sizeof(int)));
            long long* tot_d = reinterpret_cast<long long*>(base + need - 256);
            long long tot_h = 0;
            int *lq = nullptr, *lc = nullptr;
            const double* q = m->ctab + (size_t)i0 * CT + CT_RKT;
            if (int rc = build_near_list(ctx, ni, q, CT, *m, 0, np->analytic_calc, 0., counts, offs, tot_d, &tot_h,
                                         &lq, &lc))
                return rc;
            (void)hipFree(m->near_lq);
            (void)hipFree(m->near_lc);
            m->near_lq = m->near_lc = nullptr;
            m->near_tot = -1;
            if (tot_h > 0) {
                HIPCHK(hipMalloc(&m->near_lq, (size_t)tot_h * sizeof(int)));
                HIPCHK(hipMalloc(&m->near_lc, (size_t)tot_h * sizeof(int)));
                HIPCHK(hipMemcpyAsync(m->near_lq, lq, (size_t)tot_h * sizeof(int), hipMemcpyDeviceToDevice,
                                      ctx->stream));
                HIPCHK(hipMemcpyAsync(m->near_lc, lc, (size_t)tot_h * sizeof(int), hipMemcpyDeviceToDevice,
                                      ctx->stream));
            }
            m->near_tot = tot_h;
            m->near_i0 = i0;
            m->near_ni = ni;
            m->near_ac = np->analytic_calc;
        }
        if (m->near_tot > 0) {
            // the list holds point indices local to i0; the kernel adds i0
            Timer t(ctx, "near");
            HIPCHK(launch_assemble_generic(ctx->stream, m->nv, 0, m->near_tot, m->near_lq, m->near_lc, m->n, i0, ni,
                                           m->verts, m->ctab, npp, k_re, k_im, Z_d, ld));
            t.stop();
        }
    }
    return BIELCC_OK;
}

int bielcc_assemble(bielcc_ctx ctx, int mesh_id, int which, const bielcc_num_param* np, double k_re, double k_im,
                    int row0, int nrows, double* Z_d, int64_t ld) {
    return assemble_impl(ctx, mesh_id, which, np, k_re, k_im, row0, nrows, BIELCC_PREC_FP64, Z_d, ld);
}

int bielcc_assemble_ex(bielcc_ctx ctx, int mesh_id, int which, const bielcc_num_param* np, double k_re,
                       double k_im, int row0, int nrows, int precision, double* Z_d, int64_t ld) {
    return assemble_impl(ctx, mesh_id, which, np, k_re, k_im, row0, nrows, precision, Z_d, ld);
}

int bielcc_rhs_planewave(bielcc_ctx ctx, int mesh_id, int nrhs, const double* kvec, const double* e0, double* B_d,
                         int64_t ldb) {
    if (!ctx) return BIELCC_ERR_ARG;
    DevMesh* m = find_mesh(ctx, mesh_id);
    if (!m) return fail(ctx, BIELCC_ERR_NOTFOUND, "unknown mesh id");
    if (nrhs <= 0 || !kvec || !e0 || !B_d || ldb < 2 * m->n) return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    DeviceGuard g(ctx->device);
    std::vector<double> ke((size_t)nrhs * 6);
    for (int r = 0; r < nrhs; r++)
        for (int d = 0; d < 3; d++) { ke[6 * r + d] = kvec[3 * r + d]; ke[6 * r + 3 + d] = e0[3 * r + d]; }
    HIPCHK(ensure(ctx->scratch, ctx->scratch_cap, ke.size() * sizeof(double)));
    HIPCHK(hipMemcpyAsync(ctx->scratch, ke.data(), ke.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(launch_rhs_planewave(ctx->stream, m->n, m->ctab, nrhs, reinterpret_cast<double*>(ctx->scratch), B_d, ldb));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return BIELCC_OK;
}

int bielcc_lu_factor(bielcc_ctx ctx, int n, double* A_d, int64_t lda, int* piv_d, int* info) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (n <= 0 || !A_d || !piv_d || !info || lda < n) return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    DeviceGuard g(ctx->device);
    const char* msg = nullptr;
    Timer t(ctx, "lu_factor");
    hipError_t e = lu_factor(ctx->stream, ctx->lu, n, A_d, lda, piv_d, info, &msg);
    if (e != hipSuccess) return fail(ctx, BIELCC_ERR_HIP, msg ? msg : hipGetErrorString(e));
    t.stop();
    if (*info != 0) return fail(ctx, BIELCC_ERR_SINGULAR, "zero pivot at column " + std::to_string(*info));
    return BIELCC_OK;
}

int bielcc_lu_solve(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, const int* piv_d, int nrhs,
                    double* B_d, int64_t ldb) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (n <= 0 || nrhs <= 0 || !A_d || !piv_d || !B_d || lda < n || ldb < n) return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    DeviceGuard g(ctx->device);
    const char* msg = nullptr;
    Timer t(ctx, "lu_solve");
    hipError_t e = lu_solve(ctx->stream, ctx->lu, n, A_d, lda, piv_d, nrhs, B_d, ldb, &msg);
    if (e != hipSuccess) return fail(ctx, BIELCC_ERR_HIP, msg ? msg : hipGetErrorString(e));
    t.stop();
    return BIELCC_OK;
}

int bielcc_currents_by_basis(bielcc_ctx ctx, int mesh_id, int nrhs, const double* b_d, int64_t ldb, double* j_d) {
    if (!ctx) return BIELCC_ERR_ARG;
    DevMesh* m = find_mesh(ctx, mesh_id);
    if (!m) return fail(ctx, BIELCC_ERR_NOTFOUND, "unknown mesh id");
    if (nrhs <= 0 || !b_d || !j_d || ldb < 2 * m->n) return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    DeviceGuard g(ctx->device);
    HIPCHK(launch_currents(ctx->stream, m->n, m->ctab, nrhs, b_d, ldb, j_d));
    return BIELCC_OK;
}

static int field_impl(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im,
                      int nrhs, const double* jE_d, const double* jM_d, double mEr, double mEi, double mHr,
                      double mHi, const double* inc, int near_mode, int precision, int M, const double* pts_d,
                      double* E_d, double* H_d);

// Mixed field, after its FP32 sums: the points whose modelled FP32 error is not
// small against their scattered field (k_fix_flags, field.hip) are recomputed
// with FP64 sums and written over.  inc: host [nrhs][6] or null; inc_d its device copy.
static int mixed_fixup(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im, int nrhs,
                       const double* jE_d, const double* jM_d, double mEr, double mEi, double mHr, double mHi,
                       const double* inc, const double* inc_d, int near_mode, int M, const double* pts_d,
                       double* E_d, double* H_d) {
    if (!ctx->mixed_fixup || ctx->capturing || M <= 0) return BIELCC_OK;
    DevMesh* m = find_mesh(ctx, mesh_id);
    const size_t mx_b = align256((size_t)2 * nrhs * sizeof(double)), cnt_b = 256;
    const size_t flag_b = align256((size_t)nrhs * M * sizeof(int)), list_b = align256((size_t)M * sizeof(int));
    const size_t pts_b = align256((size_t)3 * M * sizeof(double)), f_b = align256((size_t)nrhs * M * 6 * sizeof(double));
    HIPCHK(ensure(ctx->fix, ctx->fix_cap, mx_b + cnt_b + flag_b + list_b + pts_b + f_b * (H_d ? 2 : 1)));
    char* b = static_cast<char*>(ctx->fix);
    double* mx = reinterpret_cast<double*>(b);
    int* count = reinterpret_cast<int*>(b + mx_b);
    int* flags = reinterpret_cast<int*>(b + mx_b + cnt_b);
    int* list = reinterpret_cast<int*>(b + mx_b + cnt_b + flag_b);
    double* sub = reinterpret_cast<double*>(b + mx_b + cnt_b + flag_b + list_b);
    double* Es = reinterpret_cast<double*>(b + mx_b + cnt_b + flag_b + list_b + pts_b);
    double* Hs = H_d ? Es + f_b / sizeof(double) : nullptr;
    FixArgs f;
    f.M = M; f.N = m->n; f.nrhs = nrhs; f.pts = pts_d; f.ctab = m->ctab; f.jE = jE_d; f.jM = jM_d; f.k_re = k_re;
    f.mE2 = mEr * mEr + mEi * mEi; f.mH2 = mHr * mHr + mHi * mHi; f.inc = inc ? inc_d : nullptr;
    f.E = E_d; f.H = H_d; f.mx = mx; f.flags = flags;
    HIPCHK(launch_fix_flags(ctx->stream, f, count, list));
    int cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, count, sizeof cnt, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (cnt <= 0) return BIELCC_OK;
    ctx->fixup_points += cnt;
    HIPCHK(launch_fix_gather(ctx->stream, cnt, list, pts_d, sub));
    if (int rc = field_impl(ctx, mesh_id, np, k_re, k_im, nrhs, jE_d, jM_d, mEr, mEi, mHr, mHi, inc, near_mode,
                            BIELCC_PREC_FP64, cnt, sub, Es, Hs))
        return rc;
    HIPCHK(launch_fix_scatter(ctx->stream, cnt, M, nrhs, list, Es, Hs, E_d, H_d));
    return BIELCC_OK;
}

// The share of points a mixed field would send back to FP64 sums, estimated on every
// stride-th point: their FP32 field (without the recomputation) and the FP32 error model.
// When the currents carry large parts whose fields cancel (near a resonance), most points
// fail the model, and the FP32 pass plus the recomputation costs more than FP64 sums
// alone (sphere of 20,000 quads at ka = 10: 262 ms FP32 + about 340 ms recomputed, against
// 347 ms in FP64).
constexpr int kProbePoints = 1024;
static int mixed_probe(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im, int nrhs,
                       const double* jE_d, const double* jM_d, double mEr, double mEi, double mHr, double mHi,
                       const double* inc, int near_mode, int M, const double* pts_d, bool has_h, double* frac) {
    DevMesh* m = find_mesh(ctx, mesh_id);
    const int stride = std::max(1, M / kProbePoints);
    const int Ms = (M + stride - 1) / stride;
    const size_t list_b = align256((size_t)Ms * sizeof(int)), pts_b = align256((size_t)3 * Ms * sizeof(double));
    const size_t f_b = align256((size_t)nrhs * Ms * 6 * sizeof(double)), flag_b = align256((size_t)nrhs * Ms * sizeof(int));
    const size_t mx_b = align256((size_t)2 * nrhs * sizeof(double)), inc_b = align256((size_t)nrhs * 6 * sizeof(double));
    HIPCHK(ensure(ctx->probe, ctx->probe_cap, list_b * 2 + pts_b + f_b * 2 + flag_b + mx_b + inc_b + 256));
    char* b = static_cast<char*>(ctx->probe);
    int* idx = reinterpret_cast<int*>(b);
    int* list = reinterpret_cast<int*>(b + list_b);
    double* sub = reinterpret_cast<double*>(b + 2 * list_b);
    double* Es = reinterpret_cast<double*>(b + 2 * list_b + pts_b);
    double* Hs = has_h ? Es + f_b / sizeof(double) : nullptr;
    int* flags = reinterpret_cast<int*>(b + 2 * list_b + pts_b + 2 * f_b);
    double* mx = reinterpret_cast<double*>(b + 2 * list_b + pts_b + 2 * f_b + flag_b);
    double* inc_d = reinterpret_cast<double*>(b + 2 * list_b + pts_b + 2 * f_b + flag_b + mx_b);
    int* count = reinterpret_cast<int*>(b + 2 * list_b + pts_b + 2 * f_b + flag_b + mx_b + inc_b);
    std::vector<int> hidx(Ms);
    for (int i = 0; i < Ms; i++) hidx[i] = i * stride;
    HIPCHK(hipMemcpyAsync(idx, hidx.data(), (size_t)Ms * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    if (inc) HIPCHK(hipMemcpyAsync(inc_d, inc, (size_t)nrhs * 6 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(launch_fix_gather(ctx->stream, Ms, idx, pts_d, sub));
    ctx->in_probe = true;
    const int rc = field_impl(ctx, mesh_id, np, k_re, k_im, nrhs, jE_d, jM_d, mEr, mEi, mHr, mHi, inc, near_mode,
                              BIELCC_PREC_MIXED, Ms, sub, Es, Hs);
    ctx->in_probe = false;
    if (rc) return rc;
    FixArgs f;
    f.M = Ms; f.N = m->n; f.nrhs = nrhs; f.pts = sub; f.ctab = m->ctab; f.jE = jE_d; f.jM = jM_d; f.k_re = k_re;
    f.mE2 = mEr * mEr + mEi * mEi; f.mH2 = mHr * mHr + mHi * mHi; f.inc = inc ? inc_d : nullptr;
    f.E = Es; f.H = Hs; f.mx = mx; f.flags = flags;
    HIPCHK(launch_fix_flags(ctx->stream, f, count, list));
    int cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, count, sizeof cnt, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *frac = (double)cnt / Ms;
    return BIELCC_OK;
}

static int field_impl(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im,
                      int nrhs, const double* jE_d, const double* jM_d, double mEr, double mEi, double mHr,
                      double mHi, const double* inc, int near_mode, int precision, int M, const double* pts_d,
                      double* E_d, double* H_d) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (int rc = check_np(ctx, np)) return rc;
    DevMesh* m = find_mesh(ctx, mesh_id);
    if (!m) return fail(ctx, BIELCC_ERR_NOTFOUND, "unknown mesh id");
    if (M < 0 || near_mode < 0 || near_mode > 1) return fail(ctx, BIELCC_ERR_ARG, "bad M or near_mode");
    if (nrhs < 1) return fail(ctx, BIELCC_ERR_ARG, "nrhs must be >= 1");
    if (precision != BIELCC_PREC_FP64 && precision != BIELCC_PREC_MIXED)
        return fail(ctx, BIELCC_ERR_ARG, "precision must be BIELCC_PREC_FP64 or BIELCC_PREC_MIXED");
    if (M == 0) return BIELCC_OK;
    if (!jE_d || !pts_d || !E_d) return fail(ctx, BIELCC_ERR_ARG, "NULL buffer");
    if (!std::isfinite(k_re) || !std::isfinite(k_im)) return fail(ctx, BIELCC_ERR_ARG, "non-finite wave number");
    const int nn = np->n_start * np->n_start;
    if (np->p_max != 1 || nn > 512 || !(np->analytic_calc > 1.))
        return fail(ctx, BIELCC_ERR_UNSUPPORTED,
                    "field driver needs p_max = 1, n_start <= 22 and analytic_calc > 1 (use bielcc_op_batch)");
    DeviceGuard g(ctx->device);
    // FP32 node sums issue at the FP64 rate (no packed math), and where the node phase is the
    // centre-phase series the real-k FP64 loop takes fewer instructions than the FP32 loop with
    // its cancellation-free r - dcen, so the mixed field sums in FP64 there.
    if (precision == BIELCC_PREC_MIXED && !ctx->mixed_far32 && k_im == 0. && !ctx->no_fast_phase &&
        phase_mode(k_re, m->grid_step, m->max_rad) >= 2) {
        precision = BIELCC_PREC_FP64;
        ctx->auto_fp64++;
    }
    if (precision == BIELCC_PREC_MIXED && ctx->mixed_fixup && ctx->mixed_probe_pct > 0 && !ctx->capturing &&
        !ctx->in_probe && M >= 4 * kProbePoints) {
        double frac = 0.;
        if (int rc = mixed_probe(ctx, mesh_id, np, k_re, k_im, nrhs, jE_d, jM_d, mEr, mEi, mHr, mHi, inc, near_mode,
                                 M, pts_d, H_d != nullptr, &frac))
            return rc;
        if (frac * 100. > ctx->mixed_probe_pct) {
            precision = BIELCC_PREC_FP64;
            ctx->probe_fp64++;
        }
    }
    if (int rc = ensure_tables(ctx, *m, np->n_start, 0)) return rc;
    if (precision == BIELCC_PREC_MIXED)
        if (int rc = ensure_tables32(ctx, *m, np->n_start)) return rc;
    const int has_jm = jM_d != nullptr;
    const int nacc = has_jm ? 24 : 12;
    const double thresh = np->analytic_calc * m->grid_step;
    // enough workgroups to fill the chip: split the cells when the points alone do not,
    // into a count whose last round of workgroups is nearly full
    const int bx = (M + 255) / 256;
    FieldArgs a;
    a.nv = m->nv; a.M = M; a.N = m->n; a.nn = nn; a.pts = pts_d; a.ctab = m->ctab; a.nodes = m->nodes;
    a.nodes32 = m->nodes32; a.nodes32_yy = m->nodes32_yy;
    a.jE = jE_d; a.jM = jM_d; a.nrhs = nrhs; a.rchunk = ctx->field_rchunk;
    a.k_re = k_re; a.k_im = k_im; a.mode = near_mode; a.ac = np->analytic_calc;
    a.thresh = thresh;
    a.fast_phase = ctx->no_fast_phase ? 0 : phase_mode(k_re, m->grid_step, m->max_rad);
    a.prec = precision;
    const int nsplit = field_far_splits(a, std::max(1, std::min((ctx->field_wg + bx - 1) / bx, (m->n + 63) / 64)));
    size_t part_b = align256((size_t)nsplit * nrhs * nacc * M * sizeof(double));
    size_t cnt_b = align256((size_t)M * sizeof(int));
    size_t off_b = align256((size_t)(M + 1) * sizeof(long long));
    size_t inc_b = align256((size_t)nrhs * 6 * sizeof(double));
    size_t need = part_b + cnt_b + off_b + inc_b + 256;
    HIPCHK(ensure(ctx->scratch, ctx->scratch_cap, need));
    char* base = reinterpret_cast<char*>(ctx->scratch);
    double* part = reinterpret_cast<double*>(base);
    int* counts = reinterpret_cast<int*>(base + part_b);
    long long* offs = reinterpret_cast<long long*>(base + part_b + cnt_b);
    double* inc_d = reinterpret_cast<double*>(base + part_b + cnt_b + off_b);
    long long* tot_d = reinterpret_cast<long long*>(base + part_b + cnt_b + off_b + inc_b);
    if (inc) HIPCHK(hipMemcpyAsync(inc_d, inc, (size_t)nrhs * 6 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const bool fixup = precision == BIELCC_PREC_MIXED && !ctx->in_probe;
    if (nrhs >= ctx->opmat_min_rhs && !jM_d) {
        if (int rc = field_opmat(ctx, *m, np, k_re, k_im, nrhs, jE_d, mEr, mEi, inc ? inc_d : nullptr, near_mode,
                                 precision, M, pts_d, E_d, H_d, counts, offs, tot_d))
            return rc;
        return fixup ? mixed_fixup(ctx, mesh_id, np, k_re, k_im, nrhs, jE_d, jM_d, mEr, mEi, mHr, mHi, inc,
                                   inc ? inc_d : nullptr, near_mode, M, pts_d, E_d, H_d)
                     : BIELCC_OK;
    }
    a.nsplit = nsplit; a.part = part;
    {
        Timer t(ctx, "field_far");
        HIPCHK(launch_field_far(ctx->stream, a));
        t.stop();
    }
    long long tot_h = 0;
    int *lq = nullptr, *lc = nullptr;
    if (int rc = build_near_list(ctx, M, pts_d, 3, *m, near_mode, np->analytic_calc, thresh, counts, offs, tot_d,
                                 &tot_h, &lq, &lc))
        return rc;
    // one current set: the operators applied to it per pair; several: per-pair 3x3 operator matrices
    const int as_matrix = nrhs > 1;
    double* near_out = nullptr;
    if (tot_h > 0) {
        HIPCHK(ensure(ctx->near_buf, ctx->near_cap, (size_t)tot_h * (as_matrix ? 36 : nacc) * sizeof(double)));
        near_out = reinterpret_cast<double*>(ctx->near_buf);
        Timer t(ctx, "near");
        HIPCHK(launch_field_near(ctx->stream, m->nv, tot_h, lq, lc, pts_d, m->verts, jE_d, jM_d, to_np(np), k_re,
                                 k_im, near_mode, as_matrix, near_out));
        t.stop();
    }
    FinalizeArgs f;
    f.M = M; f.N = m->n; f.nsplit = nsplit; f.nrhs = nrhs; f.part = part; f.jE = jE_d; f.jM = jM_d;
    f.offs = offs; f.lc = lc; f.near_out = near_out; f.near_mat = as_matrix;
    f.mEr = mEr; f.mEi = mEi; f.mHr = mHr; f.mHi = mHi;
    f.inc = inc ? inc_d : nullptr; f.pts = pts_d; f.E = E_d; f.H = H_d;
    HIPCHK(launch_field_finalize(ctx->stream, f));
    return fixup ? mixed_fixup(ctx, mesh_id, np, k_re, k_im, nrhs, jE_d, jM_d, mEr, mEi, mHr, mHi, inc,
                               inc ? inc_d : nullptr, near_mode, M, pts_d, E_d, H_d)
                 : BIELCC_OK;
}

int bielcc_field(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im,
                 const double* jE_d, const double* jM_d, double multE_re, double multE_im, double multH_re,
                 double multH_im, const double* kvec, const double* e0, int near_mode, int M, const double* pts_d,
                 double* E_d, double* H_d) {
    if ((kvec == nullptr) != (e0 == nullptr)) return fail(ctx, BIELCC_ERR_ARG, "kvec and e0 go together");
    double ke[6] = {0, 0, 0, 0, 0, 0};
    if (kvec) for (int d = 0; d < 3; d++) { ke[d] = kvec[d]; ke[3 + d] = e0[d]; }
    return field_impl(ctx, mesh_id, np, k_re, k_im, 1, jE_d, jM_d, multE_re, multE_im, multH_re, multH_im,
                      kvec ? ke : nullptr, near_mode, BIELCC_PREC_FP64, M, pts_d, E_d, H_d);
}

int bielcc_get_field_e(bielcc_ctx ctx, int mesh_id, const double* jE_d, double mult_re, double mult_im,
                       const double* kvec, const double* e0, const bielcc_num_param* np, double k_re, double k_im,
                       int M, const double* pts_d, double* E_d) {
    if (!kvec || !e0) return fail(ctx, BIELCC_ERR_ARG, "GetFieldE_Ideal needs the incident plane wave");
    double ke[6] = {kvec[0], kvec[1], kvec[2], e0[0], e0[1], e0[2]};
    return field_impl(ctx, mesh_id, np, k_re, k_im, 1, jE_d, nullptr, mult_re, mult_im, 1., 0., ke, 0,
                      BIELCC_PREC_FP64, M, pts_d, E_d, nullptr);
}

int bielcc_field_batch(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im,
                       int nrhs, const double* jE_d, const double* jM_d, double multE_re, double multE_im,
                       double multH_re, double multH_im, const double* inc, int near_mode, int precision, int M,
                       const double* pts_d, double* E_d, double* H_d) {
    return field_impl(ctx, mesh_id, np, k_re, k_im, nrhs, jE_d, jM_d, multE_re, multE_im, multH_re, multH_im, inc,
                      near_mode, precision, M, pts_d, E_d, H_d);
}

int bielcc_far_pattern(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im,
                       int nrhs, const double* jE_d, const double* jM_d, double multE_re, double multE_im,
                       double multH_re, double multH_im, int precision, int ndir, const double* dirs_d,
                       double* Einf_d, double* Hinf_d) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (int rc = check_np(ctx, np)) return rc;
    DevMesh* m = find_mesh(ctx, mesh_id);
    if (!m) return fail(ctx, BIELCC_ERR_NOTFOUND, "unknown mesh id");
    if (nrhs < 1 || ndir < 0) return fail(ctx, BIELCC_ERR_ARG, "nrhs must be >= 1 and ndir >= 0");
    if (precision != BIELCC_PREC_FP64 && precision != BIELCC_PREC_MIXED)
        return fail(ctx, BIELCC_ERR_ARG, "precision must be BIELCC_PREC_FP64 or BIELCC_PREC_MIXED");
    if (ndir == 0) return BIELCC_OK;
    if (!jE_d || !dirs_d || !Einf_d) return fail(ctx, BIELCC_ERR_ARG, "NULL buffer");
    if (!std::isfinite(k_re) || !std::isfinite(k_im)) return fail(ctx, BIELCC_ERR_ARG, "non-finite wave number");
    const int nn = np->n_start * np->n_start;
    if (nn > 512) return fail(ctx, BIELCC_ERR_UNSUPPORTED, "far pattern needs n_start <= 22");
    DeviceGuard g(ctx->device);
    if (int rc = ensure_tables(ctx, *m, np->n_start, 0)) return rc;
    if (precision == BIELCC_PREC_MIXED)
        if (int rc = ensure_tables32(ctx, *m, np->n_start)) return rc;
    const int bx = (ndir + 255) / 256;
    const int nsplit = std::max(1, std::min((ctx->field_wg + bx - 1) / bx, (m->n + 63) / 64));
    HIPCHK(ensure(ctx->scratch, ctx->scratch_cap, (size_t)nsplit * nrhs * 12 * ndir * sizeof(double)));
    PatternArgs a;
    a.N = m->n; a.nn = nn; a.ndir = ndir; a.nrhs = nrhs; a.nsplit = nsplit; a.dirs = dirs_d; a.ctab = m->ctab;
    a.nodes = m->nodes; a.nodes32 = m->nodes32; a.jE = jE_d; a.jM = jM_d; a.k_re = k_re; a.k_im = k_im;
    a.fast_phase = std::fabs(k_re) * m->grid_step <= 1.95 && !ctx->no_fast_phase;
    a.prec = precision; a.part = reinterpret_cast<double*>(ctx->scratch); a.Einf = Einf_d; a.Hinf = Hinf_d;
    Timer t(ctx, "far_pattern");
    HIPCHK(launch_far_pattern(ctx->stream, a, multE_re, multE_im, multH_re, multH_im));
    t.stop();
    return BIELCC_OK;
}

int bielcc_lu32_factor(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, float* F_d, int64_t ldf, int* piv_d,
                       int* info) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (n <= 0 || !A_d || !F_d || !piv_d || !info || lda < n || ldf < n)
        return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    DeviceGuard g(ctx->device);
    const char* msg = nullptr;
    Timer t(ctx, "lu32_factor");
    hipError_t e = lu32_factor(ctx->stream, ctx->lu, n, A_d, lda, F_d, ldf, piv_d, info, &msg);
    if (e != hipSuccess) return fail(ctx, BIELCC_ERR_HIP, msg ? msg : hipGetErrorString(e));
    t.stop();
    if (*info != 0) return fail(ctx, BIELCC_ERR_SINGULAR, "zero FP32 pivot at column " + std::to_string(*info));
    return BIELCC_OK;
}

int bielcc_lu_factor_rbt(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, double* F_d, int64_t ldf,
                         int* piv_d, double* rbt_d, uint64_t seed, int* info) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (n <= 0 || !A_d || !F_d || !piv_d || !rbt_d || !info || lda < n || ldf < n)
        return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    if ((const void*)A_d == (const void*)F_d) return fail(ctx, BIELCC_ERR_ARG, "F_d must not be A_d (A is kept)");
    DeviceGuard g(ctx->device);
    const char* msg = nullptr;
    Timer t(ctx, "lu_factor_rbt");
    hipError_t e = lu_factor_rbt(ctx->stream, ctx->lu, n, A_d, lda, F_d, ldf, piv_d, rbt_d, seed, info, &msg);
    if (e != hipSuccess) return fail(ctx, BIELCC_ERR_HIP, msg ? msg : hipGetErrorString(e));
    t.stop();
    if (*info != 0) return fail(ctx, BIELCC_ERR_SINGULAR, "zero pivot (no interchanges) at column " + std::to_string(*info));
    return BIELCC_OK;
}

int bielcc_rbt_solve(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, const double* F_d, int64_t ldf,
                     const int* piv_d, const double* rbt_d, int nrhs, double* B_d, int64_t ldb, double tol,
                     int max_iter, int* iters, double* resid) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (n <= 0 || nrhs <= 0 || !A_d || !F_d || !piv_d || !rbt_d || !B_d || !iters || !resid || lda < n ||
        ldf < n || ldb < n || !(tol > 0.) || max_iter < 0)
        return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    DeviceGuard g(ctx->device);
    const char* msg = nullptr;
    int status = 0;
    Timer t(ctx, "rbt_solve");
    hipError_t e = rbt_solve(ctx->stream, ctx->lu, n, A_d, lda, F_d, ldf, piv_d, rbt_d, nrhs, B_d, ldb, tol, max_iter,
                             iters, resid, &status, &msg);
    if (e != hipSuccess) return fail(ctx, BIELCC_ERR_HIP, msg ? msg : hipGetErrorString(e));
    t.stop();
    if (status != 0) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "refinement stopped at relative residual %.3e > tol %.3e after %d corrections",
                      *resid, tol, *iters);
        return fail(ctx, BIELCC_ERR_UNSUPPORTED, buf);
    }
    return BIELCC_OK;
}

static int refine_solve_impl(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, const float* F_d, int64_t ldf,
                             const int* piv_d, const double* rbt_d, int nrhs, double* B_d, int64_t ldb, double tol,
                             int max_iter, int* iters, double* resid) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (n <= 0 || nrhs <= 0 || !A_d || !F_d || !piv_d || !B_d || !iters || !resid || lda < n || ldf < n ||
        ldb < n || !(tol > 0.) || max_iter < 0)
        return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    DeviceGuard g(ctx->device);
    const char* msg = nullptr;
    int status = 0;
    Timer t(ctx, "refine_solve");
    hipError_t e = refine_solve(ctx->stream, ctx->lu, n, A_d, lda, F_d, ldf, piv_d, nrhs, B_d, ldb, tol, max_iter,
                                iters, resid, &status, &msg, rbt_d);
    if (e != hipSuccess) return fail(ctx, BIELCC_ERR_HIP, msg ? msg : hipGetErrorString(e));
    t.stop();
    if (status != 0) {
        char buf[200];
        if (*resid <= tol)   // the residual met tol; the forward-error estimate did not reach refine_ferr
            std::snprintf(buf, sizeof buf,
                          "refinement stalled at relative residual %.3e (<= tol %.3e) before the estimated forward "
                          "error reached %.1e, after %d corrections",
                          *resid, tol, lu_get_refine_ferr(ctx->lu), *iters);
        else
            std::snprintf(buf, sizeof buf, "refinement stopped at relative residual %.3e > tol %.3e after %d corrections",
                          *resid, tol, *iters);
        return fail(ctx, BIELCC_ERR_UNSUPPORTED, buf);
    }
    return BIELCC_OK;
}

int bielcc_refine_solve(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, const float* F_d, int64_t ldf,
                        const int* piv_d, int nrhs, double* B_d, int64_t ldb, double tol, int max_iter,
                        int* iters, double* resid) {
    return refine_solve_impl(ctx, n, A_d, lda, F_d, ldf, piv_d, nullptr, nrhs, B_d, ldb, tol, max_iter, iters, resid);
}

int bielcc_refine_solve_rbt(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, const float* F_d, int64_t ldf,
                            const int* piv_d, const double* rbt_d, int nrhs, double* B_d, int64_t ldb, double tol,
                            int max_iter, int* iters, double* resid) {
    if (ctx && !rbt_d) return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    return refine_solve_impl(ctx, n, A_d, lda, F_d, ldf, piv_d, rbt_d, nrhs, B_d, ldb, tol, max_iter, iters, resid);
}

int bielcc_lu32_factor_rbt(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, float* F_d, int64_t ldf,
                           int* piv_d, double* rbt_d, uint64_t seed, int* info) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (n <= 0 || !A_d || !F_d || !piv_d || !rbt_d || !info || lda < n || ldf < n)
        return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    DeviceGuard g(ctx->device);
    const char* msg = nullptr;
    Timer t(ctx, "lu32_factor");
    hipError_t e = lu32_factor_rbt(ctx->stream, ctx->lu, n, A_d, lda, F_d, ldf, piv_d, rbt_d, seed, info, &msg);
    if (e != hipSuccess) return fail(ctx, BIELCC_ERR_HIP, msg ? msg : hipGetErrorString(e));
    t.stop();
    if (*info != 0) return fail(ctx, BIELCC_ERR_SINGULAR, "zero FP32 pivot (no interchanges) at column " + std::to_string(*info));
    return BIELCC_OK;
}

int bielcc_residual(bielcc_ctx ctx, int n, const double* A_d, int64_t lda, const double* X_d, int64_t ldx,
                    const double* B_d, int64_t ldb, int nrhs, int doubled, double* R_d, int64_t ldr) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (n <= 0 || nrhs <= 0 || !A_d || !X_d || !B_d || !R_d || lda < n || ldx < n || ldb < n || ldr < n)
        return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    if (doubled != 0 && doubled != 1) return fail(ctx, BIELCC_ERR_ARG, "doubled must be 0 or 1");
    DeviceGuard g(ctx->device);
    const char* msg = nullptr;
    Timer t(ctx, "residual");
    hipError_t e = residual(ctx->stream, ctx->lu, n, A_d, lda, X_d, ldx, B_d, ldb, nrhs, doubled == 1, R_d, ldr, &msg);
    if (e != hipSuccess) return fail(ctx, BIELCC_ERR_HIP, msg ? msg : hipGetErrorString(e));
    t.stop();
    return BIELCC_OK;
}

int bielcc_set_tuning(bielcc_ctx ctx, const char* key, int value) {
    if (!ctx || !key) return BIELCC_ERR_ARG;
    const std::string k(key);
    if (k == "field_wg" || k == "assemble_wg") {
        if (value < 1 || value > (1 << 20)) return fail(ctx, BIELCC_ERR_ARG, "workgroup target out of range");
        (k == "field_wg" ? ctx->field_wg : ctx->assemble_wg) = value;
        return BIELCC_OK;
    }
    if (k == "field_rchunk") {
        if (value != 1 && value != 2 && value != 4) return fail(ctx, BIELCC_ERR_ARG, "field_rchunk must be 1, 2 or 4");
        ctx->field_rchunk = value;
        return BIELCC_OK;
    }
    if (k == "trsv") {
        if (value != 0 && value != 1) return fail(ctx, BIELCC_ERR_ARG, "trsv must be 0 or 1");
        lu_set_trsv(ctx->lu, value == 1);
        return BIELCC_OK;
    }
    if (k == "trsv_spin") {
        if (value < 1) return fail(ctx, BIELCC_ERR_ARG, "trsv_spin must be >= 1");
        lu_set_spin_limit(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "lu_fused_panel") {
        if (value < 0 || value > 2) return fail(ctx, BIELCC_ERR_ARG, "lu_fused_panel must be 0, 1 or 2");
        lu_set_fused_panel(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "refine_ferr_exp") {
        if (value < 0 || value > 15) return fail(ctx, BIELCC_ERR_ARG, "refine_ferr_exp must be 0..15");
        lu_set_refine_ferr(ctx->lu, value == 0 ? 0. : std::pow(10., -value));
        return BIELCC_OK;
    }
    if (k == "refine_mode") {
        if (value != 0 && value != 1) return fail(ctx, BIELCC_ERR_ARG, "refine_mode must be 0 or 1");
        lu_set_refine_mode(ctx->lu, value, 0);
        return BIELCC_OK;
    }
    if (k == "refine_resid_dd") {
        if (value < 0 || value > 2) return fail(ctx, BIELCC_ERR_ARG, "refine_resid_dd must be 0, 1 or 2");
        lu_set_resid_dd(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "refine_restart") {
        if (value < 1 || value > 64) return fail(ctx, BIELCC_ERR_ARG, "refine_restart must be 1..64");
        lu_set_refine_mode(ctx->lu, -1, value);
        return BIELCC_OK;
    }
    if (k == "trsv_max64") {
        if (value < 0 || value > 8) return fail(ctx, BIELCC_ERR_ARG, "trsv_max64 must be 0..8");
        lu_set_trsv_limits(ctx->lu, value, -1);
        return BIELCC_OK;
    }
    if (k == "trsv_max32") {
        if (value < 0 || value > 16) return fail(ctx, BIELCC_ERR_ARG, "trsv_max32 must be 0..16");
        lu_set_trsv_limits(ctx->lu, -1, -1, value);
        return BIELCC_OK;
    }
    if (k == "trsv_chunks32") {
        if (value != 0 && value != 1) return fail(ctx, BIELCC_ERR_ARG, "trsv_chunks32 must be 0 or 1");
        lu_set_trsv_limits(ctx->lu, -1, value);
        return BIELCC_OK;
    }
    if (k == "lu_trsm_lib") {
        if (value < 0 || value > 2) return fail(ctx, BIELCC_ERR_ARG, "lu_trsm_lib must be 0, 1 or 2");
        lu_set_trsm_lib_panel(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "lu_outer_nb") {
        if (value != 0 && value != 64 && value != 128 && value != 256 && value != 512)
            return fail(ctx, BIELCC_ERR_ARG, "lu_outer_nb must be 0, 64, 128, 256 or 512");
        lu_set_outer_nb(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "blas_atomics") {
        if (value != 0 && value != 1) return fail(ctx, BIELCC_ERR_ARG, "blas_atomics must be 0 or 1");
        lu_set_blas_atomics(ctx->lu, value == 1);
        return BIELCC_OK;
    }
    if (k == "mixed_probe_pct") {
        if (value < 0 || value > 100) return fail(ctx, BIELCC_ERR_ARG, "mixed_probe_pct must be 0..100");
        ctx->mixed_probe_pct = value;
        return BIELCC_OK;
    }
    if (k == "mixed_far32") {
        if (value != 0 && value != 1) return fail(ctx, BIELCC_ERR_ARG, "mixed_far32 must be 0 or 1");
        ctx->mixed_far32 = value == 1;
        return BIELCC_OK;
    }
    if (k == "mixed_fixup") {
        if (value != 0 && value != 1) return fail(ctx, BIELCC_ERR_ARG, "mixed_fixup must be 0 or 1");
        ctx->mixed_fixup = value == 1;
        return BIELCC_OK;
    }
    if (k == "lu_panel_rows64") {
        if (value != 64 && value != 128) return fail(ctx, BIELCC_ERR_ARG, "lu_panel_rows64 must be 64 or 128");
        lu_set_panel_rows64(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "lu_panel_rows32") {
        if (value != 0 && value != 64 && value != 128) return fail(ctx, BIELCC_ERR_ARG, "lu_panel_rows32 must be 0, 64 or 128");
        lu_set_panel_rows32(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "lu_panel_lds") {
        if (value < 0 || value > 2) return fail(ctx, BIELCC_ERR_ARG, "lu_panel_lds must be 0, 1 or 2");
        lu_set_panel_lds(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "lu_tournament") {
        if (value != 0 && value != 1) return fail(ctx, BIELCC_ERR_ARG, "lu_tournament must be 0 or 1");
        lu_set_tournament(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "lu32_gemm16") {
        if (value < 0 || value > 2) return fail(ctx, BIELCC_ERR_ARG, "lu32_gemm16 must be 0, 1 or 2");
        lu_set_gemm16(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "lu_look_ahead2") {
        if (value != 0 && value != 1) return fail(ctx, BIELCC_ERR_ARG, "lu_look_ahead2 must be 0 or 1");
        lu_set_look_ahead2(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "lu_panel_wait_us") {
        if (value < 0) return fail(ctx, BIELCC_ERR_ARG, "lu_panel_wait_us must be >= 0");
        lu_set_panel_wait_us(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "lu_look_ahead") {
        if (value < 0 || value > 2) return fail(ctx, BIELCC_ERR_ARG, "lu_look_ahead must be 0, 1 or 2");
        lu_set_look_ahead(ctx->lu, value);
        return BIELCC_OK;
    }
    if (k == "trsm_lib") {
        if (value != 0 && value != 1) return fail(ctx, BIELCC_ERR_ARG, "trsm_lib must be 0 or 1");
        lu_set_lib_trsm(ctx->lu, value == 1);
        return BIELCC_OK;
    }
    if (k == "plan_field_cells" || k == "plan_asm_cells") {
        if (value < 8 || value > 4096) return fail(ctx, BIELCC_ERR_ARG, "cells per tile must be in [8, 4096]");
        (k == "plan_field_cells" ? ctx->plan_field_cells : ctx->plan_asm_cells) = value;
        return BIELCC_OK;
    }
    if (k == "plan_wg_per_cu") {
        if (value < 1 || value > 32) return fail(ctx, BIELCC_ERR_ARG, "plan_wg_per_cu must be in [1, 32]");
        ctx->plan_wg_per_cu = value;
        return BIELCC_OK;
    }
    if (k == "plan_ws_mb") {
        if (value < 16) return fail(ctx, BIELCC_ERR_ARG, "plan_ws_mb must be >= 16");
        ctx->plan_ws_mb = value;
        return BIELCC_OK;
    }
    if (k == "opmat_min_rhs") {
        if (value < 1) return fail(ctx, BIELCC_ERR_ARG, "opmat_min_rhs must be >= 1");
        ctx->opmat_min_rhs = value;
        return BIELCC_OK;
    }
    return fail(ctx, BIELCC_ERR_ARG, "unknown tuning key " + k);
}

int bielcc_get_tuning(bielcc_ctx ctx, const char* key, int* value) {
    if (!ctx || !key || !value) return BIELCC_ERR_ARG;
    const std::string k(key);
    if (k == "field_wg") *value = ctx->field_wg;
    else if (k == "assemble_wg") *value = ctx->assemble_wg;
    else if (k == "field_rchunk") *value = ctx->field_rchunk;
    else if (k == "trsv") *value = lu_get_trsv(ctx->lu) ? 1 : 0;
    else if (k == "trsm_lib") *value = lu_get_lib_trsm(ctx->lu) ? 1 : 0;
    else if (k == "trsv_spin") *value = lu_get_spin_limit(ctx->lu);
    else if (k == "lu_fused_panel") *value = lu_get_fused_panel(ctx->lu);
    else if (k == "lu_look_ahead") *value = lu_get_look_ahead(ctx->lu);
    else if (k == "lu_panel_lds") *value = lu_get_panel_lds(ctx->lu);
    else if (k == "lu_panel_rows32") *value = lu_get_panel_rows32(ctx->lu);
    else if (k == "lu_panel_rows64") *value = lu_get_panel_rows64(ctx->lu);
    else if (k == "lu_tournament") *value = lu_get_tournament(ctx->lu);
    else if (k == "lu32_gemm16") *value = lu_get_gemm16(ctx->lu);
    else if (k == "lu_look_ahead2") *value = lu_get_look_ahead2(ctx->lu);
    else if (k == "mixed_fixup") *value = ctx->mixed_fixup ? 1 : 0;
    else if (k == "mixed_fixup_points") *value = (int)std::min<long long>(ctx->fixup_points, INT_MAX);
    else if (k == "mixed_probe_pct") *value = ctx->mixed_probe_pct;
    else if (k == "mixed_probe_fp64") *value = (int)std::min<long long>(ctx->probe_fp64, INT_MAX);
    else if (k == "mixed_far32") *value = ctx->mixed_far32 ? 1 : 0;
    else if (k == "mixed_auto_fp64") *value = (int)std::min<long long>(ctx->auto_fp64, INT_MAX);
    else if (k == "blas_atomics") *value = lu_get_blas_atomics(ctx->lu) ? 1 : 0;
    else if (k == "lu_outer_nb") *value = lu_get_outer_nb(ctx->lu);
    else if (k == "lu_trsm_lib") *value = lu_get_trsm_lib_panel(ctx->lu);
    else if (k == "trsv_max64") *value = lu_get_trsv_max64(ctx->lu);
    else if (k == "trsv_max32") *value = lu_get_trsv_max32(ctx->lu);
    else if (k == "refine_ferr_exp") {
        const double f = lu_get_refine_ferr(ctx->lu);
        *value = f > 0. ? (int)std::lround(-std::log10(f)) : 0;
    }
    else if (k == "refine_mode") *value = lu_get_refine_mode(ctx->lu);
    else if (k == "refine_restart") *value = lu_get_refine_restart(ctx->lu);
    else if (k == "refine_resid_dd") *value = lu_get_resid_dd(ctx->lu);
    else if (k == "trsv_chunks32") *value = lu_get_trsv_chunks32(ctx->lu) ? 1 : 0;
    else if (k == "trsv_expired") *value = lu_trsv_retries(ctx->lu, 0);
    else if (k == "trsv_ticket_faults") *value = lu_trsv_retries(ctx->lu, 1);
    else if (k == "lu_panel_wait_us") *value = lu_get_panel_wait_us(ctx->lu);
    else if (k == "lu_panel_aborts") *value = lu_panel_aborts(ctx->lu);
    else if (k == "lu_panel_late_admits") *value = lu_late_admits(ctx->lu);
    else if (k == "plan_field_cells") *value = ctx->plan_field_cells;
    else if (k == "plan_asm_cells") *value = ctx->plan_asm_cells;
    else if (k == "plan_wg_per_cu") *value = ctx->plan_wg_per_cu;
    else if (k == "plan_ws_mb") *value = ctx->plan_ws_mb;
    else if (k == "opmat_min_rhs") *value = ctx->opmat_min_rhs;
    else if (k == "n_cu") *value = ctx->n_cu;
    else if (k == "scratch_kb") {   // the context's grow-only device buffers, KiB rounded up
        size_t b = ctx->scratch_cap + ctx->lists_cap + ctx->near_cap + ctx->opmat_ws_cap + ctx->opmat_meta_cap +
                   ctx->opmat_part_cap + ctx->small_cap + ctx->fix_cap + ctx->probe_cap + lu_work_bytes(ctx->lu);
        for (const auto& kv : ctx->meshes)   // the meshes' kept near-pair lists
            if (kv.second.near_tot > 0) b += (size_t)kv.second.near_tot * 2 * sizeof(int);
        *value = (int)std::min<size_t>((b + 1023) >> 10, (size_t)INT_MAX);
    }
    else return fail(ctx, BIELCC_ERR_ARG, "unknown tuning key " + k);
    return BIELCC_OK;
}

int bielcc_timing_enable(bielcc_ctx ctx, int on) {
    if (!ctx) return BIELCC_ERR_ARG;
    ctx->timing = on != 0;
    return BIELCC_OK;
}

int bielcc_timing_reset(bielcc_ctx ctx) {
    if (!ctx) return BIELCC_ERR_ARG;
    ctx->times.clear();
    return BIELCC_OK;
}

int bielcc_timing(bielcc_ctx ctx, const char* name, double* total_ms, int* count) {
    if (!ctx || !name || !total_ms || !count) return BIELCC_ERR_ARG;
    auto it = ctx->times.find(name);
    if (it == ctx->times.end()) { *total_ms = 0.; *count = 0; return BIELCC_OK; }
    *total_ms = it->second.first;
    *count = it->second.second;
    return BIELCC_OK;
}

}  // extern "C"

namespace bielcc_capi {
int field_mixed_fixup(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im, int nrhs,
                      const double* jE_d, double mEr, double mEi, const double* inc, const double* inc_d, int near_mode,
                      int M, const double* pts_d, double* E_d, double* H_d) {
    return mixed_fixup(ctx, mesh_id, np, k_re, k_im, nrhs, jE_d, nullptr, mEr, mEi, 1., 0., inc, inc_d, near_mode, M,
                       pts_d, E_d, H_d);
}
}  // namespace bielcc_capi
