// Plans: the collocation matrices or the fields of several meshes at once.
//
// A plan fixes its items (mesh, wave number, caller buffers) at creation and
// precomputes everything that needs the host: the tile list of the persistent
// kernel, the near-pair lists, the operator-matrix workspace and the BLAS
// handle.  bielcc_plan_run then only enqueues kernels and GEMMs on the
// context's stream -- no allocation, no host synchronisation, no memset -- so
// bielcc_plan_capture can record it into a HIP graph that later runs replay.
//
// Field plans evaluate bielcc_field_batch (electric currents only) through the
// operator matrices WK, WR of each (mesh, point range) segment: one persistent
// launch writes the far-pair blocks of every segment of a workspace group, the
// exact kernel writes the near-pair blocks, and E = multE WK J + E_inc,
// H = WR J are complex GEMMs on the matrix cores over all current sets at once.
#include "ctx_internal.h"

using namespace bielcc_impl;
using namespace bielcc_capi;

namespace {

struct NearSet {
    long long n = 0;
    int* lq = nullptr;   // point index within the segment (field) or item (assembly)
    int* lc = nullptr;   // cell index
};

struct SegInfo {
    int item, p0, M;     // points p0 .. p0+M-1 of the item
    size_t ws_off;       // byte offset of WK in the workspace (WR follows)
    int parts;           // split-K parts of its GEMMs
};

struct Group {
    int t0, nt;          // tiles
    int s0, s1;          // segments
    int grid;
};

}  // namespace

struct bielcc_plan_s {
    int kind = 0;        // 0 assembly, 1 field
    int which = 0;
    int precision = BIELCC_PREC_FP64;
    int near_mode = 0;
    bool ck = false;
    NumP np{};
    std::vector<void*> allocs;     // device memory owned by the plan
    // assembly
    std::vector<AssembleArgs> aitems;
    std::vector<const DevMesh*> ameshes;
    AssembleArgs* aitems_d = nullptr;
    // field
    std::vector<bielcc_field_item> fitems;
    std::vector<int> fN;
    std::vector<OpSeg> segs;
    std::vector<SegInfo> seginfo;
    OpSeg* segs_d = nullptr;
    char* ws = nullptr;
    size_t ws_bytes = 0;
    double* part = nullptr;        // split-K partial products (the largest segment's)
    // both
    std::vector<NearSet> near;     // per segment (field) / per item (assembly)
    std::vector<Group> groups;
    int4* tiles_d = nullptr;
    long long ntiles = 0, pairs = 0, near_pairs = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int runs = 0;
    // mixed field plans: FP64 far sums from the run after one in which the FP32 error
    // control recomputed more than mixed_probe_pct percent of the points (the points,
    // meshes and wavenumbers that make FP32 sums fail are the plan's own); a captured
    // graph is captured again at that next run
    bool fp64_sums = false;
    bool recapture = false;
};

namespace {

void plan_free(bielcc_plan p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (void* a : p->allocs) (void)hipFree(a);
    delete p;
}

// plan-owned device allocation
template <typename T>
hipError_t palloc(bielcc_plan p, size_t bytes, T** out) {
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, bytes ? bytes : 1);
    if (e != hipSuccess) return e;
    p->allocs.push_back(d);
    *out = reinterpret_cast<T*>(d);
    return hipSuccess;
}

// Near-pair list of nq query points (stride q_stride doubles) against mesh m,
// in plan memory.  Host-synchronising: plan creation only.
int plan_near(bielcc_ctx ctx, bielcc_plan p, int nq, const double* q, int q_stride, const DevMesh& m, int mode,
              double ac, double thresh, NearSet* out) {
    *out = NearSet{};
    if (nq <= 0) return BIELCC_OK;
    void* tmp = nullptr;
    const size_t cnt_b = align256((size_t)nq * sizeof(int)), off_b = align256((size_t)(nq + 1) * sizeof(long long));
    HIPCHK(hipMalloc(&tmp, cnt_b + off_b + 256));
    int* counts = reinterpret_cast<int*>(tmp);
    long long* offs = reinterpret_cast<long long*>(reinterpret_cast<char*>(tmp) + cnt_b);
    long long* tot_d = reinterpret_cast<long long*>(reinterpret_cast<char*>(tmp) + cnt_b + off_b);
    long long tot = 0;
    const NearFar nf = near_far(m, mode, ac, thresh);
    hipError_t e = launch_near_count(ctx->stream, nq, q, q_stride, m.n, m.ctab, mode, ac, thresh, counts, nf.v);
    if (e == hipSuccess) e = launch_scan(ctx->stream, nq, counts, offs, tot_d);
    if (e == hipSuccess) e = hipMemcpyAsync(&tot, tot_d, sizeof(long long), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && tot > 0) {
        e = palloc(p, (size_t)tot * sizeof(int), &out->lq);
        if (e == hipSuccess) e = palloc(p, (size_t)tot * sizeof(int), &out->lc);
        if (e == hipSuccess)
            e = launch_near_fill(ctx->stream, nq, q, q_stride, m.n, m.ctab, mode, ac, thresh, offs, out->lq, out->lc,
                                 nf.v);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        out->n = tot;
    }
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(ctx, e, "near-pair list");
    return BIELCC_OK;
}

int upload_tiles(bielcc_ctx ctx, bielcc_plan p, const std::vector<int4>& tiles) {
    p->ntiles = (long long)tiles.size();
    HIPCHK(palloc(p, tiles.size() * sizeof(int4), &p->tiles_d));
    // on the context's stream (a null-stream copy is not ordered with it), complete before
    // the host vector goes away
    HIPCHK(hipMemcpyAsync(p->tiles_d, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return BIELCC_OK;
}

int grid_for(bielcc_ctx ctx, int ntiles) {
    return std::max(1, std::min(ntiles, ctx->n_cu * ctx->plan_wg_per_cu));
}

int run_assemble(bielcc_ctx ctx, bielcc_plan p) {
    const Group& g = p->groups[0];
    Timer t(ctx, "plan_assemble_tiles");
    HIPCHK(launch_assemble_multi(ctx->stream, p->aitems_d, p->tiles_d, g.nt, g.grid, p->which, p->ck,
                                 p->precision == BIELCC_PREC_MIXED));
    t.stop();
    for (size_t i = 0; i < p->aitems.size(); i++) {
        const NearSet& ns = p->near[i];
        if (ns.n == 0) continue;
        const AssembleArgs& a = p->aitems[i];
        const DevMesh& m = *p->ameshes[i];
        HIPCHK(launch_assemble_generic(ctx->stream, m.nv, 0, ns.n, ns.lq, ns.lc, m.n, 0, m.n, m.verts, m.ctab, p->np,
                                       a.k_re, a.k_im, a.Z, a.ld));
    }
    return BIELCC_OK;
}

int run_field(bielcc_ctx ctx, bielcc_plan p) {
    for (const bielcc_field_item& it : p->fitems)
        HIPCHK(launch_init_incident(ctx->stream, it.M, it.nrhs, it.pts_d, it.inc_d, it.E_d));
    const bool p32 = p->precision == BIELCC_PREC_MIXED && !p->fp64_sums;
    for (const Group& g : p->groups) {
        {
            Timer t(ctx, "plan_field_tiles");
            HIPCHK(launch_field_opmat(ctx->stream, p->segs_d, p->tiles_d + g.t0, g.nt, g.grid, p->ck, p32));
            t.stop();
        }
        for (int s = g.s0; s < g.s1; s++) {
            const NearSet& ns = p->near[s];
            if (ns.n > 0) HIPCHK(launch_opmat_near(ctx->stream, ns.n, ns.lq, ns.lc, p->segs[s], p->np, p->near_mode));
        }
        Timer tg(ctx, "plan_field_gemm");
        for (int s = g.s0; s < g.s1; s++) {
            const SegInfo& si = p->seginfo[s];
            const bielcc_field_item& it = p->fitems[si.item];
            if (int rc = opmat_gemms(ctx, p->segs[s], it.nrhs, it.jE_d, make_double2(it.multE_re, it.multE_im),
                                     it.E_d + 6 * (size_t)si.p0, it.H_d ? it.H_d + 6 * (size_t)si.p0 : nullptr,
                                     3LL * it.M, p->part, si.parts))
                return rc;
        }
        tg.stop();
    }
    return BIELCC_OK;
}

int run_impl(bielcc_ctx ctx, bielcc_plan p) {
    return p->kind == 0 ? run_assemble(ctx, p) : run_field(ctx, p);
}

// Captures one run of the plan into a graph (warm: one eager run first, so that libraries
// finish their lazy set-up outside the capture).
int capture_impl(bielcc_ctx ctx, bielcc_plan p, bool warm) {
    if (warm) {
        if (int rc = run_impl(ctx, p)) return rc;
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    int rc = run_impl(ctx, p);
    ctx->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(ctx, e, "hipGraphInstantiate");
    }
    p->graph = g;
    p->exec = x;
    return BIELCC_OK;
}

}  // namespace

namespace bielcc_capi {

int opmat_gemms(bielcc_ctx ctx, const OpSeg& sg, int nrhs, const double* jE, double2 mE, double* E, double* H,
                long long ldo, double* part, int nparts) {
    const long long m3 = 3LL * sg.M, n3 = 3LL * sg.N;
    const char* msg = nullptr;
    hipError_t e = hipSuccess;
    if (nparts <= 1) {
        e = blas_zgemm(ctx->stream, ctx->lu, m3, nrhs, n3, mE, sg.WK, m3, jE, n3, make_double2(1., 0.), E, ldo, &msg);
        if (e == hipSuccess && H)
            e = blas_zgemm(ctx->stream, ctx->lu, m3, nrhs, n3, make_double2(1., 0.), sg.WR, m3, jE, n3,
                           make_double2(0., 0.), H, ldo, &msg);
        if (e != hipSuccess) return fail(ctx, BIELCC_ERR_HIP, msg ? msg : hipGetErrorString(e));
        return BIELCC_OK;
    }
    const int cpp = (sg.N + nparts - 1) / nparts;   // cells per part
    const int full = sg.N / cpp, rest = sg.N - full * cpp;
    const long long kc = 3LL * cpp, sc = m3 * nrhs;
    for (int which = 0; which < (H ? 2 : 1); which++) {
        const double* W = which == 0 ? sg.WK : sg.WR;
        const double2 al = which == 0 ? mE : make_double2(1., 0.);
        e = blas_zgemm_strided(ctx->stream, ctx->lu, m3, nrhs, kc, al, W, m3, kc * m3, jE, n3, kc,
                               make_double2(0., 0.), part, m3, sc, full, &msg);
        if (e == hipSuccess && rest > 0)
            e = blas_zgemm(ctx->stream, ctx->lu, m3, nrhs, 3LL * rest, al, W + 2 * (size_t)(full * kc * m3), m3,
                           jE + 2 * (size_t)(full * kc), n3, make_double2(0., 0.), part + 2 * (size_t)(full * sc), m3,
                           &msg);
        if (e == hipSuccess)
            e = launch_sum_partials(ctx->stream, m3, nrhs, full + (rest > 0), part, which == 0 ? E : H, ldo,
                                    which == 0);
        if (e != hipSuccess) return fail(ctx, BIELCC_ERR_HIP, msg ? msg : hipGetErrorString(e));
    }
    return BIELCC_OK;
}

int field_opmat(bielcc_ctx ctx, DevMesh& m, const bielcc_num_param* np, double k_re, double k_im, int nrhs,
                const double* jE_d, double mEr, double mEi, const double* inc_d, int near_mode, int precision, int M,
                const double* pts_d, double* E_d, double* H_d, int* counts, long long* offs, long long* tot_d) {
    const int nn = np->n_start * np->n_start;
    const int N = m.n;
    // point chunks whose matrices fit a 4 GiB (at most plan_ws_mb) workspace
    const size_t per_pt = (H_d ? 2 : 1) * (size_t)9 * 16 * N;
    const size_t cap = std::min((size_t)ctx->plan_ws_mb << 20, (size_t)4 << 30);
    const int chunk = (int)std::max<size_t>(256, cap / per_pt / 256 * 256);
    const int Mc = std::min(M, chunk);
    const size_t wk = align256((size_t)9 * 16 * Mc * N);
    HIPCHK(ensure(ctx->opmat_ws, ctx->opmat_ws_cap, wk * (H_d ? 2 : 1)));
    char* ws = reinterpret_cast<char*>(ctx->opmat_ws);
    std::vector<OpSeg> segs;
    std::vector<int4> tiles;
    std::vector<int> t0;
    const int cpt = ctx->plan_field_cells;
    for (int p0 = 0; p0 < M; p0 += chunk) {
        OpSeg sg;
        sg.M = std::min(chunk, M - p0); sg.N = N; sg.nn = nn; sg.nv = m.nv;
        sg.pts = pts_d + 3 * (size_t)p0; sg.ctab = m.ctab; sg.nodes = m.nodes;
        sg.nodes32 = m.nodes32; sg.nodes32_yy = m.nodes32_yy; sg.verts = m.verts;
        sg.k_re = k_re; sg.k_im = k_im;
        sg.fast_phase = fast_phase_ok(k_re, m.grid_step) && !ctx->no_fast_phase;
        sg.mode = near_mode; sg.ac = np->analytic_calc; sg.thresh = np->analytic_calc * m.grid_step;
        // every chunk reuses the workspace: the stream orders chunk i's GEMMs before chunk i+1's tiles
        sg.WK = reinterpret_cast<double*>(ws);
        sg.WR = H_d ? reinterpret_cast<double*>(ws + align256((size_t)9 * 16 * sg.M * N)) : nullptr;
        t0.push_back((int)tiles.size());
        const int s = (int)segs.size();
        for (int pb = 0; pb < sg.M; pb += 256)
            for (int c0 = 0; c0 < N; c0 += cpt) tiles.push_back(make_int4(s, pb, c0, std::min(N, c0 + cpt)));
        segs.push_back(sg);
    }
    t0.push_back((int)tiles.size());
    const size_t seg_b = align256(segs.size() * sizeof(OpSeg));
    HIPCHK(ensure(ctx->opmat_meta, ctx->opmat_meta_cap, seg_b + tiles.size() * sizeof(int4)));
    OpSeg* segs_d = reinterpret_cast<OpSeg*>(ctx->opmat_meta);
    int4* tiles_d = reinterpret_cast<int4*>(reinterpret_cast<char*>(ctx->opmat_meta) + seg_b);
    HIPCHK(hipMemcpyAsync(segs_d, segs.data(), segs.size() * sizeof(OpSeg), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(tiles_d, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));   // the host vectors go out of scope
    HIPCHK(launch_init_incident(ctx->stream, M, nrhs, pts_d, inc_d, E_d));
    const NumP npp = to_np(np);
    const int grid = grid_for(ctx, t0[1] - t0[0]);
    for (size_t s = 0; s < segs.size(); s++) {
        const OpSeg& sg = segs[s];
        const int p0 = (int)(s * (size_t)chunk);
        long long tot_h = 0;
        int *lq = nullptr, *lc = nullptr;
        // (synchronises: the previous chunk's GEMMs are done before its lists are overwritten)
        if (int rc = build_near_list(ctx, sg.M, sg.pts, 3, m, near_mode, np->analytic_calc, sg.thresh, counts, offs,
                                     tot_d, &tot_h, &lq, &lc))
            return rc;
        {
            Timer t(ctx, "field_opmat");
            HIPCHK(launch_field_opmat(ctx->stream, segs_d, tiles_d + t0[s], t0[s + 1] - t0[s], grid, k_im != 0.,
                                      precision == BIELCC_PREC_MIXED));
            t.stop();
        }
        if (tot_h > 0) {
            Timer t(ctx, "near");
            HIPCHK(launch_opmat_near(ctx->stream, tot_h, lq, lc, sg, npp, near_mode));
            t.stop();
        }
        const int parts = opmat_parts(ctx, sg.M, N);
        if (parts > 1)
            HIPCHK(ensure(ctx->opmat_part, ctx->opmat_part_cap, (size_t)parts * 3 * sg.M * nrhs * 16));
        if (int rc = opmat_gemms(ctx, sg, nrhs, jE_d, make_double2(mEr, mEi), E_d + 6 * (size_t)p0,
                                 H_d ? H_d + 6 * (size_t)p0 : nullptr, 3LL * M,
                                 reinterpret_cast<double*>(ctx->opmat_part), parts))
            return rc;
    }
    return BIELCC_OK;
}

}  // namespace bielcc_capi

extern "C" {

int bielcc_plan_assemble(bielcc_ctx ctx, int nitems, const bielcc_assemble_item* items, int which,
                         const bielcc_num_param* np, int precision, bielcc_plan* out) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (!out || nitems < 1 || !items) return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    *out = nullptr;
    if (int rc = check_np(ctx, np)) return rc;
    if (which != 0 && which != 1) return fail(ctx, BIELCC_ERR_ARG, "which must be 0 (K) or 1 (R)");
    if (precision != BIELCC_PREC_FP64 && precision != BIELCC_PREC_MIXED)
        return fail(ctx, BIELCC_ERR_ARG, "precision must be BIELCC_PREC_FP64 or BIELCC_PREC_MIXED");
    const int nn = np->n_start * np->n_start;
    if (np->p_max != 1 || np->p_max_seg != 1 || nn > 512 || !(np->analytic_calc > 1.))
        return fail(ctx, BIELCC_ERR_UNSUPPORTED,
                    "plans need p_max = p_max_seg = 1, n_start <= 22 and analytic_calc > 1 (use bielcc_assemble)");
    DeviceGuard dg(ctx->device);
    bielcc_plan p = new bielcc_plan_s();
    p->kind = 0;
    p->which = which;
    p->precision = precision;
    p->np = to_np(np);
    std::vector<int4> tiles;
    const int cpt = ctx->plan_asm_cells;
    for (int i = 0; i < nitems; i++) {
        const bielcc_assemble_item& it = items[i];
        DevMesh* m = find_mesh(ctx, it.mesh_id);
        if (!m) { plan_free(p); return fail(ctx, BIELCC_ERR_NOTFOUND, "unknown mesh id in item " + std::to_string(i)); }
        if (!it.Z_d || it.ld < 2LL * m->n || !std::isfinite(it.k_re) || !std::isfinite(it.k_im)) {
            plan_free(p);
            return fail(ctx, BIELCC_ERR_ARG, "item " + std::to_string(i) + ": Z is NULL, ld < 2N or k not finite");
        }
        if (which == 0 && m->nv * np->n_start_seg > 256) {
            plan_free(p);
            return fail(ctx, BIELCC_ERR_UNSUPPORTED, "n_start_seg too large for the tile kernel");
        }
        int rc = ensure_tables(ctx, *m, np->n_start, which == 0 ? np->n_start_seg : 0);
        if (!rc && precision == BIELCC_PREC_MIXED) rc = ensure_tables32(ctx, *m, np->n_start);
        if (rc) { plan_free(p); return rc; }
        AssembleArgs a;
        a.nv = m->nv; a.N = m->n; a.nn = nn; a.ns = np->n_start_seg; a.which = which;
        a.i0 = 0; a.ni = m->n; a.ctab = m->ctab; a.nodes = m->nodes; a.segs = m->segs;
        a.k_re = it.k_re; a.k_im = it.k_im; a.ac = np->analytic_calc; a.Z = it.Z_d; a.ld = it.ld;
        a.fast_phase = fast_phase_ok(it.k_re, m->grid_step) && !ctx->no_fast_phase;
        a.prec = precision; a.nodes32 = m->nodes32; a.nodes32_yy = m->nodes32_yy; a.wg_target = 0;
        p->aitems.push_back(a);
        p->ameshes.push_back(m);
        p->ck = p->ck || it.k_im != 0.;
        for (int bx = 0; bx < (m->n + 255) / 256; bx++)
            for (int c0 = 0; c0 < m->n; c0 += cpt) tiles.push_back(make_int4(i, bx, c0, std::min(m->n, c0 + cpt)));
        p->pairs += (long long)m->n * m->n;
        NearSet ns;
        if (which == 0) {
            if (int rc2 = plan_near(ctx, p, m->n, m->ctab + CT_RKT, CT, *m, 0, np->analytic_calc, 0., &ns)) {
                plan_free(p);
                return rc2;
            }
        }
        p->near.push_back(ns);
        p->near_pairs += ns.n;
    }
    hipError_t e = palloc(p, p->aitems.size() * sizeof(AssembleArgs), &p->aitems_d);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->aitems_d, p->aitems.data(), p->aitems.size() * sizeof(AssembleArgs),
                           hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) { plan_free(p); return hip_fail(ctx, e, "plan upload"); }
    if (int rc = upload_tiles(ctx, p, tiles)) { plan_free(p); return rc; }
    p->groups.push_back(Group{0, (int)tiles.size(), 0, nitems, grid_for(ctx, (int)tiles.size())});
    *out = p;
    return BIELCC_OK;
}

int bielcc_plan_field(bielcc_ctx ctx, int nitems, const bielcc_field_item* items, const bielcc_num_param* np,
                      int near_mode, int precision, bielcc_plan* out) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (!out || nitems < 1 || !items) return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    *out = nullptr;
    if (int rc = check_np(ctx, np)) return rc;
    if (near_mode < 0 || near_mode > 1) return fail(ctx, BIELCC_ERR_ARG, "near_mode must be 0 or 1");
    if (precision != BIELCC_PREC_FP64 && precision != BIELCC_PREC_MIXED)
        return fail(ctx, BIELCC_ERR_ARG, "precision must be BIELCC_PREC_FP64 or BIELCC_PREC_MIXED");
    const int nn = np->n_start * np->n_start;
    if (np->p_max != 1 || nn > 512 || !(np->analytic_calc > 1.))
        return fail(ctx, BIELCC_ERR_UNSUPPORTED, "plans need p_max = 1, n_start <= 22 and analytic_calc > 1");
    DeviceGuard dg(ctx->device);
    bielcc_plan p = new bielcc_plan_s();
    p->kind = 1;
    p->precision = precision;
    p->near_mode = near_mode;
    p->np = to_np(np);
    const size_t cap = (size_t)ctx->plan_ws_mb << 20;
    // segments: items split into point ranges whose operator matrices fit the workspace
    std::vector<const DevMesh*> meshes;
    for (int i = 0; i < nitems; i++) {
        const bielcc_field_item& it = items[i];
        DevMesh* m = find_mesh(ctx, it.mesh_id);
        if (!m) { plan_free(p); return fail(ctx, BIELCC_ERR_NOTFOUND, "unknown mesh id in item " + std::to_string(i)); }
        if (it.nrhs < 1 || it.M < 0 || (it.M > 0 && (!it.jE_d || !it.pts_d || !it.E_d)) ||
            !std::isfinite(it.k_re) || !std::isfinite(it.k_im)) {
            plan_free(p);
            return fail(ctx, BIELCC_ERR_ARG, "item " + std::to_string(i) + ": bad nrhs, M, buffer or wave number");
        }
        int rc = ensure_tables(ctx, *m, np->n_start, 0);
        if (!rc && precision == BIELCC_PREC_MIXED) rc = ensure_tables32(ctx, *m, np->n_start);
        if (rc) { plan_free(p); return rc; }
        meshes.push_back(m);
        p->fitems.push_back(it);
        p->fN.push_back(m->n);
        p->ck = p->ck || it.k_im != 0.;
        const size_t per_pt = (it.H_d ? 2 : 1) * (size_t)9 * 16 * m->n;   // bytes of WK (+WR) per point
        long long chunk = (long long)(cap / per_pt) / 256 * 256;
        if (it.M > 0 && chunk < 256) {
            plan_free(p);
            return fail(ctx, BIELCC_ERR_UNSUPPORTED, "operator-matrix workspace (plan_ws_mb) below 256 points of item " +
                                                         std::to_string(i));
        }
        for (int p0 = 0; p0 < it.M; p0 += (int)std::min<long long>(chunk, it.M)) {
            const int Ms = (int)std::min<long long>(chunk, it.M - p0);
            OpSeg sg;
            sg.M = Ms; sg.N = m->n; sg.nn = nn; sg.nv = m->nv;
            sg.pts = it.pts_d + 3 * (size_t)p0; sg.ctab = m->ctab; sg.nodes = m->nodes;
            sg.nodes32 = m->nodes32; sg.nodes32_yy = m->nodes32_yy; sg.verts = m->verts;
            sg.k_re = it.k_re; sg.k_im = it.k_im;
            sg.fast_phase = fast_phase_ok(it.k_re, m->grid_step) && !ctx->no_fast_phase;
            sg.mode = near_mode; sg.ac = np->analytic_calc; sg.thresh = np->analytic_calc * m->grid_step;
            sg.WK = sg.WR = nullptr;
            p->segs.push_back(sg);
            p->seginfo.push_back(SegInfo{i, p0, Ms, 0, opmat_parts(ctx, Ms, m->n)});
            p->pairs += (long long)Ms * m->n;
        }
    }
    // workspace groups: consecutive segments whose matrices fit the cap together
    std::vector<int4> tiles;
    const int cpt = ctx->plan_field_cells;
    size_t off = 0;
    int s0 = 0;
    auto close_group = [&](int s1) {
        if (s1 <= s0) return;
        const int t0 = (int)tiles.size();
        for (int s = s0; s < s1; s++) {
            const OpSeg& sg = p->segs[s];
            for (int pb = 0; pb < sg.M; pb += 256)
                for (int c0 = 0; c0 < sg.N; c0 += cpt) tiles.push_back(make_int4(s, pb, c0, std::min(sg.N, c0 + cpt)));
        }
        const int nt = (int)tiles.size() - t0;
        p->groups.push_back(Group{t0, nt, s0, s1, grid_for(ctx, nt)});
        p->ws_bytes = std::max(p->ws_bytes, off);
        off = 0;
        s0 = s1;
    };
    for (int s = 0; s < (int)p->segs.size(); s++) {
        const SegInfo& si = p->seginfo[s];
        const size_t wk = align256((size_t)9 * 16 * si.M * p->segs[s].N);
        const size_t need = wk * (p->fitems[si.item].H_d ? 2 : 1);
        if (off + need > cap) close_group(s);
        p->seginfo[s].ws_off = off;
        off += need;
    }
    close_group((int)p->segs.size());
    hipError_t e = hipSuccess;
    if (p->ws_bytes > 0) e = palloc(p, p->ws_bytes, &p->ws);
    size_t part_bytes = 0;
    for (const SegInfo& si : p->seginfo)
        if (si.parts > 1)
            part_bytes = std::max(part_bytes, (size_t)si.parts * 3 * si.M * p->fitems[si.item].nrhs * 16);
    if (e == hipSuccess && part_bytes > 0) e = palloc(p, part_bytes, &p->part);
    if (e != hipSuccess) { plan_free(p); return hip_fail(ctx, e, "operator-matrix workspace"); }
    for (size_t s = 0; s < p->segs.size(); s++) {
        const SegInfo& si = p->seginfo[s];
        const size_t wk = align256((size_t)9 * 16 * si.M * p->segs[s].N);
        p->segs[s].WK = reinterpret_cast<double*>(p->ws + si.ws_off);
        p->segs[s].WR = p->fitems[si.item].H_d ? reinterpret_cast<double*>(p->ws + si.ws_off + wk) : nullptr;
    }
    // near-pair lists per segment
    for (size_t s = 0; s < p->segs.size(); s++) {
        const OpSeg& sg = p->segs[s];
        NearSet ns;
        if (int rc = plan_near(ctx, p, sg.M, sg.pts, 3, *meshes[p->seginfo[s].item], near_mode, sg.ac, sg.thresh, &ns)) {
            plan_free(p);
            return rc;
        }
        p->near.push_back(ns);
        p->near_pairs += ns.n;
    }
    if (!p->segs.empty()) {
        e = palloc(p, p->segs.size() * sizeof(OpSeg), &p->segs_d);
        if (e == hipSuccess)
            e = hipMemcpyAsync(p->segs_d, p->segs.data(), p->segs.size() * sizeof(OpSeg), hipMemcpyHostToDevice,
                               ctx->stream);
        if (e != hipSuccess) { plan_free(p); return hip_fail(ctx, e, "plan upload"); }
        if (int rc = upload_tiles(ctx, p, tiles)) { plan_free(p); return rc; }
    }
    // the BLAS handle (and its workspace) exists before any capture
    const char* msg = nullptr;
    e = blas_zgemm(ctx->stream, ctx->lu, 0, 0, 0, make_double2(0., 0.), nullptr, 1, nullptr, 1, make_double2(0., 0.),
                   nullptr, 1, &msg);
    if (e != hipSuccess) { plan_free(p); return fail(ctx, BIELCC_ERR_HIP, msg ? msg : hipGetErrorString(e)); }
    *out = p;
    return BIELCC_OK;
}

int bielcc_plan_run(bielcc_ctx ctx, bielcc_plan p) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (!p) return fail(ctx, BIELCC_ERR_ARG, "plan is NULL");
    DeviceGuard dg(ctx->device);
    if (p->recapture) {   // the plan changed to FP64 sums after its graph was captured
        p->recapture = false;
        if (int rc = capture_impl(ctx, p, false)) return rc;
    }
    Timer t(ctx, p->kind == 0 ? "plan_assemble" : "plan_field");
    if (p->exec) {
        HIPCHK(hipGraphLaunch(p->exec, ctx->stream));
    } else if (int rc = run_impl(ctx, p)) {
        return rc;
    }
    // mixed field: FP32 error control per item, after the launch or replay (it reads a
    // count back, so it stays out of the captured graph)
    if (p->kind == 1 && p->precision == BIELCC_PREC_MIXED && !p->fp64_sums && ctx->mixed_fixup) {
        const bielcc_num_param np{p->np.eps, p->np.rs, p->np.rs_seg, p->np.analytic_calc, p->np.n_start,
                                  p->np.n_start_seg, p->np.p_max, p->np.p_max_seg};
        const long long fixed0 = ctx->fixup_points;
        long long points = 0;
        for (const bielcc_field_item& it : p->fitems) {
            if (it.M <= 0) continue;
            points += it.M;
            std::vector<double> inc;
            if (it.inc_d) {
                inc.resize((size_t)it.nrhs * 6);
                HIPCHK(hipMemcpyAsync(inc.data(), it.inc_d, inc.size() * sizeof(double), hipMemcpyDeviceToHost,
                                      ctx->stream));
                HIPCHK(hipStreamSynchronize(ctx->stream));
            }
            if (int rc = field_mixed_fixup(ctx, it.mesh_id, &np, it.k_re, it.k_im, it.nrhs, it.jE_d, it.multE_re,
                                           it.multE_im, it.inc_d ? inc.data() : nullptr, it.inc_d, p->near_mode, it.M,
                                           it.pts_d, it.E_d, it.H_d))
                return rc;
        }
        if (ctx->mixed_probe_pct > 0 && (ctx->fixup_points - fixed0) * 100 > (long long)ctx->mixed_probe_pct * points) {
            p->fp64_sums = true;
            if (p->exec) {
                HIPCHK(hipStreamSynchronize(ctx->stream));
                (void)hipGraphExecDestroy(p->exec);
                (void)hipGraphDestroy(p->graph);
                p->exec = nullptr;
                p->graph = nullptr;
                p->recapture = true;
            }
        }
    }
    t.stop();
    p->runs++;
    return BIELCC_OK;
}

int bielcc_plan_capture(bielcc_ctx ctx, bielcc_plan p) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (!p) return fail(ctx, BIELCC_ERR_ARG, "plan is NULL");
    DeviceGuard dg(ctx->device);
    if (p->exec) return BIELCC_OK;
    p->recapture = false;
    return capture_impl(ctx, p, true);
}

int bielcc_plan_destroy(bielcc_ctx ctx, bielcc_plan p) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (!p) return BIELCC_OK;
    DeviceGuard dg(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    plan_free(p);
    return BIELCC_OK;
}

int bielcc_plan_info(bielcc_ctx ctx, bielcc_plan p, const char* key, double* value) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (!p || !key || !value) return fail(ctx, BIELCC_ERR_ARG, "bad argument");
    const std::string k(key);
    if (k == "tiles") *value = (double)p->ntiles;
    else if (k == "launches") *value = (double)p->groups.size();
    else if (k == "grid") *value = p->groups.empty() ? 0. : (double)p->groups[0].grid;
    else if (k == "pairs") *value = (double)p->pairs;
    else if (k == "near_pairs") *value = (double)p->near_pairs;
    else if (k == "segments") *value = (double)(p->kind == 1 ? p->segs.size() : p->aitems.size());
    else if (k == "workspace_bytes") *value = (double)p->ws_bytes;
    else if (k == "captured") *value = (p->exec || p->recapture) ? 1. : 0.;
    else if (k == "fp64_sums") *value = p->fp64_sums ? 1. : 0.;
    else return fail(ctx, BIELCC_ERR_ARG, "unknown plan info key " + k);
    return BIELCC_OK;
}

}  // extern "C"
