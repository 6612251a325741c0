// bielcc_field_noalloc: the Stratton-Chu field driver of bielcc_field with all
// of its work memory (cell-split partial sums, near-pair counts, offsets and
// lists, per-pair near results, the incident wave) in one file-scope device
// array, so that a call allocates nothing on the device.  For callers that
// cannot allocate (tight serving loops, code that forbids hipMalloc).
//
// The array exists once per GPU and lives until the process ends.  Host code
// never uses its name as an address: at context creation, once per device and
// with that device current, hipGetSymbolAddress gives the device's own address,
// kept per device and passed to the kernels like any work buffer.  Every call
// on a GPU shares that GPU's array: calls that overlap on one GPU (two
// contexts, two streams) race.  Nothing in it is accumulated into: every slot a
// call reads was written earlier in the same call.
#include "ctx_internal.h"
#include <mutex>

using namespace bielcc_impl;
using namespace bielcc_capi;

namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB per GPU
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) char scratch[kScratchBytes];

namespace {
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};   // per-device address of `scratch`
}  // namespace

namespace bielcc_capi {
// Called by bielcc_ctx_create with the context's device current.
int noalloc_setup(bielcc_ctx ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(ctx, BIELCC_ERR_UNSUPPORTED, "device id above 63");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[ctx->device]) return BIELCC_OK;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(ctx, BIELCC_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch[ctx->device] = p;
    return BIELCC_OK;
}
}  // namespace bielcc_capi

extern "C" {

int bielcc_field_noalloc(bielcc_ctx ctx, int mesh_id, const bielcc_num_param* np, double k_re, double k_im,
                         const double* jE_d, const double* jM_d, double multE_re, double multE_im, double multH_re,
                         double multH_im, const double* kvec, const double* e0, int near_mode, int M,
                         const double* pts_d, double* E_d, double* H_d) {
    if (!ctx) return BIELCC_ERR_ARG;
    if (int rc = check_np(ctx, np)) return rc;
    DevMesh* m = find_mesh(ctx, mesh_id);
    if (!m) return fail(ctx, BIELCC_ERR_NOTFOUND, "unknown mesh id");
    if (M < 0 || near_mode < 0 || near_mode > 1) return fail(ctx, BIELCC_ERR_ARG, "bad M or near_mode");
    if ((kvec == nullptr) != (e0 == nullptr)) return fail(ctx, BIELCC_ERR_ARG, "kvec and e0 go together");
    if (M == 0) return BIELCC_OK;
    if (!jE_d || !pts_d || !E_d) return fail(ctx, BIELCC_ERR_ARG, "NULL buffer");
    if (!std::isfinite(k_re) || !std::isfinite(k_im)) return fail(ctx, BIELCC_ERR_ARG, "non-finite wave number");
    const int nn = np->n_start * np->n_start;
    if (np->p_max != 1 || nn > 512 || !(np->analytic_calc > 1.))
        return fail(ctx, BIELCC_ERR_UNSUPPORTED,
                    "field driver needs p_max = 1, n_start <= 22 and analytic_calc > 1 (use bielcc_op_batch)");
    if (m->nodes_n != np->n_start)
        return fail(ctx, BIELCC_ERR_UNSUPPORTED,
                    "the mesh's quadrature nodes for this n_start are not built: they are built (allocated) by "
                    "one bielcc_field or bielcc_assemble call on the mesh");
    char* base;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = static_cast<char*>(ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr);
    }
    if (!base) return fail(ctx, BIELCC_ERR_HIP, "no scratch address for this device");
    DeviceGuard g(ctx->device);
    const int nacc = jM_d ? 24 : 12;
    const double thresh = np->analytic_calc * m->grid_step;
    // the cell split of bielcc_field (same sums, same order), fewer splits if the
    // partial sums would not leave half of the scratch for the near pairs
    const int bx = (M + 255) / 256;
    FieldArgs a;
    a.nv = m->nv; a.M = M; a.N = m->n; a.nn = nn; a.pts = pts_d; a.ctab = m->ctab; a.nodes = m->nodes;
    a.nodes32 = nullptr; a.nodes32_yy = nullptr;
    a.jE = jE_d; a.jM = jM_d; a.nrhs = 1; a.rchunk = ctx->field_rchunk;
    a.k_re = k_re; a.k_im = k_im; a.mode = near_mode; a.ac = np->analytic_calc;
    a.thresh = thresh;
    a.fast_phase = ctx->no_fast_phase ? 0 : phase_mode(k_re, m->grid_step, m->max_rad);
    a.prec = BIELCC_PREC_FP64;
    int nsplit = field_far_splits(a, std::max(1, std::min((ctx->field_wg + bx - 1) / bx, (m->n + 63) / 64)));
    const size_t cnt_b = align256((size_t)M * sizeof(int));
    const size_t off_b = align256((size_t)(M + 1) * sizeof(long long));
    const size_t inc_b = 256, tot_b = 256;
    auto part_bytes = [&](int s) { return align256((size_t)s * nacc * M * sizeof(double)); };
    while (nsplit > 1 && part_bytes(nsplit) + cnt_b + off_b + inc_b + tot_b > kScratchBytes / 2) nsplit--;
    const size_t part_b = part_bytes(nsplit);
    const size_t fixed = part_b + cnt_b + off_b + inc_b + tot_b;
    if (fixed > kScratchBytes)
        return fail(ctx, BIELCC_ERR_UNSUPPORTED,
                    "M = " + std::to_string(M) + " points need " + std::to_string(fixed) +
                        " bytes of work memory, beyond the 64 MiB of bielcc_field_noalloc (use bielcc_field)");
    double* part = reinterpret_cast<double*>(base);
    int* counts = reinterpret_cast<int*>(base + part_b);
    long long* offs = reinterpret_cast<long long*>(base + part_b + cnt_b);
    double* inc_d = reinterpret_cast<double*>(base + part_b + cnt_b + off_b);
    long long* tot_d = reinterpret_cast<long long*>(base + part_b + cnt_b + off_b + inc_b);
    double ke[6] = {0, 0, 0, 0, 0, 0};
    if (kvec) {
        for (int d = 0; d < 3; d++) { ke[d] = kvec[d]; ke[3 + d] = e0[d]; }
        HIPCHK(hipMemcpyAsync(inc_d, ke, sizeof ke, hipMemcpyHostToDevice, ctx->stream));
    }
    a.nsplit = nsplit; a.part = part;
    HIPCHK(launch_field_far(ctx->stream, a));
    // near pairs: count, scan, then the lists and results in the rest of the scratch
    const NearFar nf = near_far(*m, near_mode, np->analytic_calc, thresh);
    HIPCHK(launch_near_count(ctx->stream, M, pts_d, 3, m->n, m->ctab, near_mode, np->analytic_calc, thresh, counts,
                             nf.v));
    HIPCHK(launch_scan(ctx->stream, M, counts, offs, tot_d));
    long long T = 0;
    HIPCHK(hipMemcpyAsync(&T, tot_d, sizeof T, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int *lq = nullptr, *lc = nullptr;
    double* near_out = nullptr;
    if (T > 0) {
        const size_t l_b = align256((size_t)T * sizeof(int));
        const size_t need = fixed + 2 * l_b + (size_t)T * nacc * sizeof(double);
        if (need > kScratchBytes)
            return fail(ctx, BIELCC_ERR_UNSUPPORTED,
                        std::to_string(T) + " near pairs need " + std::to_string(need) +
                            " bytes of work memory, beyond the 64 MiB of bielcc_field_noalloc (use bielcc_field)");
        lq = reinterpret_cast<int*>(base + fixed);
        lc = reinterpret_cast<int*>(base + fixed + l_b);
        near_out = reinterpret_cast<double*>(base + fixed + 2 * l_b);
        HIPCHK(launch_near_fill(ctx->stream, M, pts_d, 3, m->n, m->ctab, near_mode, np->analytic_calc, thresh, offs,
                                lq, lc, nf.v));
        HIPCHK(launch_field_near(ctx->stream, m->nv, T, lq, lc, pts_d, m->verts, jE_d, jM_d, to_np(np), k_re, k_im,
                                 near_mode, 0, near_out));
    }
    FinalizeArgs f;
    f.M = M; f.N = m->n; f.nsplit = nsplit; f.nrhs = 1; f.part = part; f.jE = jE_d; f.jM = jM_d;
    f.offs = offs; f.lc = lc; f.near_out = near_out; f.near_mat = 0;
    f.mEr = multE_re; f.mEi = multE_im; f.mHr = multH_re; f.mHi = multH_im;
    f.inc = kvec ? inc_d : nullptr; f.pts = pts_d; f.E = E_d; f.H = H_d;
    HIPCHK(launch_field_finalize(ctx->stream, f));
    return BIELCC_OK;
}

}  // extern "C"
