// Dense complex FP64 LU with partial pivoting and the multi-RHS solve.
//
// The reference links LAPACK/BLAS but never calls them (CMakeLists.txt:18-19,56),
// so this is the library's own solver.  Right-looking blocked LU, panel width
// kNB, column-major:
//   panel  : per column, an argmax of |re|+|im| over the remaining rows (the
//            pivot rule of LAPACK's izamax, first index on ties), the row swap
//            inside the panel, scaling by the reciprocal pivot and the rank-1
//            update of the rest of the panel (own kernels);
//   laswp  : the panel's row interchanges applied to the other columns;
//   trsm   : U12 = L11^{-1} A12 with L11 and a 64-column tile of A12 in LDS;
//   update : A22 -= L21 U12, a plain ZGEMM (rocBLAS, FP64 matrix cores).
// The solve applies the composed row permutation, then blocked forward and
// back substitution (LDS triangular tiles + ZGEMM for the off-diagonal blocks).
#include "bielcc_internal.h"

#include <rocblas/rocblas.h>
#include <vector>

using namespace bielcc_dev;

namespace bielcc_impl {

constexpr int kNB = 64;
constexpr int kPivBlocks = 256;

struct LuWork {
    rocblas_handle blas = nullptr;
    double* pval = nullptr;     // [pcap] partial maxima
    int* pidx = nullptr;        // [pcap] partial argmax
    int pcap = 0;
    double* pinv = nullptr;     // reciprocal pivot (re, im) + flag
    int* info = nullptr;        // first zero pivot, 1-based (0 if none)
    int* perm = nullptr;        // composed permutation for the solve
    size_t perm_cap = 0;
};

LuWork* lu_work_create() { return new LuWork(); }

void lu_work_destroy(LuWork* w) {
    if (!w) return;
    if (w->blas) rocblas_destroy_handle(w->blas);
    (void)hipFree(w->pval); (void)hipFree(w->pidx); (void)hipFree(w->pinv); (void)hipFree(w->info);
    (void)hipFree(w->perm);
    delete w;
}

static hipError_t lu_work_init(LuWork* w, const char** msg, int n = 0) {
    hipError_t e;
    const int need = std::max(kPivBlocks, (n + 255) / 256 + 1);
    if (w->pcap < need) {
        (void)hipFree(w->pval);
        (void)hipFree(w->pidx);
        w->pval = nullptr;
        w->pidx = nullptr;
        w->pcap = 0;
        if ((e = hipMalloc(&w->pval, need * sizeof(double))) != hipSuccess) return e;
        if ((e = hipMalloc(&w->pidx, need * sizeof(int))) != hipSuccess) return e;
        w->pcap = need;
    }
    if (!w->pinv) {
        if ((e = hipMalloc(&w->pinv, 4 * sizeof(double))) != hipSuccess) return e;
        if ((e = hipMalloc(&w->info, sizeof(int))) != hipSuccess) return e;
    }
    if (!w->blas) {
        if (rocblas_create_handle(&w->blas) != rocblas_status_success) {
            *msg = "rocblas_create_handle failed";
            return hipErrorUnknown;
        }
    }
    return hipSuccess;
}

__device__ inline double cabs1(double2 v) { return fabs(v.x) + fabs(v.y); }

// Phase 1: partial argmax of |re|+|im| over rows j..n-1 of column j.
__global__ __launch_bounds__(256) void k_piv_partial(int n, int j, const double2* __restrict__ A, long long lda,
                                                     double* __restrict__ pval, int* __restrict__ pidx) {
    __shared__ double sv[256];
    __shared__ int si[256];
    const double2* col = A + (size_t)j * lda;
    double best = -1.;
    int bi = 0x7fffffff;
    for (int i = j + blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double v = cabs1(col[i]);
        if (v > best) { best = v; bi = i; }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            double v = sv[threadIdx.x + off];
            int ix = si[threadIdx.x + off];
            if (v > sv[threadIdx.x] || (v == sv[threadIdx.x] && ix < si[threadIdx.x])) {
                sv[threadIdx.x] = v;
                si[threadIdx.x] = ix;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { pval[blockIdx.x] = sv[0]; pidx[blockIdx.x] = si[0]; }
}

// Phase 2: final argmax, record the pivot, swap rows j and p inside the
// panel columns [c0, c1), store the reciprocal pivot.
__global__ __launch_bounds__(256) void k_piv_final(int nparts, int j, int c0, int c1, double2* __restrict__ A,
                                                   long long lda, const double* __restrict__ pval,
                                                   const int* __restrict__ pidx, int* __restrict__ piv,
                                                   double* __restrict__ pinv, int* __restrict__ info) {
    __shared__ double sv[256];
    __shared__ int si[256];
    __shared__ int sp;
    double best = -1.;
    int bi = 0x7fffffff;
    for (int t = threadIdx.x; t < nparts; t += 256) {
        double v = pval[t];
        int ix = pidx[t];
        if (v > best || (v == best && ix < bi)) { best = v; bi = ix; }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            double v = sv[threadIdx.x + off];
            int ix = si[threadIdx.x + off];
            if (v > sv[threadIdx.x] || (v == sv[threadIdx.x] && ix < si[threadIdx.x])) {
                sv[threadIdx.x] = v;
                si[threadIdx.x] = ix;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int p = (si[0] == 0x7fffffff) ? j : si[0];
        piv[j] = p;
        sp = p;
        if (!(sv[0] > 0.) && *info == 0) *info = j + 1;
    }
    __syncthreads();
    const int p = sp;
    if (p != j) {
        for (int c = c0 + threadIdx.x; c < c1; c += 256) {
            double2* colp = A + (size_t)c * lda;
            double2 t = colp[j];
            colp[j] = colp[p];
            colp[p] = t;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 a = A[(size_t)j * lda + j];
        double den = a.x * a.x + a.y * a.y;
        if (den > 0.) { pinv[0] = a.x / den; pinv[1] = -a.y / den; pinv[2] = 1.; }
        else { pinv[0] = 0.; pinv[1] = 0.; pinv[2] = 0.; }
    }
}

// Scale column j below the diagonal and apply the rank-1 update to panel
// columns j+1 .. c1-1 (one lane per row; row j staged in LDS).  When column
// j+1 is in the panel, each workgroup also reduces the updated |re|+|im| of
// its rows of column j+1 to a partial argmax, so the next column's pivot
// search needs only the final reduction.
__global__ __launch_bounds__(256) void k_panel_update(int n, int j, int c1, double2* __restrict__ A, long long lda,
                                                      const double* __restrict__ pinv, double* __restrict__ pval,
                                                      int* __restrict__ pidx) {
    __shared__ double2 urow[kNB];
    __shared__ double sv[256];
    __shared__ int si[256];
    const int nc = c1 - j - 1;
    for (int t = threadIdx.x; t < nc; t += 256) urow[t] = A[(size_t)(j + 1 + t) * lda + j];
    __syncthreads();
    const int i = j + 1 + blockIdx.x * 256 + threadIdx.x;
    double best = -1.;
    int bi = 0x7fffffff;
    if (i < n) {
        double2 l = A[(size_t)j * lda + i];
        if (pinv[2] != 0.) {
            double2 s = make_double2(l.x * pinv[0] - l.y * pinv[1], l.x * pinv[1] + l.y * pinv[0]);
            l = s;
            A[(size_t)j * lda + i] = l;
        }
        for (int t = 0; t < nc; t++) {
            double2* p = A + (size_t)(j + 1 + t) * lda + i;
            double2 u = urow[t];
            double2 a = *p;
            a.x -= l.x * u.x - l.y * u.y;
            a.y -= l.x * u.y + l.y * u.x;
            *p = a;
            if (t == 0) { best = cabs1(a); bi = i; }
        }
    }
    if (nc <= 0) return;   // uniform: no next column inside the panel
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            double v = sv[threadIdx.x + off];
            int ix = si[threadIdx.x + off];
            if (v > sv[threadIdx.x] || (v == sv[threadIdx.x] && ix < si[threadIdx.x])) {
                sv[threadIdx.x] = v;
                si[threadIdx.x] = ix;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { pval[blockIdx.x] = sv[0]; pidx[blockIdx.x] = si[0]; }
}

// Apply the interchanges piv[k0 .. k0+nb) to columns [cbeg, cend).
__global__ __launch_bounds__(256) void k_laswp(int k0, int nb, int cbeg, int cend, double2* __restrict__ A,
                                               long long lda, const int* __restrict__ piv) {
    const int c = cbeg + blockIdx.x * 256 + threadIdx.x;
    if (c >= cend) return;
    double2* col = A + (size_t)c * lda;
    for (int t = 0; t < nb; t++) {
        int r = k0 + t, p = piv[r];
        if (p != r) {
            double2 v = col[r];
            col[r] = col[p];
            col[p] = v;
        }
    }
}

// X = T^{-1} X for the nb x nb triangle T at (t0, t0) of A (lower unit when
// lower != 0, else upper non-unit), X = columns [cbeg, cend) of B rows t0..t0+nb.
// One workgroup of 256 lanes per 64 columns; T and the X tile live in LDS.
// Column sweep: lane (r = tid % 64, column group tid / 64) owns row r of
// every fourth column; step q publishes x_q and every lane updates its rows.
__global__ __launch_bounds__(256) void k_trsm(int nb, int lower, const double2* __restrict__ A, long long lda,
                                              long long t0, double2* __restrict__ B, long long ldb, long long brow,
                                              int cbeg, int cend) {
    __shared__ double2 T[kNB][kNB + 1];
    __shared__ double2 X[kNB][kNB + 1];
    __shared__ double2 dinv[kNB];
    const int tid = threadIdx.x;
    const int r = tid & 63, cg = tid >> 6;
    for (int q = cg; q < nb; q += 4)
        if (r < nb) T[r][q] = A[(size_t)(t0 + q) * lda + t0 + r];
    const int c0 = cbeg + blockIdx.x * 64;
    const int ncol = min(64, cend - c0);
    for (int cc = cg; cc < ncol; cc += 4)
        if (r < nb) X[r][cc] = B[(size_t)(c0 + cc) * ldb + brow + r];
    __syncthreads();
    if (!lower && tid < nb) {
        double2 d = T[tid][tid];
        double den = d.x * d.x + d.y * d.y;
        dinv[tid] = make_double2(d.x / den, -d.y / den);
    }
    __syncthreads();
    if (lower) {
        for (int q = 0; q < nb - 1; q++) {
            if (r > q && r < nb) {
                const double2 l = T[r][q];
                for (int cc = cg; cc < ncol; cc += 4) {
                    const double2 x = X[q][cc];
                    double2 a = X[r][cc];
                    a.x -= l.x * x.x - l.y * x.y;
                    a.y -= l.x * x.y + l.y * x.x;
                    X[r][cc] = a;
                }
            }
            __syncthreads();
        }
    } else {
        // X[q] is final once step q begins; rows < q are updated with x_q = X[q] / T[q][q]
        for (int q = nb - 1; q > 0; q--) {
            if (r < q) {
                const double2 u = T[r][q], di = dinv[q];
                for (int cc = cg; cc < ncol; cc += 4) {
                    const double2 b = X[q][cc];
                    const double2 x = make_double2(b.x * di.x - b.y * di.y, b.x * di.y + b.y * di.x);
                    double2 a = X[r][cc];
                    a.x -= u.x * x.x - u.y * x.y;
                    a.y -= u.x * x.y + u.y * x.x;
                    X[r][cc] = a;
                }
            }
            __syncthreads();
        }
        if (r < nb) {
            const double2 di = dinv[r];
            for (int cc = cg; cc < ncol; cc += 4) {
                const double2 b = X[r][cc];
                X[r][cc] = make_double2(b.x * di.x - b.y * di.y, b.x * di.y + b.y * di.x);
            }
        }
        __syncthreads();
    }
    for (int cc = cg; cc < ncol; cc += 4)
        if (r < nb) B[(size_t)(c0 + cc) * ldb + brow + r] = X[r][cc];
}

static hipError_t zgemm_update(LuWork* w, hipStream_t s, long long m, long long n, long long k, const double2* Aa,
                               long long lda, const double2* Bb, long long ldb, double2* Cc, long long ldc,
                               const char** msg) {
    if (m <= 0 || n <= 0 || k <= 0) return hipSuccess;
    rocblas_set_stream(w->blas, s);
    rocblas_double_complex alpha = {-1., 0.}, beta = {1., 0.};
    rocblas_status st = rocblas_zgemm_64(w->blas, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha,
                                         reinterpret_cast<const rocblas_double_complex*>(Aa), lda,
                                         reinterpret_cast<const rocblas_double_complex*>(Bb), ldb, &beta,
                                         reinterpret_cast<rocblas_double_complex*>(Cc), ldc);
    if (st != rocblas_status_success) {
        *msg = "rocblas_zgemm_64 failed";
        return hipErrorUnknown;
    }
    return hipGetLastError();
}

hipError_t lu_factor(hipStream_t s, LuWork* w, int n, double* Ad, long long lda, int* piv, int* info,
                     const char** msg) {
    hipError_t e = lu_work_init(w, msg, n);
    if (e != hipSuccess) return e;
    double2* A = reinterpret_cast<double2*>(Ad);
    if ((e = hipMemsetAsync(w->info, 0, sizeof(int), s)) != hipSuccess) return e;
    for (int k0 = 0; k0 < n; k0 += kNB) {
        const int nb = min(kNB, n - k0);
        const int c1 = k0 + nb;
        int nparts = 0;
        for (int jj = 0; jj < nb; jj++) {
            const int j = k0 + jj;
            const int m = n - j;
            if (jj == 0) {   // first panel column: fresh search (later ones come from k_panel_update)
                nparts = min(kPivBlocks, (m + 255) / 256);
                hipLaunchKernelGGL(k_piv_partial, dim3(nparts), dim3(256), 0, s, n, j, A, lda, w->pval, w->pidx);
            }
            hipLaunchKernelGGL(k_piv_final, dim3(1), dim3(256), 0, s, nparts, j, k0, c1, A, lda, w->pval, w->pidx,
                               piv, w->pinv, w->info);
            if (m > 1) {
                nparts = (m - 1 + 255) / 256;
                hipLaunchKernelGGL(k_panel_update, dim3(nparts), dim3(256), 0, s, n, j, c1, A, lda, w->pinv,
                                   w->pval, w->pidx);
            }
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (k0 > 0) hipLaunchKernelGGL(k_laswp, dim3((k0 + 255) / 256), dim3(256), 0, s, k0, nb, 0, k0, A, lda, piv);
        if (c1 < n) {
            hipLaunchKernelGGL(k_laswp, dim3((n - c1 + 255) / 256), dim3(256), 0, s, k0, nb, c1, n, A, lda, piv);
            hipLaunchKernelGGL(k_trsm, dim3((n - c1 + 63) / 64), dim3(256), 0, s, nb, 1, A, lda, (long long)k0, A,
                               lda, (long long)k0, c1, n);
            e = zgemm_update(w, s, n - c1, n - c1, nb, A + (size_t)k0 * lda + c1, lda, A + (size_t)c1 * lda + k0,
                             lda, A + (size_t)c1 * lda + c1, lda, msg);
            if (e != hipSuccess) return e;
        }
    }
    if ((e = hipMemcpyAsync(info, w->info, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

// B(i, :) <- B(perm(i), :) with perm the composition of the interchanges
__global__ void k_gather_rows(int n, int nrhs, const int* __restrict__ perm, const double2* __restrict__ src,
                              double2* __restrict__ B, long long ldb) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * nrhs) return;
    int i = (int)(idx % n), r = (int)(idx / n);
    B[(size_t)r * ldb + i] = src[(size_t)r * n + perm[i]];
}

__global__ void k_copy_cols(int n, int nrhs, const double2* __restrict__ B, long long ldb, double2* __restrict__ dst) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * nrhs) return;
    int i = (int)(idx % n), r = (int)(idx / n);
    dst[(size_t)r * n + i] = B[(size_t)r * ldb + i];
}

hipError_t lu_solve(hipStream_t s, LuWork* w, int n, const double* Ad, long long lda, const int* piv, int nrhs,
                    double* Bd, long long ldb, const char** msg) {
    hipError_t e = lu_work_init(w, msg);
    if (e != hipSuccess) return e;
    const double2* A = reinterpret_cast<const double2*>(Ad);
    double2* B = reinterpret_cast<double2*>(Bd);
    // compose the interchanges into a gather permutation (host, O(n))
    std::vector<int> hp(n), perm(n);
    if ((e = hipMemcpyAsync(hp.data(), piv, n * sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = 0; i < n; i++) {
        int p = hp[i];
        if (p < 0 || p >= n) { *msg = "pivot index out of range"; return hipErrorInvalidValue; }
        int t = perm[i]; perm[i] = perm[p]; perm[p] = t;
    }
    const size_t perm_bytes = ((size_t)n * sizeof(int) + 255) / 256 * 256;
    const size_t need = perm_bytes + (size_t)n * nrhs * sizeof(double2);
    if (w->perm_cap < need) {
        (void)hipFree(w->perm);
        w->perm = nullptr;
        w->perm_cap = 0;
        if ((e = hipMalloc(&w->perm, need)) != hipSuccess) return e;
        w->perm_cap = need;
    }
    int* dperm = w->perm;
    double2* tmp = reinterpret_cast<double2*>(reinterpret_cast<char*>(w->perm) + perm_bytes);
    if ((e = hipMemcpyAsync(dperm, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    long long tot = (long long)n * nrhs;
    hipLaunchKernelGGL(k_copy_cols, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, nrhs, B, ldb, tmp);
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, nrhs, dperm, tmp, B, ldb);
    // forward: L y = P b
    for (int k0 = 0; k0 < n; k0 += kNB) {
        const int nb = min(kNB, n - k0);
        hipLaunchKernelGGL(k_trsm, dim3((nrhs + 63) / 64), dim3(256), 0, s, nb, 1, A, lda, (long long)k0, B, ldb,
                           (long long)k0, 0, nrhs);
        if (k0 + nb < n) {
            e = zgemm_update(w, s, n - k0 - nb, nrhs, nb, A + (size_t)k0 * lda + k0 + nb, lda, B + k0, ldb,
                             B + k0 + nb, ldb, msg);
            if (e != hipSuccess) return e;
        }
    }
    // back: U x = y
    int last = ((n - 1) / kNB) * kNB;
    for (int k0 = last; k0 >= 0; k0 -= kNB) {
        const int nb = min(kNB, n - k0);
        hipLaunchKernelGGL(k_trsm, dim3((nrhs + 63) / 64), dim3(256), 0, s, nb, 0, A, lda, (long long)k0, B, ldb,
                           (long long)k0, 0, nrhs);
        if (k0 > 0) {
            e = zgemm_update(w, s, k0, nrhs, nb, A + (size_t)k0 * lda, lda, B + k0, ldb, B, ldb, msg);
            if (e != hipSuccess) return e;
        }
    }
    return hipGetLastError();
}

}  // namespace bielcc_impl
