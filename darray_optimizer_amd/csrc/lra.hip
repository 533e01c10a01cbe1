// CholeskyQR2 and the randomised range finder (DArray::LRA + the
// orthogonalisation and power iterations around it).
//
// Reference pipeline (testsvm.cpp:86-122): LRA (K*Omega, dkernel.h:12-46) ->
// tall_skinny_qr_factorize (Householder TSQR, lapack_like.h:281-379) -> LRA
// again for K*Q.  Here the orthogonalisation is CholeskyQR2 (Gram on MFMA,
// small blocked Cholesky, right triangular solve, twice) and the K-passes
// never materialise K.
#include <utility>

#include "dh_common.h"

int dh_project_f64_split(dh_handle_t h, int m, int n, int d, int l, const double* Xi, int ldxi,
                        const double* Xj, int ldxj, const double* ni, const double* nj,
                        const double* yi, const double* yj, double gamma, const double* O,
                        int ldo, double beta, double* Y, int ldy);
int dh_lra_bf16_passes(dh_handle_t h, int n, int d, int l, int q, const double* X, int ldx,
                       const double* y, double gamma, const double* Omega, int ldo, double* Q,
                       int ldq, double* W, int ldw);
int dh_lra_f16_passes(dh_handle_t h, int n, int d, int l, int q, const double* X, int ldx,
                      const double* y, double gamma, const double* Omega, int ldo, double* Q,
                      int ldq, double* W, int ldw);
int dh_lra_fp8_passes(dh_handle_t h, int n, int d, int l, int q, const double* X, int ldx,
                      const double* y, double gamma, const double* Omega, int ldo, double* Q,
                      int ldq, double* W, int ldw);

// dst (n x n, ld n) = triu(src); strict lower set to zero.
__global__ __launch_bounds__(256) void triu_copy_kernel(int n, const double* __restrict__ src,
                                                         int lds, double* __restrict__ dst) {
    size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)n * n) return;
    int i = (int)(idx % n), j = (int)(idx / n);
    dst[idx] = (i <= j) ? src[i + (size_t)j * lds] : 0.0;
}

// Device scalars of the CholeskyQR checks live after the potrf status word.
constexpr size_t kCqDmaxOffset = 128;  // bytes into DH_WS_INFO: double max diag(G)

// dmax = max_k G[k][k]; with shift != 0 also G += s*I where
// s = 11 (m n + n (n+1)) u trace(G) -- the shift of shifted CholeskyQR3
// (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa, SIAM J. Sci. Comput.
// 2020) with ||Y||_2^2 bounded by ||Y||_F^2 = trace(G).
__global__ __launch_bounds__(256) void gram_diag_kernel(int m, int n, double* __restrict__ G,
                                                         int ldg, double* __restrict__ dmax,
                                                         int shift) {
    __shared__ double smax[256], ssum[256];
    double mx = 0.0, sm = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) {
        double v = G[k + (size_t)k * ldg];
        mx = v > mx ? v : mx;
        sm += v;
    }
    smax[threadIdx.x] = mx;
    ssum[threadIdx.x] = sm;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + st]);
            ssum[threadIdx.x] += ssum[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (shift) {
        double s = 11.0 * ((double)m * n + (double)n * (n + 1)) * 1.1102230246251565e-16 * ssum[0];
        for (int k = threadIdx.x; k < n; k += 256) G[k + (size_t)k * ldg] += s;
    }
    if (threadIdx.x == 0) *dmax = smax[0];
}

// Numerical breakdown test after a successful potrf: a pivot with
// U_kk^2 <= n u max_j G_jj means Y is rank deficient to working precision
// (kappa(Y) beyond ~u^{-1/2}); record the first such k (1-based) in *info.
__global__ __launch_bounds__(256) void cholqr_check_kernel(int n, const double* __restrict__ U,
                                                            int ldu,
                                                            const double* __restrict__ dmax,
                                                            int* __restrict__ info) {
    double tol = (double)n * 2.220446049250313e-16 * (*dmax);
    for (int k = threadIdx.x; k < n; k += 256) {
        double u = U[k + (size_t)k * ldu];
        if (!(u * u > tol)) atomicMin(info, k + 1);
    }
}

__global__ void reset_check_kernel(int* info) { *info = 0x7fffffff; }

// One CholeskyQR sweep on Y (m x n): G = upper(Y^T Y) (SYRK split over m),
// optional shift, G = U^T U, breakdown check, Y := Y U^{-1}.  On return
// *info is 0 or the 1-based index of the first failing pivot; U stays in G.
static int cholqr_sweep(dh_handle_t h, int m, int n, double* Y, int ldy, double* G, bool shift,
                        int* info) {
    char* ws = (char*)dh_ws(h, DH_WS_INFO, 256);
    if (!ws) return DH_ERR_ALLOC;
    double* dmax = (double*)(ws + kCqDmaxOffset);
    int* chk = (int*)(ws + kCqDmaxOffset + 64);
    DH_CHECK_STATUS(dh_gemm_impl<double>(h, true, false, n, n, m, 1.0, Y, ldy, Y, ldy, 0.0, G, n, 1));
    gram_diag_kernel<<<1, 256, 0, h->stream>>>(m, n, G, n, dmax, shift ? 1 : 0);
    DH_CHECK_LAUNCH();
    DH_CHECK_STATUS(dh_potrf_impl<double>(h, n, G, n, 64, info));
    if (*info) return 0;
    reset_check_kernel<<<1, 1, 0, h->stream>>>(chk);
    DH_CHECK_LAUNCH();
    cholqr_check_kernel<<<1, 256, 0, h->stream>>>(n, G, n, dmax, chk);
    DH_CHECK_LAUNCH();
    int hchk = 0;
    DH_CHECK_HIP(hipMemcpyAsync(&hchk, chk, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    if (hchk != 0x7fffffff) {
        *info = hchk;
        return 0;
    }
    return dh_trsm_impl<double>(h, 'R', 'N', false, m, n, 1.0, G, n, Y, ldy);
}

int dh_chol_small_max_order();
int dh_gram_small_launch(dh_handle_t h, int m, int n, const double* Y, int ldy, double* G,
                         int ldg);

// G = upper(Y^T Y): the register-blocked Gram kernel where it applies
// (128 < n <= 272), else the tiled GEMM.
static int cholqr_gram(dh_handle_t h, int m, int n, const double* Y, int ldy, double* G) {
    const int st = dh_gram_small_launch(h, m, n, Y, ldy, G, n);
    if (st != 1) return st;
    return dh_gemm_impl<double>(h, true, false, n, n, m, 1.0, Y, ldy, Y, ldy, 0.0, G, n, 1);
}
int dh_chol_small_launch(dh_handle_t h, int n, double* G, int ldg, double* Dinv, int m_rows,
                         bool shift, bool rel_check, int* info_dev);
int dh_trsm_rows_launch(dh_handle_t h, int m, int n, double* Y, int ldy, const double* U,
                        int ldu, const double* Dinv, const int* guard, int nguard);

// One CholeskyQR sweep in place for orders <= 544: G = upper(Y^T Y) (SYRK),
// one-workgroup Cholesky with the relative breakdown test, Y := Y U^{-1} by
// the row-parallel MFMA triangular solve.
static int cholqr_sweep_small(dh_handle_t h, int m, int n, double* Y, int ldy, double* G,
                              double* Dinv, bool shift, int* info) {
    int* dinfo = (int*)dh_ws(h, DH_WS_INFO, 256);
    if (!dinfo) return DH_ERR_ALLOC;
    DH_CHECK_STATUS(cholqr_gram(h, m, n, Y, ldy, G));
    DH_CHECK_STATUS(dh_chol_small_launch(h, n, G, n, Dinv, m, shift, true, dinfo));
    DH_CHECK_HIP(hipMemcpyAsync(info, dinfo, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    if (*info) return 0;
    return dh_trsm_rows_launch(h, m, n, Y, ldy, G, n, Dinv, nullptr, 0);
}

// The same sweep with no host synchronisation: the Cholesky status goes to
// dinfo[slot] and the solve is skipped on the device unless dinfo[0 .. slot]
// are all zero, so a failed sweep leaves Y untouched.
static int cholqr_sweep_small_async(dh_handle_t h, int m, int n, double* Y, int ldy, double* G,
                                    double* Dinv, int* dinfo, int slot) {
    DH_CHECK_STATUS(cholqr_gram(h, m, n, Y, ldy, G));
    DH_CHECK_STATUS(dh_chol_small_launch(h, n, G, n, Dinv, m, false, true, dinfo + slot));
    return dh_trsm_rows_launch(h, m, n, Y, ldy, G, n, Dinv, dinfo, slot + 1);
}

// R accumulation across sweeps: Racc = U_sweep * Racc (both upper).
static int accumulate_r(dh_handle_t h, int n, const double* G, double* Racc, double* Rtmp,
                        bool first, double* out, int ldo) {
    size_t nn = (size_t)n * n;
    if (first) {
        triu_copy_kernel<<<dh_cdiv((long)nn, 256), 256, 0, h->stream>>>(n, G, n, Racc);
        DH_CHECK_LAUNCH();
        if (out)
            DH_CHECK_HIP(hipMemcpy2DAsync(out, sizeof(double) * ldo, Racc, sizeof(double) * n,
                                          sizeof(double) * n, n, hipMemcpyDeviceToDevice,
                                          h->stream));
        return 0;
    }
    triu_copy_kernel<<<dh_cdiv((long)nn, 256), 256, 0, h->stream>>>(n, G, n, Rtmp);
    DH_CHECK_LAUNCH();
    double* t2 = Rtmp + nn;
    DH_CHECK_STATUS(
        dh_gemm_impl<double>(h, false, false, n, n, n, 1.0, Rtmp, n, Racc, n, 0.0, t2, n, 0));
    DH_CHECK_HIP(hipMemcpyAsync(Racc, t2, sizeof(double) * nn, hipMemcpyDeviceToDevice,
                                h->stream));
    if (out)
        DH_CHECK_HIP(hipMemcpy2DAsync(out, sizeof(double) * ldo, Racc, sizeof(double) * n,
                                      sizeof(double) * n, n, hipMemcpyDeviceToDevice, h->stream));
    return 0;
}

// CholeskyQR2 (shifted CholeskyQR3 if the first Gram matrix is numerically
// singular).  Orders up to dh_chol_small_max_order() use the one-workgroup
// Cholesky and the row-parallel solve; larger orders use the blocked potrf
// and the recursive TRSM.
int dh_cholqr2_impl(dh_handle_t h, int m, int n, double* Y, int ldy, double* R, int ldr) {
    if (n == 0) return 0;
    size_t nn = (size_t)n * n;
    if (n > dh_chol_small_max_order()) {
        double* G = (double*)dh_ws(h, DH_WS_GRAM, sizeof(double) * nn * 4);
        if (!G) return DH_ERR_ALLOC;
        double *Racc = G + nn, *Rtmp = G + 2 * nn;
        int info = 0, sweeps = 2;
        DH_CHECK_STATUS(cholqr_sweep(h, m, n, Y, ldy, G, false, &info));
        if (info) {
            DH_CHECK_STATUS(cholqr_sweep(h, m, n, Y, ldy, G, true, &info));
            if (info)
                return dh_fail(info, "CholeskyQR: shifted Gram matrix not positive definite "
                                     "(pivot %d)", info);
            sweeps = 3;
        }
        if (R) DH_CHECK_STATUS(accumulate_r(h, n, G, Racc, Rtmp, true, sweeps == 1 ? R : nullptr, ldr));
        for (int sw = 2; sw <= sweeps; ++sw) {
            DH_CHECK_STATUS(cholqr_sweep(h, m, n, Y, ldy, G, false, &info));
            if (info)
                return dh_fail(1000000 * (sw - 1) + info,
                               "CholeskyQR: Gram matrix of sweep %d is numerically singular "
                               "(pivot %d)", sw, info);
            if (R)
                DH_CHECK_STATUS(accumulate_r(h, n, G, Racc, Rtmp, false,
                                             sw == sweeps ? R : nullptr, ldr));
        }
        return 0;
    }
    double* G = (double*)dh_ws(h, DH_WS_GRAM, sizeof(double) * nn * 5);
    if (!G) return DH_ERR_ALLOC;
    double *Racc = G + nn, *Rtmp = G + 2 * nn;  // Rtmp spans 2 nn
    double* Dbuf = (double*)dh_ws(h, DH_WS_RBUF, sizeof(double) * 256 * (size_t)dh_cdiv(n, 16));
    if (!Dbuf) return DH_ERR_ALLOC;
    // Both sweeps queued back to back, one host synchronisation for the two
    // statuses.
    int* dinfo = (int*)dh_ws(h, DH_WS_INFO, 256);
    if (!dinfo) return DH_ERR_ALLOC;
    DH_CHECK_STATUS(cholqr_sweep_small_async(h, m, n, Y, ldy, G, Dbuf, dinfo, 0));
    if (R) DH_CHECK_STATUS(accumulate_r(h, n, G, Racc, Rtmp, true, nullptr, ldr));
    DH_CHECK_STATUS(cholqr_sweep_small_async(h, m, n, Y, ldy, G, Dbuf, dinfo, 1));
    if (R) DH_CHECK_STATUS(accumulate_r(h, n, G, Racc, Rtmp, false, R, ldr));
    int hbuf[2] = {0, 0};
    int* pinned = dh_host_status(h);
    int* hinfo = pinned ? pinned : hbuf;
    DH_CHECK_HIP(hipMemcpyAsync(hinfo, dinfo, 2 * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    if (hinfo[0] == 0) {
        if (hinfo[1] == 0) return 0;
        return dh_fail(1000000 + hinfo[1],
                       "CholeskyQR: Gram matrix of sweep 2 is numerically singular (pivot %d)",
                       hinfo[1]);
    }
    int info = 0, sweeps = 3;
    {
        // The first sweep broke down and both solves were skipped, so Y is
        // untouched: restart with the shifted Gram (shifted CholeskyQR3).
        DH_CHECK_STATUS(cholqr_sweep_small(h, m, n, Y, ldy, G, Dbuf, true, &info));
        if (info)
            return dh_fail(info, "CholeskyQR: shifted Gram matrix not positive definite (pivot %d)",
                           info);
    }
    if (R) DH_CHECK_STATUS(accumulate_r(h, n, G, Racc, Rtmp, true, nullptr, ldr));
    for (int sw = 2; sw <= sweeps; ++sw) {
        DH_CHECK_STATUS(cholqr_sweep_small(h, m, n, Y, ldy, G, Dbuf, false, &info));
        if (info)
            return dh_fail(1000000 * (sw - 1) + info,
                           "CholeskyQR: Gram matrix of sweep %d is numerically singular (pivot %d)",
                           sw, info);
        if (R)
            DH_CHECK_STATUS(
                accumulate_r(h, n, G, Racc, Rtmp, false, sw == sweeps ? R : nullptr, ldr));
    }
    return 0;
}

// col[i] = uniform pseudo-random in [-1, 1) from a counter hash (splitmix64).
__global__ __launch_bounds__(256) void random_column_kernel(int m, double* __restrict__ col,
                                                            uint64_t seed) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        col[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
}

// Orthonormal basis of the range finder.  CholeskyQR2 (shifted CholeskyQR3
// fallback) fails for Y of rank < n -- a kernel of rank < l, e.g. fewer
// distinct points than basis columns -- where the reference's Householder
// TSQR (lapack_like.h:281-379) would still return orthonormal columns.  Then
// the column at the failing pivot is replaced by a pseudo-random vector and
// the orthogonalisation repeated: the basis spans range(Y) plus random
// directions, which can only enlarge the subspace the range finder keeps.
static int replace_column(dh_handle_t h, int m, double* Y, int ldy, int attempt, int piv) {
    uint64_t seed = 0xD1B54A32D192ED03ull * (uint64_t)(attempt + 1) + (uint64_t)piv;
    random_column_kernel<<<max(1, min(dh_cdiv(m, 256), 1024)), 256, 0, h->stream>>>(
        m, Y + (size_t)(piv - 1) * ldy, seed);
    DH_CHECK_LAUNCH();
    return 0;
}

static int range_orth_from(dh_handle_t h, int m, int n, double* Y, int ldy, int first) {
    for (int attempt = first; attempt <= n; ++attempt) {
        int st = dh_cholqr2_impl(h, m, n, Y, ldy, nullptr, 0);
        if (st <= 0) return st;
        int piv = st % 1000000;
        if (piv < 1 || piv > n) return st;
        DH_CHECK_STATUS(replace_column(h, m, Y, ldy, attempt, piv));
    }
    return dh_fail(n, "range finder: basis of %d columns could not be completed", n);
}

int dh_range_orth(dh_handle_t h, int m, int n, double* Y, int ldy) {
    return range_orth_from(h, m, n, Y, ldy, 0);
}

// dh_range_orth with the status check deferred, so that the caller can queue
// the next K-pass before the host waits: both CholeskyQR sweeps and the copy of
// their statuses to pinned host memory are queued, then an event.  Orders
// above dh_chol_small_max_order() (or no pinned memory) take the synchronous
// path and leave nothing pending.
int dh_range_orth_begin(dh_handle_t h, int m, int n, double* Y, int ldy, int* pending) {
    *pending = 0;
    int* pinned = dh_host_status(h);
    if (n == 0 || n > dh_chol_small_max_order() || !pinned) return dh_range_orth(h, m, n, Y, ldy);
    if (!h->ev_orth) DH_CHECK_HIP(hipEventCreateWithFlags(&h->ev_orth, hipEventDisableTiming));
    const size_t nn = (size_t)n * n;
    double* G = (double*)dh_ws(h, DH_WS_GRAM, sizeof(double) * nn * 5);
    if (!G) return DH_ERR_ALLOC;
    double* Dbuf = (double*)dh_ws(h, DH_WS_RBUF, sizeof(double) * 256 * (size_t)dh_cdiv(n, 16));
    if (!Dbuf) return DH_ERR_ALLOC;
    int* dinfo = (int*)dh_ws(h, DH_WS_INFO, 256);
    if (!dinfo) return DH_ERR_ALLOC;
    DH_CHECK_STATUS(cholqr_sweep_small_async(h, m, n, Y, ldy, G, Dbuf, dinfo, 0));
    DH_CHECK_STATUS(cholqr_sweep_small_async(h, m, n, Y, ldy, G, Dbuf, dinfo, 1));
    DH_CHECK_HIP(hipMemcpyAsync(pinned + 2, dinfo, 2 * sizeof(int), hipMemcpyDeviceToHost,
                                h->stream));
    DH_CHECK_HIP(hipEventRecord(h->ev_orth, h->stream));
    *pending = 1;
    return 0;
}

// Completes dh_range_orth_begin: waits for the statuses only.  After a
// breakdown it waits for the stream and finishes on the synchronous path as
// dh_range_orth would have (a failed first sweep left Y untouched; a failed
// second sweep left the first sweep's Q, as after attempt 0), and sets
// *redone: work queued after the begin read a Y that has changed since.
int dh_range_orth_finish(dh_handle_t h, int m, int n, double* Y, int ldy, int* pending,
                         int* redone) {
    *redone = 0;
    if (!*pending) return 0;
    *pending = 0;
    DH_CHECK_HIP(hipEventSynchronize(h->ev_orth));
    const int* hinfo = dh_host_status(h) + 2;
    const int s0 = hinfo[0], s1 = hinfo[1];
    if (s0 == 0 && s1 == 0) return 0;
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    *redone = 1;
    if (s0 != 0) return range_orth_from(h, m, n, Y, ldy, 0);
    DH_CHECK_STATUS(replace_column(h, m, Y, ldy, 0, s1));
    return range_orth_from(h, m, n, Y, ldy, 1);
}

extern "C" int dh_cholqr2_f64(dh_handle_t h, int m, int n, double* Y, int ldy, double* R,
                              int ldr) {
    const char* name = "dh_cholqr2_f64";
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (m < 0) return dh_fail(-2, "%s: m < 0", name);
    if (n < 0 || n > m) return dh_fail(-3, "%s: need 0 <= n <= m", name);
    if (ldy < max(1, m)) return dh_fail(-5, "%s: ldy < m", name);
    if (R && ldr < max(1, n)) return dh_fail(-7, "%s: ldr < n", name);
    int st = dh_cholqr2_impl(h, m, n, Y, ldy, R, ldr);
    if (st) return st;
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

static int copy_panel(dh_handle_t h, int n, int l, const double* src, int lds, double* dst,
                      int ldd) {
    DH_CHECK_HIP(hipMemcpy2DAsync(dst, sizeof(double) * (size_t)ldd, src,
                                  sizeof(double) * (size_t)lds, sizeof(double) * (size_t)n, l,
                                  hipMemcpyDeviceToDevice, h->stream));
    return 0;
}

// FP64 passes: Y = K*Omega -> Q; q x { K*Q -> orthonormal basis }; W = K*Q.
static int lra_f64(dh_handle_t h, int n, int d, int l, int q, const double* X, int ldx,
                   const double* y, double gamma, const double* Omega, int ldo, double* Q,
                   int ldq, double* W, int ldw) {
    double* nrm = (double*)dh_ws(h, DH_WS_NORMS, sizeof(double) * (size_t)n);
    if (!nrm) return DH_ERR_ALLOC;
    DH_CHECK_STATUS(dh_sqnorms_f64(h, n, d, X, ldx, nrm));
    DH_CHECK_STATUS(dh_project_f64_split(h, n, n, d, l, X, ldx, X, ldx, nrm, nrm, y, y, gamma,
                                        Omega, ldo, 0.0, Q, ldq));
    DH_CHECK_STATUS(dh_range_orth(h, n, l, Q, ldq));
    // Power iterations alternate the basis between Q and W (no copy for even q).
    double *cur = Q, *nxt = W;
    int ldc = ldq, ldn = ldw;
    for (int it = 0; it < q; ++it) {
        DH_CHECK_STATUS(dh_project_f64_split(h, n, n, d, l, X, ldx, X, ldx, nrm, nrm, y, y, gamma,
                                            cur, ldc, 0.0, nxt, ldn));
        DH_CHECK_STATUS(dh_range_orth(h, n, l, nxt, ldn));
        std::swap(cur, nxt);
        std::swap(ldc, ldn);
    }
    if (cur != Q) DH_CHECK_STATUS(copy_panel(h, n, l, cur, ldc, Q, ldq));
    return dh_project_f64_split(h, n, n, d, l, X, ldx, X, ldx, nrm, nrm, y, y, gamma, Q, ldq, 0.0,
                               W, ldw);
}

extern "C" int dh_lra(dh_handle_t h, int prec, int n, int d, int l, int q, const double* X,
                      int ldx, const double* y, double gamma, const double* Omega, int ldo,
                      double* Q, int ldq, double* W, int ldw) {
    const char* name = "dh_lra";
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (prec < DH_PREC_F64 || prec > DH_PREC_FP8) return dh_fail(-2, "%s: bad prec %d", name, prec);
    if (n < 0) return dh_fail(-3, "%s: n < 0", name);
    if (d < 1) return dh_fail(-4, "%s: d < 1", name);
    if (l < 0 || l > n) return dh_fail(-5, "%s: need 0 <= l <= n", name);
    if (q < 0) return dh_fail(-6, "%s: q < 0", name);
    if (ldx < d) return dh_fail(-8, "%s: ldx < d", name);
    if (ldo < max(1, n)) return dh_fail(-12, "%s: ldo < n", name);
    if (ldq < max(1, n)) return dh_fail(-14, "%s: ldq < n", name);
    if (ldw < max(1, n)) return dh_fail(-16, "%s: ldw < n", name);
    if (n == 0 || l == 0) return 0;
    // The low-precision kernels hold at most 256 features: above that every
    // precision runs the FP64 passes (materialised K row blocks).
    if (d > 256) prec = DH_PREC_F64;
    int st;
    if (prec == DH_PREC_F64)
        st = lra_f64(h, n, d, l, q, X, ldx, y, gamma, Omega, ldo, Q, ldq, W, ldw);
    else if (prec == DH_PREC_BF16)
        st = dh_lra_bf16_passes(h, n, d, l, q, X, ldx, y, gamma, Omega, ldo, Q, ldq, W, ldw);
    else if (prec == DH_PREC_F16)
        st = dh_lra_f16_passes(h, n, d, l, q, X, ldx, y, gamma, Omega, ldo, Q, ldq, W, ldw);
    else
        st = dh_lra_fp8_passes(h, n, d, l, q, X, ldx, y, gamma, Omega, ldo, Q, ldq, W, ldw);
    if (st) return st;
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    return 0;
}
