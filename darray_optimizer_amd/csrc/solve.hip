// Solves with the regularised core of the RBF kernel, A X = B, A SPD.
//
// dh_posv_f64: cholesky_factorize (lapack_like.h:9-56) then the two triangular
// solves U^T Z = B (triangular_solve "L","U","T", blas_like.h:95-128) and
// U X = Z (the L,U,N case the reference leaves unimplemented).
//
// dh_posv_mixed: the same factorisation in FP32 on a rounded copy of A, used
// inside FP64 iterative refinement (the scheme of LAPACK's dsposv):
//   X0 = A32^{-1} B;  repeat: R = B - A X (FP64, A symmetric from its upper
//   triangle), stop when max_j ||r_j|| / (||A||_F ||x_j|| + ||b_j||) <= tol,
//   else X += A32^{-1} R.
// If the FP32 factorisation breaks down, the call falls back to dh_posv_f64 on
// a copy of A and reports iterations = -1 (dsposv's convention).
#include <algorithm>
#include <cmath>
#include <vector>

#include "dh_common.h"

namespace {

__global__ __launch_bounds__(256) void d2s_upper_kernel(int n, const double* __restrict__ A, int lda,
                                                        float* __restrict__ S, int lds) {
    size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)n * n) return;
    int i = (int)(idx % n), j = (int)(idx / n);
    if (i <= j) S[i + (size_t)j * lds] = (float)A[i + (size_t)j * lda];
}

__global__ __launch_bounds__(256) void d2s_kernel(int m, int n, const double* __restrict__ A,
                                                  int lda, float* __restrict__ S, int lds) {
    size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)m * n) return;
    int i = (int)(idx % m), j = (int)(idx / m);
    S[i + (size_t)j * lds] = (float)A[i + (size_t)j * lda];
}

// X = S (accumulate == 0) or X += S.
__global__ __launch_bounds__(256) void s2d_kernel(int m, int n, const float* __restrict__ S, int lds,
                                                  double* __restrict__ X, int ldx, int accumulate) {
    size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)m * n) return;
    int i = (int)(idx % m), j = (int)(idx / m);
    double v = (double)S[i + (size_t)j * lds];
    double* x = X + i + (size_t)j * ldx;
    *x = accumulate ? *x + v : v;
}

__global__ __launch_bounds__(256) void copy_kernel(int m, int n, const double* __restrict__ A,
                                                   int lda, double* __restrict__ B, int ldb) {
    size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)m * n) return;
    int i = (int)(idx % m), j = (int)(idx / m);
    B[i + (size_t)j * ldb] = A[i + (size_t)j * lda];
}

// out[j] = ||A[:, j]||_2^2, one workgroup per column.
__global__ __launch_bounds__(256) void colnorm2_kernel(int m, const double* __restrict__ A, int lda,
                                                       double* __restrict__ out) {
    __shared__ double red[4];
    const double* a = A + (size_t)blockIdx.x * lda;
    double s = 0.0;
    for (int i = threadIdx.x; i < m; i += 256) s += a[i] * a[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// ||A||_F^2 of the symmetric matrix stored in the upper triangle: partial sums
// per workgroup (off-diagonal entries counted twice).
__global__ __launch_bounds__(256) void fro2_sym_upper_kernel(int n, const double* __restrict__ A,
                                                             int lda, double* __restrict__ partial) {
    __shared__ double red[4];
    double s = 0.0;
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < (size_t)n * n;
         idx += (size_t)gridDim.x * 256) {
        int i = (int)(idx % n), j = (int)(idx / n);
        if (i <= j) {
            double v = A[i + (size_t)j * lda];
            s += (i == j) ? v * v : 2.0 * v * v;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void sum_kernel(int count, const double* __restrict__ partial, double* __restrict__ out) {
    double s = 0.0;
    for (int b = 0; b < count; ++b) s += partial[b];
    *out = s;
}

int launch_fro2_sym(dh_handle_t h, int n, const double* A, int lda, double* scratch, double* out) {
    int nb = 512;
    fro2_sym_upper_kernel<<<nb, 256, 0, h->stream>>>(n, A, lda, scratch);
    DH_CHECK_LAUNCH();
    sum_kernel<<<1, 1, 0, h->stream>>>(nb, scratch, out);
    DH_CHECK_LAUNCH();
    return 0;
}

template <typename T>
int two_solves(dh_handle_t h, int n, int nrhs, const T* U, int ldu, T* B, int ldb) {
    DH_CHECK_STATUS(dh_trsm_impl<T>(h, 'L', 'T', false, n, nrhs, T(1), U, ldu, B, ldb));
    return dh_trsm_impl<T>(h, 'L', 'N', false, n, nrhs, T(1), U, ldu, B, ldb);
}

}  // namespace

extern "C" int dh_posv_f64(dh_handle_t h, char uplo, int n, int nrhs, double* A, int lda,
                           double* B, int ldb, int bs) {
    const char* name = "dh_posv_f64";
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (uplo != 'U' && uplo != 'u')
        return dh_fail(DH_ERR_UNSUPPORTED, "%s: only uplo='U' is implemented", name);
    if (n < 0) return dh_fail(-3, "%s: n < 0", name);
    if (nrhs < 0) return dh_fail(-4, "%s: nrhs < 0", name);
    if (lda < std::max(1, n)) return dh_fail(-6, "%s: lda < n", name);
    if (ldb < std::max(1, n)) return dh_fail(-8, "%s: ldb < n", name);
    int info = 0;
    DH_CHECK_STATUS(dh_potrf_impl<double>(h, n, A, lda, bs, &info));
    if (info) return dh_fail(info, "%s: leading minor of order %d is not positive definite", name,
                             info);
    if (nrhs == 0) return 0;
    DH_CHECK_STATUS(two_solves<double>(h, n, nrhs, A, lda, B, ldb));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int dh_posv_mixed(dh_handle_t h, char uplo, int n, int nrhs, const double* A, int lda,
                             const double* B, int ldb, double* X, int ldx, int bs, double tol,
                             int max_iter, int* iters, double* berr) {
    const char* name = "dh_posv_mixed";
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (uplo != 'U' && uplo != 'u')
        return dh_fail(DH_ERR_UNSUPPORTED, "%s: only uplo='U' is implemented", name);
    if (n < 0) return dh_fail(-3, "%s: n < 0", name);
    if (nrhs < 0) return dh_fail(-4, "%s: nrhs < 0", name);
    if (lda < std::max(1, n)) return dh_fail(-6, "%s: lda < n", name);
    if (ldb < std::max(1, n)) return dh_fail(-8, "%s: ldb < n", name);
    if (ldx < std::max(1, n)) return dh_fail(-10, "%s: ldx < n", name);
    if (max_iter < 0) return dh_fail(-13, "%s: max_iter < 0", name);
    if (iters) *iters = 0;
    if (berr) *berr = 0.0;
    if (n == 0 || nrhs == 0) return 0;
    const size_t nn = (size_t)n * n, nr = (size_t)n * nrhs;
    float* S = (float*)dh_ws(h, DH_WS_SOLVE_A, sizeof(float) * nn);
    if (!S) return DH_ERR_ALLOC;
    // R (FP64 residual), Rf (FP32 right-hand side), norms and scalars.
    const size_t rf_off = sizeof(double) * nr;
    const size_t nrm_off = (rf_off + sizeof(float) * nr + 255) & ~size_t(255);
    char* base = (char*)dh_ws(h, DH_WS_SOLVE_R, nrm_off + sizeof(double) * (3 * nrhs + 1024));
    if (!base) return DH_ERR_ALLOC;
    double* R = (double*)base;
    float* Rf = (float*)(base + rf_off);
    double* nrm = (double*)(base + nrm_off);  // b2[nrhs], x2[nrhs], r2[nrhs], scratch...
    double *b2 = nrm, *x2 = nrm + nrhs, *r2 = nrm + 2 * nrhs, *scr = nrm + 3 * nrhs;

    DH_CHECK_STATUS(launch_fro2_sym(h, n, A, lda, scr, scr + 1000));
    colnorm2_kernel<<<nrhs, 256, 0, h->stream>>>(n, B, ldb, b2);
    DH_CHECK_LAUNCH();
    std::vector<double> hn(3 * nrhs + 1);
    // Normwise backward error max_j ||b_j - A x_j|| / (||A||_F ||x_j|| + ||b_j||) of X,
    // leaving the residual in R; +inf when any column's error is not finite.
    auto backward_error = [&](double* out) -> int {
        copy_kernel<<<dh_cdiv((long)nr, 256), 256, 0, h->stream>>>(n, nrhs, B, ldb, R, n);
        DH_CHECK_LAUNCH();
        DH_CHECK_STATUS(dh_symm_upper_impl<double>(h, n, nrhs, -1.0, A, lda, X, ldx, 1.0, R, n));
        colnorm2_kernel<<<nrhs, 256, 0, h->stream>>>(n, X, ldx, x2);
        DH_CHECK_LAUNCH();
        colnorm2_kernel<<<nrhs, 256, 0, h->stream>>>(n, R, n, r2);
        DH_CHECK_LAUNCH();
        DH_CHECK_HIP(hipMemcpyAsync(hn.data(), nrm, sizeof(double) * 3 * nrhs,
                                    hipMemcpyDeviceToHost, h->stream));
        double anorm2 = 0.0;
        DH_CHECK_HIP(hipMemcpyAsync(&anorm2, scr + 1000, sizeof(double), hipMemcpyDeviceToHost,
                                    h->stream));
        DH_CHECK_HIP(hipStreamSynchronize(h->stream));
        const double anorm = std::sqrt(anorm2);
        double be = 0.0;
        for (int j = 0; j < nrhs; ++j) {
            const double den = anorm * std::sqrt(hn[nrhs + j]) + std::sqrt(hn[j]);
            const double rn = std::sqrt(hn[2 * nrhs + j]);
            // A NaN anywhere (residual, X or B) must not compare as converged.
            double e = den > 0.0 ? rn / den : (rn == 0.0 ? 0.0 : INFINITY);
            if (!std::isfinite(e)) e = INFINITY;
            be = std::max(be, e);
        }
        *out = be;
        return 0;
    };
    // FP64 Cholesky solve on a copy of A, as LAPACK dsposv falls back when the
    // FP32 factorisation breaks down or the refinement does not converge.
    auto fp64_solve = [&]() -> int {
        double* A64 = (double*)dh_ws(h, DH_WS_SOLVE_A, sizeof(double) * nn);
        if (!A64) return DH_ERR_ALLOC;
        copy_kernel<<<dh_cdiv((long)nn, 256), 256, 0, h->stream>>>(n, n, A, lda, A64, n);
        DH_CHECK_LAUNCH();
        copy_kernel<<<dh_cdiv((long)nr, 256), 256, 0, h->stream>>>(n, nrhs, B, ldb, X, ldx);
        DH_CHECK_LAUNCH();
        int info64 = 0;
        DH_CHECK_STATUS(dh_potrf_impl<double>(h, n, A64, n, bs, &info64));
        if (info64)
            return dh_fail(info64, "%s: matrix not positive definite (minor %d)", name, info64);
        DH_CHECK_STATUS(two_solves<double>(h, n, nrhs, A64, n, X, ldx));
        double be = 0.0;
        DH_CHECK_STATUS(backward_error(&be));
        if (iters) *iters = -1;
        if (berr) *berr = be;
        if (!(be <= tol) && !(be <= 1e3 * n * 1.1102230246251565e-16))
            return dh_fail(DH_ERR_NOT_CONVERGED,
                           "%s: FP64 fallback solve has backward error %.3e (tol %.3e)", name, be,
                           tol);
        return 0;
    };

    d2s_upper_kernel<<<dh_cdiv((long)nn, 256), 256, 0, h->stream>>>(n, A, lda, S, n);
    DH_CHECK_LAUNCH();
    int info = 0;
    DH_CHECK_STATUS(dh_potrf_impl<float>(h, n, S, n, bs, &info));
    if (info) return fp64_solve();  // FP32 factorisation broke down
    // X0 = A32^{-1} B
    d2s_kernel<<<dh_cdiv((long)nr, 256), 256, 0, h->stream>>>(n, nrhs, B, ldb, Rf, n);
    DH_CHECK_LAUNCH();
    DH_CHECK_STATUS(two_solves<float>(h, n, nrhs, S, n, Rf, n));
    s2d_kernel<<<dh_cdiv((long)nr, 256), 256, 0, h->stream>>>(n, nrhs, Rf, n, X, ldx, 0);
    DH_CHECK_LAUNCH();
    double be = INFINITY, prev = INFINITY;
    int it = 0;
    for (;; ++it) {
        DH_CHECK_STATUS(backward_error(&be));
        if (be <= tol) break;
        // Not converged within max_iter corrections, diverged (non-finite), or
        // stalled (a correction that does not halve the error: kappa * u32
        // near 1): solve in FP64 instead.
        if (it >= max_iter || !std::isfinite(be) || (it > 0 && be > 0.5 * prev))
            return fp64_solve();
        prev = be;
        d2s_kernel<<<dh_cdiv((long)nr, 256), 256, 0, h->stream>>>(n, nrhs, R, n, Rf, n);
        DH_CHECK_LAUNCH();
        DH_CHECK_STATUS(two_solves<float>(h, n, nrhs, S, n, Rf, n));
        s2d_kernel<<<dh_cdiv((long)nr, 256), 256, 0, h->stream>>>(n, nrhs, Rf, n, X, ldx, 1);
        DH_CHECK_LAUNCH();
    }
    if (iters) *iters = it;
    if (berr) *berr = be;
    return 0;
}
