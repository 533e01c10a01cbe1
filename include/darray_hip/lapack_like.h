// DArray's Cholesky factorisations on the GPU, plus the solves and the
// CholeskyQR2 of the low-rank path.
//
// cholesky_factorize (lapack_like.h:9-56) and cholesky_factorize_recursive
// (lapack_like.h:58-90) factor the upper triangle of A[gi:gi+n, gi:gi+n] in
// place on A's device copy.  Failure handling is the reference's: a
// non-positive pivot prints "<s|d>potrf error code <info> at i= <block start>.
// Stopping" and returns with A partially factored (lapack_like.h:25-28);
// last_potrf_info() reports the global 1-based order of the failing pivot
// (0 on success).
#pragma once

#include <vector>

#include "blas_like.h"

namespace DArray {

inline int& last_potrf_info() {
    static thread_local int info = 0;
    return info;
}

namespace detail {
inline int potrf(int n, double* A, int lda, int bs) {
    return dh_potrf_f64(handle(), 'U', n, A, lda, bs);
}
inline int potrf(int n, float* A, int lda, int bs) {
    return dh_potrf_f32(handle(), 'U', n, A, lda, bs);
}
template <typename T>
constexpr const char* potrf_name() {
    return std::is_same<T, float>::value ? "spotrf" : "dpotrf";
}
inline void require_upper(char uplo, const char* fn) {
    if (uplo != 'U' && uplo != 'u') {
        std::fprintf(stderr, "%s: only uplo 'U' is implemented\n", fn);
        std::exit(1);
    }
}
}  // namespace detail

// Right-looking blocked Cholesky, block bs: potrf(A11) -> trsm(L,U,T) on the
// block row -> SYRK of the trailing matrix.  The strict lower triangle is
// left as it was.
template <typename T>
void cholesky_factorize(DMatrix<T>& A, char uplo, int n, int gi, int bs = 256) {
    detail::require_upper(uplo, "cholesky_factorize");
    last_potrf_info() = 0;
    if (n <= 0) return;
    int info = detail::potrf(n, A.device_at(gi, gi), A.lld(), bs);
    if (info < 0) check(info, "cholesky_factorize");
    if (info > 0) {
        int blk = (info - 1) / bs * bs;
        std::printf("%s error code %d at i= %d. Stopping\n", detail::potrf_name<T>(),
                    info - blk, gi + blk);
        last_potrf_info() = info;
    }
}

namespace detail {
template <typename T>
void chol_recursive(DMatrix<T>& A, int n, int i, int bs) {
    if (n <= 0 || last_potrf_info() != 0) return;
    if (n <= bs) {
        int info = potrf(n, A.device_at(i, i), A.lld(), std::max(n, 1));
        if (info < 0) check(info, "cholesky_factorize_recursive");
        if (info > 0) {
            std::printf("%s error code %d at i= %d. Stopping\n", potrf_name<T>(), info, i);
            last_potrf_info() = i + info;
        }
        return;
    }
    const int n1 = n / 2;
    chol_recursive(A, n1, i, bs);
    if (last_potrf_info() != 0) return;
    DMatrixDescriptor<T> G11{A, i, i, i + n1, i + n1};
    DMatrixDescriptor<T> G12{A, i, i + n1, i + n1, i + n};
    DMatrixDescriptor<T> A22{A, i + n1, i + n1, i + n, i + n};
    triangular_solve("L", "U", "T", "N", n1, n - n1, T(1), G11, G12, bs);
    matrix_multiply("T", "N", G12, G12, A22, T(-1), T(1));
    chol_recursive(A, n - n1, i + n1, bs);
}
}  // namespace detail

// Recursive Cholesky: factor the leading half, solve U11^T A12 = A12
// (triangular_solve L,U,T), update the trailing block A22 -= A12^T A12 as a
// full GEMM (matrix_multiply T,N, so the strict lower triangle of A22 takes
// the update too, as in the reference), recurse on A22; blocks of at most bs
// are factored directly.  The reference ignores the base case's potrf info
// (lapack_like.h:68); here a failing pivot stops the recursion with the
// blocked version's message.
template <typename T>
void cholesky_factorize_recursive(DMatrix<T>& A, char uplo, int n, int i, int bs = 32) {
    detail::require_upper(uplo, "cholesky_factorize_recursive");
    last_potrf_info() = 0;
    detail::chol_recursive(A, n, i, std::max(bs, 1));
}

// ---- not in the reference: the solves and the orthogonalisation of the path

// Solve A X = B for SPD A[n x n] (upper triangle) and B[n x nrhs] in place:
// FP64 Cholesky (block bs) then U^T Z = B, U X = Z.  Returns 0 or the potrf info.
inline int posv(DMatrix<double>& A, DMatrix<double>& B, int n, int nrhs, int bs = 256) {
    int st = dh_posv_f64(handle(), 'U', n, nrhs, A.device_data(), A.lld(), B.device_data(),
                         B.lld(), bs);
    if (st < 0) check(st, "posv");
    return st;
}

// Mixed-precision solve: FP32 Cholesky of a rounded copy of A and FP64
// iterative refinement to a normwise backward error <= tol; A and B are
// unchanged, X receives the solution.  Returns 0 or the potrf info of the
// FP64 fallback; iters / berr as in dh_posv_mixed.
inline int posv_mixed(DMatrix<double>& A, DMatrix<double>& B, DMatrix<double>& X, int n, int nrhs,
                      int bs = 256, double tol = 1e-10, int max_iter = 50, int* iters = nullptr,
                      double* berr = nullptr) {
    int it = 0;
    double be = 0.0;
    const double* a = A.device_read();
    const double* b = B.device_read();
    int st = dh_posv_mixed(handle(), 'U', n, nrhs, a, A.lld(), b, B.lld(), X.device_data(), X.lld(),
                           bs, tol, max_iter, &it, &be);
    if (st < 0) check(st, "posv_mixed");
    if (iters) *iters = it;
    if (berr) *berr = be;
    return st;
}

// CholeskyQR2 of Y[m x n] in FP64: Y := Q, R[n x n] := upper factor (R may be
// empty).  Returns 0 or the breakdown code of dh_cholqr2_f64.
inline int cholesky_qr2(DMatrix<double>& Y, DMatrix<double>* R = nullptr) {
    const int m = Y.dims()[0], n = Y.dims()[1];
    int st = dh_cholqr2_f64(handle(), m, n, Y.device_data(), Y.lld(),
                            R ? R->device_data() : nullptr, R ? R->lld() : 1);
    if (st < 0) check(st, "cholesky_qr2");
    return st;
}

// Singular values of the square block LA, descending, as an m x 1 LMatrix
// (DArray::svd, lapack_like.h:388-413, which calls sgesvd_ "A","A" on the
// host and returns S).  Computed on the device in FP64 by the one-sided
// Jacobi SVD (dh_gesvj_f64); as in the reference the index arguments are
// not used.
template <typename T>
LMatrix<T> svd(int rank, LMatrix<T> LA, int i1, int j1, int i2, int j2) {
    (void)rank;
    (void)i1;
    (void)j1;
    (void)i2;
    (void)j2;
    const int m = LA.dims()[0], n = LA.dims()[1];
    if (m != n) {
        std::fprintf(stderr, "svd: expects a square block (%d x %d)\n", m, n);
        std::exit(1);
    }
    LMatrix<T> S(m, 1);
    if (m == 0) return S;
    std::vector<double> a((size_t)m * n), s(n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) a[i + (size_t)j * m] = (double)LA(i, j);
    DeviceBuffer<double> dA((size_t)m * n), dV((size_t)n * n), dS(n);
    check(dh_copy_2d(handle(), DH_COPY_H2D, m, n, sizeof(double), a.data(), m, dA.get(), m), "svd");
    int st = dh_gesvj_f64(handle(), m, n, dA.get(), m, dS.get(), dV.get(), n, nullptr);
    if (st < 0) check(st, "svd");
    if (st > 0) std::fprintf(stderr, "svd: Jacobi SVD did not converge\n");
    check(dh_copy_2d(handle(), DH_COPY_D2H, n, 1, sizeof(double), dS.get(), n, s.data(), n), "svd");
    for (int i = 0; i < m; ++i) S(i, 0) = (T)s[i];
    return S;
}

// ---- not in the reference: the Nystrom factor of testsvm.cpp:126-172.
// From the range finder's Q and W = K Q (n x l): Z (n x k) with
// Z Z^T = Q V_k Lam_k V_k^T Q^T for the k largest eigenpairs of Q^T K Q
// (where the reference takes Z = Q diag(S) from the singular values only).
// lambda (k x 1, may be NULL) receives the eigenvalues.  Returns 0, or 1 if
// the Jacobi iteration did not converge.
inline int nystrom_factor(Grid g, int n, int l, int k, DMatrix<double>& Q, DMatrix<double>& W,
                          DMatrix<double>& Z, DMatrix<double>* lambda = nullptr) {
    (void)g;
    int st = dh_nystrom_f64(handle(), n, l, k, Q.device_read(), Q.lld(), W.device_read(), W.lld(),
                            Z.device_data(), Z.lld(), lambda ? lambda->device_data() : nullptr);
    if (st < 0) check(st, "nystrom_factor");
    return st;
}

}  // namespace DArray
