// DArray's BLAS-like operations on DMatrix sub-blocks, run on the GPU.
//
// matrix_multiply (blas_like.h:62-93) and triangular_solve
// (blas_like.h:95-128) act on the device copies of their DMatrix operands
// through the C ABI (dh_gemm_*, dh_trsm_*); multiply_matrix and
// symmetric_update (cblas_like.h:39-69) act on host LMatrix blocks through a
// device round trip.  Differences from the reference:
//   * matrix_multiply implements all four transpose cases (the reference
//     only T,N) with one GEMM instead of k-blocks of bs (bs is accepted and
//     ignored: the product is the same up to rounding);
//   * triangular_solve applies alpha (the reference solves with 1,
//     blas_like.h:98,110) and supports L,U,N and R,U,N besides L,U,T.
#pragma once

#include "darray.h"

namespace DArray {

namespace detail {

inline int gemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda,
                const double* B, int ldb, double beta, double* C, int ldc) {
    return dh_gemm_f64(handle(), ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
inline int gemm(char ta, char tb, int m, int n, int k, float alpha, const float* A, int lda,
                const float* B, int ldb, float beta, float* C, int ldc) {
    return dh_gemm_f32(handle(), ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
inline int trsm(char side, char uplo, char ta, char diag, int m, int n, double alpha,
                const double* A, int lda, double* B, int ldb) {
    return dh_trsm_f64(handle(), side, uplo, ta, diag, m, n, alpha, A, lda, B, ldb);
}
inline int trsm(char side, char uplo, char ta, char diag, int m, int n, float alpha, const float* A,
                int lda, float* B, int ldb) {
    return dh_trsm_f32(handle(), side, uplo, ta, diag, m, n, alpha, A, lda, B, ldb);
}
inline int syrk_ut(int n, int k, double alpha, const double* A, int lda, double beta, double* C,
                   int ldc) {
    return dh_syrk_ut_f64(handle(), n, k, alpha, A, lda, beta, C, ldc);
}
inline int syrk_ut(int n, int k, float alpha, const float* A, int lda, float beta, float* C,
                   int ldc) {
    return dh_syrk_ut_f32(handle(), n, k, alpha, A, lda, beta, C, ldc);
}

inline char upper(const char* s) { return (char)std::toupper((unsigned char)*s); }

// Host LMatrix <-> packed device buffer.
template <typename T>
void upload(LMatrix<T>& A, DeviceBuffer<T>& d) {
    check(dh_copy_2d(handle(), DH_COPY_H2D, A.dims()[0], A.dims()[1], sizeof(T), A.data(), A.ld(),
                     d.get(), std::max(A.dims()[0], 1)),
          "upload");
}
template <typename T>
void download(DeviceBuffer<T>& d, LMatrix<T>& A) {
    check(dh_copy_2d(handle(), DH_COPY_D2H, A.dims()[0], A.dims()[1], sizeof(T), d.get(),
                     std::max(A.dims()[0], 1), A.data(), A.ld()),
          "download");
}

}  // namespace detail

// C[i1:i2, j1:j2] = alpha * op(A) * op(B) + beta * C on the device.
template <typename T>
void matrix_multiply(const char* transa, const char* transb, DMatrixDescriptor<T> Adesc,
                     DMatrixDescriptor<T> Bdesc, DMatrixDescriptor<T> Cdesc, T alpha, T beta,
                     int bs = 128) {
    (void)bs;
    const char ta = detail::upper(transa), tb = detail::upper(transb);
    const int m = Cdesc.m(), n = Cdesc.n();
    const int k = (ta == 'T') ? Adesc.m() : Adesc.n();
    const int am = (ta == 'T') ? Adesc.n() : Adesc.m();
    const int bk = (tb == 'T') ? Bdesc.n() : Bdesc.m();
    const int bn = (tb == 'T') ? Bdesc.m() : Bdesc.n();
    if (am != m || bn != n || bk != k) {
        std::fprintf(stderr, "matrix_multiply: shapes do not conform (%dx%d = %dx%d * %dx%d)\n", m,
                     n, am, k, bk, bn);
        std::exit(1);
    }
    // Inputs first (read-only), then the output: the output may share a
    // DMatrix with an input.
    const T* A = Adesc.device_read();
    const T* B = Bdesc.device_read();
    T* C = Cdesc.device_data();
    check(detail::gemm(ta, tb, m, n, k, alpha, A, Adesc.ld(), B, Bdesc.ld(), beta, C, Cdesc.ld()),
          "matrix_multiply");
}

// Triangular solve with the upper triangle of A (m x m for side L, n x n for
// side R) on B[m x n]:  L,U,T: B := alpha U^-T B;  L,U,N: B := alpha U^-1 B;
// R,U,N: B := alpha B U^-1.
template <typename T>
void triangular_solve(const char* side, const char* uplo, const char* transa, const char* diag,
                      int m, int n, T alpha, DMatrixDescriptor<T> Adesc,
                      DMatrixDescriptor<T> Bdesc, int bs = 64) {
    (void)bs;
    const T* A = Adesc.device_read();
    T* B = Bdesc.device_data();
    check(detail::trsm(detail::upper(side), detail::upper(uplo), detail::upper(transa),
                       detail::upper(diag), m, n, alpha, A, Adesc.ld(), B, Bdesc.ld()),
          "triangular_solve");
}

// Local block GEMM C = alpha op(A) op(B) + beta C on host LMatrix blocks
// (cblas_like.h:39-60), computed on the device.
template <typename T>
void multiply_matrix(const char* ta, const char* tb, T alpha, T beta, LMatrix<T>& A, LMatrix<T>& B,
                     LMatrix<T>& C) {
    const char a = detail::upper(ta), b = detail::upper(tb);
    const int m = C.dims()[0], n = C.dims()[1];
    const int k = (a == 'T') ? A.dims()[0] : A.dims()[1];
    DeviceBuffer<T> dA((size_t)std::max(A.dims()[0], 1) * std::max(A.dims()[1], 1));
    DeviceBuffer<T> dB((size_t)std::max(B.dims()[0], 1) * std::max(B.dims()[1], 1));
    DeviceBuffer<T> dC((size_t)std::max(m, 1) * std::max(n, 1));
    detail::upload(A, dA);
    detail::upload(B, dB);
    detail::upload(C, dC);
    check(detail::gemm(a, b, m, n, k, alpha, dA.get(), std::max(A.dims()[0], 1), dB.get(),
                       std::max(B.dims()[0], 1), beta, dC.get(), std::max(m, 1)),
          "multiply_matrix");
    detail::download(dC, C);
}

// Upper SYRK A22 := beta A22 + alpha A12^T A12 on host blocks
// (symmetric_update, cblas_like.h:61-69); the strict lower triangle of A22
// is not written.
template <typename T>
void symmetric_update(const char* uplo, const char* trans, T alpha, LMatrix<T>& A12, T beta,
                      LMatrix<T>& A22) {
    if (detail::upper(uplo) != 'U' || detail::upper(trans) != 'T') {
        std::fprintf(stderr, "symmetric_update: only uplo U, trans T\n");
        std::exit(1);
    }
    const int n = A22.dims()[0], k = A12.dims()[0];
    DeviceBuffer<T> dA((size_t)std::max(k, 1) * std::max(n, 1));
    DeviceBuffer<T> dC((size_t)std::max(n, 1) * std::max(n, 1));
    detail::upload(A12, dA);
    detail::upload(A22, dC);
    check(detail::syrk_ut(n, k, alpha, dA.get(), std::max(k, 1), beta, dC.get(), std::max(n, 1)),
          "symmetric_update");
    detail::download(dC, A22);
}

}  // namespace DArray
