// The reference's object-level seam, called exactly as DArray's drivers call
// it: the C++-linkage host-pointer functions rbf_kernel, tc_gemm and tc_sgemm
// (dkernel.h:6-8, dkernel.cu:132-275) between cuda_init and cuda_finalize
// (testsvm.cpp:7-8,41,326), with DArray's column-major host layouts.  Each
// result is checked against a host evaluation of the same sums in double
// from the same float inputs.  Covered: Xi / Xj with ld larger than the row
// count (the reference ignores ldxi / ldxj, dkernel.cu:139-141), K with
// ldk > m and untouched padding rows, labels, an inner dimension > 8192 (the
// reference overwrites C per 8192-wide block, dkernel.cu:204-223), beta != 0
// (the reference zeroes C, dkernel.cu:201), beta = 0 with NaN in C, all four
// tc_sgemm transposes (the reference has N,N and T,N), and a second
// cuda_init after cuda_finalize.
#include <cmath>
#include <cstdio>
#include <limits>
#include <random>
#include <vector>

#include "darray_hip/dkernel.h"

namespace {

std::mt19937 rng(2024);

std::vector<float> rand_mat(int rows, int cols, int ld, float lo, float hi, float pad) {
    std::uniform_real_distribution<float> u(lo, hi);
    std::vector<float> a((size_t)ld * cols, pad);
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i) a[i + (size_t)j * ld] = u(rng);
    return a;
}

int report(const char* what, double err, double tol) {
    std::printf("%-58s err %.3e (tol %.0e) %s\n", what, err, tol, err <= tol ? "ok" : "BAD");
    return err <= tol ? 0 : 1;
}

int check_rbf(int m, int n, int d, int pad_i, int pad_j, int pad_k, bool labels) {
    const int ldxi = m + pad_i, ldxj = n + pad_j, ldk = m + pad_k;
    const float sentinel = -7.0f;
    auto Xi = rand_mat(m, d, ldxi, 0.f, 1.f, 1e30f);  // padding poisoned: must not be read
    auto Xj = rand_mat(n, d, ldxj, 0.f, 1.f, 1e30f);
    std::vector<float> Yi(m), Yj(n);
    for (int i = 0; i < m; ++i) Yi[i] = (i % 3 == 0) ? -1.f : 1.f;
    for (int j = 0; j < n; ++j) Yj[j] = (j % 4 == 1) ? -1.f : 1.f;
    std::vector<float> K((size_t)ldk * n, sentinel);
    const float gamma = 1.0f / d;
    rbf_kernel(0, m, n, d, Xi.data(), ldxi, Xj.data(), ldxj, labels ? Yi.data() : nullptr,
               labels ? Yj.data() : nullptr, K.data(), ldk, gamma);
    double err = 0.0;
    int bad_pad = 0;
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < m; ++i) {
            double d2 = 0.0;
            for (int f = 0; f < d; ++f) {
                double t = (double)Xi[i + (size_t)f * ldxi] - (double)Xj[j + (size_t)f * ldxj];
                d2 += t * t;
            }
            double want = std::exp(-(double)gamma * d2);
            if (labels) want *= (double)Yi[i] * Yj[j];
            err = std::max(err, std::fabs(K[i + (size_t)j * ldk] - want));
        }
        for (int i = m; i < ldk; ++i) bad_pad += K[i + (size_t)j * ldk] != sentinel;
    }
    char what[160];
    std::snprintf(what, sizeof what, "rbf_kernel m=%d n=%d d=%d ldxi=%d ldxj=%d ldk=%d%s", m, n, d,
                  ldxi, ldxj, ldk, labels ? " labels" : "");
    int bad = report(what, err, 5e-6);
    if (bad_pad) std::printf("  %d padding entries of K were written\n", bad_pad);
    return bad + (bad_pad != 0);
}

// op(A) m x n, op(B) n x k, C m x k, all column-major with the given lds.
int check_gemm(bool sgemm, char ta, char tb, int m, int k, int n, float alpha, float beta,
               bool nan_c) {
    const bool TA = ta == 'T', TB = tb == 'T';
    const int ar = TA ? n : m, ac = TA ? m : n, br = TB ? k : n, bc = TB ? n : k;
    const int lda = ar + 3, ldb = br + 1, ldc = m + 2;
    auto A = rand_mat(ar, ac, lda, -1.f, 1.f, 1e30f);
    auto B = rand_mat(br, bc, ldb, -1.f, 1.f, 1e30f);
    auto C = rand_mat(m, k, ldc, -1.f, 1.f, -7.f);
    if (nan_c)
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < m; ++i) C[i + (size_t)j * ldc] = std::numeric_limits<float>::quiet_NaN();
    std::vector<double> want((size_t)m * k);
    double scale = 0.0;
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < m; ++i) {
            double s = 0.0, sa = 0.0;
            for (int p = 0; p < n; ++p) {
                double a = TA ? A[p + (size_t)i * lda] : A[i + (size_t)p * lda];
                double b = TB ? B[j + (size_t)p * ldb] : B[p + (size_t)j * ldb];
                s += a * b;
                sa += std::fabs(a * b);
            }
            double c0 = beta != 0.f ? (double)C[i + (size_t)j * ldc] : 0.0;
            want[i + (size_t)j * m] = (double)alpha * s + (double)beta * c0;
            scale = std::max(scale, std::fabs(alpha) * sa + std::fabs(beta * c0));
        }
    if (sgemm)
        tc_sgemm(0, ta, tb, m, k, n, A.data(), lda, B.data(), ldb, C.data(), ldc, alpha, beta);
    else if (alpha == 1.f && beta == 0.f)
        tc_gemm(0, m, k, n, A.data(), lda, B.data(), ldb, C.data(), ldc);  // header defaults
    else
        tc_gemm(0, m, k, n, A.data(), lda, B.data(), ldb, C.data(), ldc, alpha, beta);
    double err = 0.0;
    int bad_pad = 0;
    for (int j = 0; j < k; ++j) {
        for (int i = 0; i < m; ++i) {
            double v = C[i + (size_t)j * ldc];
            double e = std::fabs(v - want[i + (size_t)j * m]);
            err = std::max(err, std::isfinite(v) ? e : INFINITY);
        }
        for (int i = m; i < ldc; ++i) bad_pad += C[i + (size_t)j * ldc] != -7.f;
    }
    char what[160];
    std::snprintf(what, sizeof what, "%s(%c,%c) m=%d k=%d n=%d alpha=%g beta=%g%s",
                  sgemm ? "tc_sgemm" : "tc_gemm", ta, tb, m, k, n, alpha, beta,
                  nan_c ? " C=NaN" : "");
    // FP32 products with FP32 accumulation: bounded by a few n * u32 of sum |a b|.
    int bad = report(what, err / scale, 4e-6 * std::sqrt((double)n) + 1e-6);
    if (bad_pad) std::printf("  %d padding entries of C were written\n", bad_pad);
    return bad + (bad_pad != 0);
}

}  // namespace

int main() {
    cuda_init();
    int bad = 0;
    bad += check_rbf(300, 517, 37, 5, 3, 7, true);
    bad += check_rbf(1, 1, 1, 0, 0, 0, false);
    bad += check_rbf(129, 65, 130, 0, 11, 0, false);
    bad += check_gemm(false, 'N', 'N', 200, 70, 9000, 1.f, 0.f, true);
    bad += check_gemm(false, 'N', 'N', 200, 70, 9000, 0.5f, -1.5f, false);
    bad += check_gemm(false, 'N', 'N', 33, 1, 17, 2.f, 1.f, false);
    for (char ta : {'N', 'T'})
        for (char tb : {'N', 'T'}) {
            bad += check_gemm(true, ta, tb, 150, 90, 8300, 1.f, 0.f, true);
            bad += check_gemm(true, ta, tb, 77, 45, 300, -0.75f, 0.5f, false);
        }
    cuda_finalize();
    cuda_init();  // a second init after finalize, as a driver run twice in one process would
    bad += check_gemm(true, 'T', 'N', 64, 64, 64, 1.f, 2.f, false);
    cuda_finalize();
    std::printf("%s\n", bad ? "FAIL" : "PASS");
    return bad ? 1 : 0;
}
