// Known-answer Cholesky, after perftests/perftest_cholesky.cpp:46-84:
// A(i,j) = min(i+1, j+1) = C C^T with C lower-triangular ones, so the upper
// factor U is all ones on and above the diagonal.  The reference probes
// A(n-1,n-1) against 1 (relative 1e-5); this driver checks every upper
// entry, and that the blocked version leaves the strict lower triangle as
// it was.  Usage: chol_ones <n> <bs> <blocked|recursive> <float|double>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "darray_hip/lapack_like.h"

template <typename T>
int run(int n, int bs, bool recursive, double tol) {
    DArray::Grid g;
    DArray::DMatrix<T> A(g, n, n);
    A.set_function([](int i, int j) -> T { return (T)std::min(i + 1, j + 1); });
    DArray::ElapsedTimer timer;
    if (recursive)
        DArray::cholesky_factorize_recursive(A, 'U', n, 0, bs);
    else
        DArray::cholesky_factorize(A, 'U', n, 0, bs);
    const T* a = A.host_read();  // waits for the factorisation
    double s = timer.elapsed_s();
    int info = DArray::last_potrf_info();
    double worst = 0.0, worst_lower = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double v = a[i + (size_t)j * A.lld()];
            if (i <= j)
                worst = std::max(worst, std::fabs(v - 1.0));
            else if (!recursive)
                worst_lower = std::max(worst_lower, std::fabs(v - (double)std::min(i + 1, j + 1)));
        }
    std::printf("chol_ones n=%d bs=%d %s %s: info=%d  max|U-1|=%.3e  lower changed by %.3e  "
                "%.3f s  %.1f GFLOP/s  A[%d,%d]=%g\n",
                n, bs, recursive ? "recursive" : "blocked", sizeof(T) == 4 ? "float" : "double", info,
                worst, worst_lower, s, (double)n * n * n / 3.0 / s / 1e9, n - 1, n - 1,
                (double)a[(n - 1) + (size_t)(n - 1) * A.lld()]);
    bool ok = info == 0 && worst <= tol && worst_lower == 0.0;
    std::printf("%s\n", ok ? "PASS" : "FAIL");
    return ok ? 0 : 1;
}

int main(int argc, char** argv) {
    int n = argc > 1 ? std::atoi(argv[1]) : 1024;
    int bs = argc > 2 ? std::atoi(argv[2]) : 256;
    bool recursive = argc > 3 && std::strcmp(argv[3], "recursive") == 0;
    bool fp32 = argc > 4 && std::strcmp(argv[4], "float") == 0;
    // Entries of U are exactly 1; FP32 rounding of the trailing updates grows with n.
    return fp32 ? run<float>(n, bs, recursive, 1e-5 * std::max(1.0, n / 64.0))
                : run<double>(n, bs, recursive, 1e-12 * std::max(1.0, n / 64.0));
}
