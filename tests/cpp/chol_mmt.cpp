// Cholesky of A = m m^T with m(i,j) = 1+i+j for i >= j (tests/test3.cpp:23-46),
// n = 10, recursive with base block 4 as the reference runs it, then the
// blocked version.  The upper factor is U = m^T exactly (m has a positive
// diagonal), which the driver checks; the reference only prints it.
#include <cstdio>

#include "darray_hip/lapack_like.h"

template <typename T>
int run(bool recursive, double tol) {
    const int n = 10;
    DArray::Grid g;
    DArray::DMatrix<T> A(g, n, n);
    DArray::LMatrix<T> m(n, n);
    m.set_function([](int i, int j) -> T { return i >= j ? (T)(1 + i + j) : (T)0; });
    DArray::LMatrix<T> C(n, n);
    C.set_constant(0);
    DArray::multiply_matrix<T>("N", "T", 1, 1, m, m, C);  // C = m m^T
    A.dereplicate_in_all(C, 0, 0, n, n);
    if (recursive)
        DArray::cholesky_factorize_recursive(A, 'U', n, 0, 4);
    else
        DArray::cholesky_factorize(A, 'U', n, 0, 4);
    auto U = A.collect(0);
    double worst = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) worst = std::max(worst, std::fabs((double)U(i, j) - m(j, i)) / m(j, i));
    std::printf("chol_mmt %s %s: max rel |U - m^T| = %.3e (info %d)\n",
                recursive ? "recursive" : "blocked", sizeof(T) == 4 ? "float" : "double", worst,
                DArray::last_potrf_info());
    return (worst <= tol && DArray::last_potrf_info() == 0) ? 0 : 1;
}

int main() {
    int bad = run<double>(true, 1e-12) + run<double>(false, 1e-12) + run<float>(true, 1e-4) +
              run<float>(false, 1e-4);
    std::printf("%s\n", bad ? "FAIL" : "PASS");
    return bad ? 1 : 0;
}
