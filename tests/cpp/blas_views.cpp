// matrix_multiply and triangular_solve on DMatrixDescriptor views, after
// tests/test5.cpp:15-47 (10 x 10 matrices, views of the same DMatrix), with
// expected results computed on the host in double and compared.
#include <cstdio>
#include <vector>

#include "darray_hip/blas_like.h"

using Mat = std::vector<double>;  // 10 x 10 column-major, ld 10

template <typename T>
Mat snapshot(DArray::DMatrix<T>& A) {
    Mat r(100);
    const T* a = A.host_read();
    for (int j = 0; j < 10; ++j)
        for (int i = 0; i < 10; ++i) r[i + 10 * j] = a[i + (size_t)j * A.lld()];
    return r;
}

double maxdiff(const Mat& a, const Mat& b) {
    double w = 0.0, s = 0.0;
    for (size_t e = 0; e < a.size(); ++e) {
        w = std::max(w, std::fabs(a[e] - b[e]));
        s = std::max(s, std::fabs(b[e]));
    }
    return w / std::max(s, 1e-300);
}

template <typename T>
int run(double tol) {
    DArray::Grid g;
    int bad = 0;
    const char* tn = sizeof(T) == 4 ? "float" : "double";
    // C = alpha op(A) op(B) + beta C on views, every transpose pair.
    for (char ta : {'N', 'T'})
        for (char tb : {'N', 'T'}) {
            DArray::DMatrix<T> A(g, 10, 10);
            A.set_function([](int gi, int gj) -> T { return (T)(gi + gj / 10.0); });
            Mat a = snapshot(A), want = a;
            // A2 = A[0:5, 5:10], A1 = A[5:10, 0:5], A3 = A[5:10, 5:10]
            DArray::DMatrixDescriptor<T> A1{A, 5, 0, 10, 5}, A2{A, 0, 5, 5, 10}, A3{A, 5, 5, 10, 10};
            auto el = [&](const Mat& M, int r0, int c0, char t, int i, int k) {
                return t == 'N' ? M[(r0 + i) + 10 * (c0 + k)] : M[(r0 + k) + 10 * (c0 + i)];
            };
            for (int j = 0; j < 5; ++j)
                for (int i = 0; i < 5; ++i) {
                    double s = 0.0;
                    for (int k = 0; k < 5; ++k) s += el(a, 0, 5, ta, i, k) * el(a, 5, 0, tb, k, j);
                    want[(5 + i) + 10 * (5 + j)] = -1.0 * s + 0.5 * a[(5 + i) + 10 * (5 + j)];
                }
            const char sa[2] = {ta, 0}, sb[2] = {tb, 0};
            DArray::matrix_multiply<T>(sa, sb, A2, A1, A3, T(-1), T(0.5), 2);
            double e = maxdiff(snapshot(A), want);
            std::printf("matrix_multiply %c%c %s: rel diff %.3e\n", ta, tb, tn, e);
            bad += e > tol;
        }
    // triangular_solve with the upper triangle of A1 = A[5:10, 0:5] on B2 = B[0:5, 5:10]
    // (test5.cpp:36-47), and the L,U,N and R,U,N cases.
    for (int c = 0; c < 3; ++c) {
        const char* side = c == 2 ? "R" : "L";
        const char* trans = c == 0 ? "T" : "N";
        DArray::DMatrix<T> A(g, 10, 10), B(g, 10, 10);
        A.set_function([](int i, int j) -> T { return (T)(1 + i + j); });
        B.set_function([](int gi, int gj) -> T { return (T)(gi + gj / 10.0); });
        Mat a = snapshot(A), b = snapshot(B), want = b;
        auto U = [&](int i, int j) { return i <= j ? a[(5 + i) + 10 * j] : 0.0; };
        const double alpha = 2.0;
        // Solve column by column (L) or row by row (R) by substitution.
        if (c == 0) {  // U^T X = alpha B: forward substitution
            for (int j = 0; j < 5; ++j)
                for (int i = 0; i < 5; ++i) {
                    double s = alpha * b[i + 10 * (5 + j)];
                    for (int k = 0; k < i; ++k) s -= U(k, i) * want[k + 10 * (5 + j)];
                    want[i + 10 * (5 + j)] = s / U(i, i);
                }
        } else if (c == 1) {  // U X = alpha B: back substitution
            for (int j = 0; j < 5; ++j)
                for (int i = 4; i >= 0; --i) {
                    double s = alpha * b[i + 10 * (5 + j)];
                    for (int k = i + 1; k < 5; ++k) s -= U(i, k) * want[k + 10 * (5 + j)];
                    want[i + 10 * (5 + j)] = s / U(i, i);
                }
        } else {  // X U = alpha B: forward over columns
            for (int i = 0; i < 5; ++i)
                for (int j = 0; j < 5; ++j) {
                    double s = alpha * b[i + 10 * (5 + j)];
                    for (int k = 0; k < j; ++k) s -= want[i + 10 * (5 + k)] * U(k, j);
                    want[i + 10 * (5 + j)] = s / U(j, j);
                }
        }
        DArray::DMatrixDescriptor<T> A1{A, 5, 0, 10, 5}, B2{B, 0, 5, 5, 10};
        DArray::triangular_solve<T>(side, "U", trans, "N", 5, 5, T(alpha), A1, B2, 3);
        double e = maxdiff(snapshot(B), want);
        std::printf("triangular_solve %s,U,%s %s: rel diff %.3e\n", side, trans, tn, e);
        bad += e > tol;
    }
    return bad;
}

int main() {
    int bad = run<double>(1e-13) + run<float>(1e-5);
    std::printf("%s\n", bad ? "FAIL" : "PASS");
    return bad ? 1 : 0;
}
