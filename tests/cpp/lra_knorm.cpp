// DArray::LRA and DArray::Knorm (dkernel.h:12-76) through the device path,
// against a host evaluation of the same sums in double, plus the fused range
// finder on a point set with r distinct points (rank K <= r, so the
// approximation error of a rank-l basis, l >= r, is at rounding level).
#include <cstdio>
#include <vector>

#include "darray_hip/dkernel.h"

// A[i, c] = sum_j y_i y_j exp(-gamma |x_i - x_j|^2) O[j, c] on the host.
std::vector<double> host_lra(int n, int d, int k, const double* X, const double* y, const double* O,
                             double gamma, double* knorm2) {
    std::vector<double> A((size_t)n * k, 0.0), krow(n);
    double s2 = 0.0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            double d2 = 0.0;
            for (int f = 0; f < d; ++f) {
                double t = X[f + (size_t)i * d] - X[f + (size_t)j * d];
                d2 += t * t;
            }
            krow[j] = y[i] * y[j] * std::exp(-gamma * d2);
            s2 += krow[j] * krow[j];
        }
        for (int c = 0; c < k; ++c) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) s += krow[j] * O[j + (size_t)c * n];
            A[i + (size_t)c * n] = s;
        }
    }
    *knorm2 = s2;
    return A;
}

template <typename T>
double rel_err(DArray::DMatrix<T>& A, const std::vector<double>& want, int n, int k) {
    const T* a = A.host_read();
    double e = 0.0, s = 0.0;
    for (int c = 0; c < k; ++c)
        for (int i = 0; i < n; ++i) {
            double w = want[i + (size_t)c * n], v = a[i + (size_t)c * A.lld()];
            e += (v - w) * (v - w);
            s += w * w;
        }
    return std::sqrt(e / s);
}

const char* prec_name(DArray::Precision p) {
    switch (p) {
        case DArray::Precision::FP64: return "fp64";
        case DArray::Precision::Mixed: return "mixed (fp16)";
        case DArray::Precision::BF16: return "bf16";
        default: return "fp8";
    }
}

template <typename T>
int run_lra(int n, int d, int k, DArray::Precision prec, double tol) {
    DArray::Grid g;
    DArray::DMatrix<T> X(g, d, n), Y(g, 1, n), O(g, n, k), A(g, n, k);
    X.set_uniform_seed(0, 1, 0);
    O.set_normal_seed(0, 1, 1);
    Y.set_function([](int, int j) -> T { return (j % 3 == 0) ? T(-1) : T(1); });
    const double gamma = 1.0 / d;
    std::vector<double> x((size_t)d * n), y(n), o((size_t)n * k);
    for (int j = 0; j < n; ++j) {
        for (int f = 0; f < d; ++f) x[f + (size_t)j * d] = X(f, j);
        y[j] = Y(0, j);
        for (int c = 0; c < k; ++c) o[j + (size_t)c * n] = O(j, c);
    }
    double k2 = 0.0;
    auto want = host_lra(n, d, k, x.data(), y.data(), o.data(), gamma, &k2);
    DArray::LRA(g, n, d, k, X, Y, O, A, (float)gamma, prec);
    double e = rel_err(A, want, n, k);
    std::printf("LRA<%s> %s n=%d d=%d k=%d: rel Frobenius error vs host %.3e (tol %.0e)\n",
                sizeof(T) == 4 ? "float" : "double", prec_name(prec),
                n, d, k, e, tol);
    int bad = e > tol;
    // Knorm with Z = 0: ||K - 0|| = ||K||, ratio exactly 1.
    DArray::DMatrix<T> Z(g, n, k);
    Z.set_constant(0);
    double ratio = DArray::Knorm(g, n, d, k, X, Y, Z, (float)gamma);
    std::printf("Knorm(Z=0) ratio %.15f, host ||K||_F %.6f\n", ratio, std::sqrt(k2));
    bad += std::fabs(ratio - 1.0) > 1e-12;
    return bad;
}

int run_range_finder() {
    // 8 distinct points repeated over n = 4096 columns: rank(K) <= 8 < l.
    const int n = 4096, d = 12, l = 16, q = 1, r = 8;
    DArray::Grid g;
    DArray::DMatrix<double> P(g, d, r), X(g, d, n), Om(g, n, l), Q(g, n, l), W(g, n, l);
    P.set_uniform_seed(-1, 1, 3);
    X.set_function([&](int f, int j) { return P(f, (j * 5) % r); });
    Om.set_normal_seed(0, 1, 1);
    const double gamma = 0.5 / d;
    int bad = 0;
    for (auto prec : {DArray::Precision::FP64, DArray::Precision::Mixed}) {
        int st = DArray::randomized_range_finder(g, n, d, l, q, X, nullptr, Om, Q, W, gamma, prec);
        double e = DArray::lra_error(g, n, d, l, X, nullptr, Q, W, gamma);
        double tol = prec == DArray::Precision::FP64 ? 1e-10 : 1e-2;
        std::printf("range finder %s, rank-%d kernel, l=%d: status %d, ||K-QB||/||K|| = %.3e (tol %.0e)\n",
                    prec_name(prec), r, l, st, e, tol);
        bad += (e > tol);
    }
    return bad;
}

int main() {
    int bad = 0;
    bad += run_lra<double>(1024, 16, 40, DArray::Precision::FP64, 1e-12);
    bad += run_lra<double>(1024, 16, 40, DArray::Precision::Mixed, 2e-2);
    bad += run_lra<double>(1024, 16, 40, DArray::Precision::BF16, 2e-2);
    bad += run_lra<float>(1024, 16, 40, DArray::Precision::FP64, 1e-6);
    bad += run_range_finder();
    std::printf("%s\n", bad ? "FAIL" : "PASS");
    return bad ? 1 : 0;
}
