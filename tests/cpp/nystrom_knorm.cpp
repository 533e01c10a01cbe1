// The consumer step of testsvm.cpp:86-177 through the DArray header mirror:
// range finder -> C = Q^T K Q -> eigenpairs -> Z -> Knorm, plus DArray::svd
// (lapack_like.h:388-413) on matrices with known singular values.
// Checks: svd of a diagonal and of a rotated diagonal matrix, Z^T Z = diag(lam)
// (Z = Q V_k Lam_k^1/2 with Q, V_k orthonormal), ||K - Z Z^T|| no smaller than
// the range finder's ||K - Q W^T||, and an exact rank-r kernel (r distinct
// points, r <= k) reproduced to rounding.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "darray_hip/dkernel.h"
#include "darray_hip/lapack_like.h"

namespace {

int check(const char* what, double err, double tol) {
    std::printf("%-64s %.3e (tol %.0e) %s\n", what, err, tol, err <= tol ? "ok" : "BAD");
    return err <= tol ? 0 : 1;
}

int svd_known_answers() {
    int bad = 0;
    // diag(3, -7, 0.5, 2): singular values 7, 3, 2, 0.5 exactly
    DArray::LMatrix<double> D(4, 4);
    D.set_constant(0.0);
    D(0, 0) = 3.0;
    D(1, 1) = -7.0;
    D(2, 2) = 0.5;
    D(3, 3) = 2.0;
    auto S = DArray::svd<double>(0, D, 0, 0, 4, 4);
    const double want[4] = {7.0, 3.0, 2.0, 0.5};
    double e = 0.0;
    for (int i = 0; i < 4; ++i) e = std::max(e, std::fabs(S(i, 0) - want[i]));
    bad += check("svd(diag(3,-7,0.5,2)) vs 7,3,2,0.5", e, 0.0);
    // G diag(s) G^T with G a product of plane rotations: singular values s
    const int n = 96;
    DArray::LMatrix<double> A(n, n);
    A.set_constant(0.0);
    for (int i = 0; i < n; ++i) A(i, i) = std::pow(10.0, -6.0 * i / (n - 1));
    std::mt19937 gen(7);
    std::uniform_real_distribution<double> ang(0.0, 6.283185307179586);
    for (int r = 0; r < 4 * n; ++r) {  // A <- G A G^T for random rotations G(p, q)
        int p = gen() % n, q = gen() % n;
        if (p == q) continue;
        double t = ang(gen), c = std::cos(t), s = std::sin(t);
        for (int j = 0; j < n; ++j) {
            double x = A(p, j), y = A(q, j);
            A(p, j) = c * x - s * y;
            A(q, j) = s * x + c * y;
        }
        for (int i = 0; i < n; ++i) {
            double x = A(i, p), y = A(i, q);
            A(i, p) = c * x - s * y;
            A(i, q) = s * x + c * y;
        }
    }
    auto S2 = DArray::svd<double>(0, A, 0, 0, n, n);
    e = 0.0;
    for (int i = 0; i < n; ++i) e = std::max(e, std::fabs(S2(i, 0) - std::pow(10.0, -6.0 * i / (n - 1))));
    bad += check("svd(G diag(10^-6i/95) G^T) max abs error", e, 1e-13);
    DArray::LMatrix<float> F(1, 1);
    F(0, 0) = -2.5f;
    auto S3 = DArray::svd<float>(0, F, 0, 0, 1, 1);
    bad += check("svd<float> of [-2.5]", std::fabs(S3(0, 0) - 2.5f), 0.0);
    return bad;
}

int nystrom_on_mixture(int n, int d, int k, int p, int q, bool labels) {
    DArray::Grid g;
    const int l = k + p;
    DArray::DMatrix<double> X(g, d, n), Y(g, 1, n), Om(g, n, l), Q(g, n, l), W(g, n, l);
    DArray::DMatrix<double> Z(g, n, k), lam(g, k, 1);
    X.set_uniform_seed(-1, 1, 11);
    Om.set_normal_seed(0, 1, 12);
    Y.set_function([](int, int j) -> double { return (j % 5 == 2) ? -1.0 : 1.0; });
    const double gamma = 1.0 / d;
    int st = DArray::randomized_range_finder(g, n, d, l, q, X, labels ? &Y : nullptr, Om, Q, W,
                                             gamma);
    double eq = DArray::lra_error(g, n, d, l, X, labels ? &Y : nullptr, Q, W, gamma);
    int sn = DArray::nystrom_factor(g, n, l, k, Q, W, Z, &lam);
    double ratio = DArray::Knorm(g, n, d, k, X, Y, Z, (float)gamma);
    if (!labels) ratio = DArray::lra_error(g, n, d, k, X, nullptr, Z, Z, gamma);
    const double* z = Z.host_read();
    const double* lm = lam.host_read();
    // Z^T Z = diag(lam)
    double e = 0.0;
    for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) {
            double s = 0.0;
            for (int i = 0; i < n; ++i) s += z[i + (size_t)a * Z.lld()] * z[i + (size_t)b * Z.lld()];
            e = std::max(e, std::fabs(s - (a == b ? std::max(lm[a], 0.0) : 0.0)));
        }
    char what[160];
    std::snprintf(what, sizeof what, "n=%d d=%d k=%d l=%d q=%d%s: status %d/%d, |Z^T Z - diag(lam)| / lam_1",
                  n, d, k, l, q, labels ? " labels" : "", st, sn);
    int bad = check(what, e / lm[0], 1e-12);
    std::printf("  ||K-QW^T||/||K|| = %.4e, ||K-ZZ^T||/||K|| = %.4e, lam_1 = %.4e, lam_k = %.4e\n",
                eq, ratio, lm[0], lm[k - 1]);
    bad += (st != 0) + (sn != 0);
    bad += !(ratio >= eq * (1 - 1e-9) && ratio < 0.5);
    for (int a = 1; a < k; ++a) bad += lm[a] > lm[a - 1];  // descending
    return bad;
}

int nystrom_exact_low_rank() {
    const int n = 3000, d = 6, r = 10, l = 24, k = 16;
    DArray::Grid g;
    DArray::DMatrix<double> P(g, d, r), X(g, d, n), Om(g, n, l), Q(g, n, l), W(g, n, l), Z(g, n, k);
    P.set_uniform_seed(-1, 1, 5);
    X.set_function([&](int f, int j) { return P(f, (j * 7) % r); });
    Om.set_normal_seed(0, 1, 6);
    DArray::randomized_range_finder(g, n, d, l, 1, X, nullptr, Om, Q, W, 0.4);
    DArray::nystrom_factor(g, n, l, k, Q, W, Z);
    double e = DArray::lra_error(g, n, d, k, X, nullptr, Z, Z, 0.4);
    return check("rank-10 kernel (10 distinct points), k=16: ||K-ZZ^T||/||K||", e, 1e-11);
}

}  // namespace

int main() {
    int bad = svd_known_answers();
    bad += nystrom_on_mixture(4096, 16, 64, 10, 1, false);
    bad += nystrom_on_mixture(3000, 32, 100, 12, 2, true);
    bad += nystrom_exact_low_rank();
    std::printf("%s\n", bad ? "FAIL" : "PASS");
    return bad ? 1 : 0;
}
