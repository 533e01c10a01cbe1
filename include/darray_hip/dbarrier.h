// DArray's barrier-method SVM solver (dbarrier.h, testsvm.cpp:191-321) and
// LIBSVM model writer (dbarrier.h:394-470) on one MI355X.
//
// The reference runs the solver on the host over an n x k LMatrix Z with
// dgemm_ / ddot_; here Z, the labels and the dual variables live on the
// device (DMatrix device copies) and the whole loop runs there
// (dh_svm_barrier): Newton steps with Jacobi-preconditioned CG
// (pre_conjgrad), backtracking line search, t *= 20 until the duality-gap
// bound is met.  Differences from the reference (bug fixes): the Newton
// system's right-hand side is -grad (the reference uses e_1), the operator
// has no added n I, the preconditioner includes the barrier Hessian, and
// the model's bias uses (Z Z^T)_ij with LIBSVM's label and nr_sv fields.
#pragma once

#include "dkernel.h"
#include "lapack_like.h"

namespace DArray {

struct BarrierStats {
    int barrier_iterations = 0, newton_steps = 0, cg_iterations = 0;
    double f = 0.0;  // a^T Z Z^T a / 2 - sum(a) at the solution
};

// Solve the SVM dual on the Nystrom factor Z (n x k) of diag(y) K diag(y)
// for labels Y (1 x n, +-1) and box bound C; a (n x 1) receives the
// solution.  Returns the iteration counts and the objective.
inline BarrierStats barrier_svm(Grid g, int n, int k, DMatrix<double>& Z, DMatrix<double>& Y,
                                double C, DMatrix<double>& a, double cg_tol = 1e-2,
                                int max_outer = 50, int max_newton = 100, int max_cg = -1) {
    (void)g;
    int stats[3] = {0, 0, 0};
    BarrierStats out;
    check(dh_svm_barrier(handle(), n, k, Z.device_read(), Z.lld(), Y.device_read(), C, cg_tol,
                         max_outer, max_newton, max_cg < 0 ? 10 * n : max_cg, a.device_data(),
                         stats, &out.f),
          "barrier_svm");
    out.barrier_iterations = stats[0];
    out.newton_steps = stats[1];
    out.cg_iterations = stats[2];
    return out;
}

// pre_conjgrad (dbarrier.h:83-139): (Z Z^T + diag(D)) x = b from the initial
// guess x by CG preconditioned with diag(Z Z^T) + D, until ||r|| <= tol (the
// reference's 1e-2 by default).  Pass D + n to reproduce the reference's
// operator Z Z^T + D + n I.  Returns the iterations.
inline int pre_conjgrad(int n, int k, DMatrix<double>& Z, DMatrix<double>& D, DMatrix<double>& b,
                        DMatrix<double>& x, double tol = 1e-2, int max_iter = -1) {
    int it = 0;
    check(dh_svm_pcg(handle(), n, k, Z.device_read(), Z.lld(), D.device_read(), b.device_read(),
                     x.device_data(), tol, max_iter < 0 ? 10 * n : max_iter, &it, nullptr),
          "pre_conjgrad");
    return it;
}

// writemodel (dbarrier.h:394-470): LIBSVM model file of the solution a for
// the Nystrom factor dZ (n x k), labels dY (n x 1), points dX (d x n); host
// LMatrix arguments as in the reference.  Prints the support-vector counts
// and the bias as the reference does.
template <typename T>
void writemodel(const std::string filepath, int n, int d, int k, LMatrix<T>& a, T C,
                LMatrix<T>& dZ, LMatrix<T> dY, LMatrix<T> dX, T gamma, int minus, int plus) {
    (void)minus;
    (void)plus;
    std::vector<double> av(n), yv(n), zv((size_t)n * k), xv((size_t)d * n);
    for (int i = 0; i < n; ++i) {
        av[i] = (double)a(i, 0);
        yv[i] = (double)dY(i, 0);
        for (int c = 0; c < k; ++c) zv[i + (size_t)c * n] = (double)dZ(i, c);
        for (int f = 0; f < d; ++f) xv[f + (size_t)i * d] = (double)dX(f, i);
    }
    double b = 0.0;
    int nsv = dh_write_libsvm_model(filepath.c_str(), n, d, k, av.data(), (double)C, zv.data(), n,
                                    yv.data(), xv.data(), d, (double)gamma, &b);
    check(nsv < 0 ? nsv : 0, "writemodel");
    int nfree = 0;
    for (int i = 0; i < n; ++i) nfree += av[i] > 1e-6 && av[i] < (double)C - 1e-6;
    std::printf("#BSV %d, #SV %d\n", nfree, nsv);
    std::printf("approx mean b=%.6e\n", b);
}

}  // namespace DArray
