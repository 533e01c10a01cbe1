// testsvm.cpp's whole flow (tests/testsvm.cpp:44-325) on the DArray header
// mirror: LIBSVM file -> read_input_file -> labelled range finder (DArray::LRA's
// role, FP16 K-passes) -> Nystrom factor Z and Knorm -> barrier SVM solver ->
// writemodel.  The data set is written by this program: two Gaussian blobs in
// d = 8 (labels +-1 by blob, 30% of the features of each point set to zero so
// the file is sparse).  Checks: the file round-trips through the reader, the
// solver ends inside the box with its gap bound met, the model file parses
// with consistent counts, and the model's decision values sum_i y_i a_i K_ij + b
// (through Z) classify the training points (accuracy >= 0.95).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "darray_hip/dbarrier.h"
#include "darray_hip/read_input.h"

int main(int argc, char** argv) {
    const int n = 4096, d = 8, k = 128, p = 16, qq = 1;
    const double gamma = 0.25, C = 10.0;
    const std::string dir = argc > 1 ? argv[1] : ".";
    const std::string data = dir + "/svm_pipeline.libsvm", model = dir + "/svm_pipeline.model";
    std::vector<double> xs((size_t)d * n), ys(n);
    {
        std::mt19937 gen(17);
        std::normal_distribution<double> nd(0.0, 0.6);
        std::uniform_real_distribution<double> u(0.0, 1.0);
        std::ofstream f(data);
        f.precision(17);
        for (int j = 0; j < n; ++j) {
            const int cls = (j % 3 == 0) ? 1 : -1;  // unbalanced classes
            ys[j] = cls;
            f << (cls > 0 ? "+1" : "-1");
            for (int t = 0; t < d; ++t) {
                double v = (t < 2 ? 1.2 * cls : 0.0) + nd(gen);
                if (u(gen) < 0.3) v = 0.0;
                xs[t + (size_t)j * d] = v;
                if (v != 0.0) f << ' ' << t + 1 << ':' << v;
            }
            f << '\n';
        }
    }
    int bad = 0;
    std::vector<double> x((size_t)d * n), y(n);
    std::string fname = data;
    read_input_file<double>(&fname[0], n, d, x.data(), d, y.data());
    double rd = 0.0;
    for (size_t e = 0; e < x.size(); ++e) rd = std::max(rd, std::fabs(x[e] - xs[e]));
    for (int j = 0; j < n; ++j) rd = std::max(rd, std::fabs(y[j] - ys[j]));
    std::printf("read_input_file: max |difference| %.3e\n", rd);
    bad += rd != 0.0;

    DArray::Grid g;
    DArray::DMatrix<double> X(g, d, n), Y(g, 1, n), Om(g, n, k + p), Q(g, n, k + p), W(g, n, k + p);
    DArray::DMatrix<double> Z(g, n, k), A(g, n, 1);
    X.set_value_from_matrix(x.data(), d);
    Y.set_value_from_matrix(y.data(), 1);
    Om.set_normal_seed(0, 1, 1);
    int st = DArray::randomized_range_finder(g, n, d, k + p, qq, X, &Y, Om, Q, W, gamma,
                                             DArray::Precision::Mixed);
    bad += st != 0;
    st = DArray::nystrom_factor(g, n, k + p, k, Q, W, Z);
    bad += st != 0;
    double ratio = DArray::Knorm(g, n, d, k, X, Y, Z, (float)gamma);
    bad += !(ratio < 0.05);

    auto stats = DArray::barrier_svm(g, n, k, Z, Y, C, A);
    std::printf("barrier: %d barrier iterations, %d Newton steps, %d CG iterations, f(a) = %.8e\n",
                stats.barrier_iterations, stats.newton_steps, stats.cg_iterations, stats.f);
    const double* a = A.host_read();
    const double* z = Z.host_read();
    double amin = 1e300, amax = -1e300;
    for (int i = 0; i < n; ++i) {
        amin = std::min(amin, a[i]);
        amax = std::max(amax, a[i]);
    }
    std::printf("a in [%.3e, %.3e] (C = %g)\n", amin, amax, C);
    bad += !(amin > 0.0 && amax < C);

    DArray::LMatrix<double> la(n, 1), ldz(n, k), ldy(n, 1), ldx(d, n);
    for (int i = 0; i < n; ++i) {
        la(i, 0) = a[i];
        ldy(i, 0) = y[i];
        for (int c = 0; c < k; ++c) ldz(i, c) = z[i + (size_t)c * Z.lld()];
        for (int f = 0; f < d; ++f) ldx(f, i) = x[f + (size_t)i * d];
    }
    DArray::writemodel<double>(model, n, d, k, la, C, ldz, ldy, ldx, gamma, 0, 0);

    // Parse the model back: header fields and the support vectors.
    std::ifstream mf(model);
    std::string line, key;
    double rho = 0.0;
    int total = -1, npos = -1, nneg = -1, lines = 0;
    bool sv = false;
    while (std::getline(mf, line)) {
        if (sv) {
            ++lines;
            continue;
        }
        std::istringstream ls(line);
        ls >> key;
        if (key == "rho") ls >> rho;
        if (key == "total_sv") ls >> total;
        if (key == "nr_sv") ls >> npos >> nneg;
        if (key == "SV") sv = true;
    }
    std::printf("model: total_sv %d, nr_sv %d %d, %d SV lines, rho %.6e\n", total, npos, nneg,
                lines, rho);
    bad += !(total == lines && npos + nneg == total && total > 0);

    // Decision values through Z: y_j Z_j . (Z^T (a o 1)) - rho ... with
    // K_ij ~ y_i y_j (Z Z^T)_ij: sum_i a_i y_i K_ij = y_j sum_i a_i (Z Z^T)_ij.
    std::vector<double> v(k, 0.0);
    for (int c = 0; c < k; ++c)
        for (int i = 0; i < n; ++i) v[c] += z[i + (size_t)c * Z.lld()] * a[i];
    int correct = 0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int c = 0; c < k; ++c) s += z[j + (size_t)c * Z.lld()] * v[c];
        const double dec = y[j] * s - rho;
        correct += (dec > 0) == (y[j] > 0);
    }
    const double acc = (double)correct / n;
    std::printf("training accuracy %.4f\n", acc);
    bad += !(acc >= 0.95);
    std::printf("%s\n", bad ? "FAIL" : "PASS");
    return bad ? 1 : 0;
}
