// LIBSVM-format input and output of the kernel-SVM path (host code, no GPU).
//
// Reader: read_input_file (read_input.h:8-38) -- one point per line,
// "<label> <index>:<value> ...", label 1 -> +1, anything else -> -1,
// 1-based feature indices into the column of a d x n point-per-column X.
// Writer: writemodel (dbarrier.h:394-470) -- the LIBSVM "c_svc" / "rbf" model
// file with the support vectors of a dual solution a (0 <= a_i <= C) and the
// bias estimated from the free support vectors.
//
// Deviations from the reference (bug fixes): feature indices above d are
// reported instead of written past the column (read_input.h:29-31; indices
// <= 0 are skipped as there), a token that is not index:value is reported, ldx
// is honoured (the reference indexes with d), X and y are zeroed first (the
// reference's caller does it, testsvm.cpp:62-69); the model's bias uses the
// kernel entry (Z Z^T)_ij between support vector i and free vector j (the
// reference sums the diagonal (Z Z^T)_ii, dbarrier.h:428-434), and the header
// carries the class labels "1 -1" and per-class support-vector counts with
// the support vectors grouped by class, as LIBSVM's svm_load_model reads
// them (the reference writes the class sizes and the free / bound counts,
// dbarrier.h:458-459).
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/darray_hip.h"

int dh_fail(int code, const char* fmt, ...);

namespace {

struct File {
    FILE* f;
    explicit File(FILE* p) : f(p) {}
    ~File() {
        if (f) fclose(f);
    }
};

}  // namespace

extern "C" int dh_read_libsvm(const char* path, int n, int d, double* X, int ldx, double* y) {
    const char* name = "dh_read_libsvm";
    if (!path) return dh_fail(-1, "%s: path is NULL", name);
    if (n < 0) return dh_fail(-2, "%s: n < 0", name);
    if (d < 0) return dh_fail(-3, "%s: d < 0", name);
    if (n > 0 && d > 0 && !X) return dh_fail(-4, "%s: X is NULL", name);
    if (ldx < (d > 1 ? d : 1)) return dh_fail(-5, "%s: ldx < d", name);
    if (n > 0 && !y) return dh_fail(-6, "%s: y is NULL", name);
    File in(std::fopen(path, "r"));
    if (!in.f) return dh_fail(-1, "%s: cannot open %s: %s", name, path, std::strerror(errno));
    for (int j = 0; j < n; ++j) {
        y[j] = 0.0;
        for (int f = 0; f < d; ++f) X[f + (size_t)j * ldx] = 0.0;
    }
    std::string line;
    int j = 0;
    for (; j < n; ++j) {
        // one line, any length
        line.clear();
        int c;
        bool any = false;
        while ((c = std::fgetc(in.f)) != EOF) {
            any = true;
            if (c == '\n') break;
            line.push_back((char)c);
        }
        if (!any) break;  // end of file: fewer than n points
        // tokens are separated by blanks (spaces, tabs, a trailing CR)
        const char* s = line.c_str();
        auto skip = [&]() {
            while (*s == ' ' || *s == '\t' || *s == '\r') ++s;
        };
        skip();
        char* end = nullptr;
        double label = std::strtod(s, &end);
        if (end == s) label = 0.0;  // empty or non-numeric: atof gives 0 -> class -1
        y[j] = (label == 1.0) ? 1.0 : -1.0;
        s = end == s ? s : end;
        for (;;) {
            while (*s && *s != ' ' && *s != '\t' && *s != '\r') ++s;  // rest of this token
            skip();
            if (!*s) break;
            char* colon = nullptr;
            long index = std::strtol(s, &colon, 10);
            if (colon == s || *colon != ':') {
                if (*s == '#') break;  // comment to the end of the line
                return dh_fail(-1, "%s: %s:%d: malformed feature \"%.20s\"", name, path, j + 1, s);
            }
            const char* vs = colon + 1;
            char* vend = nullptr;
            double v = std::strtod(vs, &vend);
            if (vend == vs)
                return dh_fail(-1, "%s: %s:%d: malformed value \"%.20s\"", name, path, j + 1, s);
            if (index > d)
                return dh_fail(DH_ERR_UNSUPPORTED, "%s: %s:%d: feature index %ld > d = %d", name,
                               path, j + 1, index, d);
            if (index >= 1) X[(index - 1) + (size_t)j * ldx] = v;  // index <= 0 skipped, as the reference
            s = vend;
        }
    }
    return j;
}

extern "C" int dh_write_libsvm_model(const char* path, int n, int d, int k, const double* a,
                                     double C, const double* Z, int ldz, const double* y,
                                     const double* X, int ldx, double gamma, double* bias) {
    const char* name = "dh_write_libsvm_model";
    if (!path) return dh_fail(-1, "%s: path is NULL", name);
    if (n < 0) return dh_fail(-2, "%s: n < 0", name);
    if (d < 0) return dh_fail(-3, "%s: d < 0", name);
    if (k < 0) return dh_fail(-4, "%s: k < 0", name);
    if (n > 0 && (!a || !y || (k > 0 && !Z) || (d > 0 && !X)))
        return dh_fail(-5, "%s: NULL array", name);
    if (ldz < (n > 1 ? n : 1)) return dh_fail(-8, "%s: ldz < n", name);
    if (ldx < (d > 1 ? d : 1)) return dh_fail(-11, "%s: ldx < d", name);
    const double tiny = 1e-6;  // dbarrier.h:398-401
    std::vector<int> sv, free_sv;
    int npos = 0, nneg = 0;
    for (int i = 0; i < n; ++i) {
        if (a[i] > tiny) {
            sv.push_back(i);
            if (a[i] < C - tiny) free_sv.push_back(i);
            (y[i] > 0 ? npos : nneg) += 1;
        }
    }
    // b from up to 50 free support vectors j: y_j - sum_i a_i y_i K_ij with
    // K_ij ~ y_i y_j (Z Z^T)_ij, i.e. y_j (1 - sum_i a_i (Z Z^T)_ij).
    const int nb = std::min<int>((int)free_sv.size(), 50);
    double b = 0.0;
    std::vector<double> bs(nb);
    for (int t = 0; t < nb; ++t) {
        const int j = free_sv[t];
        double s = 0.0;
        for (int i : sv) {
            double zz = 0.0;
            for (int c = 0; c < k; ++c) zz += Z[i + (size_t)c * ldz] * Z[j + (size_t)c * ldz];
            s += a[i] * zz;
        }
        bs[t] = y[j] * (1.0 - s);
        b += bs[t];
    }
    if (nb > 0) b /= nb;
    if (bias) *bias = b;
    File out(std::fopen(path, "w"));
    if (!out.f) return dh_fail(-1, "%s: cannot create %s: %s", name, path, std::strerror(errno));
    std::fprintf(out.f, "svm_type c_svc\nkernel_type rbf\ngamma %.17g\nnr_class 2\ntotal_sv %d\n",
                 gamma, (int)sv.size());
    std::fprintf(out.f, "rho %.17g\nlabel 1 -1\nnr_sv %d %d\nSV\n", -b, npos, nneg);
    for (int cls = 0; cls < 2; ++cls)  // class +1 first, as "label 1 -1" orders them
        for (int i : sv) {
            if ((y[i] > 0) != (cls == 0)) continue;
            std::fprintf(out.f, "%.17g", y[i] * a[i]);
            const double* x = X + (size_t)i * ldx;
            for (int f = 0; f < d; ++f)
                if (x[f] != 0.0) std::fprintf(out.f, " %d:%.17g", f + 1, x[f]);
            std::fputc('\n', out.f);
        }
    if (std::ferror(out.f)) return dh_fail(-1, "%s: write to %s failed", name, path);
    return (int)sv.size();
}
