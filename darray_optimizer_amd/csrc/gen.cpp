// Seeded host-side input generators (no GPU use).
//
// DArray fills a column-major local block row-outer / column-inner from one
// engine (set_uniform darray.h:513-522, set_normal_seed darray.h:535-544); the
// same loop order and the same libstdc++ engines and distributions are used
// here so a DArray program seeded the same way sees the same arrays.
#include <cstdint>
#include <random>
#include <vector>

#include "../../include/darray_hip.h"

extern "C" int dh_mt19937_raw(uint32_t seed, int count, uint32_t* out) {
    if (count < 0) return -2;
    if (count > 0 && !out) return -3;
    std::mt19937 g(seed);
    for (int i = 0; i < count; ++i) out[i] = (uint32_t)g();
    return 0;
}

extern "C" int dh_fill_uniform(uint32_t seed, double a, double b, int m, int n, double* A,
                               int lda) {
    if (m < 0) return -4;
    if (n < 0) return -5;
    if (lda < (m > 1 ? m : 1)) return -7;
    if ((size_t)m * n > 0 && !A) return -6;
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> dist(a, b);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) A[i + (size_t)j * lda] = dist(g);
    return 0;
}

extern "C" int dh_fill_normal(uint32_t seed, double mean, double stddev, int m, int n, double* A,
                              int lda) {
    if (m < 0) return -4;
    if (n < 0) return -5;
    if (lda < (m > 1 ? m : 1)) return -7;
    if ((size_t)m * n > 0 && !A) return -6;
    std::mt19937 g(seed);
    std::normal_distribution<double> dist(mean, stddev);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) A[i + (size_t)j * lda] = dist(g);
    return 0;
}

extern "C" int dh_fill_mixture(uint64_t seed, int d, int n, int ncentres, double sigma, double* X,
                               int ldx, double* labels, double* centres) {
    if (d < 1) return -2;
    if (n < 0) return -3;
    if (ncentres < 1) return -4;
    if (ldx < d) return -7;
    if (n > 0 && !X) return -6;
    std::mt19937_64 g(seed);
    std::vector<double> c((size_t)d * ncentres);
    std::uniform_real_distribution<double> ud(-1.0, 1.0);
    for (int k = 0; k < ncentres; ++k)
        for (int f = 0; f < d; ++f) c[f + (size_t)k * d] = ud(g);
    std::vector<int> assign(n > 0 ? n : 1);
    std::uniform_int_distribution<int> cd(0, ncentres - 1);
    for (int i = 0; i < n; ++i) assign[i] = cd(g);
    std::normal_distribution<double> nd(0.0, sigma);
    for (int i = 0; i < n; ++i) {
        const double* ci = &c[(size_t)assign[i] * d];
        double* xi = X + (size_t)i * ldx;
        for (int f = 0; f < d; ++f) xi[f] = ci[f] + nd(g);
    }
    if (labels) {
        for (int i = 0; i < n; ++i) {
            const double* xi = X + (size_t)i * ldx;
            int best = 0;
            double bd = 0.0;
            for (int k = 0; k < ncentres; ++k) {
                double s = 0.0;
                for (int f = 0; f < d; ++f) {
                    double t = xi[f] - c[f + (size_t)k * d];
                    s += t * t;
                }
                if (k == 0 || s < bd) {
                    bd = s;
                    best = k;
                }
            }
            labels[i] = (best % 2 == 0) ? 1.0 : -1.0;
        }
    }
    if (centres)
        for (size_t e = 0; e < c.size(); ++e) centres[e] = c[e];
    return 0;
}
