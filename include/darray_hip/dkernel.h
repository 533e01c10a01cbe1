// DArray's RBF low-rank entry points on one MI355X.
//
// Same declarations as DArray's dkernel.h:6-8 (the host-pointer drop-ins
// rbf_kernel / tc_gemm / tc_sgemm, exported with C++ linkage by
// libdarray_hip.so) and the templates LRA (dkernel.h:12-46) and Knorm
// (dkernel.h:49-76), which here run the fused device kernels on the DMatrix
// device copies instead of materialising K in NN x NN host tiles.
// Differences from the reference (bug fixes):
//   * LRA fills all of A (the reference returns after the first tile,
//     dkernel.h:43) and accumulates every column chunk (beta, dkernel.h:41);
//   * Knorm prints the true ||K||_F and ||K - Z Z^T||_F (the reference adds
//     per-tile norms and keeps only the last tile's residual, dkernel.h:67-72).
#pragma once

#include "darray.h"

extern void rbf_kernel(int rank, int m, int n, int d, float* Xi, int ldxi, float* Xj, int ldxj,
                       float* Yi, float* Yj, float* K, int ldk, float gamma);
extern void tc_gemm(int rank, int m, int k, int n, float* A, int lda, float* BB, int ldb, float* C,
                    int ldc, float alpha = 1, float beta = 0);
extern void tc_sgemm(int rank, char transA, char transB, int m, int k, int n, float* A, int lda,
                     float* BB, int ldb, float* C, int ldc, float alpha = 1, float beta = 0);
extern void cuda_init();
extern void cuda_finalize();

namespace DArray {

// Arithmetic of the K-passes: FP64 MFMA throughout; Mixed = FP16 MFMA with
// FP32 accumulation (DH_PREC_F16: points centred, scaled by a power of two and
// rounded once; the mixed mode whose approximation error stays within 1.1x of
// FP64); BF16 (DH_PREC_BF16, same kernels with 3 fewer mantissa bits); FP8
// (DH_PREC_FP8: e4m3 distance, BF16 projection).  Above 256 features every
// mode runs the FP64 passes (the 16-bit and FP8 kernels hold d <= 256).
enum class Precision {
    FP64 = DH_PREC_F64,
    Mixed = DH_PREC_F16,
    BF16 = DH_PREC_BF16,
    FP8 = DH_PREC_FP8
};

inline Precision& lra_precision() {
    static thread_local Precision p = Precision::FP64;
    return p;
}
// Precision used by LRA when none is passed.
inline void set_lra_precision(Precision p) { lra_precision() = p; }

namespace detail {

// Read-only FP64 device image of a DMatrix block: the device copy itself for
// double, a converted copy for float.
template <typename T>
struct F64In;
template <>
struct F64In<double> {
    const double* p;
    int ld;
    explicit F64In(DMatrix<double>& M) : p(M.device_read()), ld(M.lld()) {}
};
template <>
struct F64In<float> {
    DeviceBuffer<double> buf;
    const double* p;
    int ld;
    explicit F64In(DMatrix<float>& M)
        : buf((size_t)std::max(M.dims()[0], 1) * std::max(M.dims()[1], 1)), p(buf.get()),
          ld(std::max(M.dims()[0], 1)) {
        check(dh_convert_f32_f64(handle(), M.dims()[0], M.dims()[1], M.device_read(), M.lld(),
                                 buf.get(), ld),
              "convert to FP64");
    }
};

// Writable FP64 device image of a DMatrix block; commit() stores it back.
template <typename T>
struct F64Out;
template <>
struct F64Out<double> {
    double* p;
    int ld;
    explicit F64Out(DMatrix<double>& M) : p(M.device_data()), ld(M.lld()) {}
    void commit() {}
};
template <>
struct F64Out<float> {
    DMatrix<float>& M;
    DeviceBuffer<double> buf;
    double* p;
    int ld;
    explicit F64Out(DMatrix<float>& Mo)
        : M(Mo), buf((size_t)std::max(Mo.dims()[0], 1) * std::max(Mo.dims()[1], 1)), p(buf.get()),
          ld(std::max(Mo.dims()[0], 1)) {}
    void commit() {
        check(dh_convert_f64_f32(handle(), M.dims()[0], M.dims()[1], p, ld, M.device_data(),
                                 M.lld()),
              "convert to FP32");
    }
};

}  // namespace detail

// A[0:n, 0:k] = (y y^T o K) * O[0:n, 0:k], K[i,j] = exp(-gamma |x_i - x_j|^2),
// for X d x n (one point per column), labels Y 1 x n, O and A n x k.
template <typename T>
void LRA(Grid g, int n, int d, int k, DMatrix<T> X, DMatrix<T>& Y, DMatrix<T>& O, DMatrix<T>& A,
         float gamma, Precision prec) {
    (void)g;
    detail::F64In<T> x(X), y(Y), o(O);
    detail::F64Out<T> a(A);
    int st;
    switch (prec) {
        case Precision::FP64:
            st = dh_rbf_project_f64(handle(), n, n, d, k, x.p, x.ld, x.p, x.ld, y.p, y.p, gamma,
                                    o.p, o.ld, 0.0, a.p, a.ld);
            break;
        case Precision::Mixed:
            st = dh_rbf_project_f16(handle(), n, n, d, k, x.p, x.ld, x.p, x.ld, y.p, y.p, gamma,
                                    o.p, o.ld, a.p, a.ld);
            break;
        case Precision::BF16:
            st = dh_rbf_project_bf16(handle(), n, n, d, k, x.p, x.ld, x.p, x.ld, y.p, y.p, gamma,
                                     o.p, o.ld, a.p, a.ld);
            break;
        default:
            st = dh_rbf_project_fp8(handle(), n, n, d, k, x.p, x.ld, x.p, x.ld, y.p, y.p, gamma,
                                    o.p, o.ld, a.p, a.ld);
            break;
    }
    check(st, "LRA");
    a.commit();
}

template <typename T>
void LRA(Grid g, int n, int d, int k, DMatrix<T> X, DMatrix<T>& Y, DMatrix<T>& O, DMatrix<T>& A,
         float gamma) {
    LRA(g, n, d, k, X, Y, O, A, gamma, lra_precision());
}

// Prints ||K||_F, ||K - Z Z^T||_F and their ratio for the labelled kernel
// (y y^T o K) of X (d x n), Z n x k, in the reference's format; returns the ratio.
template <typename T>
double Knorm(Grid g, int n, int d, int k, DMatrix<T>& X, DMatrix<T>& Y, DMatrix<T>& Z, float gamma) {
    detail::F64In<T> x(X), y(Y), z(Z);
    double e2[2] = {0.0, 0.0};
    check(dh_lra_error(handle(), n, d, k, x.p, x.ld, y.p, gamma, z.p, z.ld, z.p, z.ld, e2), "Knorm");
    double knorm = std::sqrt(std::max(e2[1], 0.0)), zknorm = std::sqrt(std::max(e2[0], 0.0));
    double ratio = knorm > 0.0 ? zknorm / knorm : 0.0;
    if (g.rank() == 0)
        std::printf("fnorm(K)=%f fnorm(K-U*U)=%f Residual fnorm=%f\n", knorm, zknorm, ratio);
    return ratio;
}

// ---- not in the reference: the whole range finder of testsvm.cpp:86-120, fused.
// Omega n x l; Q, W n x l receive the orthonormal basis and W = K Q (B = W^T).
// Returns 0 or a CholeskyQR breakdown code.
inline int randomized_range_finder(Grid g, int n, int d, int l, int q, DMatrix<double>& X,
                                   DMatrix<double>* Y, DMatrix<double>& Omega, DMatrix<double>& Q,
                                   DMatrix<double>& W, double gamma,
                                   Precision prec = Precision::FP64) {
    (void)g;
    const double* y = Y ? Y->device_read() : nullptr;
    int st = dh_lra(handle(), (int)prec, n, d, l, q, X.device_read(), X.lld(), y, gamma,
                    Omega.device_read(), Omega.lld(), Q.device_data(), Q.lld(), W.device_data(),
                    W.lld());
    if (st < 0) check(st, "randomized_range_finder");
    return st;
}

// ||K - Q W^T||_F / ||K||_F in FP64 with K regenerated on the device.
inline double lra_error(Grid g, int n, int d, int l, DMatrix<double>& X, DMatrix<double>* Y,
                        DMatrix<double>& Q, DMatrix<double>& W, double gamma) {
    (void)g;
    double e2[2] = {0.0, 0.0};
    const double* y = Y ? Y->device_read() : nullptr;
    check(dh_lra_error(handle(), n, d, l, X.device_read(), X.lld(), y, gamma, Q.device_read(),
                       Q.lld(), W.device_read(), W.lld(), e2),
          "lra_error");
    return e2[1] > 0.0 ? std::sqrt(std::max(e2[0], 0.0) / e2[1]) : 0.0;
}

}  // namespace DArray
