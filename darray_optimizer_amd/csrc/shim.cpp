// C++-linkage drop-ins for the entry points of DArray's CUDA object
// (include/dkernel.o built from dkernel.cu, linked by CMakeLists.txt:40-41,49-50):
//   rbf_kernel, tc_gemm, tc_sgemm  (declared dkernel.h:6-8)
//   cuda_init, cuda_finalize       (dkernel.cu:28-39; declared extern by testsvm.cpp:7-8)
// Same signatures and argument meaning, host pointers in and out, synchronous
// on return (callers read the results immediately, dkernel.h:35).  Device
// buffers are cached between calls instead of cudaMallocManaged per call.
//
// Intentional deviations (bug fixes): tc_gemm / tc_sgemm honour beta (the
// reference zeroes C on the device, dkernel.cu:201, 241) and are correct for
// any inner dimension (the reference overwrites C per 8192-wide inner block,
// dkernel.cu:204-223); tc_sgemm implements all four transpose cases; the RBF
// tile clamps squared distances at zero; rbf_kernel honours ldxi / ldxj.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/darray_hip.h"

namespace {

dh_handle_t g_handle = nullptr;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void* get(size_t want) {
        if (want <= bytes) return p;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, want) != hipSuccess) {
            fprintf(stderr, "GPUassert: hipMalloc of %zu bytes failed\n", want);
            exit(1);
        }
        bytes = want;
        return p;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

DevBuf g_buf[6];

// gpuAssert semantics (dkernel.cu:20-26): report and exit.
void check(int status, const char* where) {
    if (status != 0) {
        fprintf(stderr, "GPUassert: %s: %s (status %d)\n", where, dh_last_error_string(), status);
        exit(status < 0 ? 1 : status);
    }
}

void hcheck(hipError_t e, const char* where) {
    if (e != hipSuccess) {
        fprintf(stderr, "GPUassert: %s %s\n", hipGetErrorString(e), where);
        exit((int)e);
    }
}

dh_handle_t handle() {
    if (!g_handle) check(dh_create(0, nullptr, &g_handle), "cuda_init");
    return g_handle;
}

// Host m x n block (ld) -> device packed m x n.
template <typename T>
T* upload(DevBuf& b, const T* src, int m, int n, int ld) {
    T* d = (T*)b.get(sizeof(T) * (size_t)(m > 0 ? m : 1) * (n > 0 ? n : 1));
    if (m > 0 && n > 0)
        hcheck(hipMemcpy2D(d, sizeof(T) * m, src, sizeof(T) * (size_t)ld, sizeof(T) * m, n,
                           hipMemcpyHostToDevice),
               "upload");
    return d;
}

template <typename T>
void download(T* dst, int ld, const T* src, int m, int n) {
    if (m > 0 && n > 0)
        hcheck(hipMemcpy2D(dst, sizeof(T) * (size_t)ld, src, sizeof(T) * m, sizeof(T) * m, n,
                           hipMemcpyDeviceToHost),
               "download");
}

}  // namespace

void cuda_init() { (void)handle(); }

void cuda_finalize() {
    for (auto& b : g_buf) b.release();
    if (g_handle) dh_destroy(g_handle);
    g_handle = nullptr;
}

// Xi: m x d column-major (one point per row, ld ldxi); Xj: n x d (ld ldxj);
// K: m x n (ld ldk).  K[i,j] = Yi[i]*Yj[j]*exp(-gamma*|xi - xj|^2).
void rbf_kernel(int rank, int m, int n, int d, float* Xi, int ldxi, float* Xj, int ldxj,
                float* Yi, float* Yj, float* K, int ldk, float gamma) {
    (void)rank;
    dh_handle_t h = handle();
    if (m <= 0 || n <= 0) return;
    // The device kernels take points as columns (d x m); transpose on the host.
    std::vector<float> xi((size_t)d * m), xj((size_t)d * n);
    for (int f = 0; f < d; ++f)
        for (int i = 0; i < m; ++i) xi[f + (size_t)i * d] = Xi[i + (size_t)f * ldxi];
    for (int f = 0; f < d; ++f)
        for (int j = 0; j < n; ++j) xj[f + (size_t)j * d] = Xj[j + (size_t)f * ldxj];
    float* dxi = upload(g_buf[0], xi.data(), d, m, d);
    float* dxj = upload(g_buf[1], xj.data(), d, n, d);
    float* dyi = Yi ? upload(g_buf[2], Yi, m, 1, m) : nullptr;
    float* dyj = Yj ? upload(g_buf[3], Yj, n, 1, n) : nullptr;
    float* dk = (float*)g_buf[4].get(sizeof(float) * (size_t)m * n);
    check(dh_rbf_tile_f32(h, m, n, d, dxi, d, dxj, d, dyi, dyj, gamma, dk, m), "rbf_kernel");
    check(dh_synchronize(h), "rbf_kernel");
    download(K, ldk, dk, m, n);
}

// C[m x k] = alpha * A[m x n] * B[n x k] + beta * C.
void tc_gemm(int rank, int m, int k, int n, float* A, int lda, float* BB, int ldb, float* C,
             int ldc, float alpha, float beta) {
    (void)rank;
    dh_handle_t h = handle();
    if (m <= 0 || k <= 0) return;
    float* da = upload(g_buf[0], A, m, n, lda);
    float* db = upload(g_buf[1], BB, n, k, ldb);
    float* dc = beta != 0.0f ? upload(g_buf[4], C, m, k, ldc)
                             : (float*)g_buf[4].get(sizeof(float) * (size_t)m * k);
    check(dh_gemm_f32(h, 'N', 'N', m, k, n, alpha, da, m > 0 ? m : 1, db, n > 0 ? n : 1, beta, dc,
                      m),
          "tc_gemm");
    check(dh_synchronize(h), "tc_gemm");
    download(C, ldc, dc, m, k);
}

// C[m x k] = alpha * op(A) * op(B) + beta * C, op(A) m x n, op(B) n x k.
void tc_sgemm(int rank, char transA, char transB, int m, int k, int n, float* A, int lda,
              float* BB, int ldb, float* C, int ldc, float alpha, float beta) {
    (void)rank;
    dh_handle_t h = handle();
    if (m <= 0 || k <= 0) return;
    bool ta = (transA == 'T' || transA == 't'), tb = (transB == 'T' || transB == 't');
    int ar = ta ? n : m, ac = ta ? m : n;
    int br = tb ? k : n, bc = tb ? n : k;
    float* da = upload(g_buf[0], A, ar, ac, lda);
    float* db = upload(g_buf[1], BB, br, bc, ldb);
    float* dc = beta != 0.0f ? upload(g_buf[4], C, m, k, ldc)
                             : (float*)g_buf[4].get(sizeof(float) * (size_t)m * k);
    check(dh_gemm_f32(h, ta ? 'T' : 'N', tb ? 'T' : 'N', m, k, n, alpha, da, ar > 0 ? ar : 1, db,
                      br > 0 ? br : 1, beta, dc, m),
          "tc_sgemm");
    check(dh_synchronize(h), "tc_sgemm");
    download(C, ldc, dc, m, k);
}
