// Nystrom factor of the range finder (the consumer step of testsvm.cpp:126-177)
// and the one-sided Jacobi SVD it runs on.
//
// Reference: C = Q^T (K Q) by tc_gemm (testsvm.cpp:126-137), its singular
// values by sgesvd_ on the host (DArray::svd, lapack_like.h:388-413), then
// Z = Q diag(S) (testsvm.cpp:146-172) -- which drops the singular vectors.
// Here C = (Q^T W + W^T Q)/2 with W = K Q from dh_lra, its eigenpairs from a
// one-sided (Hestenes) Jacobi SVD on the device in FP64, and
// Z = Q V_k diag(sqrt(max(lam_k, 0))), so Z Z^T = Q V_k Lam_k V_k^T Q^T is
// the rank-k truncation of Q Q^T K Q Q^T.
//
// Jacobi: all column pairs of one round of a round-robin tournament are
// disjoint, so a round is one launch with a wave per pair (three dot products
// by wave reduction, then the rotation of the pair's columns of A and V).
// A device flag records whether a sweep rotated anything; the rounds of later
// sweeps return at once after a sweep that did not, and the host reads the
// flag once per sweep.
#include <algorithm>
#include <cmath>

#include "dh_common.h"

namespace {

constexpr double kJacobiTol = 1e-14;  // rotate when |a_p.a_q| > tol |a_p| |a_q|
constexpr int kMaxSweeps = 60;

__device__ __forceinline__ int rr_player(int pos, int round, int N) {
    return pos == 0 ? 0 : 1 + (pos - 1 + round) % (N - 1);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// flags[0]: a pair rotated in the current sweep; flags[1]: converged.
__global__ __launch_bounds__(64) void jacobi_round_kernel(int m, int n, int N, int round,
                                                          double* __restrict__ A, int lda,
                                                          double* __restrict__ V, int ldv,
                                                          double tol, int* __restrict__ flags) {
    if (flags[1]) return;
    int p = rr_player(blockIdx.x, round, N), q = rr_player(N - 1 - blockIdx.x, round, N);
    if (p >= n || q >= n) return;  // the padding player of an odd n
    if (p > q) {
        const int t = p;
        p = q;
        q = t;
    }
    double* ap = A + (size_t)p * lda;
    double* aq = A + (size_t)q * lda;
    const int lane = threadIdx.x;
    double al = 0.0, be = 0.0, ga = 0.0;
    for (int e = lane; e < m; e += 64) {
        const double x = ap[e], y = aq[e];
        al += x * x;
        be += y * y;
        ga += x * y;
    }
    al = wave_sum(al);
    be = wave_sum(be);
    ga = wave_sum(ga);
    if (!(fabs(ga) > tol * sqrt(al * be))) return;
    // Rutishauser's rotation: the smaller root t of t^2 + 2 zeta t - 1 = 0.
    const double zeta = (be - al) / (2.0 * ga);
    const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
    for (int e = lane; e < m; e += 64) {
        const double x = ap[e], y = aq[e];
        ap[e] = c * x - s * y;
        aq[e] = s * x + c * y;
    }
    double* vp = V + (size_t)p * ldv;
    double* vq = V + (size_t)q * ldv;
    for (int e = lane; e < n; e += 64) {
        const double x = vp[e], y = vq[e];
        vp[e] = c * x - s * y;
        vq[e] = s * x + c * y;
    }
    if (lane == 0) atomicOr(&flags[0], 1);
}

// End of a sweep: converged if nothing rotated; clear the rotation flag.
__global__ void jacobi_sweep_end_kernel(int* __restrict__ flags, int* __restrict__ sweeps) {
    if (threadIdx.x != 0 || flags[1]) return;
    sweeps[0] += 1;
    if (flags[0] == 0) flags[1] = 1;
    flags[0] = 0;
}

__global__ void jacobi_init_flags_kernel(int* __restrict__ flags, int* __restrict__ sweeps) {
    if (threadIdx.x == 0) {
        flags[0] = 0;
        flags[1] = 0;
        sweeps[0] = 0;
    }
}

// V = I (n x n).
__global__ __launch_bounds__(256) void identity_kernel(int n, double* __restrict__ V, int ldv) {
    const size_t total = (size_t)n * n;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (size_t)gridDim.x * 256) {
        const int i = (int)(e % n), j = (int)(e / n);
        V[i + (size_t)j * ldv] = (i == j) ? 1.0 : 0.0;
    }
}

// key[j] = |A[:, j]| (mode 0) or V[:, j] . A[:, j] (mode 1: the signed
// eigenvalue when A = C V for symmetric C); one wave per column.
__global__ __launch_bounds__(256) void column_key_kernel(int m, int n, const double* __restrict__ A,
                                                         int lda, const double* __restrict__ V,
                                                         int ldv, int mode,
                                                         double* __restrict__ key) {
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (j >= n) return;
    const double* a = A + (size_t)j * lda;
    double s = 0.0;
    if (mode == 0) {
        for (int e = lane; e < m; e += 64) s += a[e] * a[e];
    } else {
        const double* v = V + (size_t)j * ldv;
        for (int e = lane; e < m; e += 64) s += v[e] * a[e];
    }
    s = wave_sum(s);
    if (lane == 0) key[j] = mode == 0 ? sqrt(s) : s;
}

// ord[rank] = j for keys in descending order, ties by column index (one
// workgroup; rank by counting, n <= a few thousand).
__global__ __launch_bounds__(1024) void rank_desc_kernel(int n, const double* __restrict__ key,
                                                         int* __restrict__ ord) {
    for (int j = threadIdx.x; j < n; j += 1024) {
        const double kj = key[j];
        int r = 0;
        for (int i = 0; i < n; ++i) {
            const double ki = key[i];
            r += (ki > kj) || (ki == kj && i < j);
        }
        ord[r] = j;
    }
}

// dst[:, j] = scale_j * src[:, ord[j]] for j < cols (m rows); scale_j =
// sqrt(max(key[ord[j]], 0)) when sqrt_scale, else 1; keys_out[j] = key[ord[j]].
__global__ __launch_bounds__(256) void gather_columns_kernel(
    int m, int cols, const double* __restrict__ src, int lds, const int* __restrict__ ord,
    const double* __restrict__ key, int sqrt_scale, double* __restrict__ dst, int ldd,
    double* __restrict__ keys_out) {
    const size_t total = (size_t)m * cols;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (size_t)gridDim.x * 256) {
        const int i = (int)(e % m), j = (int)(e / m);
        const int o = ord[j];
        double sc = 1.0;
        if (sqrt_scale) {
            const double v = key[o];
            sc = v > 0.0 ? sqrt(v) : 0.0;
        }
        dst[i + (size_t)j * ldd] = sc * src[i + (size_t)o * lds];
        if (i == 0 && keys_out) keys_out[j] = key[o];
    }
}

// C = (C + C^T) / 2 on the n x n C (ld ldc).
__global__ __launch_bounds__(256) void symmetrize_kernel(int n, double* __restrict__ C, int ldc) {
    const size_t total = (size_t)n * n;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (size_t)gridDim.x * 256) {
        const int i = (int)(e % n), j = (int)(e / n);
        if (i >= j) continue;
        const double s = 0.5 * (C[i + (size_t)j * ldc] + C[j + (size_t)i * ldc]);
        C[i + (size_t)j * ldc] = s;
        C[j + (size_t)i * ldc] = s;
    }
}

__global__ __launch_bounds__(256) void copy_cols_kernel(int m, int n, const double* __restrict__ src,
                                                        int lds, double* __restrict__ dst, int ldd) {
    const size_t total = (size_t)m * n;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (size_t)gridDim.x * 256) {
        const int i = (int)(e % m), j = (int)(e / m);
        dst[i + (size_t)j * ldd] = src[i + (size_t)j * lds];
    }
}

int grid_for(size_t total) { return (int)std::min<size_t>((total + 255) / 256, 4096); }

// Jacobi sweeps on A (m x n) accumulating V (n x n, set to I here); flags and
// the sweep counter live in `words` (3 ints of device memory).  *sweeps_out
// receives the sweeps run, or -1 if kMaxSweeps did not converge.
int jacobi_sweeps(dh_handle_t h, int m, int n, double* A, int lda, double* V, int ldv,
                  int* words, int* sweeps_out) {
    hipStream_t s = h->stream;
    identity_kernel<<<grid_for((size_t)n * n), 256, 0, s>>>(n, V, ldv);
    DH_CHECK_LAUNCH();
    jacobi_init_flags_kernel<<<1, 64, 0, s>>>(words, words + 2);
    DH_CHECK_LAUNCH();
    *sweeps_out = 0;
    if (n < 2) return 0;
    int* host = dh_host_status(h);
    if (!host) return dh_fail(DH_ERR_ALLOC, "dh_gesvj: pinned status words unavailable");
    const int N = n + (n & 1);
    // Sweeps are queued back to back; the flags are read from sweep 6 on
    // (fewer never suffice in FP64 for n > 16), and once a sweep rotates no
    // pair the remaining queued rounds return at once.
    for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
        for (int r = 0; r < N - 1; ++r)
            jacobi_round_kernel<<<N / 2, 64, 0, s>>>(m, n, N, r, A, lda, V, ldv, kJacobiTol,
                                                     words);
        DH_CHECK_LAUNCH();
        jacobi_sweep_end_kernel<<<1, 64, 0, s>>>(words, words + 2);
        DH_CHECK_LAUNCH();
        if (sweep >= 5 || sweep == kMaxSweeps - 1) {
            DH_CHECK_HIP(hipMemcpyAsync(host, words, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
            DH_CHECK_HIP(hipStreamSynchronize(s));
            if (host[1]) {
                *sweeps_out = host[2];
                return 0;
            }
        }
    }
    *sweeps_out = -1;
    return 0;
}

}  // namespace

extern "C" int dh_gesvj_f64(dh_handle_t h, int m, int n, double* A, int lda, double* sva,
                            double* V, int ldv, int* sweeps) {
    const char* name = "dh_gesvj_f64";
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (m < 0) return dh_fail(-2, "%s: m < 0", name);
    if (n < 0 || n > m) return dh_fail(-3, "%s: need 0 <= n <= m", name);
    if (lda < std::max(1, m)) return dh_fail(-5, "%s: lda < m", name);
    if (ldv < std::max(1, n)) return dh_fail(-8, "%s: ldv < n", name);
    if (sweeps) *sweeps = 0;
    if (n == 0) return 0;
    DH_CHECK_HIP(hipSetDevice(h->device));
    const size_t mn = (size_t)m * n, nn = (size_t)n * n;
    char* ws = (char*)dh_ws(h, DH_WS_EIG, sizeof(double) * (mn + nn + n) + sizeof(int) * (n + 64));
    if (!ws) return DH_ERR_ALLOC;
    double* tA = (double*)ws;
    double* tV = tA + mn;
    double* key = tV + nn;
    int* ord = (int*)(key + n);
    int* words = ord + n;
    int sw = 0;
    DH_CHECK_STATUS(jacobi_sweeps(h, m, n, A, lda, V, ldv, words, &sw));
    if (sweeps) *sweeps = sw;
    // Singular values = column norms, columns of A and V permuted to descending order.
    column_key_kernel<<<dh_cdiv(n, 4), 256, 0, h->stream>>>(m, n, A, lda, V, ldv, 0, key);
    DH_CHECK_LAUNCH();
    rank_desc_kernel<<<1, 1024, 0, h->stream>>>(n, key, ord);
    DH_CHECK_LAUNCH();
    gather_columns_kernel<<<grid_for(mn), 256, 0, h->stream>>>(m, n, A, lda, ord, key, 0, tA, m,
                                                               sva);
    DH_CHECK_LAUNCH();
    gather_columns_kernel<<<grid_for(nn), 256, 0, h->stream>>>(n, n, V, ldv, ord, key, 0, tV, n,
                                                               nullptr);
    DH_CHECK_LAUNCH();
    copy_cols_kernel<<<grid_for(mn), 256, 0, h->stream>>>(m, n, tA, m, A, lda);
    DH_CHECK_LAUNCH();
    copy_cols_kernel<<<grid_for(nn), 256, 0, h->stream>>>(n, n, tV, n, V, ldv);
    DH_CHECK_LAUNCH();
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    if (sw < 0) return dh_fail(1, "%s: no convergence in %d sweeps", name, kMaxSweeps);
    return 0;
}

extern "C" int dh_nystrom_f64(dh_handle_t h, int n, int l, int k, const double* Q, int ldq,
                              const double* W, int ldw, double* Z, int ldz, double* lambda) {
    const char* name = "dh_nystrom_f64";
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (n < 0) return dh_fail(-2, "%s: n < 0", name);
    if (l < 0 || l > n) return dh_fail(-3, "%s: need 0 <= l <= n", name);
    if (k < 0 || k > l) return dh_fail(-4, "%s: need 0 <= k <= l", name);
    if (ldq < std::max(1, n)) return dh_fail(-6, "%s: ldq < n", name);
    if (ldw < std::max(1, n)) return dh_fail(-8, "%s: ldw < n", name);
    if (ldz < std::max(1, n)) return dh_fail(-10, "%s: ldz < n", name);
    if (k == 0 || n == 0) return 0;
    DH_CHECK_HIP(hipSetDevice(h->device));
    const size_t ll = (size_t)l * l;
    char* ws = (char*)dh_ws(h, DH_WS_EIG,
                            sizeof(double) * (2 * ll + (size_t)l * k + l) + sizeof(int) * (l + 64));
    if (!ws) return DH_ERR_ALLOC;
    double* C = (double*)ws;
    double* V = C + ll;
    double* Vk = V + ll;
    double* lam = Vk + (size_t)l * k;
    int* ord = (int*)(lam + l);
    int* words = ord + l;
    // C = Q^T W (testsvm.cpp:126-137), symmetrised.
    DH_CHECK_STATUS(dh_gemm_impl<double>(h, true, false, l, l, n, 1.0, Q, ldq, W, ldw, 0.0, C, l, 0));
    symmetrize_kernel<<<grid_for(ll), 256, 0, h->stream>>>(l, C, l);
    DH_CHECK_LAUNCH();
    int sw = 0;
    DH_CHECK_STATUS(jacobi_sweeps(h, l, l, C, l, V, l, words, &sw));
    // C now holds C0 V: lam_j = v_j . (C0 v_j); top k in descending order.
    column_key_kernel<<<dh_cdiv(l, 4), 256, 0, h->stream>>>(l, l, C, l, V, l, 1, lam);
    DH_CHECK_LAUNCH();
    rank_desc_kernel<<<1, 1024, 0, h->stream>>>(l, lam, ord);
    DH_CHECK_LAUNCH();
    // Vk = V[:, top k] diag(sqrt(max(lam, 0))); lambda receives the k eigenvalues.
    gather_columns_kernel<<<grid_for((size_t)l * k), 256, 0, h->stream>>>(l, k, V, l, ord, lam, 1,
                                                                          Vk, l, lambda);
    DH_CHECK_LAUNCH();
    DH_CHECK_STATUS(dh_gemm_impl<double>(h, false, false, n, k, l, 1.0, Q, ldq, Vk, l, 0.0, Z, ldz, 0));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    if (sw < 0) return dh_fail(1, "%s: Jacobi did not converge in %d sweeps", name, kMaxSweeps);
    return 0;
}
