// Dense building blocks on gfx950 MFMA: GEMM / SYRK (f32 and f64), 64 x 64
// diagonal-block Cholesky / inverse kernels, and the blocked drivers of
// the Cholesky and TRSM that DArray runs through BLAS/LAPACK:
//   multiply_matrix / matrix_multiply  (cblas_like.h:39-60, blas_like.h:62-93)
//   symmetric_update -> ssyrk_         (cblas_like.h:61-69)
//   triangular_solve                   (blas_like.h:95-128)
//   cholesky_factorize                 (lapack_like.h:9-56)
//
// One template serves both precisions: the f64 and f32 16x16x4 MFMAs share
// an operand layout (mfma_tile.h).
#include <cstdio>

#include "dh_common.h"
#include "mfma_tile.h"

// ------------------------------------------------------------------ GEMM
//
// C tile BM x BN per 256-thread workgroup (tile_mma in mfma_tile.h).
// blockIdx.z selects a k-range of `kchunk`; with `part` != nullptr each
// split writes its raw product to part[z] (m x n, ld m) and a second kernel
// reduces.  uplo_upper skips tiles strictly below the diagonal and never
// writes an element with i > j (SYRK semantics).  `stop` (nullable) is a
// device status word: nonzero makes the kernel return at once.
template <typename T, int BM, int BN, bool TA, bool TB, bool SYMA = false>
__global__ __launch_bounds__(256) void gemm_kernel(int m, int n, int k, T alpha,
                                                    const T* __restrict__ A, int lda,
                                                    const T* __restrict__ B, int ldb, T beta,
                                                    T* __restrict__ C, int ldc, int kchunk,
                                                    int uplo_upper, T* __restrict__ part,
                                                    const int* __restrict__ stop) {
    constexpr int FM = BM / 32, FN = BN / 32;
    __shared__ T sA[kTileBK][BM + kTilePad];
    __shared__ T sB[kTileBK][BN + kTilePad];
    if (stop && *stop) return;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    if (uplo_upper && m0 >= n0 + BN) return;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(k, kbeg + kchunk);
    typename Mfma<T>::acc_t acc[FM][FN];
    zero_acc<T, FM, FN>(acc);
    tile_mma<T, BM, BN, TA, TB, SYMA>(acc, sA, sB, m, n, m0, n0, kbeg, kend, A, lda, B, ldb);
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int i = m0 + tile_row<T, BM, BN>(a, r);
                int j = n0 + tile_col<BM, BN>(b);
                if (i < m && j < n && (!uplo_upper || i <= j)) {
                    T v = acc[a][b][r];
                    if (part) {
                        part[(size_t)blockIdx.z * m * n + i + (size_t)j * m] = v;
                    } else {
                        T* c = C + i + (size_t)j * ldc;
                        *c = (beta == T(0)) ? alpha * v : alpha * v + beta * (*c);
                    }
                }
            }
}

template <typename T>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(int m, int n, int nsplit,
                                                             const T* __restrict__ part, T alpha,
                                                             T beta, T* __restrict__ C, int ldc,
                                                             int uplo_upper,
                                                             const int* __restrict__ stop) {
    if (stop && *stop) return;
    size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t mn = (size_t)m * n;
    if (idx >= mn) return;
    int i = (int)(idx % m), j = (int)(idx / m);
    if (uplo_upper && i > j) return;
    T s = 0;
    for (int z = 0; z < nsplit; ++z) s += part[(size_t)z * mn + idx];
    T* c = C + i + (size_t)j * ldc;
    *c = (beta == T(0)) ? alpha * s : alpha * s + beta * (*c);
}

template <typename T, int BM, int BN>
static int launch_gemm(dh_handle_t h, bool ta, bool tb, int m, int n, int k, T alpha, const T* A,
                       int lda, const T* B, int ldb, T beta, T* C, int ldc, int uplo,
                       const int* stop) {
    int tm = dh_cdiv(m, BM), tn = dh_cdiv(n, BN);
    long tiles = (long)tm * tn;
    if (uplo) {  // tiles strictly below the diagonal return at once
        long below = 0;
        for (int a = 0; a < tm; ++a) below += max(0, min(tn, (a * BM) / BN));
        tiles -= below;
    }
    int nsplit = 1;
    // Tall-skinny products (CholeskyQR Gram, TRSM updates with short outputs)
    // have few output tiles and a long k: split k for ~4 workgroups per CU.
    if (tiles < 4L * h->num_cu && k >= 1024) {
        nsplit = (int)((4L * h->num_cu + tiles - 1) / tiles);
        nsplit = min(nsplit, k / 256);
        nsplit = max(nsplit, 1);
        nsplit = min(nsplit, 64);
    }
    int kchunk = dh_cdiv(dh_cdiv(k, nsplit), 16) * 16;
    if (kchunk <= 0) kchunk = 16;
    nsplit = dh_cdiv(k, kchunk);
    if (nsplit < 1) nsplit = 1;
    T* part = nullptr;
    if (nsplit > 1) {
        // GEMMs queued on the look-ahead stream run concurrently with those of
        // the handle's stream: each stream has its own partials slot (a slot
        // only grows after the stream that uses it has drained).
        const dh_slot slot = (h->aux && h->stream == h->aux) ? DH_WS_SPLITK_AUX : DH_WS_SPLITK;
        part = (T*)dh_ws(h, slot, sizeof(T) * (size_t)nsplit * m * n);
        if (!part) return DH_ERR_ALLOC;
    }
    dim3 grid(tm, tn, nsplit);
    if (!ta && !tb)
        gemm_kernel<T, BM, BN, false, false><<<grid, 256, 0, h->stream>>>(
            m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kchunk, uplo, part, stop);
    else if (!ta && tb)
        gemm_kernel<T, BM, BN, false, true><<<grid, 256, 0, h->stream>>>(
            m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kchunk, uplo, part, stop);
    else if (ta && !tb)
        gemm_kernel<T, BM, BN, true, false><<<grid, 256, 0, h->stream>>>(
            m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kchunk, uplo, part, stop);
    else
        gemm_kernel<T, BM, BN, true, true><<<grid, 256, 0, h->stream>>>(
            m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kchunk, uplo, part, stop);
    DH_CHECK_LAUNCH();
    if (nsplit > 1) {
        size_t mn = (size_t)m * n;
        splitk_reduce_kernel<T><<<dh_cdiv((long)mn, 256), 256, 0, h->stream>>>(
            m, n, nsplit, part, alpha, beta, C, ldc, uplo, stop);
        DH_CHECK_LAUNCH();
    }
    return 0;
}

template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(int m, int n, T alpha, T* __restrict__ C,
                                                     int ldc, int uplo_upper) {
    size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)m * n) return;
    int i = (int)(idx % m), j = (int)(idx / m);
    if (uplo_upper && i > j) return;
    T* c = C + i + (size_t)j * ldc;
    *c = (alpha == T(0)) ? T(0) : alpha * (*c);
}

// Internal GEMM with an optional device stop word.
template <typename T>
static int gemm_stop(dh_handle_t h, bool ta, bool tb, int m, int n, int k, T alpha, const T* A,
                     int lda, const T* B, int ldb, T beta, T* C, int ldc, int uplo,
                     const int* stop) {
    if (m == 0 || n == 0) return 0;
    if (k == 0 || alpha == T(0)) {
        if (beta == T(1)) return 0;
        scale_kernel<T><<<dh_cdiv((long)m * n, 256), 256, 0, h->stream>>>(m, n, beta, C, ldc,
                                                                           uplo);
        DH_CHECK_LAUNCH();
        return 0;
    }
    // FP64 keeps 64 x 64 tiles (90 registers, five waves per SIMD): its 128 x 128
    // tile (238 registers, two waves) measured slower in the blocked Cholesky
    // (FP64 posv at n = 16384: 57.3 ms against 51.6 ms).
    if (sizeof(T) == 4 && (long)m * n >= 2048L * 2048L)
        return launch_gemm<T, 128, 128>(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                                        uplo, stop);
    return launch_gemm<T, 64, 64>(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, uplo,
                                  stop);
}

template <typename T>
int dh_gemm_impl(dh_handle_t h, bool ta, bool tb, int m, int n, int k, T alpha, const T* A,
                 int lda, const T* B, int ldb, T beta, T* C, int ldc, int uplo_upper) {
    return gemm_stop<T>(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, uplo_upper,
                        nullptr);
}
template <typename T>
int dh_gemm_stop_impl(dh_handle_t h, bool ta, bool tb, int m, int n, int k, T alpha, const T* A,
                      int lda, const T* B, int ldb, T beta, T* C, int ldc, const int* stop) {
    return gemm_stop<T>(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0, stop);
}
template int dh_gemm_stop_impl<double>(dh_handle_t, bool, bool, int, int, int, double,
                                       const double*, int, const double*, int, double, double*,
                                       int, const int*);
template int dh_gemm_impl<double>(dh_handle_t, bool, bool, int, int, int, double, const double*,
                                  int, const double*, int, double, double*, int, int);
template int dh_gemm_impl<float>(dh_handle_t, bool, bool, int, int, int, float, const float*,
                                 int, const float*, int, float, float*, int, int);

// ------------------------------------------------------ small triangles

constexpr int kSmall = 64;

// Diagonal blocks are stored as Linv = U_bb^{-T} (lower triangular),
// column-major with ld 64, one 4096-element block per 64 rows of the
// triangle; blocks of order nb < 64 are padded with the identity.  With it
// the three solves of a diagonal block are products:
//   U^{-T} B = Linv B,   U^{-1} B = Linv^T B,   B U^{-1} = B Linv^T.

// Cholesky of the nb <= 64 diagonal block at A (upper triangle, in place)
// together with Linv.  Symmetric Gaussian elimination on [A | I]: step k
// subtracts m_i = A(k,i) / A(k,k) times row k from every row i > k, so the
// left part ends as D Ut (Ut unit upper, A = Ut^T D Ut) and the right part
// as Ut^{-T}; then U = D^{1/2} Ut and U^{-T} = D^{-1/2} Ut^{-T}.  Both parts
// share one 64 x 64 register grid over kPotrfThreads threads (NW waves):
// thread t holds row i = t % 64 at columns j = t / 64 + NW q (q < E = 64/NW),
// the left part where j >= i and the strictly lower right part where j < i.
// Row k is broadcast through LDS (double buffered: one barrier per step).
// The step is a latency chain (write, barrier, read, divide, update), so
// 512 threads -- 8 values each -- beat 256 with 16 (about 1300 against 1800
// cycles per step).  A non-positive (or NaN) pivot at local step k stores
// base + k + 1 in *info (first failure wins) and leaves A as it was.
constexpr int kPotrfThreads = 512;
template <typename T>
__global__ __launch_bounds__(kPotrfThreads) void potrf_inv_small_kernel(int nb,
                                                                         T* __restrict__ A,
                                                                         int lda,
                                                                         T* __restrict__ Linv,
                                                                         int* __restrict__ info,
                                                                         int base) {
    constexpr int NW = kPotrfThreads / 64, E = kSmall / NW;
    // Row k is broadcast permuted, column j at (j % NW) * E + j / NW, so the
    // E columns a thread holds are contiguous (vector LDS reads/writes).
    __shared__ T rows[2][kSmall];
    __shared__ T diag[kSmall];
    if (*info != 0) return;
    const int t = threadIdx.x, i = t & (kSmall - 1), w = t >> 6;
    const int pi = (i % NW) * E + i / NW;
    T a[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int j = w + NW * q;
        a[q] = (i < nb && j < nb) ? (i <= j ? A[i + (size_t)j * lda] : T(0))
                                  : (i == j ? T(1) : T(0));
    }
    int fail = 0;
    for (int k = 0; k < kSmall; ++k) {
        T* rb = rows[k & 1];
        if (i == k) {
#pragma unroll
            for (int q = 0; q < E; ++q) rb[w * E + q] = a[q];
        }
        __syncthreads();
        // All reads of the step at once; the pivot test comes after the
        // update (a failed factorisation writes nothing back).
        const T piv = rb[(k % NW) * E + k / NW];
        const T rbi = rb[pi];
        T rj[E];
#pragma unroll
        for (int q = 0; q < E; ++q) rj[q] = rb[w * E + q];
        // Branch-free update (selects keep a[] in place): rows i <= k keep
        // their values.
        const T mi = rbi / piv;
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int j = w + NW * q;
            const T f = (j >= i || j < k) ? rj[q] : (j == k ? T(1) : T(0));
            const T upd = a[q] - mi * f;
            a[q] = (i > k) ? upd : a[q];
        }
        if (!(piv > T(0))) {  // the same value in every thread: a uniform exit
            fail = k + 1;
            break;
        }
        if (t == 0) diag[k] = piv;
    }
    if (fail) {
        if (t == 0) atomicCAS(info, 0, base + fail);
        return;
    }
    __syncthreads();
    const T si = sqrt(diag[i]);
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int j = w + NW * q;
        if (i <= j && j < nb) A[i + (size_t)j * lda] = a[q] / si;
        // Linv(i, j) = Ut^{-T}(i, j) / sqrt(d_i), unit diagonal of Ut^{-T}.
        Linv[i + kSmall * j] = (j < i) ? a[q] / si : (j == i ? T(1) / si : T(0));
    }
}

// Linv blocks of an existing upper triangle of order m: workgroup b inverts
// diagonal block b (order nb = min(64, m - 64 b), unit diagonal if `unit`)
// by back substitution, thread j solving U_bb x = e_j, and stores the
// transpose through LDS so the global stores are coalesced.
template <typename T>
__global__ __launch_bounds__(64) void inv_upper_blocks_kernel(int m, const T* __restrict__ A,
                                                               int lda, T* __restrict__ Linv,
                                                               int unit,
                                                               const int* __restrict__ stop) {
    __shared__ T s[kSmall][kSmall + 1];
    if (stop && *stop) return;
    const int b = blockIdx.x, nb = min(kSmall, m - kSmall * b);
    const T* Ab = A + (size_t)kSmall * b * (1 + (size_t)lda);
#pragma unroll 16
    for (int u = 0; u < kSmall; ++u) {
        const int e = threadIdx.x + 64 * u;
        const int i = e % kSmall, c = e / kSmall;
        s[i][c] = (i <= c && c < nb) ? Ab[i + (size_t)c * lda] : T(0);
    }
    __syncthreads();
    const int j = threadIdx.x;
    T x[kSmall];
#pragma unroll
    for (int i = kSmall - 1; i >= 0; --i) {
        T acc = (i == j) ? T(1) : T(0);
#pragma unroll
        for (int k = i + 1; k < kSmall; ++k) acc -= s[i][k] * x[k];
        // Rows past nb are the identity padding.
        x[i] = (i < nb && j < nb) ? (unit ? acc : acc / s[i][i]) : acc;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSmall; ++i) s[i][j] = x[i];  // s = U_bb^{-1}
    __syncthreads();
    T* Lb = Linv + (size_t)4096 * b;
#pragma unroll 16
    for (int u = 0; u < kSmall; ++u) {
        const int e = threadIdx.x + 64 * u;
        const int r = e % kSmall, c = e / kSmall;
        Lb[e] = s[c][r];
    }
}

// Leaf of the blocked triangular solve, C = op(A) op(B) on one slab of at
// most 64 rows (side L) or 64 columns (side R), in place: C is the same
// memory as B (side L) or A (side R).  That is safe because the slab's
// 64-wide dimension is one tile: each workgroup reads its whole input tile
// (tile_mma ends in a barrier after its last load) before writing the same
// tile, and no other workgroup touches it.  The pointers are deliberately
// not __restrict__.
template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(256) void trsm_leaf_kernel(int m, int n, int k, const T* A, int lda,
                                                         const T* B, int ldb, T* C, int ldc,
                                                         const int* stop) {
    __shared__ T sA[kTileBK][kSmall + kTilePad];
    __shared__ T sB[kTileBK][kSmall + kTilePad];
    if (stop && *stop) return;
    const int m0 = blockIdx.x * kSmall, n0 = blockIdx.y * kSmall;
    typename Mfma<T>::acc_t acc[2][2];
    zero_acc<T, 2, 2>(acc);
    tile_mma<T, kSmall, kSmall, TA, TB>(acc, sA, sB, m, n, m0, n0, 0, k, A, lda, B, ldb);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = m0 + tile_row<T, kSmall, kSmall>(a, r);
                const int j = n0 + tile_col<kSmall, kSmall>(b);
                if (i < m && j < n) C[i + (size_t)j * ldc] = acc[a][b][r];
            }
}

// B slab := (Linv block) applied as the solve with diagonal block `blk`;
// `order` <= 64 is the block's order, `count` the other dimension.
template <typename T>
static int trsm_leaf(dh_handle_t h, char side, char trans, int order, int count, const T* Lb,
                     T* B, int ldb, const int* stop) {
    if (side == 'L') {
        dim3 grid(1, dh_cdiv(count, kSmall));
        if (trans == 'T')
            trsm_leaf_kernel<T, false, false><<<grid, 256, 0, h->stream>>>(
                order, count, order, Lb, kSmall, B, ldb, B, ldb, stop);
        else
            trsm_leaf_kernel<T, true, false><<<grid, 256, 0, h->stream>>>(
                order, count, order, Lb, kSmall, B, ldb, B, ldb, stop);
    } else {
        dim3 grid(dh_cdiv(count, kSmall), 1);
        trsm_leaf_kernel<T, false, true><<<grid, 256, 0, h->stream>>>(
            count, order, order, B, ldb, Lb, kSmall, B, ldb, stop);
    }
    DH_CHECK_LAUNCH();
    return 0;
}

// --------------------------------------------------- block-sweep TRSM
//
// For few right-hand sides the recursion's leaves and updates are too
// small to fill the GPU, so side-L solves with up to kSweepMaxCols columns
// run as a sweep over the 64-row blocks, one launch per block:
//   forward  (U^T X = B): X_i = Linv_i B_i,    B_j -= U_ij^T X_i  (j > i)
//   backward (U X = B):   X_i = Linv_i^T B_i,  B_j -= U_ji X_i    (j < i)
// Workgroup (b, c) owns block j = i +- b and the 64 columns of tile c; every
// workgroup recomputes X_i for its tile (64^3 multiply-adds, cheaper than a
// second launch), workgroup (0, c) stores it to X, and the others update
// their B_j.  B_i itself is read again by no later launch, so nothing
// waits on a store made in the same launch.  X is copied back to B at the
// end.

constexpr int kSweepMaxCols = 512;
constexpr int kLdsPad = kSmall + 1;

// acc += op(sA) sB over k = 0..63, both operands in LDS:
// op(sA)[r][k] = TA ? sA[k][r] : sA[r][k].  Waves in a 2 x 2 grid, each
// owning a 32 x 32 quarter as 2 x 2 MFMA blocks (as tile_mma).
template <typename T, bool TA>
__device__ __forceinline__ void lds_mma64(typename Mfma<T>::acc_t (&acc)[2][2],
                                          const T (*sA)[kLdsPad], const T (*sB)[kLdsPad]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1, g = lane >> 4, li = lane & 15;
#pragma unroll 4
    for (int ks = 0; ks < kSmall / 4; ++ks) {
        const int k = 4 * ks + g;
        T av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int r = wm * 32 + a * 16 + li;
            av[a] = TA ? sA[k][r] : sA[r][k];
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) bv[b] = sB[k][wn * 32 + b * 16 + li];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = Mfma<T>::op(av[a], bv[b], acc[a][b]);
    }
}

template <typename T, bool BACK>
__global__ __launch_bounds__(256) void trsm_sweep_kernel(int m, int n, int i,
                                                          const T* __restrict__ U, int ldu,
                                                          const T* __restrict__ Linv,
                                                          T* __restrict__ B, int ldb,
                                                          T* __restrict__ X, int ldx,
                                                          const int* __restrict__ stop) {
    __shared__ T sA[kSmall][kLdsPad];  // Linv_i
    __shared__ T sB[kSmall][kLdsPad];  // B_i, then X_i
    __shared__ T sU[kSmall][kLdsPad];  // coupling block
    if (stop && *stop) return;
    const int b = blockIdx.x, c0 = blockIdx.y * kSmall;
    const int j = BACK ? i - b : i + b;
    const int r0 = kSmall * i, rj = kSmall * j;
    const T* Lb = Linv + (size_t)4096 * i;
    // Every operand is loaded up front (one memory round trip per launch):
    // Linv_i (identity-padded), B_i (zero past row m / column n), the
    // coupling block -- forward U_ij (rows of block i, columns of block j,
    // used transposed), backward U_ji -- and this workgroup's B_j tile in
    // the accumulator layout.  All loads are issued into registers first,
    // from addresses clamped into the matrices, and stored to LDS after.
    T la[16], lb[16], lu[16], bj[2][2][4];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int e = threadIdx.x + 256 * u;
        const int r = e % kSmall, c = e / kSmall;
        la[u] = Lb[e];
        lb[u] = B[min(r0 + r, m - 1) + (size_t)min(c0 + c, n - 1) * ldb];
    }
    if (b != 0) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = threadIdx.x + 256 * u;
            const int r = e % kSmall, c = e / kSmall;
            const int ur = min(BACK ? rj + r : r0 + r, m - 1);
            const int uc = min(BACK ? r0 + c : rj + c, m - 1);
            lu[u] = U[ur + (size_t)uc * ldu];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = min(rj + tile_row<T, kSmall, kSmall>(a, r), m - 1);
                    const int col = min(c0 + tile_col<kSmall, kSmall>(bb), n - 1);
                    bj[a][bb][r] = B[row + (size_t)col * ldb];
                }
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int e = threadIdx.x + 256 * u;
        const int r = e % kSmall, c = e / kSmall;
        sA[r][c] = la[u];
        sB[r][c] = (r0 + r < m && c0 + c < n) ? lb[u] : T(0);
        if (b != 0) {
            const bool ok = BACK ? (rj + r < m && r0 + c < m) : (r0 + r < m && rj + c < m);
            sU[r][c] = ok ? lu[u] : T(0);
        }
    }
    __syncthreads();
    typename Mfma<T>::acc_t acc[2][2];
    zero_acc<T, 2, 2>(acc);
    lds_mma64<T, BACK>(acc, sA, sB);
    __syncthreads();
    // X_i tile into sB (block row major, as the next product reads it).
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                sB[tile_row<T, kSmall, kSmall>(a, r)][tile_col<kSmall, kSmall>(bb)] =
                    acc[a][bb][r];
    __syncthreads();
    if (b == 0) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = threadIdx.x + 256 * u;
            const int r = e % kSmall, c = e / kSmall;
            if (r0 + r < m && c0 + c < n) X[(r0 + r) + (size_t)(c0 + c) * ldx] = sB[r][c];
        }
        return;
    }
    zero_acc<T, 2, 2>(acc);
    lds_mma64<T, !BACK>(acc, sU, sB);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = rj + tile_row<T, kSmall, kSmall>(a, r);
                const int col = c0 + tile_col<kSmall, kSmall>(bb);
                if (row < m && col < n) B[row + (size_t)col * ldb] = bj[a][bb][r] - acc[a][bb][r];
            }
}

template <typename T>
__global__ __launch_bounds__(256) void copy_cols_kernel(int m, int n, const T* __restrict__ S,
                                                         int lds, T* __restrict__ D, int ldd,
                                                         const int* __restrict__ stop) {
    if (stop && *stop) return;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)m * n) return;
    const int r = (int)(idx % m), c = (int)(idx / m);
    D[r + (size_t)c * ldd] = S[r + (size_t)c * lds];
}

// Side-L solve of U^T X = B (trans 'T') or U X = B (trans 'N'), U of order
// m with Linv blocks in D, B m x n with n <= kSweepMaxCols.
template <typename T>
static int trsm_sweep(dh_handle_t h, char trans, int m, int n, const T* A, int lda, const T* D,
                      T* B, int ldb, const int* stop) {
    const int nblk = dh_cdiv(m, kSmall), tiles = dh_cdiv(n, kSmall);
    T* X = (T*)dh_ws(h, DH_WS_TRSM_X, sizeof(T) * (size_t)m * n);
    if (!X) return DH_ERR_ALLOC;
    for (int s = 0; s < nblk; ++s) {
        if (trans == 'T') {
            const int i = s;
            trsm_sweep_kernel<T, false><<<dim3(nblk - i, tiles), 256, 0, h->stream>>>(
                m, n, i, A, lda, D, B, ldb, X, m, stop);
        } else {
            const int i = nblk - 1 - s;
            trsm_sweep_kernel<T, true><<<dim3(i + 1, tiles), 256, 0, h->stream>>>(
                m, n, i, A, lda, D, B, ldb, X, m, stop);
        }
        DH_CHECK_LAUNCH();
    }
    copy_cols_kernel<T><<<dh_cdiv((long)m * n, 256), 256, 0, h->stream>>>(m, n, X, m, B, ldb,
                                                                          stop);
    DH_CHECK_LAUNCH();
    return 0;
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }

// --------------------------------------------------------- TRSM drivers
//
// Recursive halving as triangular_solve (blas_like.h:113-122), split at a
// multiple of 64 so every leaf is one 64-row (64-column) diagonal block,
// and each leaf is an MFMA product with that block's Linv instead of a
// substitution.  `D` holds the Linv blocks of the whole triangle passed at
// the top; `blk` is the index of the current sub-triangle's first block.

template <typename T>
static int trsm_rec(dh_handle_t h, char side, char trans, int m, int n, const T* A, int lda,
                    T* B, int ldb, const T* D, int blk, const int* stop) {
    if (m == 0 || n == 0) return 0;
    const int order = side == 'L' ? m : n;
    if (order <= kSmall)
        return trsm_leaf<T>(h, side, trans, order, side == 'L' ? n : m, D + (size_t)4096 * blk, B,
                            ldb, stop);
    const int h1 = ((order / 2 + kSmall - 1) / kSmall) * kSmall, h2 = order - h1;
    const int b2 = blk + h1 / kSmall;
    if (side == 'L' && trans == 'T') {
        // U^T X = B: top half, update, bottom half.
        DH_CHECK_STATUS(trsm_rec<T>(h, side, trans, h1, n, A, lda, B, ldb, D, blk, stop));
        DH_CHECK_STATUS(gemm_stop<T>(h, true, false, h2, n, h1, T(-1), A + (size_t)h1 * lda, lda,
                                     B, ldb, T(1), B + h1, ldb, 0, stop));
        return trsm_rec<T>(h, side, trans, h2, n, A + h1 + (size_t)h1 * lda, lda, B + h1, ldb, D,
                           b2, stop);
    }
    if (side == 'L') {
        // U X = B: bottom half first.
        DH_CHECK_STATUS(trsm_rec<T>(h, side, trans, h2, n, A + h1 + (size_t)h1 * lda, lda, B + h1,
                                    ldb, D, b2, stop));
        DH_CHECK_STATUS(gemm_stop<T>(h, false, false, h1, n, h2, T(-1), A + (size_t)h1 * lda,
                                     lda, B + h1, ldb, T(1), B, ldb, 0, stop));
        return trsm_rec<T>(h, side, trans, h1, n, A, lda, B, ldb, D, blk, stop);
    }
    // side R, no transpose: X U = B, triangle of order n, B m x n.
    DH_CHECK_STATUS(trsm_rec<T>(h, side, trans, m, h1, A, lda, B, ldb, D, blk, stop));
    DH_CHECK_STATUS(gemm_stop<T>(h, false, false, m, h2, h1, T(-1), B, ldb, A + (size_t)h1 * lda,
                                 lda, T(1), B + (size_t)h1 * ldb, ldb, 0, stop));
    return trsm_rec<T>(h, side, trans, m, h2, A + h1 + (size_t)h1 * lda, lda,
                       B + (size_t)h1 * ldb, ldb, D, b2, stop);
}

// Triangular solve with the upper triangle A (order m for side L, n for
// side R): Linv of its diagonal blocks, then the recursion.
template <typename T>
static int trsm_blocked(dh_handle_t h, char side, char trans, bool unit, int m, int n, const T* A,
                        int lda, T* B, int ldb, const int* stop) {
    if (m == 0 || n == 0) return 0;
    const int order = side == 'L' ? m : n;
    const int nblk = dh_cdiv(order, kSmall);
    T* D = (T*)dh_ws(h, DH_WS_TRSM_INV, sizeof(T) * 4096 * (size_t)nblk);
    if (!D) return DH_ERR_ALLOC;
    inv_upper_blocks_kernel<T><<<nblk, 64, 0, h->stream>>>(order, A, lda, D, unit ? 1 : 0, stop);
    DH_CHECK_LAUNCH();
    if (side == 'L' && m > kSmall && n <= kSweepMaxCols)
        return trsm_sweep<T>(h, trans, m, n, A, lda, D, B, ldb, stop);
    return trsm_rec<T>(h, side, trans, m, n, A, lda, B, ldb, D, 0, stop);
}

template <typename T>
int dh_trsm_impl(dh_handle_t h, char side, char trans, bool unit, int m, int n, T alpha,
                 const T* A, int lda, T* B, int ldb) {
    if (m == 0 || n == 0) return 0;
    if (alpha != T(1)) {
        scale_kernel<T><<<dh_cdiv((long)m * n, 256), 256, 0, h->stream>>>(m, n, alpha, B, ldb, 0);
        DH_CHECK_LAUNCH();
        if (alpha == T(0)) return 0;
    }
    return trsm_blocked<T>(h, side, trans, unit, m, n, A, lda, B, ldb, nullptr);
}
template int dh_trsm_impl<double>(dh_handle_t, char, char, bool, int, int, double, const double*,
                                  int, double*, int);
template int dh_trsm_impl<float>(dh_handle_t, char, char, bool, int, int, float, const float*,
                                 int, float*, int);

// --------------------------------------------------------- Cholesky driver

// Factor one diagonal block of order b (<= any): right-looking in 64-wide
// steps entirely on the device, leaving the Linv of each 64-block in D;
// failures land in *dinfo with their global index.
template <typename T>
static int potrf_block(dh_handle_t h, int b, T* A11, int lda, T* D, int* dinfo, int base) {
    for (int ii = 0; ii < b; ii += kSmall) {
        int bb = min(kSmall, b - ii);
        T* Aii = A11 + ii + (size_t)ii * lda;
        T* Lb = D + (size_t)4096 * (ii / kSmall);
        potrf_inv_small_kernel<T><<<1, kPotrfThreads, 0, h->stream>>>(bb, Aii, lda, Lb, dinfo,
                                                                     base + ii);
        DH_CHECK_LAUNCH();
        if (ii + bb < b) {
            T* A12 = A11 + ii + (size_t)(ii + bb) * lda;
            T* A22 = A11 + (ii + bb) + (size_t)(ii + bb) * lda;
            DH_CHECK_STATUS(trsm_leaf<T>(h, 'L', 'T', bb, b - ii - bb, Lb, A12, lda, dinfo));
            DH_CHECK_STATUS(gemm_stop<T>(h, true, false, b - ii - bb, b - ii - bb, bb, T(-1), A12,
                                         lda, A12, lda, T(1), A22, lda, 1, dinfo));
        }
    }
    return 0;
}

template <typename T>
int dh_potrf_impl(dh_handle_t h, int n, T* A, int lda, int bs, int* info_host) {
    *info_host = 0;
    if (n == 0) return 0;
    if (bs <= 0) bs = 256;
    int* dinfo = (int*)dh_ws(h, DH_WS_INFO, 64);
    if (!dinfo) return DH_ERR_ALLOC;
    // Linv blocks of the current panel's diagonal block.
    T* D = (T*)dh_ws(h, DH_WS_TRSM_INV, sizeof(T) * 4096 * (size_t)dh_cdiv(min(bs, n), kSmall));
    if (!D) return DH_ERR_ALLOC;
    set_int_kernel<<<1, 1, 0, h->stream>>>(dinfo, 0);
    DH_CHECK_LAUNCH();
    // cholesky_factorize (lapack_like.h:17-52): potrf(A11); trsm("L","U","T") on
    // A12; upper SYRK A22 -= A12^T A12.  The SYRK is split with a look-ahead
    // of one panel: the rows of the next panel are updated on the handle's
    // stream, which then factors that panel (latency-bound one-workgroup
    // kernels and its panel TRSM) while the rest of the trailing matrix is
    // updated on a second stream.
    //   stream:  ... potrf/trsm(p), record P | wait U(p-1), update rows of p+1 | potrf/trsm(p+1) ...
    //   aux:     wait P, update the rest after p, record U(p) | ...
    // Regions: the aux update writes rows past panel p+1; panel p+1's
    // factorisation reads and writes rows of p+1 only, and U12 of panel p is
    // no longer written by anyone.
    const bool lookahead = n > bs;
    if (lookahead) {
        DH_CHECK_STATUS(dh_aux_stream(h));
        DH_CHECK_HIP(hipEventRecord(h->ev_panel, h->stream));  // aux after prior work
        DH_CHECK_HIP(hipStreamWaitEvent(h->aux, h->ev_panel, 0));
    }
    bool pending = false;  // an aux update of an earlier panel was queued
    for (int i = 0; i < n; i += bs) {
        int b = min(bs, n - i);
        T* A11 = A + i + (size_t)i * lda;
        DH_CHECK_STATUS(potrf_block<T>(h, b, A11, lda, D, dinfo, i));
        if (i + b >= n) break;
        T* A12 = A + i + (size_t)(i + b) * lda;
        T* A22 = A + (i + b) + (size_t)(i + b) * lda;
        const int nt = n - i - b;
        DH_CHECK_STATUS(trsm_rec<T>(h, 'L', 'T', b, nt, A11, lda, A12, lda, D, 0, dinfo));
        if (!lookahead) {
            DH_CHECK_STATUS(gemm_stop<T>(h, true, false, nt, nt, b, T(-1), A12, lda, A12, lda,
                                         T(1), A22, lda, 1, dinfo));
            continue;
        }
        const int b2 = min(bs, nt), nr = nt - b2;
        DH_CHECK_HIP(hipEventRecord(h->ev_panel, h->stream));
        if (pending) DH_CHECK_HIP(hipStreamWaitEvent(h->stream, h->ev_update, 0));
        // Rows of the next panel, all trailing columns (upper part only).
        DH_CHECK_STATUS(gemm_stop<T>(h, true, false, b2, nt, b, T(-1), A12, lda, A12, lda, T(1),
                                     A22, lda, 1, dinfo));
        if (nr > 0) {
            DH_CHECK_HIP(hipStreamWaitEvent(h->aux, h->ev_panel, 0));
            hipStream_t main = h->stream;
            h->stream = h->aux;  // gemm_stop launches on h->stream
            const int st = gemm_stop<T>(h, true, false, nr, nr, b, T(-1), A12 + (size_t)b2 * lda,
                                        lda, A12 + (size_t)b2 * lda, lda, T(1),
                                        A22 + b2 + (size_t)b2 * lda, lda, 1, dinfo);
            h->stream = main;
            DH_CHECK_STATUS(st);
            DH_CHECK_HIP(hipEventRecord(h->ev_update, h->aux));
            pending = true;
        }
    }
    if (pending) DH_CHECK_HIP(hipStreamWaitEvent(h->stream, h->ev_update, 0));
    int hinfo = 0;
    DH_CHECK_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    *info_host = hinfo;
    return 0;
}
template int dh_potrf_impl<double>(dh_handle_t, int, double*, int, int, int*);
template int dh_potrf_impl<float>(dh_handle_t, int, float*, int, int, int*);

// C = alpha * A * B + beta * C with A (m x m) symmetric, upper triangle stored.
template <typename T>
int dh_symm_upper_impl(dh_handle_t h, int m, int n, T alpha, const T* A, int lda, const T* B,
                       int ldb, T beta, T* C, int ldc) {
    if (m == 0 || n == 0) return 0;
    constexpr int BM = 64, BN = 64;
    dim3 grid(dh_cdiv(m, BM), dh_cdiv(n, BN), 1);
    gemm_kernel<T, BM, BN, false, false, true><<<grid, 256, 0, h->stream>>>(
        m, n, m, alpha, A, lda, B, ldb, beta, C, ldc, m, 0, nullptr, nullptr);
    DH_CHECK_LAUNCH();
    return 0;
}
template int dh_symm_upper_impl<double>(dh_handle_t, int, int, double, const double*, int,
                                        const double*, int, double, double*, int);

// ------------------------------------------------------------- C ABI

template <typename T>
static int gemm_api(const char* name, dh_handle_t h, char ta, char tb, int m, int n, int k,
                    T alpha, const T* A, int lda, const T* B, int ldb, T beta, T* C, int ldc) {
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (!dh_is_n(ta) && !dh_is_t(ta)) return dh_fail(-2, "%s: bad transa '%c'", name, ta);
    if (!dh_is_n(tb) && !dh_is_t(tb)) return dh_fail(-3, "%s: bad transb '%c'", name, tb);
    if (m < 0) return dh_fail(-4, "%s: m < 0", name);
    if (n < 0) return dh_fail(-5, "%s: n < 0", name);
    if (k < 0) return dh_fail(-6, "%s: k < 0", name);
    int arows = dh_is_t(ta) ? k : m, brows = dh_is_t(tb) ? n : k;
    if (lda < max(1, arows)) return dh_fail(-9, "%s: lda too small", name);
    if (ldb < max(1, brows)) return dh_fail(-11, "%s: ldb too small", name);
    if (ldc < max(1, m)) return dh_fail(-14, "%s: ldc too small", name);
    return dh_gemm_impl<T>(h, dh_is_t(ta), dh_is_t(tb), m, n, k, alpha, A, lda, B, ldb, beta, C,
                           ldc, 0);
}

extern "C" int dh_gemm_f64(dh_handle_t h, char ta, char tb, int m, int n, int k, double alpha,
                           const double* A, int lda, const double* B, int ldb, double beta,
                           double* C, int ldc) {
    return gemm_api<double>("dh_gemm_f64", h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C,
                            ldc);
}
extern "C" int dh_gemm_f32(dh_handle_t h, char ta, char tb, int m, int n, int k, float alpha,
                           const float* A, int lda, const float* B, int ldb, float beta, float* C,
                           int ldc) {
    return gemm_api<float>("dh_gemm_f32", h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C,
                           ldc);
}

template <typename T>
static int syrk_api(const char* name, dh_handle_t h, int n, int k, T alpha, const T* A, int lda,
                    T beta, T* C, int ldc) {
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (n < 0) return dh_fail(-2, "%s: n < 0", name);
    if (k < 0) return dh_fail(-3, "%s: k < 0", name);
    if (lda < max(1, k)) return dh_fail(-6, "%s: lda too small", name);
    if (ldc < max(1, n)) return dh_fail(-9, "%s: ldc too small", name);
    return dh_gemm_impl<T>(h, true, false, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, 1);
}
extern "C" int dh_syrk_ut_f64(dh_handle_t h, int n, int k, double alpha, const double* A, int lda,
                              double beta, double* C, int ldc) {
    return syrk_api<double>("dh_syrk_ut_f64", h, n, k, alpha, A, lda, beta, C, ldc);
}
extern "C" int dh_syrk_ut_f32(dh_handle_t h, int n, int k, float alpha, const float* A, int lda,
                              float beta, float* C, int ldc) {
    return syrk_api<float>("dh_syrk_ut_f32", h, n, k, alpha, A, lda, beta, C, ldc);
}

template <typename T>
static int trsm_api(const char* name, dh_handle_t h, char side, char uplo, char trans, char diag,
                    int m, int n, T alpha, const T* A, int lda, T* B, int ldb) {
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    char s = (side == 'l' ? 'L' : side == 'r' ? 'R' : side);
    if (s != 'L' && s != 'R') return dh_fail(-2, "%s: bad side '%c'", name, side);
    if (uplo != 'U' && uplo != 'u')
        return dh_fail(DH_ERR_UNSUPPORTED, "%s: only uplo='U' is implemented", name);
    if (!dh_is_n(trans) && !dh_is_t(trans)) return dh_fail(-4, "%s: bad transa '%c'", name, trans);
    if (diag != 'N' && diag != 'n' && diag != 'U' && diag != 'u')
        return dh_fail(-5, "%s: bad diag '%c'", name, diag);
    if (s == 'R' && dh_is_t(trans))
        return dh_fail(DH_ERR_UNSUPPORTED, "%s: side='R' with transa='T' is not implemented",
                       name);
    if (m < 0) return dh_fail(-6, "%s: m < 0", name);
    if (n < 0) return dh_fail(-7, "%s: n < 0", name);
    int ka = (s == 'L') ? m : n;
    if (lda < max(1, ka)) return dh_fail(-10, "%s: lda too small", name);
    if (ldb < max(1, m)) return dh_fail(-12, "%s: ldb too small", name);
    return dh_trsm_impl<T>(h, s, dh_is_t(trans) ? 'T' : 'N', diag == 'U' || diag == 'u', m, n,
                           alpha, A, lda, B, ldb);
}
extern "C" int dh_trsm_f64(dh_handle_t h, char side, char uplo, char trans, char diag, int m,
                           int n, double alpha, const double* A, int lda, double* B, int ldb) {
    return trsm_api<double>("dh_trsm_f64", h, side, uplo, trans, diag, m, n, alpha, A, lda, B,
                            ldb);
}
extern "C" int dh_trsm_f32(dh_handle_t h, char side, char uplo, char trans, char diag, int m,
                           int n, float alpha, const float* A, int lda, float* B, int ldb) {
    return trsm_api<float>("dh_trsm_f32", h, side, uplo, trans, diag, m, n, alpha, A, lda, B,
                           ldb);
}

template <typename T>
static int potrf_api(const char* name, dh_handle_t h, char uplo, int n, T* A, int lda, int bs) {
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (uplo != 'U' && uplo != 'u')
        return dh_fail(DH_ERR_UNSUPPORTED, "%s: only uplo='U' is implemented", name);
    if (n < 0) return dh_fail(-3, "%s: n < 0", name);
    if (lda < max(1, n)) return dh_fail(-5, "%s: lda too small", name);
    int info = 0;
    DH_CHECK_STATUS(dh_potrf_impl<T>(h, n, A, lda, bs, &info));
    if (info) dh_fail(info, "%s: leading minor of order %d is not positive definite", name, info);
    return info;
}
extern "C" int dh_potrf_f64(dh_handle_t h, char uplo, int n, double* A, int lda, int bs) {
    return potrf_api<double>("dh_potrf_f64", h, uplo, n, A, lda, bs);
}
extern "C" int dh_potrf_f32(dh_handle_t h, char uplo, int n, float* A, int lda, int bs) {
    return potrf_api<float>("dh_potrf_f32", h, uplo, n, A, lda, bs);
}
