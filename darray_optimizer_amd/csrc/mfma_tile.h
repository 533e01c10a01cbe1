// MFMA tile accumulation shared by the GEMM, RBF-tile and error-pass kernels.
//
// v_mfma_f64_16x16x4_f64 and v_mfma_f32_16x16x4_f32 share the input layout:
// for a 16x16x4 block, lane L holds A[L%16][L/16] and B[L/16][L%16].  They
// differ in where the four accumulators of lane L sit: f32 holds
// D[4*(L/16)+r][L%16], f64 holds D[L/16+4*r][L%16], r = 0..3 (Mfma<T>::row).
#pragma once

#include <hip/hip_runtime.h>

typedef double dh_d4 __attribute__((ext_vector_type(4)));
typedef float dh_f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    using acc_t = dh_d4;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // Row of accumulator r of a lane in lane group g = L/16.
    static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};
template <>
struct Mfma<float> {
    using acc_t = dh_f4;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

constexpr int kTileBK = 16;
constexpr int kTilePad = 4;

// Accumulate op(A)[m0:m0+BM, kbeg:kend] * op(B)[kbeg:kend, n0:n0+BN] into acc
// with a 256-thread workgroup: waves in a 2 x 2 grid, each owning a
// (BM/2) x (BN/2) sub-tile as (BM/32) x (BN/32) MFMA blocks.  K advances in
// steps of 16 staged through LDS as op(A)^T and op(B) (k-major rows), so each
// MFMA operand read is 16 consecutive LDS elements; the global loads of step
// k+1 are issued before the MFMAs of step k.  Out-of-range elements are
// staged as zero.  Caller provides sA[16][BM+4] and sB[16][BN+4].
//   op(A)[i][k] = TA ? A[k + i*lda] : A[i + k*lda]
//                 (SYMA: A symmetric, upper triangle stored, op(A)[i][k] = A[min][max])
//   op(B)[k][j] = TB ? B[j + k*ldb] : B[k + j*ldb]
template <typename T, int BM, int BN, bool TA, bool TB, bool SYMA = false>
__device__ __forceinline__ void tile_mma(typename Mfma<T>::acc_t (&acc)[BM / 32][BN / 32],
                                         T (*sA)[BM + kTilePad], T (*sB)[BN + kTilePad], int m,
                                         int n, int m0, int n0, int kbeg, int kend,
                                         const T* __restrict__ A, int lda,
                                         const T* __restrict__ B, int ldb) {
    constexpr int BK = kTileBK;
    constexpr int FM = BM / 32, FN = BN / 32;
    constexpr int PA = BM * BK / 256, PB = BN * BK / 256;  // staged elements per thread
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int g = lane >> 4, li = lane & 15;
    T ra[PA], rb[PB];
    // Element u of this thread's share of the A / B tile at k-offset k0.
    auto a_at = [&](int u, int& i, int& kk) {
        int e = tid + u * 256;
        if (!TA) { i = e % BM; kk = e / BM; } else { kk = e % BK; i = e / BK; }
    };
    auto b_at = [&](int u, int& j, int& kk) {
        int e = tid + u * 256;
        if (!TB) { kk = e % BK; j = e / BK; } else { j = e % BN; kk = e / BN; }
    };
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < PA; ++u) {
            int i, kk;
            a_at(u, i, kk);
            int gi = m0 + i, gk = k0 + kk;
            // SYMA: A is symmetric with only its upper triangle stored.
            ra[u] = (gi < m && gk < kend)
                        ? (SYMA ? (gi <= gk ? A[gi + (size_t)gk * lda] : A[gk + (size_t)gi * lda])
                                : (TA ? A[gk + (size_t)gi * lda] : A[gi + (size_t)gk * lda]))
                        : T(0);
        }
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            int j, kk;
            b_at(u, j, kk);
            int gj = n0 + j, gk = k0 + kk;
            rb[u] = (gj < n && gk < kend)
                        ? (TB ? B[gj + (size_t)gk * ldb] : B[gk + (size_t)gj * ldb])
                        : T(0);
        }
    };
    if (kbeg >= kend) return;
    fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
#pragma unroll
        for (int u = 0; u < PA; ++u) {
            int i, kk;
            a_at(u, i, kk);
            sA[kk][i] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            int j, kk;
            b_at(u, j, kk);
            sB[kk][j] = rb[u];
        }
        __syncthreads();
        // The next tile's global loads are in flight while this one computes.
        if (k0 + BK < kend) fetch(k0 + BK);
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            T av[FM], bv[FN];
#pragma unroll
            for (int a = 0; a < FM; ++a) av[a] = sA[4 * ks + g][wm * (BM / 2) + a * 16 + li];
#pragma unroll
            for (int b = 0; b < FN; ++b) bv[b] = sB[4 * ks + g][wn * (BN / 2) + b * 16 + li];
#pragma unroll
            for (int a = 0; a < FM; ++a)
#pragma unroll
                for (int b = 0; b < FN; ++b) acc[a][b] = Mfma<T>::op(av[a], bv[b], acc[a][b]);
        }
        __syncthreads();
    }
}

// Row / column of accumulator element (a, b, r) for this lane within the tile.
template <typename T, int BM, int BN>
__device__ __forceinline__ int tile_row(int a, int r) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    return (wave & 1) * (BM / 2) + a * 16 + Mfma<T>::row(lane >> 4, r);
}
template <int BM, int BN>
__device__ __forceinline__ int tile_col(int b) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    return (wave >> 1) * (BN / 2) + b * 16 + (lane & 15);
}

template <typename T, int FM, int FN>
__device__ __forceinline__ void zero_acc(typename Mfma<T>::acc_t (&acc)[FM][FN]) {
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[a][b] = typename Mfma<T>::acc_t{0, 0, 0, 0};
}
