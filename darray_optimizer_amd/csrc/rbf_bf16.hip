// Mixed-precision K-passes: BF16 MFMA distance and projection with FP32
// accumulation and an FP32 exp epilogue; FP64 in and out.
//
// Precision scheme (per pass):
//   * X is centred by the mean of the column point set (distances are
//     translation invariant) and rounded to BF16 once per solve;
//   * the squared norms are FP32 sums of the ROUNDED coordinates, so
//     |xi|^2 + |xj|^2 - 2 xi.xj is the squared distance of the rounded
//     points up to FP32 accumulation error;
//   * K = exp2(-gamma log2(e) d2) on v_exp_f32, clamped to [0, 1] and rounded
//     to BF16 in registers;
//   * Omega (or Q) is rounded to BF16 with the labels y_j folded in, stored
//     in the operand order of the projection MFMA ("packed");
//   * Y accumulates in FP32 and is written as FP64 for CholeskyQR2.
// Reference: DArray::LRA (dkernel.h:12-46) over rbf_kernel + tc_gemm
// (dkernel.cu:132-231), whose tensor-core path is an FP16 GemmEx with FP32
// accumulation (test_cuda.cu:111-132, dkernel.cu:161-163).
#include <cmath>
#include <cstring>

#include "dh_common.h"

// 16-byte groups of eight 16-bit values.  The same kernels serve BF16 and
// FP16 (template flag F16): the bits are stored as bf16x8 and reinterpreted
// for the FP16 MFMA.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr float kLog2e = 1.4426950408889634f;

template <bool F16>
__device__ __forceinline__ __bf16 to16(float x) {
    if constexpr (F16)
        return __builtin_bit_cast(__bf16, (_Float16)x);
    else
        return (__bf16)x;
}

template <bool F16>
__device__ __forceinline__ float from16(__bf16 v) {
    if constexpr (F16)
        return (float)__builtin_bit_cast(_Float16, v);
    else
        return (float)v;
}

template <bool F16>
__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
    if constexpr (F16)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                       __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ----------------------------------------------------------------- prep

// partial[b][f] = sum of X[f, i] over the points of slice b (FP64).
__global__ __launch_bounds__(256) void colsum_partial_kernel(int n, int d, const double* __restrict__ X,
                                                             int ldx, int slice,
                                                             double* __restrict__ partial) {
    int b = blockIdx.x;
    int i1 = min(n, (b + 1) * slice);
    for (int f = threadIdx.x; f < d; f += 256) {
        double s = 0.0;
        for (int i = b * slice; i < i1; ++i) s += X[f + (size_t)i * ldx];
        partial[(size_t)b * d + f] = s;
    }
}

__global__ __launch_bounds__(256) void colmean_final_kernel(int n, int d, int nb,
                                                            const double* __restrict__ partial,
                                                            double* __restrict__ mean) {
    for (int f = blockIdx.x * 256 + threadIdx.x; f < d; f += gridDim.x * 256) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += partial[(size_t)b * d + f];
        mean[f] = n > 0 ? s / n : 0.0;
    }
}

// One wave per point row (rows >= n are written as zeros):
// Xb = round16((X - mean) * scale), padded with zeros to dpad features, and
// a[i] = -gamma log2(e) |Xb_i|^2 (gamma of the scaled points).
template <bool F16>
__device__ __forceinline__ void convert_row(int row, int lane, int n, int d, int dpad,
                                            const double* __restrict__ X, int ldx,
                                            const double* __restrict__ mean, double scale,
                                            float neg_gl2e, __bf16* __restrict__ Xb,
                                            float* __restrict__ a) {
    float s = 0.f;
    for (int f = lane; f < dpad; f += 64) {
        __bf16 v = to16<F16>(0.0f);
        if (row < n && f < d) v = to16<F16>((float)((X[f + (size_t)row * ldx] - mean[f]) * scale));
        Xb[(size_t)row * dpad + f] = v;
        float fv = from16<F16>(v);
        s += fv * fv;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) a[row] = (row < n) ? neg_gl2e * s : 0.f;
}

template <bool F16>
__global__ __launch_bounds__(256) void convert_rows_kernel(int n, int nrows, int d, int dpad,
                                                           const double* __restrict__ X, int ldx,
                                                           const double* __restrict__ mean,
                                                           double scale, float neg_gl2e,
                                                           __bf16* __restrict__ Xb,
                                                           float* __restrict__ a) {
    int row = (blockIdx.x * 256 + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (row >= nrows) return;
    convert_row<F16>(row, lane, n, d, dpad, X, ldx, mean, scale, neg_gl2e, Xb, a);
}

// *maxabs = max |X[f, i] - mean[f]| as float bits (non-negative floats order
// as unsigned integers); *maxabs must start at 0.
__global__ __launch_bounds__(256) void maxabs_centred_kernel(int n, int d,
                                                             const double* __restrict__ X, int ldx,
                                                             const double* __restrict__ mean,
                                                             unsigned* __restrict__ maxabs) {
    float mx = 0.f;
    const long total = (long)n * d;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        int f = (int)(e % d), i = (int)(e / d);
        mx = fmaxf(mx, (float)fabs(X[f + (size_t)i * ldx] - mean[f]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) atomicMax(maxabs, __float_as_uint(mx));
}

__global__ void zero_word_kernel(unsigned* w) { *w = 0u; }

// Packed operand of the projection MFMA.  For chunk t (32 points), column
// block cb (16 columns) and lane L = 16 g + li, the 8 BF16 values are
//   Op[t][cb][L][e] = y[j] * O[j][16 cb + li],  j = 32 t + 16 (e / 4) + 4 g + e % 4,
// zero outside n x l.  One thread per (chunk, column): it reads 32 contiguous
// doubles of the column and writes four 16-byte groups.
template <bool F16>
__device__ __forceinline__ void pack_one(long idx, int n, int l, int nchunks, int ncb_total,
                                         const double* __restrict__ O, int ldo,
                                         const double* __restrict__ y, bf16x8* __restrict__ Op) {
    long ncols = (long)ncb_total * 16;
    if (idx >= (long)nchunks * ncols) return;
    int t = (int)(idx / ncols), c = (int)(idx % ncols);
    int cb = c >> 4, li = c & 15;
    float v[32];
#pragma unroll
    for (int jj = 0; jj < 32; ++jj) {
        int j = 32 * t + jj;
        double o = 0.0;
        if (j < n && c < l) {
            o = O[j + (size_t)c * ldo];
            if (y) o *= y[j];
        }
        v[jj] = (float)o;
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        bf16x8 w;
#pragma unroll
        for (int e = 0; e < 8; ++e) w[e] = to16<F16>(v[16 * (e >> 2) + 4 * g + (e & 3)]);
        Op[((size_t)t * ncb_total + cb) * 64 + 16 * g + li] = w;
    }
}

template <bool F16>
__global__ __launch_bounds__(256) void pack_omega_kernel(int n, int l, int nchunks, int ncb_total,
                                                         const double* __restrict__ O, int ldo,
                                                         const double* __restrict__ y,
                                                         bf16x8* __restrict__ Op) {
    pack_one<F16>((long)blockIdx.x * 256 + threadIdx.x, n, l, nchunks, ncb_total, O, ldo, y, Op);
}

// ------------------------------------------------------------ main kernel
//
// Workgroup = WAVES waves; wave w owns 32 rows i (two 16-row MFMA blocks).
// The j loop walks chunks of 32 points through a 3-deep LDS ring: while
// chunk t is projected, chunk t+1 is already resident (its distance tile is
// computed in the same step) and chunk t+2 is being staged, so one barrier
// per chunk suffices and the loop body is a single basic block (staging and
// look-ahead indices are clamped instead of branched).  Per chunk a wave:
//   1. S^T = Xj_chunk * Xi^T: 2 x 2 blocks of 16x16, DK MFMA 16x16x32 each;
//      A operand (Xj rows) from LDS, B operand (Xi rows) from registers.  Lane
//      L holds S^T[16 jb + 4 g + r][16 ib + li] (g = L/16, li = L%16).
//   2. exp epilogue in registers; the 8 values of a lane (jb, r) become the
//      8 BF16 A-operand elements e = 4 jb + r of the projection MFMA, i.e.
//      k-slot 8 g + e <-> j = 16 (e/4) + 4 g + e%4 -- the order Op is packed in;
//   3. Y[i, c] += K[i, j] Op[j, c]: NCB column blocks x 2 row blocks, B operand
//      = one 16-byte LDS read per column block shared by both row blocks.
// Step t runs the epilogue of chunk t next to the distance MFMAs of chunk t+1
// (independent work the scheduler interleaves), then the projection of t.
// Accumulators stay in registers for the whole j loop; Y is written once.
//
// The distance accumulators start at a_i / c2 = -|x_i|^2 / 2, so the
// exponent is one FMA, c2 * s + b_j, and the exp's clamp to [0, 1] stands in
// for clamping d2 at 0.
//
// PP (ping-pong): the waves form two groups, one wave of each per SIMD
// (waves w and w + WAVES/2), running half a step apart.  Each step is split
// by a barrier into V (epilogue of t + distance of t+1: VALU-heavy) and M
// (projection of t: MFMA and LDS).  Group B starts one half-step late, so
// on every SIMD one wave is in V while the other is in M.  With half-steps
// h, group A runs V(t) at h = 2t and M(t) at 2t+1; group B runs V(t) at
// 2t+1 and M(t) at 2t+2.  Both groups write the staging buffer of chunk
// t+2 at h = 2t+1: the last read of that buffer (chunk t-1, B's M(t-1)) is
// at 2t and its first read (A's V(t+1)) at 2t+2, so three buffers suffice.
// Every wave executes the same number of barriers.
//
// project_rows is the body of one workgroup (row block `rblk`, column group
// `cg`); rbf_project_bf16_kernel maps the grid onto it, and the batched
// persistent kernel hands it work items from a queue.
template <int DK, int NCB, int WAVES, bool PP, bool F16>
__device__ __forceinline__ void project_rows(
    int rblk, int cg, int m, int nchunks, int l, int ncb_total, const bf16x8* __restrict__ Xi,
    const bf16x8* __restrict__ Xj, const float* __restrict__ ai, const float* __restrict__ bj,
    float c2, float inv_c2, const bf16x8* __restrict__ Op, const double* __restrict__ yi,
    double* __restrict__ Y, int ldy) {
    constexpr int NT = WAVES * 64;
    constexpr int XG = DK * 4;     // 16-byte groups per point row
    constexpr int XROW = XG + 1;   // LDS row pitch (one group of padding)
    constexpr int XGRP = 32 * XG;  // groups of Xj per chunk
    constexpr int OGRP = NCB * 64; // groups of Op per chunk (this column range)
    constexpr int XPT = (XGRP + NT - 1) / NT;
    constexpr int OPT = (OGRP + NT - 1) / NT;
    constexpr int NBUF = 3;
    __shared__ bf16x8 sX[NBUF][32 * XROW];
    __shared__ bf16x8 sO[NBUF][OGRP];
    __shared__ float sB[NBUF][32];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, li = lane & 15;
    const int cb0 = cg * NCB;
    const int i0 = rblk * (WAVES * 32) + wave * 32;

    bf16x8 xb[2][DK];
    f32x4 sinit[2];
#pragma unroll
    for (int ib = 0; ib < 2; ++ib) {
        int row = i0 + 16 * ib + li;
#pragma unroll
        for (int ks = 0; ks < DK; ++ks) {
            bf16x8 z = {};
            xb[ib][ks] = (row < m) ? Xi[(size_t)row * XG + ks * 4 + g] : z;
        }
        float a0 = (row < m) ? ai[row] * inv_c2 : 0.f;
        sinit[ib] = f32x4{a0, a0, a0, a0};
    }
    f32x4 acc[2][NCB];
#pragma unroll
    for (int ib = 0; ib < 2; ++ib)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) acc[ib][cb] = f32x4{0.f, 0.f, 0.f, 0.f};

    bf16x8 rx[XPT], ro[OPT];
    float rb = 0.f;
    auto gload = [&](int t) {
#pragma unroll
        for (int u = 0; u < XPT; ++u) {
            int e = tid + u * NT;
            if (XGRP % NT == 0 || e < XGRP) rx[u] = Xj[(size_t)t * XGRP + e];
        }
#pragma unroll
        for (int u = 0; u < OPT; ++u) {
            int e = tid + u * NT;
            if (OGRP % NT == 0 || e < OGRP) ro[u] = Op[((size_t)t * ncb_total + cb0) * 64 + e];
        }
        if (tid < 32) rb = bj[t * 32 + tid];
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < XPT; ++u) {
            int e = tid + u * NT;
            if (XGRP % NT == 0 || e < XGRP) sX[buf][(e / XG) * XROW + (e % XG)] = rx[u];
        }
#pragma unroll
        for (int u = 0; u < OPT; ++u) {
            int e = tid + u * NT;
            if (OGRP % NT == 0 || e < OGRP) sO[buf][e] = ro[u];
        }
        if (tid < 32) sB[buf][tid] = rb;
    };
    auto distance = [&](int buf, f32x4 (&s)[2][2]) {
#pragma unroll
        for (int jb = 0; jb < 2; ++jb)
#pragma unroll
            for (int ib = 0; ib < 2; ++ib) s[jb][ib] = sinit[ib];
#pragma unroll
        for (int ks = 0; ks < DK; ++ks)
#pragma unroll
            for (int jb = 0; jb < 2; ++jb) {
                bf16x8 a = sX[buf][(16 * jb + li) * XROW + ks * 4 + g];
#pragma unroll
                for (int ib = 0; ib < 2; ++ib) s[jb][ib] = mfma16<F16>(a, xb[ib][ks], s[jb][ib]);
            }
    };
    // K = clamp(exp2(c2 s + b_j), 0, 1) as the projection's A operand.
    auto epilogue = [&](int buf, const f32x4 (&s)[2][2], bf16x8 (&kb)[2]) {
#pragma unroll
        for (int jb = 0; jb < 2; ++jb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float bv = sB[buf][16 * jb + 4 * g + r];
#pragma unroll
                for (int ib = 0; ib < 2; ++ib) {
                    float e = __builtin_amdgcn_exp2f(fmaf(c2, s[jb][ib][r], bv));
                    kb[ib][4 * jb + r] = to16<F16>(__builtin_amdgcn_fmed3f(e, 0.f, 1.f));
                }
            }
    };
    auto project = [&](int buf, const bf16x8 (&kb)[2]) {
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            bf16x8 bo = sO[buf][cb * 64 + lane];
#pragma unroll
            for (int ib = 0; ib < 2; ++ib) acc[ib][cb] = mfma16<F16>(kb[ib], bo, acc[ib][cb]);
        }
    };

    // Prologue: chunks 0 and 1 resident, distance tile of chunk 0 computed.
    const int last = nchunks - 1;
    gload(0);
    sstore(0);
    gload(min(1, last));
    sstore(1);
    __syncthreads();
    f32x4 s_cur[2][2], s_next[2][2];
    bf16x8 kb[2];
    distance(0, s_cur);
    int buf = 0;  // buffer of chunk t
    if (!PP) {
        for (int t = 0; t < nchunks; ++t) {
            const int nbuf = (buf == NBUF - 1) ? 0 : buf + 1;    // chunk t+1
            const int sbuf = (nbuf == NBUF - 1) ? 0 : nbuf + 1;  // chunk t+2 (being staged)
            gload(min(t + 2, last));
            epilogue(buf, s_cur, kb);
            distance(nbuf, s_next);
            project(buf, kb);
            sstore(sbuf);
            __syncthreads();
#pragma unroll
            for (int jb = 0; jb < 2; ++jb)
#pragma unroll
                for (int ib = 0; ib < 2; ++ib) s_cur[jb][ib] = s_next[jb][ib];
            buf = nbuf;
        }
    } else if (wave < WAVES / 2) {  // group A
        for (int t = 0; t < nchunks; ++t) {
            const int nbuf = (buf == NBUF - 1) ? 0 : buf + 1;
            const int sbuf = (nbuf == NBUF - 1) ? 0 : nbuf + 1;
            gload(min(t + 2, last));
            epilogue(buf, s_cur, kb);
            distance(nbuf, s_next);
            __syncthreads();  // end of V(t), h = 2t
            project(buf, kb);
            sstore(sbuf);
            __syncthreads();  // end of M(t), h = 2t+1
#pragma unroll
            for (int jb = 0; jb < 2; ++jb)
#pragma unroll
                for (int ib = 0; ib < 2; ++ib) s_cur[jb][ib] = s_next[jb][ib];
            buf = nbuf;
        }
        __syncthreads();  // matches group B's leading half-step
    } else {  // group B
        gload(min(2, last));
        __syncthreads();  // h = 0: start half a step behind
        for (int t = 0; t < nchunks; ++t) {
            const int nbuf = (buf == NBUF - 1) ? 0 : buf + 1;
            const int sbuf = (nbuf == NBUF - 1) ? 0 : nbuf + 1;
            epilogue(buf, s_cur, kb);
            distance(nbuf, s_next);
            sstore(sbuf);
            __syncthreads();  // end of V(t), h = 2t+1
            gload(min(t + 3, last));
            project(buf, kb);
            __syncthreads();  // end of M(t), h = 2t+2
#pragma unroll
            for (int jb = 0; jb < 2; ++jb)
#pragma unroll
                for (int ib = 0; ib < 2; ++ib) s_cur[jb][ib] = s_next[jb][ib];
            buf = nbuf;
        }
    }

#pragma unroll
    for (int ib = 0; ib < 2; ++ib)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int i = i0 + 16 * ib + 4 * g + r;
                int c = 16 * (cb0 + cb) + li;
                if (i < m && c < l) {
                    double v = (double)acc[ib][cb][r];
                    if (yi) v *= yi[i];
                    Y[i + (size_t)c * ldy] = v;
                }
            }
}

template <int DK, int NCB, int WAVES, bool PP, bool F16>
__global__ __launch_bounds__(WAVES * 64, 1) void rbf_project_bf16_kernel(
    int m, int nchunks, int l, int ncb_total, const bf16x8* __restrict__ Xi,
    const bf16x8* __restrict__ Xj, const float* __restrict__ ai, const float* __restrict__ bj,
    float c2, float inv_c2, const bf16x8* __restrict__ Op, const double* __restrict__ yi,
    double* __restrict__ Y, int ldy) {
    project_rows<DK, NCB, WAVES, PP, F16>(blockIdx.x, blockIdx.y, m, nchunks, l, ncb_total, Xi, Xj,
                                          ai, bj, c2, inv_c2, Op, yi, Y, ldy);
}

// Waves per workgroup: 8 (two per SIMD, 256 registers each) while the Xi
// fragments are small; 4 (one per SIMD, 512 registers) for d > 128.  The
// ping-pong schedule needs two waves per SIMD.
template <int DK>
constexpr int waves_for() { return DK >= 8 ? 4 : 8; }

#ifndef DH_BF16_PINGPONG
#define DH_BF16_PINGPONG 1
#endif

// Ping-pong or plain schedule for this shape: the autotuner's recorded
// choice, else (outside graph capture, tuning enabled) time both on this
// call's operands -- they produce identical results -- and record the faster.
template <int DK, int NCB, int W, bool F16>
int launch_tuned(dh_handle_t h, dim3 grid, int m, int nchunks, int l, int ncb_total,
                 const bf16x8* Xi, const bf16x8* Xj, const float* ai, const float* bj, float c2,
                 float inv_c2, const bf16x8* Op, const double* yi, double* Y, int ldy) {
    auto run = [&](bool pp) {
        if (pp)
            rbf_project_bf16_kernel<DK, NCB, W, true, F16><<<grid, W * 64, 0, h->stream>>>(
                m, nchunks, l, ncb_total, Xi, Xj, ai, bj, c2, inv_c2, Op, yi, Y, ldy);
        else
            rbf_project_bf16_kernel<DK, NCB, W, false, F16><<<grid, W * 64, 0, h->stream>>>(
                m, nchunks, l, ncb_total, Xi, Xj, ai, bj, c2, inv_c2, Op, yi, Y, ldy);
    };
    const uint64_t key = dh_tune_key_bf16(DK, NCB, W, m) ^ (F16 ? (uint64_t)1 << 61 : 0);
    int pp = dh_tune_lookup(h, key);
    if (pp < 0) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        DH_CHECK_HIP(hipStreamIsCapturing(h->stream, &cs));
        if (!dh_tune_enabled() || cs != hipStreamCaptureStatusNone) {
            pp = DH_BF16_PINGPONG;
        } else {
            hipEvent_t ev[3];
            for (auto& e : ev) DH_CHECK_HIP(hipEventCreate(&e));
            DH_CHECK_HIP(hipEventRecord(ev[0], h->stream));
            run(true);
            DH_CHECK_HIP(hipEventRecord(ev[1], h->stream));
            run(false);
            DH_CHECK_HIP(hipEventRecord(ev[2], h->stream));
            DH_CHECK_HIP(hipEventSynchronize(ev[2]));
            float t_pp = 0.f, t_plain = 0.f;
            DH_CHECK_HIP(hipEventElapsedTime(&t_pp, ev[0], ev[1]));
            DH_CHECK_HIP(hipEventElapsedTime(&t_plain, ev[1], ev[2]));
            for (auto& e : ev) (void)hipEventDestroy(e);
            pp = t_pp <= t_plain ? 1 : 0;
            dh_tune_store(h, key, pp);
        }
    }
    run(pp != 0);
    DH_CHECK_LAUNCH();
    return 0;
}

template <int DK, int NCB, bool F16>
int launch_bf16(dh_handle_t h, int m, int groups, int nchunks, int l, int ncb_total,
                const bf16x8* Xi, const bf16x8* Xj, const float* ai, const float* bj, float c2,
                const bf16x8* Op, const double* yi, double* Y, int ldy) {
    constexpr int W = waves_for<DK>();
    const float inv_c2 = c2 > 0.f ? 1.f / c2 : 0.f;
    dim3 grid(dh_cdiv(m, W * 32), groups);
    if (W == 8)
        return launch_tuned<DK, NCB, W, F16>(h, grid, m, nchunks, l, ncb_total, Xi, Xj, ai, bj, c2,
                                             inv_c2, Op, yi, Y, ldy);
    rbf_project_bf16_kernel<DK, NCB, W, false, F16><<<grid, W * 64, 0, h->stream>>>(
        m, nchunks, l, ncb_total, Xi, Xj, ai, bj, c2, inv_c2, Op, yi, Y, ldy);
    DH_CHECK_LAUNCH();
    return 0;
}

int ncb_instance_bf16(int want) {
    static const int inst[] = {1, 2, 3, 5, 9, 17};
    for (int v : inst)
        if (v >= want) return v;
    return 17;
}

template <int DK, bool F16>
int dispatch_bf16_ncb(dh_handle_t h, int ncb, int groups, int m, int nchunks, int l,
                      int ncb_total, const bf16x8* Xi, const bf16x8* Xj, const float* ai,
                      const float* bj, float c2, const bf16x8* Op, const double* yi, double* Y,
                      int ldy) {
#define DH_BF16_NCB(N) \
    case N:            \
        return launch_bf16<DK, N, F16>(h, m, groups, nchunks, l, ncb_total, Xi, Xj, ai, bj, c2, Op, yi, Y, ldy);
    switch (ncb) {
        DH_BF16_NCB(1)
        DH_BF16_NCB(2)
        DH_BF16_NCB(3)
        DH_BF16_NCB(5)
        DH_BF16_NCB(9)
        DH_BF16_NCB(17)
        default:
            return dh_fail(DH_ERR_UNSUPPORTED, "rbf_project_bf16: no kernel for %d column blocks",
                           ncb);
    }
#undef DH_BF16_NCB
    DH_CHECK_LAUNCH();
    return 0;
}

// Low-precision copy of one point set inside a workspace.
struct LowpPoints {
    __bf16* xb = nullptr;  // nrows_pad x dpad
    float* a = nullptr;    // nrows_pad
    int nrows_pad = 0;
};

int dpad_of(int d) { return d <= 32 ? 32 : d <= 64 ? 64 : d <= 128 ? 128 : 256; }

// Column mean of the point set X (d x n) into mean (device, d doubles).
int point_mean(dh_handle_t h, int n, int d, const double* X, int ldx, double* mean) {
    int slice = 256;
    int nb = dh_cdiv(max(n, 1), slice);
    double* part = (double*)dh_ws(h, DH_WS_PARTIAL, sizeof(double) * (size_t)nb * d);
    if (!part) return DH_ERR_ALLOC;
    colsum_partial_kernel<<<nb, 256, 0, h->stream>>>(n, d, X, ldx, slice, part);
    DH_CHECK_LAUNCH();
    colmean_final_kernel<<<dh_cdiv(d, 256), 256, 0, h->stream>>>(n, d, nb, part, mean);
    DH_CHECK_LAUNCH();
    return 0;
}

template <bool F16>
int convert_points(dh_handle_t h, int n, int d, const double* X, int ldx, const double* mean,
                   double scale, float neg_gl2e, LowpPoints& p) {
    int dp = dpad_of(d);
    convert_rows_kernel<F16><<<dh_cdiv(max(p.nrows_pad, 1), 4), 256, 0, h->stream>>>(
        n, p.nrows_pad, d, dp, X, ldx, mean, scale, neg_gl2e, p.xb, p.a);
    DH_CHECK_LAUNCH();
    return 0;
}

// Column blocking of a BF16 pass: `groups` launches-in-y of `ncb` 16-wide
// column blocks; Op is packed `ncb_pack` = groups * ncb blocks wide (zero
// beyond l) so every block a kernel reads exists.
struct Bf16Cols {
    int ncb, groups, ncb_pack;
};

Bf16Cols bf16_cols(int l) {
    int ncb_total = dh_cdiv(l, 16);
    int groups = dh_cdiv(ncb_total, 17);
    int ncb = ncb_instance_bf16(dh_cdiv(ncb_total, groups));
    groups = dh_cdiv(ncb_total, ncb);
    return {ncb, groups, groups * ncb};
}

template <bool F16>
int pack_omega(dh_handle_t h, int n, int l, const double* O, int ldo, const double* y,
               bf16x8* Op) {
    int nchunks = dh_cdiv(n, 32);
    Bf16Cols cols = bf16_cols(l);
    long threads = (long)nchunks * cols.ncb_pack * 16;
    pack_omega_kernel<F16><<<dh_cdiv(threads, 256), 256, 0, h->stream>>>(
        n, l, nchunks, cols.ncb_pack, O, ldo, y, Op);
    DH_CHECK_LAUNCH();
    return 0;
}

size_t packed_omega_bytes(int n, int l) {
    return (size_t)dh_cdiv(n, 32) * bf16_cols(l).ncb_pack * 64 * sizeof(bf16x8);
}

// Y (m x l, FP64) = K(Xi, Xj) * Op from prepared low-precision operands.
template <bool F16>
int project_16(dh_handle_t h, int m, int n, int d, int l, const LowpPoints& pi,
                 const LowpPoints& pj, float c2, const bf16x8* Op, const double* yi, double* Y,
                 int ldy) {
    if (m == 0 || l == 0) return 0;
    int dp = dpad_of(d);
    int nchunks = dh_cdiv(n, 32);
    if (nchunks == 0) return dh_fail(DH_ERR_UNSUPPORTED, "rbf_project_bf16: n = 0");
    Bf16Cols cols = bf16_cols(l);
    const int grid = cols.groups;
    const bf16x8* xi = (const bf16x8*)pi.xb;
    const bf16x8* xj = (const bf16x8*)pj.xb;
    switch (dp) {
        case 32:
            return dispatch_bf16_ncb<1, F16>(h, cols.ncb, grid, m, nchunks, l, cols.ncb_pack, xi, xj,
                                        pi.a, pj.a, c2, Op, yi, Y, ldy);
        case 64:
            return dispatch_bf16_ncb<2, F16>(h, cols.ncb, grid, m, nchunks, l, cols.ncb_pack, xi, xj,
                                        pi.a, pj.a, c2, Op, yi, Y, ldy);
        case 128:
            return dispatch_bf16_ncb<4, F16>(h, cols.ncb, grid, m, nchunks, l, cols.ncb_pack, xi, xj,
                                        pi.a, pj.a, c2, Op, yi, Y, ldy);
        default:
            return dispatch_bf16_ncb<8, F16>(h, cols.ncb, grid, m, nchunks, l, cols.ncb_pack, xi, xj,
                                        pi.a, pj.a, c2, Op, yi, Y, ldy);
    }
}

// Carve the low-precision copies of one or two point sets out of DH_WS_LOWP_X
// and fill them (centred by the mean of the column set Xj).  FP16 copies are
// also scaled by a power of two 2^-e with max |x - mean| * 2^-e in (1/2, 1],
// so the coordinates use FP16's full mantissa and stay far from its range
// limits; distances scale by 2^-2e, which *gamma_eff (the gamma of the
// stored points) absorbs exactly.  The FP16 scale costs one synchronisation.
template <bool F16>
int prepare_points(dh_handle_t h, int m, int n, int d, const double* Xi, int ldxi,
                   const double* Xj, int ldxj, double gamma, bool same, LowpPoints& pi,
                   LowpPoints& pj, double* gamma_eff) {
    int dp = dpad_of(d);
    int mpad = dh_cdiv(m, 32) * 32, npad = dh_cdiv(n, 32) * 32;
    size_t bytes_j = (size_t)npad * dp * 2 + (size_t)npad * 4;
    size_t bytes_i = same ? 0 : (size_t)mpad * dp * 2 + (size_t)mpad * 4;
    char* base = (char*)dh_ws(h, DH_WS_LOWP_X, bytes_j + bytes_i + 512);
    if (!base) return DH_ERR_ALLOC;
    double* mean = (double*)dh_ws(h, DH_WS_NORMS, sizeof(double) * (size_t)d + 64);
    if (!mean) return DH_ERR_ALLOC;
    pj.xb = (__bf16*)base;
    pj.a = (float*)(base + (size_t)npad * dp * 2);
    pj.nrows_pad = npad;
    if (same) {
        pi = pj;
    } else {
        char* bi = base + ((bytes_j + 255) & ~size_t(255));
        pi.xb = (__bf16*)bi;
        pi.a = (float*)(bi + (size_t)mpad * dp * 2);
        pi.nrows_pad = mpad;
    }
    DH_CHECK_STATUS(point_mean(h, n, d, Xj, ldxj, mean));
    double scale = 1.0;
    if (F16) {
        unsigned* word = (unsigned*)(mean + d + ((8 - d % 8) % 8));  // 64-byte tail of the slot
        zero_word_kernel<<<1, 1, 0, h->stream>>>(word);
        DH_CHECK_LAUNCH();
        int blocks = max(1, min(dh_cdiv((long)n * d, 256), 4 * h->num_cu));
        maxabs_centred_kernel<<<blocks, 256, 0, h->stream>>>(n, d, Xj, ldxj, mean, word);
        DH_CHECK_LAUNCH();
        if (!same) {
            blocks = max(1, min(dh_cdiv((long)m * d, 256), 4 * h->num_cu));
            maxabs_centred_kernel<<<blocks, 256, 0, h->stream>>>(m, d, Xi, ldxi, mean, word);
            DH_CHECK_LAUNCH();
        }
        unsigned bits = 0;
        DH_CHECK_HIP(hipMemcpyAsync(&bits, word, sizeof(bits), hipMemcpyDeviceToHost, h->stream));
        DH_CHECK_HIP(hipStreamSynchronize(h->stream));
        float mx;
        std::memcpy(&mx, &bits, sizeof(mx));
        if (mx > 0.f && std::isfinite(mx)) {
            int e;
            std::frexp((double)mx, &e);  // mx = f * 2^e, f in [1/2, 1)
            scale = std::ldexp(1.0, -e);
        }
    }
    *gamma_eff = gamma / (scale * scale);
    float neg = (float)(-*gamma_eff) * kLog2e;
    DH_CHECK_STATUS(convert_points<F16>(h, n, d, Xj, ldxj, mean, scale, neg, pj));
    if (!same) DH_CHECK_STATUS(convert_points<F16>(h, m, d, Xi, ldxi, mean, scale, neg, pi));
    return 0;
}

}  // namespace

namespace {

template <bool F16>
int rbf_project_16(const char* name, dh_handle_t h, int m, int n, int d, int l, const double* Xi,
                   int ldxi, const double* Xj, int ldxj, const double* yi, const double* yj,
                   double gamma, const double* O, int ldo, double* Y, int ldy) {
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (m < 0) return dh_fail(-2, "%s: m < 0", name);
    if (n < 0) return dh_fail(-3, "%s: n < 0", name);
    if (d < 1 || d > 256) return dh_fail(-4, "%s: need 1 <= d <= 256", name);
    if (l < 0) return dh_fail(-5, "%s: l < 0", name);
    if (ldxi < d) return dh_fail(-7, "%s: ldxi < d", name);
    if (ldxj < d) return dh_fail(-9, "%s: ldxj < d", name);
    if (ldo < max(1, n)) return dh_fail(-14, "%s: ldo < n", name);
    if (ldy < max(1, m)) return dh_fail(-16, "%s: ldy < m", name);
    if (m == 0 || l == 0) return 0;
    if (n == 0) {
        DH_CHECK_HIP(hipMemset2DAsync(Y, sizeof(double) * (size_t)ldy, 0, sizeof(double) * (size_t)m,
                                      l, h->stream));
        return 0;
    }
    LowpPoints pi, pj;
    bool same = (Xi == Xj && m == n && ldxi == ldxj);
    double g = gamma;
    DH_CHECK_STATUS(prepare_points<F16>(h, m, n, d, Xi, ldxi, Xj, ldxj, gamma, same, pi, pj, &g));
    bf16x8* Op = (bf16x8*)dh_ws(h, DH_WS_LOWP_O, packed_omega_bytes(n, l));
    if (!Op) return DH_ERR_ALLOC;
    DH_CHECK_STATUS(pack_omega<F16>(h, n, l, O, ldo, yj, Op));
    return project_16<F16>(h, m, n, d, l, pi, pj, (float)(2.0 * g) * kLog2e, Op, yi, Y, ldy);
}

}  // namespace

extern "C" int dh_rbf_project_bf16(dh_handle_t h, int m, int n, int d, int l, const double* Xi,
                                   int ldxi, const double* Xj, int ldxj, const double* yi,
                                   const double* yj, double gamma, const double* O, int ldo,
                                   double* Y, int ldy) {
    return rbf_project_16<false>("dh_rbf_project_bf16", h, m, n, d, l, Xi, ldxi, Xj, ldxj, yi, yj,
                                 gamma, O, ldo, Y, ldy);
}

extern "C" int dh_rbf_project_f16(dh_handle_t h, int m, int n, int d, int l, const double* Xi,
                                  int ldxi, const double* Xj, int ldxj, const double* yi,
                                  const double* yj, double gamma, const double* O, int ldo,
                                  double* Y, int ldy) {
    return rbf_project_16<true>("dh_rbf_project_f16", h, m, n, d, l, Xi, ldxi, Xj, ldxj, yi, yj,
                                gamma, O, ldo, Y, ldy);
}

int dh_range_orth(dh_handle_t h, int m, int n, double* Y, int ldy);

// Mixed-precision passes of dh_lra: 16-bit (BF16 or FP16) K-passes, FP64 CholeskyQR2.
template <bool F16>
int lra_16_passes(dh_handle_t h, int n, int d, int l, int q, const double* X, int ldx,
                  const double* y, double gamma, const double* Omega, int ldo, double* Q, int ldq,
                  double* W, int ldw) {
    if (d > 256) return dh_fail(DH_ERR_UNSUPPORTED, "dh_lra: 16-bit passes need d <= 256");
    LowpPoints p, p2;
    double g = gamma;
    DH_CHECK_STATUS(prepare_points<F16>(h, n, n, d, X, ldx, X, ldx, gamma, true, p, p2, &g));
    bf16x8* Op = (bf16x8*)dh_ws(h, DH_WS_LOWP_O, packed_omega_bytes(n, l));
    if (!Op) return DH_ERR_ALLOC;
    const float c2 = (float)(2.0 * g) * kLog2e;
    DH_CHECK_STATUS(pack_omega<F16>(h, n, l, Omega, ldo, y, Op));
    DH_CHECK_STATUS(project_16<F16>(h, n, n, d, l, p, p, c2, Op, y, Q, ldq));
    DH_CHECK_STATUS(dh_range_orth(h, n, l, Q, ldq));
    for (int it = 0; it < q; ++it) {
        DH_CHECK_STATUS(pack_omega<F16>(h, n, l, Q, ldq, y, Op));
        DH_CHECK_STATUS(project_16<F16>(h, n, n, d, l, p, p, c2, Op, y, W, ldw));
        DH_CHECK_STATUS(dh_range_orth(h, n, l, W, ldw));
        DH_CHECK_HIP(hipMemcpy2DAsync(Q, sizeof(double) * (size_t)ldq, W,
                                      sizeof(double) * (size_t)ldw, sizeof(double) * (size_t)n, l,
                                      hipMemcpyDeviceToDevice, h->stream));
    }
    DH_CHECK_STATUS(pack_omega<F16>(h, n, l, Q, ldq, y, Op));
    return project_16<F16>(h, n, n, d, l, p, p, c2, Op, y, W, ldw);
}

int dh_lra_bf16_passes(dh_handle_t h, int n, int d, int l, int q, const double* X, int ldx,
                       const double* y, double gamma, const double* Omega, int ldo, double* Q,
                       int ldq, double* W, int ldw) {
    return lra_16_passes<false>(h, n, d, l, q, X, ldx, y, gamma, Omega, ldo, Q, ldq, W, ldw);
}

int dh_lra_f16_passes(dh_handle_t h, int n, int d, int l, int q, const double* X, int ldx,
                      const double* y, double gamma, const double* Omega, int ldo, double* Q,
                      int ldq, double* W, int ldw) {
    return lra_16_passes<true>(h, n, d, l, q, X, ldx, y, gamma, Omega, ldo, Q, ldq, W, ldw);
}

// ------------------------------------------------------------------ batched
// Many independent point sets: blockIdx.y selects the set / job of a table,
// and the projection runs as persistent workgroups pulling work items.

#include "batch.h"

namespace {

__global__ __launch_bounds__(256) void colsum_partial_batched_kernel(const DhBf16Set* __restrict__ sets) {
    const DhBf16Set& s = sets[blockIdx.y];
    const int b = blockIdx.x;
    if (b * 256 >= s.n) return;
    const int i1 = min(s.n, (b + 1) * 256);
    for (int f = threadIdx.x; f < s.d; f += 256) {
        double acc = 0.0;
        for (int i = b * 256; i < i1; ++i) acc += s.X[f + (size_t)i * s.ldx];
        s.part[(size_t)b * s.d + f] = acc;
    }
}

__global__ __launch_bounds__(256) void colmean_final_batched_kernel(const DhBf16Set* __restrict__ sets) {
    const DhBf16Set& s = sets[blockIdx.y];
    const int nb = (s.n + 255) / 256;
    for (int f = blockIdx.x * 256 + threadIdx.x; f < s.d; f += gridDim.x * 256) {
        double acc = 0.0;
        for (int b = 0; b < nb; ++b) acc += s.part[(size_t)b * s.d + f];
        s.mean[f] = s.n > 0 ? acc / s.n : 0.0;
    }
}

__global__ __launch_bounds__(256) void convert_rows_batched_kernel(const DhBf16Set* __restrict__ sets) {
    const DhBf16Set& s = sets[blockIdx.y];
    int row = (blockIdx.x * 256 + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (row >= s.npad) return;
    convert_row<false>(row, lane, s.n, s.d, s.dpad, s.X, s.ldx, s.mean, 1.0, (float)s.neg_gl2e,
                       (__bf16*)s.xb, s.a);
}

__global__ __launch_bounds__(256) void pack_batched_kernel(const DhPackJob* __restrict__ jobs) {
    const DhPackJob& j = jobs[blockIdx.y];
    pack_one<false>((long)blockIdx.x * 256 + threadIdx.x, j.n, j.src ? j.l : 0, j.nchunks,
                    j.ncb_pack, j.src, j.lds, j.y, (bf16x8*)j.Op);
}

// Persistent workgroups: each takes the next item from *counter until the
// queue is empty; every wave of the workgroup sees the same item and leaves
// the loop together.
template <int DK, int NCB, int WAVES, bool PP>
__global__ __launch_bounds__(WAVES * 64, 1) void rbf_project_bf16_persistent(
    const DhProjJob* __restrict__ jobs, const DhProjItem* __restrict__ items, int nitems,
    int* __restrict__ counter) {
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const DhProjItem w = items[it];
        const DhProjJob& j = jobs[w.job];
        project_rows<DK, NCB, WAVES, PP, false>(w.rblk, w.cg, j.m, j.nchunks, j.l, j.ncb_pack,
                                         (const bf16x8*)j.xb, (const bf16x8*)j.xb, j.a, j.a, j.c2,
                                         j.inv_c2, (const bf16x8*)j.Op, j.y, j.Y, j.ldy);
    }
}

template <int DK, int NCB, int W, bool PP>
int launch_persistent(dh_handle_t h, const DhProjJob* jobs, const DhProjItem* items, int nitems,
                      int* counter) {
    auto kern = rbf_project_bf16_persistent<DK, NCB, W, PP>;
    static int per_cu = 0;  // resident workgroups per CU, queried once
    if (per_cu == 0) {
        DH_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, W * 64, 0));
        per_cu = max(per_cu, 1);
    }
    int grid = max(1, min(nitems, per_cu * h->num_cu));
    kern<<<grid, W * 64, 0, h->stream>>>(jobs, items, nitems, counter);
    DH_CHECK_LAUNCH();
    return 0;
}

template <int DK, int W>
int persistent_ncb(dh_handle_t h, int ncb, bool pp, const DhProjJob* jobs, const DhProjItem* items,
                   int nitems, int* counter) {
#define DH_PERS(N)                                                                          \
    case N:                                                                                 \
        return pp ? launch_persistent<DK, N, W, (W == 8)>(h, jobs, items, nitems, counter)  \
                  : launch_persistent<DK, N, W, false>(h, jobs, items, nitems, counter);
    switch (ncb) {
        DH_PERS(1)
        DH_PERS(2)
        DH_PERS(3)
        DH_PERS(5)
        DH_PERS(9)
        DH_PERS(17)
        default:
            return dh_fail(DH_ERR_UNSUPPORTED, "rbf_project_bf16: no kernel for %d column blocks",
                           ncb);
    }
#undef DH_PERS
}

}  // namespace

int dh_bf16_dpad(int d) { return dpad_of(d); }
int dh_bf16_waves(int dpad) { return dpad > 128 ? 4 : 8; }

void dh_bf16_columns(int l, int* ncb, int* groups, int* ncb_pack) {
    Bf16Cols c = bf16_cols(l);
    *ncb = c.ncb;
    *groups = c.groups;
    *ncb_pack = c.ncb_pack;
}

size_t dh_bf16_packed_bytes(int n, int ncb_pack) {
    return (size_t)dh_cdiv(n, 32) * ncb_pack * 64 * sizeof(bf16x8);
}

int dh_bf16_prepare_batched(dh_handle_t h, const DhBf16Set* sets, int count, int max_n, int max_d,
                            int max_npad) {
    if (count == 0) return 0;
    colsum_partial_batched_kernel<<<dim3(dh_cdiv(max(max_n, 1), 256), count), 256, 0, h->stream>>>(sets);
    DH_CHECK_LAUNCH();
    colmean_final_batched_kernel<<<dim3(dh_cdiv(max_d, 256), count), 256, 0, h->stream>>>(sets);
    DH_CHECK_LAUNCH();
    convert_rows_batched_kernel<<<dim3(dh_cdiv(max(max_npad, 1), 4), count), 256, 0, h->stream>>>(sets);
    DH_CHECK_LAUNCH();
    return 0;
}

int dh_bf16_pack_batched(dh_handle_t h, const DhPackJob* jobs, int count, long max_threads) {
    if (count == 0 || max_threads == 0) return 0;
    pack_batched_kernel<<<dim3(dh_cdiv(max_threads, 256), count), 256, 0, h->stream>>>(jobs);
    DH_CHECK_LAUNCH();
    return 0;
}

int dh_bf16_project_persistent(dh_handle_t h, DhBf16Class c, const DhProjJob* jobs,
                               const DhProjItem* items, int nitems, int* counter) {
    if (nitems == 0) return 0;
    bool pp = c.pp != 0;
    switch (c.dk) {
        case 1: return persistent_ncb<1, 8>(h, c.ncb, pp, jobs, items, nitems, counter);
        case 2: return persistent_ncb<2, 8>(h, c.ncb, pp, jobs, items, nitems, counter);
        case 4: return persistent_ncb<4, 8>(h, c.ncb, pp, jobs, items, nitems, counter);
        case 8: return persistent_ncb<8, 4>(h, c.ncb, false, jobs, items, nitems, counter);
        default: return dh_fail(DH_ERR_UNSUPPORTED, "rbf_project_bf16: feature blocks %d", c.dk);
    }
}
