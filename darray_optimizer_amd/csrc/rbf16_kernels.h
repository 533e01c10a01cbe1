// Fused 16-bit projection kernels: distance MFMA, exp epilogue, projection
// MFMA (templates shared by the one-call path, the per-width dispatch units
// rbf16_dk*.hip and the batched persistent kernels).
//
// Precision scheme (per pass):
//   * X is centred by the mean of the column point set (distances are
//     translation invariant) and rounded to BF16 (or FP16 after a power-of-two
//     scaling) once per solve;
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
#pragma once

#include <cmath>
#include <cstring>
#include <type_traits>

#include "dh_common.h"
#include "lds_dma.h"

// 16-byte groups of eight 16-bit values.  The same kernels serve BF16 and
// FP16 (template flag F16): the bits are stored as bf16x8 and reinterpreted
// for the FP16 MFMA.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

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

// 16x16x128 FP8 (OCP e4m3) MFMA with FP32 accumulation: lane L holds 32
// bytes of row L%16 of A (and of column L%16 of B) -- the same k-subset for
// both operands, so any consistent assignment of features to the lanes'
// bytes gives the full dot product.
__device__ __forceinline__ f32x4 mfma_fp8(i32x8 a, i32x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 0, 0, 0, 0, 0, 0);
}

__device__ __forceinline__ i32x8 join16(bf16x8 lo, bf16x8 hi) {
    typedef int i32x4 __attribute__((ext_vector_type(4)));
    i32x4 a = __builtin_bit_cast(i32x4, lo), b = __builtin_bit_cast(i32x4, hi);
    return i32x8{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
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
    const int j0 = 32 * t;
    const double* src = O + j0 + (size_t)min(c, l - 1) * ldo;
    const bool vec = c < l && j0 + 32 <= n && ((reinterpret_cast<uintptr_t>(src) & 15) == 0) &&
                     (!y || (reinterpret_cast<uintptr_t>(y + j0) & 15) == 0);
    if (vec) {
        // Whole chunk in range and 16-byte aligned: 16 unconditional 2-double
        // loads (loads under per-element branches would be issued one by one).
        double o[32];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const double2 w = reinterpret_cast<const double2*>(src)[q];
            o[2 * q] = w.x;
            o[2 * q + 1] = w.y;
        }
        if (y) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const double2 w = reinterpret_cast<const double2*>(y + j0)[q];
                o[2 * q] *= w.x;
                o[2 * q + 1] *= w.y;
            }
        }
#pragma unroll
        for (int jj = 0; jj < 32; ++jj) v[jj] = (float)o[jj];
    } else {
#pragma unroll
        for (int jj = 0; jj < 32; ++jj) {
            int j = j0 + jj;
            double o = 0.0;
            if (j < n && c < l) {
                o = O[j + (size_t)c * ldo];
                if (y) o *= y[j];
            }
            v[jj] = (float)o;
        }
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        bf16x8 w;
#pragma unroll
        for (int e = 0; e < 8; ++e) w[e] = to16<F16>(v[16 * (e >> 2) + 4 * g + (e & 3)]);
        Op[((size_t)t * ncb_total + cb) * 64 + 16 * g + li] = w;
    }
}

// ------------------------------------------------------------ main kernel
//
// Workgroup = WAVES waves; wave w owns 16*RB rows i (RB 16-row MFMA blocks).
// The j loop walks chunks of 32 points through a 3-deep LDS ring: step t
// computes the distance tile and K of chunk t while the projection of chunk
// t-1 runs and chunk t+1 is in flight, so one barrier per chunk suffices
// and the loop body is a single basic block (look-ahead indices are
// clamped instead of branched).  Staging uses direct
// global-to-LDS loads (global_load_lds_dwordx4): no registers, no LDS
// stores; each wave waits for its own loads (vmcnt) before the barrier that
// publishes them.  Per chunk a wave:
//   1. S^T = Xj_chunk * Xi^T: 2 x RB blocks of 16x16, DK MFMA 16x16x32 each;
//      A operand (Xj rows) from LDS, B operand (Xi rows) from registers.  Lane
//      L holds S^T[16 jb + 4 g + r][16 ib + li] (g = L/16, li = L%16).
//   2. exp epilogue in registers; the 8 values of a lane (jb, r) become the
//      8 16-bit A-operand elements e = 4 jb + r of the projection MFMA, i.e.
//      k-slot 8 g + e <-> j = 16 (e/4) + 4 g + e%4 -- the order Op is packed in;
//   3. Y[i, c] += K[i, j] Op[j, c]: NCB column blocks x RB row blocks, B operand
//      = one 16-byte LDS read per column block shared by the RB row blocks.
// Accumulators stay in registers for the whole j loop; Y is written once.
//
// The exponent is c2 s + (a_i + b_j) with s = x_i . x_j, c2 = 2 gamma log2(e)
// and a = -gamma log2(e) |x|^2, evaluated as c2 (a_i / c2 + s) + b_j with
// a_i / c2 the distance MFMA's starting accumulator (16-bit distances); the
// exp's clamp to [0, 1] stands in for clamping d2 at 0.
//
// D8: the distance runs in FP8 e4m3 instead (points stored as one byte per
// feature, each row scaled by a power of two 1/sinv_i so that its largest
// |coordinate| lies in [128, 256)): 16x16x128 FP8 MFMAs, two 16-byte groups
// of a row per lane and k-step, and x_i . x_j = sinv_i sinv_j (q_i . q_j);
// a_i and b_j are the squared norms of the dequantised rows (FP32), so the
// squared distance is that of the rounded points.
//
// project_rows is the body of one workgroup (row block `rblk`, column group
// `cg`); rbf_project_bf16_kernel maps the grid onto it, and the batched
// persistent kernel hands it work items from a queue.

// Padded LDS image of one chunk of Xj: 32 rows of XG 16-byte groups plus one
// group of padding per row (conflict-free A-operand reads), rounded up to
// whole 64-slot DMA instructions.
template <int DK>
struct XSlots {
    static constexpr int XG = DK * 4;
    static constexpr int XROW = XG + 1;
    static constexpr int INSTR = (32 * XROW + 63) / 64;
    static constexpr int SLOTS = INSTR * 64;
};


// LDS of one workgroup of project_rows: the 3-deep ring of Xj chunks (sX),
// packed Omega blocks (sO), norms b_j (sB) and FP8 scales (sS), and the
// exchange slots of wave pairs (sK).  Kernels declare it __shared__ or carve
// it from a larger arena.
template <int DK, int NCB, int WAVES, int RB, bool D8, bool PAIR>
struct ProjLds {
    static constexpr int NBUF = 3;
    static constexpr int NO = PAIR ? 2 * NCB : NCB;  // Omega blocks staged per chunk
    bf16x8 sX[NBUF][XSlots<DK>::SLOTS];
    bf16x8 sO[NBUF][NO * 64];
    float sB[NBUF][64];
    float sS[D8 ? NBUF : 1][64];
    bf16x8 sK[PAIR ? 2 : 1][PAIR ? WAVES / 2 : 1][PAIR ? RB : 1][PAIR ? 64 : 1];
};

template <int DK, int NCB, int WAVES, bool F16, int RB, bool D8 = false, bool PAIR = false>
__device__ __forceinline__ void project_rows(
    ProjLds<DK, NCB, WAVES, RB, D8, PAIR>& lds, int rblk, int cg, int m, int nchunks, int l,
    int ncb_total, const bf16x8* __restrict__ Xi, const bf16x8* __restrict__ Xj,
    const float* __restrict__ ai, const float* __restrict__ bj, const float* __restrict__ si,
    const float* __restrict__ sj, float c2, const bf16x8* __restrict__ Op,
    const double* __restrict__ yi, double* __restrict__ Y, int ldy) {
    using XS = XSlots<DK>;
    constexpr int XG = XS::XG, XROW = XS::XROW;
    constexpr int NBUF = 3;
    // PAIR: waves w and w + WAVES/2 share their rows and take two adjacent
    // column groups; each computes the distance and K of one 16-point half
    // of the chunk (jb = its half) and the two exchange their halves of the
    // projection's A operand through sK, so no distance is computed twice.
    constexpr int NO = PAIR ? 2 * NCB : NCB;  // Omega blocks staged per chunk
    constexpr int NJB = PAIR ? 1 : 2;         // distance blocks per wave and chunk
    static_assert(!PAIR || (RB == 2 && WAVES % 2 == 0), "pairs: RB = 2, even wave count");
    auto& sX = lds.sX;
    auto& sO = lds.sO;
    auto& sB = lds.sB;
    auto& sS = lds.sS;
    auto& sK = lds.sK;

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar addressing
    const int g = lane >> 4, li = lane & 15;
    const int half = PAIR ? wave / (WAVES / 2) : 0;         // wave-uniform
    const int wrow = PAIR ? wave % (WAVES / 2) : wave;      // row slot of the wave
    const int cb0 = PAIR ? (2 * cg + half) * NCB : cg * NCB;
    const int i0 = (rblk * (PAIR ? WAVES / 2 : WAVES) + wrow) * 16 * RB;

    // Group of a row that lane (li, g) holds at operand slot q: 16-bit
    // k-step q covers groups 4q .. 4q+3 (one per g); FP8 k-step q/2 covers
    // groups 8(q/2) .. 8(q/2)+7 (two per g).
    auto grp = [&](int q) { return D8 ? 8 * (q >> 1) + 2 * g + (q & 1) : 4 * q + g; };
    bf16x8 xb[RB][DK];
    float arow[RB], crow[RB];
#pragma unroll
    for (int ib = 0; ib < RB; ++ib) {
        int row = i0 + 16 * ib + li;
#pragma unroll
        for (int q = 0; q < DK; ++q) {
            bf16x8 z = {};
            xb[ib][q] = (row < m) ? Xi[(size_t)row * XG + grp(q)] : z;
        }
        arow[ib] = (row < m) ? ai[row] : 0.f;
        crow[ib] = D8 ? c2 * ((row < m) ? si[row] : 0.f) : c2;
    }
    // 16-bit distances start their accumulation at a_i / c2 (the first
    // MFMA's accumulator input), so the exponent is c2 s + b_j: one VALU
    // operation less per element.  FP8 distances scale s per point pair and
    // keep the sum a_i + b_j.
    f32x4 ainit[RB];
#pragma unroll
    for (int ib = 0; ib < RB; ++ib) {
        const float v = (D8 || c2 == 0.f) ? 0.f : arow[ib] / c2;
        ainit[ib] = f32x4{v, v, v, v};
    }
    f32x4 acc[RB][NCB];
#pragma unroll
    for (int ib = 0; ib < RB; ++ib)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) acc[ib][cb] = f32x4{0.f, 0.f, 0.f, 0.f};

    // Issue the loads of chunk t into buffer buf (instructions spread over the
    // waves).  Byte offsets of the wave's lanes in the padded image of Xj's
    // chunk, fixed for the whole loop (padding slots load row 0, never read).
    constexpr int NXK = (XS::INSTR + WAVES - 1) / WAVES;
    unsigned offx[NXK];
#pragma unroll
    for (int k = 0; k < XS::INSTR; k += WAVES) {
        int slot = (k + wave) * 64 + lane, row = slot / XROW, col = slot % XROW;
        if (row >= 32 || col >= XG) {
            row = 0;
            col = 0;
        }
        offx[k / WAVES] = (unsigned)(row * XG + col) * 16u;
    }
    auto stage = [&](int t, int buf) {
        const bf16x8* xc = Xj + (size_t)t * 32 * XG;
#pragma unroll
        for (int k = 0; k < XS::INSTR; k += WAVES) {
            const int kk = k + wave;
            if (XS::INSTR % WAVES == 0 || kk < XS::INSTR)
                dma16(xc, offx[k / WAVES], &sX[buf][kk * 64]);
        }
        const bf16x8* oc = Op + ((size_t)t * ncb_total + (PAIR ? 2 * cg * NCB : cb0)) * 64;
#pragma unroll
        for (int k = 0; k < NO; k += WAVES) {
            const int kk = k + wave;
            if (NO % WAVES == 0 || kk < NO) dma16(oc + kk * 64, (unsigned)lane * 16u, &sO[buf][kk * 64]);
        }
        if (wave == WAVES - 1) dma4(bj + (size_t)t * 32, (unsigned)(lane & 31) * 4u, &sB[buf][0]);
        if (D8 && wave == WAVES - 2)
            dma4(sj + (size_t)t * 32, (unsigned)(lane & 31) * 4u, &sS[D8 ? buf : 0][0]);
    };
    // Distance operands: all 2 x DK A-operand reads of a chunk are issued
    // before its MFMAs.
    // (Pairs: only the wave's half, jb = half, as s[0].)
    auto distance = [&](int buf, f32x4 (&s)[NJB][RB]) {
        bf16x8 a[DK][NJB];
#pragma unroll
        for (int q = 0; q < DK; ++q)
#pragma unroll
            for (int jj = 0; jj < NJB; ++jj) {
                const int jb = PAIR ? half : jj;
                a[q][jj] = sX[buf][(16 * jb + li) * XROW + grp(q)];
            }
#pragma unroll
        for (int jj = 0; jj < NJB; ++jj)
#pragma unroll
            for (int ib = 0; ib < RB; ++ib) s[jj][ib] = ainit[ib];
        if constexpr (D8) {
#pragma unroll
            for (int q = 0; q < DK; q += 2)
#pragma unroll
                for (int jj = 0; jj < NJB; ++jj)
#pragma unroll
                    for (int ib = 0; ib < RB; ++ib)
                        s[jj][ib] = mfma_fp8(join16(a[q][jj], a[q + 1][jj]),
                                             join16(xb[ib][q], xb[ib][q + 1]), s[jj][ib]);
        } else {
#pragma unroll
            for (int ks = 0; ks < DK; ++ks)
#pragma unroll
                for (int jj = 0; jj < NJB; ++jj)
#pragma unroll
                    for (int ib = 0; ib < RB; ++ib)
                        s[jj][ib] = mfma16<F16>(a[ks][jj], xb[ib][ks], s[jj][ib]);
        }
    };
    // K = clamp(exp2(c2 s + a_i + b_j), 0, 1) as the projection's A operand.
    // FP16 values are converted in pairs from two-element vectors, the form
    // the compiler turns into v_cvt_pk_f16_f32 (with the [0, 1] clamp folded
    // into v_exp_f32); element-wise FP16 conversions cost two more
    // instructions per pair.  BF16 pairs are packed either way.
    // (Pairs: the wave's half only, elements 0..3 of kb; the other half is
    // the partner's, see publish / gather.)
    auto epilogue = [&](int buf, const f32x4 (&s)[NJB][RB], bf16x8 (&kb)[RB]) {
        float ev[RB][8];
#pragma unroll
        for (int jj = 0; jj < NJB; ++jj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int jb = PAIR ? half : jj;
                float bv = sB[buf][16 * jb + 4 * g + r];
                float sj = D8 ? sS[D8 ? buf : 0][16 * jb + 4 * g + r] : 1.f;
#pragma unroll
                for (int ib = 0; ib < RB; ++ib) {
                    float cc = D8 ? crow[ib] * sj : crow[ib];
                    float e = __builtin_amdgcn_exp2f(
                        fmaf(cc, s[jj][ib][r], D8 ? arow[ib] + bv : bv));
                    ev[ib][4 * jj + r] = __builtin_amdgcn_fmed3f(e, 0.f, 1.f);
                }
            }
#pragma unroll
        for (int ib = 0; ib < RB; ++ib) {
            if constexpr (F16) {
                typedef float f32x2 __attribute__((ext_vector_type(2)));
                typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
                typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
                u32x4 w = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int e = 0; e < 2 * NJB; ++e)
                    w[e] = __builtin_bit_cast(
                        unsigned,
                        __builtin_convertvector((f32x2{ev[ib][2 * e], ev[ib][2 * e + 1]}), f16x2));
                kb[ib] = __builtin_bit_cast(bf16x8, w);
            } else {
#pragma unroll
                for (int e = 0; e < 4 * NJB; ++e) kb[ib][e] = to16<false>(ev[ib][e]);
            }
        }
    };
    // Pairs: the wave's half of K (elements 0..3 of kb) goes to elements
    // 4 half .. 4 half + 3 of the pair's slot in sK[par]; after the next
    // barrier both waves read the whole A operand back.
    auto publish = [&](int par, const bf16x8 (&kb)[RB]) {
        if constexpr (PAIR) {
            typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
            typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
#pragma unroll
            for (int ib = 0; ib < RB; ++ib) {
                const u32x4 w = __builtin_bit_cast(u32x4, kb[ib]);
                *(u32x2*)((unsigned*)&sK[par][wrow][ib][lane] + 2 * half) = u32x2{w[0], w[1]};
            }
        }
    };
    auto gather = [&](int par, bf16x8 (&kb)[RB]) {
        if constexpr (PAIR) {
#pragma unroll
            for (int ib = 0; ib < RB; ++ib) kb[ib] = sK[par][wrow][ib][lane];
        }
    };
    // Projection B operands go through a ring of PF registers: the read of
    // column block cb + PF is issued before the MFMAs of block cb.  Eight
    // deep: 2.48 ms against 2.54 ms for four at d = 128, l = 272 (FP16,
    // median of four rounds of 15 launches each).
    constexpr int PF = NCB < 8 ? NCB : 8;
    const int ob = half * NCB * 64;  // the wave's Omega blocks in sO
    auto project_start = [&](int buf, bf16x8 (&ring)[PF]) {
#pragma unroll
        for (int p = 0; p < PF; ++p) ring[p] = sO[buf][ob + p * 64 + lane];
    };
    auto project = [&](int buf, const bf16x8 (&kb)[RB], bf16x8 (&ring)[PF]) {
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            const bf16x8 bo = ring[cb % PF];
            if (cb + PF < NCB) ring[cb % PF] = sO[buf][ob + (cb + PF) * 64 + lane];
#pragma unroll
            for (int ib = 0; ib < RB; ++ib) acc[ib][cb] = mfma16<F16>(kb[ib], bo, acc[ib][cb]);
        }
    };

    // Step t: distance of chunk t, then its epilogue next to the projection
    // of chunk t-1 (independent work the scheduler interleaves); chunk t+1
    // is in flight.  Buffers: chunk t in t%3 (read by distance and
    // epilogue), chunk t-1 in (t-1)%3 (read by the projection), chunk t+1
    // loading into (t+1)%3, whose last reader (the projection of chunk t-2)
    // finished before the barrier that closed step t-1.
    // Pairs: the halves of chunk t are published to sK[t & 1] in step t and
    // gathered in step t + 1 (after the barrier between); sK[t & 1] is
    // rewritten in step t + 2, after both waves read it in step t + 1.
    const int last = nchunks - 1;
    stage(0, 0);
    wait_vm();
    __syncthreads();
    f32x4 s[NJB][RB];
    bf16x8 kb_prev[RB], kb[RB], ring[PF];
    // Step 0 (no projection yet), then steps 1 .. nchunks-1 as one basic block.
    stage(min(1, last), 1);
    distance(0, s);
    epilogue(0, s, kb_prev);
    publish(0, kb_prev);
    wait_vm();
    __syncthreads();
    int buf = 1;  // buffer of chunk t
    for (int t = 1; t < nchunks; ++t) {
        const int nbuf = (buf == NBUF - 1) ? 0 : buf + 1;   // chunk t+1
        const int pbuf = (buf == 0) ? NBUF - 1 : buf - 1;   // chunk t-1
        stage(min(t + 1, last), nbuf);
        project_start(pbuf, ring);
        gather((t - 1) & 1, kb_prev);
        distance(buf, s);
        epilogue(buf, s, kb);
        publish(t & 1, kb);
        project(pbuf, kb_prev, ring);
#pragma unroll
        for (int ib = 0; ib < RB; ++ib) kb_prev[ib] = kb[ib];
        wait_vm();
        __syncthreads();
        buf = nbuf;
    }
    gather(last & 1, kb_prev);
    project_start(buf == 0 ? NBUF - 1 : buf - 1, ring);
    project(buf == 0 ? NBUF - 1 : buf - 1, kb_prev, ring);  // last chunk, after the final barrier

#pragma unroll
    for (int ib = 0; ib < RB; ++ib)
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

// Grid: x = row block, y = column group, z = split of the j range: split z
// takes chunks [z cps, min((z + 1) cps, nchunks)) and writes its partial
// product to Y + z * ystride (one split: the whole range into Y).
template <int DK, int NCB, int WAVES, bool F16, int RB, bool D8 = false, bool PAIR = false>
__global__ __launch_bounds__(WAVES * 64, 1) void rbf_project_bf16_kernel(
    int m, int nchunks, int l, int ncb_total, const bf16x8* __restrict__ Xi,
    const bf16x8* __restrict__ Xj, const float* __restrict__ ai, const float* __restrict__ bj,
    const float* __restrict__ si, const float* __restrict__ sj, float c2,
    const bf16x8* __restrict__ Op, const double* __restrict__ yi, double* __restrict__ Y,
    int ldy, int cps, long ystride) {
    const int t0 = blockIdx.z * cps;
    const size_t p0 = (size_t)t0 * 32;
    __shared__ ProjLds<DK, NCB, WAVES, RB, D8, PAIR> lds;
    project_rows<DK, NCB, WAVES, F16, RB, D8, PAIR>(
        lds, blockIdx.x, blockIdx.y, m, min(cps, nchunks - t0), l, ncb_total, Xi,
        Xj + p0 * XSlots<DK>::XG, ai, bj + p0, si, sj ? sj + p0 : nullptr, c2,
        Op + (size_t)t0 * ncb_total * 64, yi, Y + blockIdx.z * ystride, ldy);
}

// Y[i + c ldy] = sum over the splits z (in order) of P[z * ystride + i + c m].
__global__ __launch_bounds__(256) void sum_splits_kernel(int m, int l, int splits,
                                                         const double* __restrict__ P,
                                                         long ystride, double* __restrict__ Y,
                                                         int ldy) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)m * l) return;
    const int i = (int)(e % m), c = (int)(e / m);
    double v = 0.0;
    for (int z = 0; z < splits; ++z) v += P[z * ystride + e];
    Y[i + (size_t)c * ldy] = v;
}

// Schedule used when tuning is off (DH_TUNE=0) or during graph capture.
#ifndef DH_BF16_DEFAULT_SCHEDULE
#define DH_BF16_DEFAULT_SCHEDULE 0
#endif

// Schedules of one projection launch (same results, different wave layout):
//   0: 8 waves x 32 rows (two waves per SIMD hide each other's latencies)
//   1: 4 waves x 64 rows (four row blocks per wave share each LDS operand read)
// and for d > 128, where the Xi fragments need the registers, 4 waves x 32 rows.
// The autotuner's recorded choice for the shape is used; without one (and
// outside graph capture, with tuning enabled) every candidate is timed on
// this call's operands and the fastest recorded.
template <int DK, int NCB, bool F16>
int launch_bf16(dh_handle_t h, int m, int groups, int nchunks, int l, int ncb_total,
                const bf16x8* Xi, const bf16x8* Xj, const float* ai, const float* bj, float c2,
                const bf16x8* Op, const double* yi, double* Y, int ldy, int splits,
                long ystride) {
    const int cps = dh_cdiv(nchunks, splits);
    const int nz = dh_cdiv(nchunks, cps);
    // An even number of 17-block column groups (l > 272) runs as wave pairs
    // (rows shared, each distance computed once); up to 128 features the
    // pair's LDS ring fits.
    if constexpr (NCB == 17 && DK <= 4) {
        if (groups % 2 == 0) {
            dh_timing_begin(h);
            rbf_project_bf16_kernel<DK, NCB, 8, F16, 2, false, true>
                <<<dim3(dh_cdiv(m, 4 * 32), groups / 2, nz), 512, 0, h->stream>>>(
                    m, nchunks, l, ncb_total, Xi, Xj, ai, bj, nullptr, nullptr, c2, Op, yi, Y,
                    ldy, cps, ystride);
            dh_timing_end(h);
            DH_CHECK_LAUNCH();
            return 0;
        }
    }
    auto run = [&](int sched) {
#define DH_RUN(W, RB)                                                                        \
    rbf_project_bf16_kernel<DK, NCB, W, F16, RB>                                             \
        <<<dim3(dh_cdiv(m, W * 16 * RB), groups, nz), W * 64, 0, h->stream>>>(               \
            m, nchunks, l, ncb_total, Xi, Xj, ai, bj, nullptr, nullptr, c2, Op, yi, Y, ldy, cps, \
            ystride)
        if constexpr (DK >= 8) {
            DH_RUN(4, 2);
        } else {
            if (sched == 0)
                DH_RUN(8, 2);
            else
                DH_RUN(4, 4);
        }
#undef DH_RUN
    };
    constexpr int ncand = DK >= 8 ? 1 : 2;
    const uint64_t key = dh_tune_key_bf16(DK, NCB, nz, m) ^ (F16 ? (uint64_t)1 << 61 : 0);
    int sched = ncand == 1 ? 0 : dh_tune_lookup(h, key);
    const int forced = dh_forced_schedule();
    if (ncand > 1 && forced >= 0 && forced < ncand) sched = forced;
    if (sched < 0) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        DH_CHECK_HIP(hipStreamIsCapturing(h->stream, &cs));
        if (!dh_tune_enabled() || cs != hipStreamCaptureStatusNone) {
            sched = DH_BF16_DEFAULT_SCHEDULE;
        } else {
            // Each candidate runs once untimed first: a kernel's first launch
            // also pays for loading its code object.
            for (int c = 0; c < ncand; ++c) {
                run(c);
                DH_CHECK_LAUNCH();
            }
            hipEvent_t ev[ncand + 1];
            for (auto& e : ev) DH_CHECK_HIP(hipEventCreate(&e));
            DH_CHECK_HIP(hipEventRecord(ev[0], h->stream));
            for (int c = 0; c < ncand; ++c) {
                run(c);
                DH_CHECK_LAUNCH();
                DH_CHECK_HIP(hipEventRecord(ev[c + 1], h->stream));
            }
            DH_CHECK_HIP(hipEventSynchronize(ev[ncand]));
            float best = 3.4e38f;
            for (int c = 0; c < ncand; ++c) {
                float t = 0.f;
                DH_CHECK_HIP(hipEventElapsedTime(&t, ev[c], ev[c + 1]));
                if (t < best) {
                    best = t;
                    sched = c;
                }
            }
            for (auto& e : ev) (void)hipEventDestroy(e);
            dh_tune_store(h, key, sched);
        }
    }
    dh_timing_begin(h);
    run(sched);
    dh_timing_end(h);
    DH_CHECK_LAUNCH();
    return 0;
}

template <int DK, bool F16>
int dispatch_bf16_ncb(dh_handle_t h, int ncb, int groups, int m, int nchunks, int l,
                      int ncb_total, const bf16x8* Xi, const bf16x8* Xj, const float* ai,
                      const float* bj, float c2, const bf16x8* Op, const double* yi, double* Y,
                      int ldy, int splits, long ystride) {
#define DH_BF16_NCB(N) \
    case N:            \
        return launch_bf16<DK, N, F16>(h, m, groups, nchunks, l, ncb_total, Xi, Xj, ai, bj, c2, Op, yi, Y, ldy, splits, ystride);
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

}  // namespace

// Per-feature-width dispatch of one projection launch (rbf16_dk<DK>.hip):
// BF16 or FP16 operands, column blocks ncb, `groups` column groups, the j
// range in `splits` parts whose partial products go to Y + z * ystride.
int dh_project16_dk1(bool f16, dh_handle_t h, int ncb, int groups, int m, int nchunks, int l, int ncb_total,
                      const bf16x8* Xi, const bf16x8* Xj, const float* ai, const float* bj, float c2,
                      const bf16x8* Op, const double* yi, double* Y, int ldy, int splits,
                      long ystride);
int dh_project16_dk2(bool f16, dh_handle_t h, int ncb, int groups, int m, int nchunks, int l, int ncb_total,
                      const bf16x8* Xi, const bf16x8* Xj, const float* ai, const float* bj, float c2,
                      const bf16x8* Op, const double* yi, double* Y, int ldy, int splits,
                      long ystride);
int dh_project16_dk4(bool f16, dh_handle_t h, int ncb, int groups, int m, int nchunks, int l, int ncb_total,
                      const bf16x8* Xi, const bf16x8* Xj, const float* ai, const float* bj, float c2,
                      const bf16x8* Op, const double* yi, double* Y, int ldy, int splits,
                      long ystride);
int dh_project16_dk8(bool f16, dh_handle_t h, int ncb, int groups, int m, int nchunks, int l, int ncb_total,
                      const bf16x8* Xi, const bf16x8* Xj, const float* ai, const float* bj, float c2,
                      const bf16x8* Op, const double* yi, double* Y, int ldy, int splits,
                      long ystride);
// FP8-distance projection (rbf16_fp8.hip): points padded to dpad8 = 128 or
// 256 one-byte features, si / sj the per-row inverse scales of Xi / Xj.
int dh_project_fp8(dh_handle_t h, int dpad8, int ncb, int groups, int m, int nchunks, int l,
                   int ncb_total, const bf16x8* Xi, const bf16x8* Xj, const float* ai,
                   const float* bj, const float* si, const float* sj, float c2,
                   const bf16x8* Op, const double* yi, double* Y, int ldy);
