// The k x k step of CholeskyQR on gfx950, the row-parallel triangular solve
// that applies its factor to the n x k panel, and (128 < k <= 272) the Gram
// product itself with the whole upper k x k result held in registers.
//
// chol_reg_kernel (one workgroup, order n <= 272: the matrix held in
// registers) and chol_small_kernel (order n <= 544): G (upper triangle read)
// is overwritten by U with G = U^T U; the strict lower triangle is untouched.
// Right-looking in 16-column panels as cholesky_factorize (lapack_like.h:17-52):
// factor the diagonal block (one wave, column per lane), solve the row panel
// U12 = U11^{-T} A12 (column per thread), update A22 -= U12^T U12 in 64 x 64
// blocks (wave per block, row per lane, panel rows cached in registers).
// It also writes the inverses of the 16 x 16 diagonal blocks of U to Dinv
// (block b at Dinv + 256 b, column-major), for trsm_rows_kernel.
//
// Breakdown: a pivot p (before the square root) fails if !(p > tol) with
// tol = 0 (plain Cholesky) or tol = n * eps * max_k G_kk (CholeskyQR: Y
// numerically rank deficient); *info receives the 1-based index of the first
// failing pivot, or 0.  With shift != 0 the diagonal is first raised by
// s = 11 (m n + n (n+1)) u trace(G) (shifted CholeskyQR3, Fukaya et al. 2020).
//
// trsm_rows_kernel: Q = Y U^{-1} in place, one wave per 16 rows of Y.  Rows
// are independent, so each wave runs the whole block-column recurrence
//   Q_J = (Y_J - sum_{K<J} Q_K U_KJ) U_JJ^{-1}
// for its rows on FP64 MFMA, its solved rows staged in LDS.
//
// Device bodies and kernels (this header), host launchers (small_spd.hip).
#pragma once

#include "dh_common.h"
#include "mfma_tile.h"

namespace {

constexpr int NB = 16;
constexpr int kMaxN = 544;
constexpr int PW = kMaxN - NB + 4;  // LDS panel pitch, multiple of 4 doubles
constexpr int NT = 512;

__device__ __forceinline__ double readlane_d(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Orders LDS traffic of one wave (no workgroup barrier: waves are independent).
// Waits for LDS traffic only (s_waitcnt lgkmcnt(0) with vmcnt / expcnt left at
// their maxima), so prefetched global loads stay in flight across it.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_wave_barrier();
}

// Breakdown tolerance and CholeskyQR3 shift of G from its diagonal (every
// thread of the workgroup gets both).
__device__ __forceinline__ void chol_scales(int n, const double* __restrict__ G, int ldg,
                                            int m_rows, int shift, int rel_check, double& tol,
                                            double& sh) {
    __shared__ double red[2][NT / 64];
    __shared__ double s_tol, s_shift;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int NW = NT / 64;
    double mx = 0.0, tr = 0.0;
    for (int k = tid; k < n; k += NT) {
        double v = G[k + (size_t)k * ldg];
        mx = fmax(mx, v);
        tr += v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, o));
        tr += __shfl_xor(tr, o);
    }
    if (lane == 0) {
        red[0][wave] = mx;
        red[1][wave] = tr;
    }
    __syncthreads();
    if (tid == 0) {
        double m2 = 0.0, t2 = 0.0;
        for (int w = 0; w < NW; ++w) {
            m2 = fmax(m2, red[0][w]);
            t2 += red[1][w];
        }
        s_tol = rel_check ? (double)n * 2.220446049250313e-16 * m2 : 0.0;
        s_shift = shift ? 11.0 * ((double)m_rows * n + (double)n * (n + 1)) *
                              1.1102230246251565e-16 * t2
                        : 0.0;
    }
    __syncthreads();
    tol = s_tol;
    sh = s_shift;
}

__device__ __forceinline__ void chol_small_body(int n, double* __restrict__ G, int ldg,
                                                double* __restrict__ Dinv, int m_rows, int shift,
                                                int rel_check, int* __restrict__ info) {
    __shared__ double sD[NB][NB + 1];            // diagonal block, sD[j][i] = A(k0+i, k0+j)
    __shared__ __align__(16) double sP[NB][PW];  // row panel U12, sP[kk][c]
    __shared__ int s_info;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, li = lane & 15;  // MFMA lane group and index
    constexpr int NW = NT / 64;

    double tol, sh;
    chol_scales(n, G, ldg, m_rows, shift, rel_check, tol, sh);
    if (tid == 0) s_info = 0;
    if (shift)
        for (int k = tid; k < n; k += NT) G[k + (size_t)k * ldg] += sh;
    __syncthreads();

    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = min(NB, n - k0), w = n - k0 - nb;
        for (int e = tid; e < nb * nb; e += NT) {
            int i = e % nb, j = e / nb;
            sD[j][i] = (i <= j) ? G[(k0 + i) + (size_t)(k0 + j) * ldg] : 0.0;
        }
        __syncthreads();
        if (wave == 0) {
            // Lane j owns column j of the block; row kk of U is broadcast with readlane.
            double a[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) a[i] = (lane < nb) ? sD[lane][i] : 0.0;
            int bad = 0;
#pragma unroll
            for (int kk = 0; kk < NB; ++kk) {
                if (kk < nb && !bad) {
                    double piv = readlane_d(a[kk], kk);
                    if (!(piv > tol)) {
                        bad = kk + 1;
                    } else {
                        double u = sqrt(piv);
                        double ukj = (lane == kk) ? u : a[kk] / u;
                        a[kk] = (lane >= kk) ? ukj : a[kk];
#pragma unroll
                        for (int i = kk + 1; i < NB; ++i) {
                            double uki = readlane_d(ukj, i);
                            if (lane > kk && i <= lane && lane < nb) a[i] -= uki * ukj;
                        }
                    }
                }
            }
            if (lane < nb)
#pragma unroll
                for (int i = 0; i < NB; ++i)
                    if (i <= lane) sD[lane][i] = a[i];
            if (lane == 0 && bad) s_info = k0 + bad;
        }
        __syncthreads();
        if (s_info) break;
        for (int e = tid; e < nb * nb; e += NT) {
            int i = e % nb, j = e / nb;
            if (i <= j) G[(k0 + i) + (size_t)(k0 + j) * ldg] = sD[j][i];
        }
        if (w == 0) break;  // last panel; nb == NB whenever w > 0
        // Row panel U12 = U11^{-T} A12, one column per thread.
        for (int t = tid; t < w; t += NT) {
            // Keep the LDS reads of U11 inside the loop (hoisting all NB^2 of
            // them out of it exhausts the register file).
            asm volatile("" ::: "memory");
            const int c = k0 + NB + t;
            double x[NB];
#pragma unroll
            for (int kk = 0; kk < NB; ++kk) x[kk] = G[(k0 + kk) + (size_t)c * ldg];
#pragma unroll
            for (int kk = 0; kk < NB; ++kk) {
                double s = x[kk];
#pragma unroll
                for (int p = 0; p < kk; ++p) s -= sD[kk][p] * x[p];
                x[kk] = s / sD[kk][kk];
            }
#pragma unroll
            for (int kk = 0; kk < NB; ++kk) {
                G[(k0 + kk) + (size_t)c * ldg] = x[kk];
                sP[kk][t] = x[kk];
            }
        }
        __syncthreads();
        // A22 -= U12^T U12 on the upper triangle in 16 x 16 blocks on FP64 MFMA
        // (k = 16: four 16x16x4 steps, both operands from the LDS panel).
        // Waves take blocks round robin; the four target values of the next
        // block are loaded (clamped addresses, no branch) before this
        // block's MFMAs.
        const int wb = (w + 15) / 16;
        const int nblk = wb * (wb + 1) / 2;
        double* A22 = G + (k0 + NB) + (size_t)(k0 + NB) * ldg;
        auto block_of = [&](int blk, int& bi, int& bj) {
            bi = 0;
            int rem = blk;
            while (rem >= wb - bi) {
                rem -= wb - bi;
                ++bi;
            }
            bj = bi + rem;
        };
        auto load_target = [&](int blk, double (&tv)[4]) {
            int bi, bj;
            block_of(min(blk, nblk - 1), bi, bj);
            const int c = min(16 * bj + li, w - 1);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = min(16 * bi + g + 4 * r, w - 1);
                tv[r] = A22[row + (size_t)c * ldg];
            }
        };
        double tv[4];
        load_target(wave, tv);
        for (int blk = wave; blk < nblk; blk += NW) {
            double tn[4];
            load_target(blk + NW, tn);
            int bi, bj;
            block_of(blk, bi, bj);
            dh_d4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int s4 = 0; s4 < NB / 4; ++s4) {
                const int kk = 4 * s4 + g;
                const double av = sP[kk][min(16 * bi + li, w - 1)];
                const double bv = sP[kk][min(16 * bj + li, w - 1)];
                acc = Mfma<double>::op(av, bv, acc);
            }
            const int c = 16 * bj + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * bi + g + 4 * r;  // accumulator r: row g + 4r, column li
                if (row < w && c < w && row <= c) A22[row + (size_t)c * ldg] = tv[r] - acc[r];
                tv[r] = tn[r];
            }
        }
        __syncthreads();
    }
    if (s_info) {
        if (tid == 0) *info = s_info;
        return;
    }
    __syncthreads();
    // Inverses of the 16 x 16 diagonal blocks: wave per block, lane per column.
    const int nb16 = (n + 15) / 16;
    for (int b = wave; b < nb16; b += NW) {
        const int c0 = 16 * b, nb = min(16, n - c0);
        // Lane j < 16 loads column j of the block, then rows are broadcast
        // from lane k with readlane: U[i][k] = column k, row i.
        double col[16];
#pragma unroll
        for (int i = 0; i < 16; ++i)
            col[i] = (lane < nb && i < nb) ? G[(c0 + i) + (size_t)(c0 + lane) * ldg] : 0.0;
        double x[16];
#pragma unroll
        for (int i = 15; i >= 0; --i) {
            double s = (i == lane) ? 1.0 : 0.0;
            if (i < nb) {
#pragma unroll
                for (int k = i + 1; k < 16; ++k)
                    if (k < nb) s -= readlane_d(col[i], k) * x[k];
                x[i] = s / readlane_d(col[i], i);
            } else {
                x[i] = 0.0;
            }
        }
        if (lane < 16)
#pragma unroll
            for (int i = 0; i < 16; ++i)
                Dinv[256 * b + i + 16 * lane] = (i <= lane && lane < nb) ? x[i] : 0.0;
    }
    if (tid == 0) *info = 0;
}

__global__ __launch_bounds__(NT) void chol_small_kernel(int n, double* __restrict__ G, int ldg,
                                                        double* __restrict__ Dinv, int m_rows,
                                                        int shift, int rel_check,
                                                        int* __restrict__ info) {
    chol_small_body(n, G, ldg, Dinv, m_rows, shift, rel_check, info);
}

// Batched: workgroup b factors problem b of the table.
__global__ __launch_bounds__(NT) void chol_small_batched_kernel(const DhSpdJob* __restrict__ jobs,
                                                                int shift, int rel_check) {
    const DhSpdJob& j = jobs[blockIdx.x];
    chol_small_body(j.n, j.G, j.ldg, j.Dinv, j.m, shift, rel_check, j.info);
    // Thread 0 wrote *j.info last inside the body.
    if (threadIdx.x == 0 && *j.info != 0) *j.sticky = 1;
}

// Q = Y U^{-1} in place for the m x n panel Y (n <= 544), one wave per 16 rows.
// The wave keeps its 16 solved rows in LDS (odd row pitch: conflict-free
// A-operand reads) and for each 16-column block J accumulates
//   acc = sum_{K<J} Q[:,K] U[K,J]      (f64 MFMA 16x16x4 steps, U from L2)
// then Q_J = (Y_J - acc) U_JJ^{-1} (four more MFMAs, U_JJ^{-1} from Dinv).
template <int WAVES>
__device__ __forceinline__ void trsm_rows_body(int blk, int m, int n, double* __restrict__ Y,
                                               int ldy, const double* __restrict__ U, int ldu,
                                               const double* __restrict__ Dinv, int pitch) {
    extern __shared__ double smem[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, li = lane & 15;
    double* Qs = smem + (size_t)wave * 16 * pitch;  // Qs[row * pitch + col]
    const int r0 = (blk * WAVES + wave) * 16;
    if (r0 >= m) return;
    const int nblk = (n + 15) / 16;
    // Y values and the U_JJ^{-1} column of block J are loaded one block ahead.
    double yv[4], dv[4];
    auto load_block = [&](int J, double (&y4)[4], double (&d4)[4]) {
        const int col = 16 * J + li;
        const bool cok = col < n;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int row = r0 + g + 4 * r;
            y4[r] = (row < m && cok) ? Y[row + (size_t)col * ldy] : 0.0;
        }
        const double* Di = Dinv + 256 * J;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) d4[ks] = Di[(4 * ks + g) + 16 * li];  // U_JJ^{-1}[4ks+g][li]
    };
    load_block(0, yv, dv);
    for (int J = 0; J < nblk; ++J) {
        const int c0 = 16 * J, col = c0 + li;
        const bool cok = col < n;
        double yn[4] = {0, 0, 0, 0}, dn[4] = {0, 0, 0, 0};
        if (J + 1 < nblk) load_block(J + 1, yn, dn);
        // acc = sum_{K<J} Q[:,K] U[K,J]: four independent accumulator chains per
        // k-step of 16; the U column is prefetched PF steps ahead through a
        // register ring (the loop is unrolled by PF so the ring index is static).
        constexpr int PF = 4;
        dh_d4 acc0 = {0, 0, 0, 0}, acc1 = acc0, acc2 = acc0, acc3 = acc0;
        const double* Ucol = U + (size_t)(cok ? col : 0) * ldu;
        const double cm = cok ? 1.0 : 0.0;
        double bq[PF][4];
#pragma unroll
        for (int p = 0; p < PF; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q) bq[p][q] = (16 * p < c0) ? Ucol[16 * p + 4 * q + g] : 0.0;
        auto step = [&](int k, double (&b)[4]) {
            const double* qa = Qs + li * pitch + k + g;  // Q[r0+li][k+g+4s]
            acc0 = Mfma<double>::op(qa[0], b[0] * cm, acc0);
            acc1 = Mfma<double>::op(qa[4], b[1] * cm, acc1);
            acc2 = Mfma<double>::op(qa[8], b[2] * cm, acc2);
            acc3 = Mfma<double>::op(qa[12], b[3] * cm, acc3);
        };
        int k = 0;
        for (; k + 16 * PF <= c0; k += 16 * PF) {
#pragma unroll
            for (int p = 0; p < PF; ++p) {
                double b[4] = {bq[p][0], bq[p][1], bq[p][2], bq[p][3]};
                const int kn = k + 16 * (p + PF);
                if (kn < c0)  // wave-uniform
#pragma unroll
                    for (int q = 0; q < 4; ++q) bq[p][q] = Ucol[kn + 4 * q + g];
                step(k + 16 * p, b);
            }
        }
#pragma unroll
        for (int p = 0; p < PF; ++p)
            if (k + 16 * p < c0) step(k + 16 * p, bq[p]);
        dh_d4 acc = (acc0 + acc1) + (acc2 + acc3);
        // T = Y_J - acc; accumulator r of lane L is row g + 4r, column li.
#pragma unroll
        for (int r = 0; r < 4; ++r) Qs[(g + 4 * r) * pitch + col] = yv[r] - acc[r];
        wave_lds_sync();
        dh_d4 q = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            q = Mfma<double>::op(Qs[li * pitch + c0 + 4 * ks + g], dv[ks], q);  // T[li][4ks+g]
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int row = r0 + g + 4 * r;
            Qs[(g + 4 * r) * pitch + col] = q[r];
            if (row < m && cok) Y[row + (size_t)col * ldy] = q[r];
        }
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            yv[r] = yn[r];
            dv[r] = dn[r];
        }
    }
}

// Q = Y U^{-1} in place for n <= 16 * NBM (the panel widths of the range
// finder up to l = 272), one wave per 16 rows, with the wave's solved blocks
// kept in registers in the MFMA A-operand layout,
//   qa[K][s] = Q[r0 + li][16 K + 4 s + g],
// instead of an LDS panel, so LDS does not cap the waves per SIMD.  For each
// block column J the workgroup stages U[0:16J, 16J:16J+16] in LDS once
// (coalesced column reads; every wave then reads its B operands there
// instead of 16 scattered columns from L2).  The products are formed
// transposed, so each solved block comes out of the MFMA already in the
// operand layout.  Both block loops are unrolled (register
// arrays need static indices); blocks past n are skipped by a uniform test.
// Every wave of the workgroup must call this (workgroup barriers); rows
// past m are neither read nor written.
template <int NBM, int WAVES>
__device__ __forceinline__ void trsm_rows_reg_body(int r0, int m, int n, double* __restrict__ Y,
                                                   int ldy, const double* __restrict__ U,
                                                   int ldu, const double* __restrict__ Dinv,
                                                   double (*sU)[17]) {
    constexpr int NT_ = WAVES * 64;
    static_assert(NT_ == 256 || NT_ == 512, "block column J of U: 256 J values over the threads");
    // Values of U's block column J staged per thread (J at 256 threads).
    auto per_thread = [](int J) { return (256 * J + NT_ - 1) / NT_; };
    const int lane = threadIdx.x & 63, g = lane >> 4, li = lane & 15;
    const int nblk = (n + 15) / 16;
    const int rowA = r0 + li;  // A-operand layout: lane (g, li) holds row li, column 4s + g
    double qa[NBM][4];
    // Loads of block column J, issued one block ahead: Y and Dinv in the
    // A-operand layout (16-row runs of a column), U's rows above the block.
    constexpr int NPF = NBM > 1 ? (256 * (NBM - 1) + NT_ - 1) / NT_ : 1;
    double yt[4], dv[4], pf[NPF];
    auto load_y = [&](int J) {
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const int col = 16 * J + 4 * s4 + g;
            const double v = Y[min(rowA, m - 1) + (size_t)min(col, n - 1) * ldy];
            yt[s4] = col < n ? v : 0.0;
            dv[s4] = Dinv[256 * J + (4 * s4 + g) + 16 * li];
        }
    };
    load_y(0);
#pragma unroll
    for (int J = 0; J < NBM; ++J) {
        if (J < nblk) {  // uniform over the workgroup
            double yc[4], dc[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                yc[s4] = yt[s4];
                dc[s4] = dv[s4];
            }
            if (J > 0) {
                // sU[k][c] = U[k][16J + c] for k < 16J (zero past column n),
                // from the registers loaded during block J - 1.
                __syncthreads();  // the previous block column's readers are done
#pragma unroll
                for (int i = 0; i < per_thread(J); ++i) {
                    const int e = threadIdx.x + i * NT_;
                    const int k = e % (16 * J), c = e / (16 * J);
                    if (NT_ == 256 || e < 256 * J) sU[k][c] = (16 * J + c < n) ? pf[i] : 0.0;
                }
                __syncthreads();
            }
            if (J + 1 < nblk) {
                load_y(J + 1);
                const int J1 = J + 1;  // a constant once the J loop is unrolled
                if (J1 < NBM) {
#pragma unroll
                    for (int i = 0; i < per_thread(J1); ++i) {
                        const int e = min(threadIdx.x + i * NT_, 256 * J1 - 1);
                        const int k = e % (16 * J1), c = e / (16 * J1);
                        pf[i] = U[k + (size_t)min(16 * J1 + c, n - 1) * ldu];
                    }
                }
            }
            // The products are formed transposed, acc^T = sum_{K<J} U_KJ^T Q_K^T
            // (one accumulator chain per k-slot s): the accumulator layout of a
            // transposed product (lane (g, li), value r: row li, column g + 4r of
            // the untransposed one) is the A-operand layout, so no LDS round trip.
            dh_d4 acc[4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
            for (int K = 0; K < J; ++K)
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4)
                    acc[s4] = Mfma<double>::op(sU[16 * K + 4 * s4 + g][li], qa[K][s4], acc[s4]);
            const dh_d4 a = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            // Q_J^T = U_JJ^{-T} (Y_J - acc)^T, again in the A-operand layout.
            dh_d4 q = {0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) q = Mfma<double>::op(dc[ks], yc[ks] - a[ks], q);
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                qa[J][s4] = q[s4];
                const int col = 16 * J + 4 * s4 + g;
                if (rowA < m && col < n) Y[rowA + (size_t)col * ldy] = qa[J][s4];
            }
        }
    }
}

constexpr int kRegBlocks = 17;  // register-resident row solve for n <= 272

template <int WAVES>
__global__ __launch_bounds__(WAVES * 64) void trsm_rows_reg_kernel(int m, int n,
                                                                   double* __restrict__ Y,
                                                                   int ldy,
                                                                   const double* __restrict__ U,
                                                                   int ldu,
                                                                   const double* __restrict__ Dinv,
                                                                   const int* __restrict__ guard,
                                                                   int nguard) {
    __shared__ double sU[16 * (kRegBlocks - 1)][17];
    // Whole-workgroup exit when an earlier Cholesky broke down.
    if (guard && (guard[0] | (nguard > 1 ? guard[1] : 0))) return;
    const int wave = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * WAVES + wave) * 16;
    trsm_rows_reg_body<kRegBlocks, WAVES>(r0, m, n, Y, ldy, U, ldu, Dinv, sU);
}

template <int WAVES>
__global__ __launch_bounds__(WAVES * 64) void trsm_rows_reg_batched_kernel(
    const DhSpdJob* __restrict__ jobs) {
    __shared__ double sU[16 * (kRegBlocks - 1)][17];
    const DhSpdJob& j = jobs[blockIdx.y];
    // Whole-workgroup exit: the body synchronises the workgroup.
    if (blockIdx.x * WAVES * 16 >= (unsigned)j.m || *j.info != 0) return;
    const int wave = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * WAVES + wave) * 16;
    trsm_rows_reg_body<kRegBlocks, WAVES>(r0, j.m, j.n, j.Y, j.ldy, j.G, j.ldg, j.Dinv, sU);
}

template <int WAVES>
__global__ __launch_bounds__(WAVES * 64) void trsm_rows_kernel(int m, int n,
                                                               double* __restrict__ Y, int ldy,
                                                               const double* __restrict__ U,
                                                               int ldu,
                                                               const double* __restrict__ Dinv,
                                                               int pitch,
                                                               const int* __restrict__ guard,
                                                               int nguard) {
    if (guard && (guard[0] | (nguard > 1 ? guard[1] : 0))) return;  // a Cholesky broke down
    trsm_rows_body<WAVES>(blockIdx.x, m, n, Y, ldy, U, ldu, Dinv, pitch);
}

// Batched: blockIdx.y selects the problem; problems whose Cholesky failed
// (*info != 0) are left untouched, as the unbatched sweep does.
template <int WAVES>
__global__ __launch_bounds__(WAVES * 64) void trsm_rows_batched_kernel(
    const DhSpdJob* __restrict__ jobs) {
    const DhSpdJob& j = jobs[blockIdx.y];
    if (blockIdx.x * WAVES * 16 >= (unsigned)j.m || *j.info != 0) return;
    trsm_rows_body<WAVES>(blockIdx.x, j.m, j.n, j.Y, j.ldy, j.G, j.ldg, j.Dinv,
                          16 * ((j.n + 15) / 16) + 1);
}

// Lane I of each 16-lane row, broadcast to the row (DPP row_newbcast): no
// LDS round trip and no scalar registers.  Every lane is written (full row and
// bank masks), so the move takes no old value: with update_dpp(0, ...) the
// compiler zeroed the destination before each of the 240 moves per 16 x 16
// factorisation (register Cholesky at order 272: 151 -> 144 us).
template <int I>
__device__ __forceinline__ double bcast16(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x150 + I, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x150 + I, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// u[I..15] -= U(KK, I) * ukj, U(KK, I) being ukj of lane I.
template <int I>
__device__ __forceinline__ void chol16_update(double (&u)[NB], double ukj) {
    if constexpr (I < NB) {
        u[I] -= bcast16<I>(ukj) * ukj;
        chol16_update<I + 1>(u, ukj);
    }
}

// Column steps KK..15 of the 16 x 16 Cholesky, lane j (mod 16) holding
// column j; bad receives the 1-based failing step.
template <int KK>
__device__ __forceinline__ void chol16_steps(double (&u)[NB], int li, int nreal, double tol,
                                             int& bad) {
    if constexpr (KK < NB) {
        if (!bad) {
            const double piv = readlane_d(u[KK], KK);
            if (KK < nreal && !(piv > tol)) {
                bad = KK + 1;
            } else {
                // 1/sqrt(piv): hardware estimate and two Newton steps.
                const double hp = 0.5 * piv;
                double y = __builtin_amdgcn_rsq(piv);
                y = fma(y, fma(-hp * y, y, 0.5), y);
                y = fma(y, fma(-hp * y, y, 0.5), y);
                const double ukj = (li == KK) ? piv * y : u[KK] * y;
                u[KK] = ukj;
                chol16_update<KK + 1>(u, ukj);
            }
        }
        chol16_steps<KK + 1>(u, li, nreal, tol, bad);
    }
}

// Register-resident Cholesky for orders n <= 16 * kRegBlocks (= 272), same
// contract as chol_small_body.  The order-16*nb matrix (nb = ceil(n/16); rows
// and columns past n hold the identity, so they factor to the identity and
// skip the pivot test) is held as 16 x 16 blocks: the off-diagonal blocks
// (bi, bj > bi) in registers, in FP64 MFMA accumulator layout (lane: rows
// g + 4r, column li), numbered row by row and dealt to the 8 waves round
// robin; the diagonal blocks in LDS (sG[i][col][row]), block i updated by
// wave i % 8.  Per 16-column panel k:
//   A  wave k % 8 factors block (k, k) (column per lane; rows of U broadcast
//      by DPP within 16-lane rows), inverts the factor by back substitution
//      (rows of U from LDS), and writes
//      U_kk to G, its inverse to Dinv and to LDS;
//   B  the owners of blocks (k, j > k) form U_kj = U_kk^{-T} A_kj (4 MFMAs,
//      A_kj's accumulator rows are the B operand) into the LDS panel, which
//      the workgroup then copies to G;
//   C  A_ij -= U_ki^T U_kj for k < i <= j (4 MFMAs per block from the LDS
//      panel).  Wave (k+1) % 8 updates block (k+1, k+1) first and runs step
//      A of panel k+1 before its other updates, so that factorisation
//      overlaps the rest of the trailing update.
// Two workgroup barriers per panel and no global loads after the first.
constexpr int kCholWaves = NT / 64;
constexpr int kCholSlots =
    (kRegBlocks * (kRegBlocks - 1) / 2 + kCholWaves - 1) / kCholWaves;  // 17

// LDS of chol_reg_body (kernels declare it __shared__ or carve it from an arena).
struct CholRegLds {
    double sG[kRegBlocks][NB][NB + 1];  // diagonal blocks, sG[i][col][row]
    double sI[NB][NB + 1];              // inverse factor of panel k, sI[j][i] = U_kk^{-1}(i, j)
    __align__(16) double sP[NB][16 * (kRegBlocks - 1) + 4];  // U_k,k+1..nb-1
    __align__(16) double sU[NB][NB];                         // U_kk^T for the back substitution
    double sRd[NB];                                          // 1 / U_kk(i, i)
    double red[2][kCholWaves];                               // diagonal max and trace
    int s_info;
};

// scales (nullable): {tol, shift} computed from a larger matrix whose trailing
// block G is (the Schur complement step of an order-544 factorisation); then
// they replace the ones from G's own diagonal.
__device__ __forceinline__ void chol_reg_body(CholRegLds& L, int n, double* __restrict__ G,
                                              int ldg, double* __restrict__ Dinv, int m_rows,
                                              int shift, int rel_check, int* __restrict__ info,
                                              const double* __restrict__ scales = nullptr) {
    auto& sG = L.sG;
    auto& sI = L.sI;
    auto& sP = L.sP;
    auto& sU = L.sU;
    auto& sRd = L.sRd;
    int& s_info = L.s_info;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane >> 4, li = lane & 15;

    static_assert(NT >= 16 * kRegBlocks, "one thread per diagonal element");
    double tol = 0.0;  // breakdown tolerance, set once the diagonal is in LDS
    const int nbk = (n + 15) / 16;
    int bI[kCholSlots], bJ[kCholSlots];  // block of each slot; bI == nbk: unused
#pragma unroll
    for (int s = 0; s < kCholSlots; ++s) {
        int rem = s * kCholWaves + wave, bi = 0;
        while (bi < nbk && rem >= nbk - 1 - bi) {
            rem -= nbk - 1 - bi;
            ++bi;
        }
        bI[s] = bi;
        bJ[s] = bi + 1 + rem;
    }
    // Diagonal blocks to LDS: upper triangle, zero below, identity padding
    // (all loads issued first, clamped addresses).
    {
        constexpr int kIt = (256 * kRegBlocks + NT - 1) / NT;
        const int ne = 256 * nbk;
        double dv[kIt];
#pragma unroll
        for (int it = 0; it < kIt; ++it) {
            const int e = min(tid + it * NT, ne - 1);
            const int i = e >> 8, row = e & 15, col = (e >> 4) & 15;
            dv[it] = G[min(16 * i + row, n - 1) + (size_t)min(16 * i + col, n - 1) * ldg];
        }
#pragma unroll
        for (int it = 0; it < kIt; ++it) {
            const int e = tid + it * NT;
            if (e < ne) {
                const int i = e >> 8, row = e & 15, col = (e >> 4) & 15;
                const int gr = 16 * i + row, gc = 16 * i + col;
                double x = (gc < n) ? dv[it] : (gr == gc ? 1.0 : 0.0);
                sG[i][col][row] = (row > col) ? 0.0 : x;
            }
        }
    }
    // Off-diagonal blocks to registers in groups of four slots: the group's
    // 16 loads are issued together (clamped addresses), then padded.
    dh_d4 a[kCholSlots];
#pragma unroll
    for (int s0 = 0; s0 < kCholSlots; s0 += 4) {
        double v[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int s = min(s0 + j, kCholSlots - 1);
            const int cc = min(16 * bJ[s] + li, n - 1);
#pragma unroll
            for (int r = 0; r < 4; ++r)
                v[j][r] = G[min(16 * bI[s] + g + 4 * r, n - 1) + (size_t)cc * ldg];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int s = s0 + j;
            if (s < kCholSlots) {
                const int c = 16 * bJ[s] + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * bI[s] + g + 4 * r;
                    a[s][r] = (row < n && c < n) ? v[j][r] : 0.0;
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);  // keep the groups apart (register pressure)
    }

    // Step A for panel k (wave k % 8).
    auto factor_diag = [&](int k) {
        // Lane j holds column j & 15 of the block (every 16-lane row a full
        // copy, so row broadcasts serve all lanes).  Entries below the
        // diagonal are not needed and take whatever the unmasked updates
        // leave there.
        double u[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) u[i] = sG[k][li][i];
        int bad = 0;
        chol16_steps<0>(u, li, n - 16 * k, tol, bad);
        const int c = 16 * k + lane;
        if (lane < NB) {
#pragma unroll
            for (int i = 0; i < NB; ++i) sU[i][lane] = u[i];
            double dg = 1.0;
#pragma unroll
            for (int i = 0; i < NB; ++i)
                if (i == lane) dg = u[i];
            sRd[lane] = 1.0 / dg;
            if (c < n)
#pragma unroll
                for (int i = 0; i < NB; ++i)
                    if (i <= lane) G[(16 * k + i) + (size_t)c * ldg] = u[i];
        }
        wave_lds_sync();
        // Column j of U_kk^{-1} by back substitution, in place: x[i] =
        // (e_j[i] - sum_{q>i} U(i, q) x[q]) / U(i, i), two partial sums.
#pragma unroll
        for (int i = NB - 1; i >= 0; --i) {
            double t0 = (i == lane) ? 1.0 : 0.0, t1 = 0.0;
#pragma unroll
            for (int q = i + 1; q < NB; ++q) {
                if ((q - i) & 1)
                    t0 -= sU[i][q] * u[q];
                else
                    t1 -= sU[i][q] * u[q];
            }
            u[i] = (t0 + t1) * sRd[i];
        }
        if (lane < NB)
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                sI[lane][i] = u[i];
                Dinv[256 * k + i + 16 * lane] = (i <= lane && c < n) ? u[i] : 0.0;
            }
        if (lane == 0 && bad) s_info = 16 * k + bad;
    };
    // Diagonal block i -= U_ki^T U_ki (panel k in sP), in LDS.
    auto update_diag = [&](int i, int k) {
        const int pi = 16 * (i - k - 1) + li;
        dh_d4 d;
#pragma unroll
        for (int r = 0; r < 4; ++r) d[r] = sG[i][li][g + 4 * r];
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const double v = sP[4 * s4 + g][pi];
            d = Mfma<double>::op(-v, v, d);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) sG[i][li][g + 4 * r] = d[r];
    };

    __syncthreads();  // diagonal blocks in LDS
    {
        // Breakdown tolerance and CholeskyQR3 shift from the diagonal, as
        // chol_scales computes them, then the shift applied in LDS.
        auto& red = L.red;
        double mx = 0.0, tr = 0.0;
        if (tid < n) {
            mx = sG[tid >> 4][tid & 15][tid & 15];
            tr = mx;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mx = fmax(mx, __shfl_xor(mx, o));
            tr += __shfl_xor(tr, o);
        }
        if (lane == 0) {
            red[0][wave] = mx;
            red[1][wave] = tr;
        }
        __syncthreads();
        double m2 = 0.0, t2 = 0.0;
#pragma unroll
        for (int w = 0; w < kCholWaves; ++w) {
            m2 = fmax(m2, red[0][w]);
            t2 += red[1][w];
        }
        tol = rel_check ? (double)n * 2.220446049250313e-16 * m2 : 0.0;
        double sh = shift ? 11.0 * ((double)m_rows * n + (double)n * (n + 1)) *
                                1.1102230246251565e-16 * t2
                          : 0.0;
        if (scales) {
            tol = scales[0];
            sh = scales[1];
        }
        if (shift && tid < n) sG[tid >> 4][tid & 15][tid & 15] += sh;
        if (tid == 0) s_info = 0;
    }
    __syncthreads();  // shifted diagonal blocks and s_info in LDS
    if (wave == 0) factor_diag(0);
    for (int k = 0;; ++k) {
        __syncthreads();  // U_kk, its inverse and s_info ready
        if (s_info || k + 1 == nbk) break;
        // ---- B: row panel U_kj = U_kk^{-T} A_kj.
#pragma unroll
        for (int s = 0; s < kCholSlots; ++s) {
            if (bI[s] == k) {
                dh_d4 acc = {0, 0, 0, 0};
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4)
                    acc = Mfma<double>::op(sI[li][4 * s4 + g], a[s][s4], acc);
                const int pc = 16 * (bJ[s] - k - 1) + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) sP[g + 4 * r][pc] = acc[r];
            }
        }
        __syncthreads();  // panel k in sP
        {  // the panel to G, 16-row runs per column
            const int cols = min(n, 16 * nbk) - 16 * (k + 1);
            for (int e = tid; e < 16 * cols; e += NT) {
                const int r = e & 15, cl = e >> 4;
                G[(16 * k + r) + (size_t)(16 * (k + 1) + cl) * ldg] = sP[r][cl];
            }
        }
        // ---- C: trailing update; block (k+1, k+1) and step A of panel k+1
        // first on wave (k+1) % 8.
        if (wave == (k + 1) % kCholWaves) {
            update_diag(k + 1, k);
            wave_lds_sync();
            factor_diag(k + 1);
        }
        for (int i = k + 2 + (wave - k - 2 + 8 * kCholWaves) % kCholWaves; i < nbk;
             i += kCholWaves)
            update_diag(i, k);
#pragma unroll
        for (int s = 0; s < kCholSlots; ++s) {
            if (bI[s] > k && bI[s] < nbk) {
                const int pi = 16 * (bI[s] - k - 1) + li, pj = 16 * (bJ[s] - k - 1) + li;
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4)
                    a[s] = Mfma<double>::op(-sP[4 * s4 + g][pi], sP[4 * s4 + g][pj], a[s]);
            }
        }
    }
    if (tid == 0) *info = s_info;
}

__global__ __launch_bounds__(NT) void chol_reg_kernel(int n, double* __restrict__ G, int ldg,
                                                      double* __restrict__ Dinv, int m_rows,
                                                      int shift, int rel_check,
                                                      int* __restrict__ info,
                                                      const double* __restrict__ scales) {
    __shared__ CholRegLds lds;
    chol_reg_body(lds, n, G, ldg, Dinv, m_rows, shift, rel_check, info, scales);
}

// {tol, shift} of the order-n G from its whole diagonal, for the two block
// factorisations of an order in (272, 544].
__global__ __launch_bounds__(NT) void chol_scales_kernel(int n, const double* __restrict__ G,
                                                         int ldg, int m_rows, int shift,
                                                         int rel_check,
                                                         double* __restrict__ scales) {
    double tol, sh;
    chol_scales(n, G, ldg, m_rows, shift, rel_check, tol, sh);
    if (threadIdx.x == 0) {
        scales[0] = tol;
        scales[1] = sh;
    }
}

// Status of the two block factorisations as one order-n status: the first
// failing pivot of the leading block, else n1 + that of the trailing one.
__global__ void chol_info_combine_kernel(const int* __restrict__ parts, int n1,
                                         int* __restrict__ info) {
    if (threadIdx.x == 0) *info = parts[0] ? parts[0] : (parts[1] ? n1 + parts[1] : 0);
}

__global__ __launch_bounds__(NT) void chol_reg_batched_kernel(const DhSpdJob* __restrict__ jobs,
                                                              int shift, int rel_check) {
    const DhSpdJob& j = jobs[blockIdx.x];
    __shared__ CholRegLds lds;
    chol_reg_body(lds, j.n, j.G, j.ldg, j.Dinv, j.m, shift, rel_check, j.info);
    if (threadIdx.x == 0 && *j.info != 0) *j.sticky = 1;
}

// Gram G = upper(Y^T Y) of an m x n panel, 128 < n <= 16 * kRegBlocks: the
// CholeskyQR product, whose 64 x 64-tiled GEMM would pad 272 to 320.  A
// workgroup (8 waves) takes a slab of rows and accumulates all upper 16 x 16
// blocks of the order-16*kRegBlocks Gram in registers (FP64 MFMA accumulator
// layout); block b (153 of them, dealt by kGramMap) belongs to wave b % 8,
// fixed at compile time so every operand index is static.  The slab streams
// through LDS in 16-row chunks (double buffered, the next chunk's loads in
// flight during the MFMAs); per 4-row step a wave reads the column operands
// its blocks use once and issues their MFMAs (A = B layout: lane (g, li)
// holds Y(k + g, 16 c + li) for column block c).  Partial sums go to
// part[z][block][r][lane] and gram_reduce_kernel adds the slabs.
constexpr int kGramBlocks = kRegBlocks * (kRegBlocks + 1) / 2;           // 153
constexpr int kGramSlots = (kGramBlocks + kCholWaves - 1) / kCholWaves;  // 20
constexpr int kGramPitch = 16 * kRegBlocks + 4;

struct GramMap {
    int bi[kGramBlocks], bj[kGramBlocks];
};
// Block b = 8 S + W is slot S of wave W.  The blocks are dealt in compact
// groups (found by a local search over partitions of the 153 blocks): each
// wave's 19 or 20 blocks use 8 of the 17 column operands, so a wave reads 8
// operands per 4-row step instead of 17 when dealt row by row.
__host__ __device__ constexpr GramMap make_gram_map() {
    constexpr int bi[kGramBlocks] = {
        0, 2, 1, 1, 1, 0, 0, 0, 0, 2, 1, 1, 1, 0, 0, 0, 0, 2, 1, 1, 1, 0, 0, 2, 0, 2, 1, 1, 1, 0,
        4, 2, 0, 2, 3, 1, 2, 0, 4, 2, 0, 4, 3, 3, 2, 1, 4, 2, 0, 4, 3, 3, 5, 1, 4, 3, 5, 4, 3, 3,
        5, 1, 4, 3, 5, 4, 7, 3, 5, 2, 6, 3, 5, 4, 7, 4, 5, 2, 6, 7, 5, 5, 7, 4, 8, 2, 6, 7, 6, 5,
        7, 4, 8, 2, 8, 7, 6, 6, 10, 6, 8, 3, 8, 7, 6, 6, 10, 6, 10, 3, 9, 8, 8, 6, 10, 9, 10, 3,
        9, 8, 8, 7, 12, 9, 10, 5, 9, 11, 10, 7, 12, 9, 11, 5, 9, 11, 12, 9, 13, 13, 11, 12, 11,
        15, 14, 14, 13, 14, 13, 12, 11, 16, 15};
    constexpr int bj[kGramBlocks] = {
        0, 4, 3, 4, 1, 1, 4, 7, 5, 6, 7, 6, 8, 2, 6, 11, 8, 9, 15, 9, 11, 3, 9, 3, 10, 14, 16, 10,
        14, 13, 6, 7, 12, 16, 3, 13, 8, 16, 8, 11, 14, 5, 7, 4, 10, 2, 11, 15, 15, 7, 10, 6, 5, 5,
        12, 8, 6, 9, 12, 9, 10, 12, 15, 11, 8, 14, 10, 14, 11, 2, 8, 15, 12, 16, 12, 4, 14, 5, 9,
        7, 15, 7, 13, 10, 8, 12, 11, 8, 6, 9, 16, 13, 10, 13, 9, 11, 12, 7, 12, 10, 13, 5, 11, 15,
        15, 14, 15, 13, 11, 13, 9, 15, 12, 16, 16, 10, 13, 16, 11, 16, 14, 9, 13, 13, 14, 13, 12,
        11, 10, 14, 15, 14, 13, 16, 15, 16, 14, 16, 15, 14, 14, 12, 12, 16, 15, 16, 16, 14, 13,
        16, 15, 16, 15};
    GramMap m{};
    for (int b = 0; b < kGramBlocks; ++b) {
        m.bi[b] = bi[b];
        m.bj[b] = bj[b];
    }
    return m;
}
// Every upper block (bi <= bj < kRegBlocks) appears exactly once.
__host__ __device__ constexpr bool gram_map_is_partition() {
    constexpr GramMap m = make_gram_map();
    for (int i = 0; i < kRegBlocks; ++i)
        for (int j = i; j < kRegBlocks; ++j) {
            int hits = 0;
            for (int b = 0; b < kGramBlocks; ++b) hits += (m.bi[b] == i && m.bj[b] == j);
            if (hits != 1) return false;
        }
    return true;
}
static_assert(gram_map_is_partition(), "Gram block map must cover the upper triangle once");
constexpr GramMap kGramMap = make_gram_map();

// acc[S] += Y_bi^T Y_bj for the blocks of wave W (indices fixed at compile time).
template <int W, int S>
__device__ __forceinline__ void gram_mfma(dh_d4 (&acc)[kGramSlots],
                                          const double (&op)[kRegBlocks], int nbk) {
    if constexpr (S < kGramSlots) {
        constexpr int b = S * kCholWaves + W;
        if constexpr (b < kGramBlocks) {
            constexpr int bi = kGramMap.bi[b], bj = kGramMap.bj[b];
            if (bj < nbk) acc[S] = Mfma<double>::op(op[bi], op[bj], acc[S]);
        }
        gram_mfma<W, S + 1>(acc, op, nbk);
    }
}

template <int W>
__device__ __forceinline__ void gram_body(int m, int n, const double* __restrict__ Y, int ldy,
                                          int r0, int r1, double* __restrict__ pz,
                                          double (*sY)[16][kGramPitch]) {
    const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, li = lane & 15;
    const int nbk = (n + 15) / 16;
    constexpr int kIt = (16 * 16 * kRegBlocks + NT - 1) / NT;
    dh_d4 acc[kGramSlots];
#pragma unroll
    for (int s = 0; s < kGramSlots; ++s) acc[s] = dh_d4{0, 0, 0, 0};
    // Chunk element e: row e & 15, column e >> 4 (16-row runs of a column).
    auto load = [&](int c0, double (&v)[kIt]) {
#pragma unroll
        for (int it = 0; it < kIt; ++it) {
            const int e = tid + it * NT, row = c0 + (e & 15), col = min(e >> 4, n - 1);
            v[it] = Y[min(row, m - 1) + (size_t)col * ldy];
        }
    };
    auto store = [&](int c0, int buf, const double (&v)[kIt]) {
#pragma unroll
        for (int it = 0; it < kIt; ++it) {
            const int e = tid + it * NT, row = c0 + (e & 15), col = e >> 4;
            if (col < 16 * kRegBlocks)
                sY[buf][e & 15][col] = (row < r1 && col < n) ? v[it] : 0.0;
        }
    };
    double v[kIt];
    load(r0, v);
    store(r0, 0, v);
    __syncthreads();
    int buf = 0;
    for (int c0 = r0; c0 < r1; c0 += 16) {
        const bool more = c0 + 16 < r1;
        if (more) load(c0 + 16, v);
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            double op[kRegBlocks];
#pragma unroll
            for (int c = 0; c < kRegBlocks; ++c) op[c] = sY[buf][4 * s4 + g][16 * c + li];
            gram_mfma<W, 0>(acc, op, nbk);
        }
        if (more) store(c0 + 16, buf ^ 1, v);
        __syncthreads();
        buf ^= 1;
    }
    // Blocks past the order (bj >= nbk) are neither written nor read back.
#pragma unroll
    for (int s = 0; s < kGramSlots; ++s) {
        const int b = s * kCholWaves + W;
        if (b < kGramBlocks && kGramMap.bj[b] < nbk)
#pragma unroll
            for (int r = 0; r < 4; ++r) pz[(size_t)b * 256 + r * 64 + lane] = acc[s][r];
    }
}

template <typename F>
__device__ __forceinline__ void gram_by_wave(F&& body) {
    switch (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) {
        case 0: body(std::integral_constant<int, 0>{}); break;
        case 1: body(std::integral_constant<int, 1>{}); break;
        case 2: body(std::integral_constant<int, 2>{}); break;
        case 3: body(std::integral_constant<int, 3>{}); break;
        case 4: body(std::integral_constant<int, 4>{}); break;
        case 5: body(std::integral_constant<int, 5>{}); break;
        case 6: body(std::integral_constant<int, 6>{}); break;
        default: body(std::integral_constant<int, 7>{}); break;
    }
}

__global__ __launch_bounds__(NT) void gram_reg_kernel(int m, int n, const double* __restrict__ Y,
                                                      int ldy, int rows, double* __restrict__ part) {
    __shared__ double sY[2][16][kGramPitch];
    const int r0 = blockIdx.x * rows, r1 = min(m, r0 + rows);
    double* pz = part + (size_t)blockIdx.x * kGramBlocks * 256;
    gram_by_wave([&](auto w) { gram_body<decltype(w)::value>(m, n, Y, ldy, r0, r1, pz, sY); });
}

// Batched: one workgroup per (problem, slab) item, the batch's problems in
// one launch; the slabs of a problem are reduced in order by the kernel below.
__global__ __launch_bounds__(NT) void gram_reg_batched_kernel(const DhGramRegJob* __restrict__ jobs,
                                                              const DhGramItem* __restrict__ items) {
    __shared__ double sY[2][16][kGramPitch];
    const DhGramItem it = items[blockIdx.x];
    const DhGramRegJob& j = jobs[it.job];
    const int r0 = it.slab * j.rows, r1 = min(j.m, r0 + j.rows);
    double* pz = j.part + (size_t)it.slab * kGramBlocks * 256;
    gram_by_wave([&](auto w) { gram_body<decltype(w)::value>(j.m, j.n, j.Y, j.ldy, r0, r1, pz, sY); });
}

// G(i, j), i <= j < n, = sum over the nz slabs of part: 64 elements per
// workgroup, four thread groups over interleaved slabs, then an LDS sum.
__device__ __forceinline__ void gram_reduce_body(int n, int nz, const double* __restrict__ part,
                                                 double* __restrict__ G, int ldg) {
    __shared__ double red[4][64];
    const int e = blockIdx.x * 64 + (threadIdx.x & 63), zg = threadIdx.x >> 6;
    double s0 = 0.0, s1 = 0.0;
    if (e < kGramBlocks * 256) {
        const size_t stride = (size_t)kGramBlocks * 256;
        int z = zg;
        for (; z + 4 < nz; z += 8) {
            s0 += part[(size_t)z * stride + e];
            s1 += part[(size_t)(z + 4) * stride + e];
        }
        if (z < nz) s0 += part[(size_t)z * stride + e];
    }
    red[zg][threadIdx.x & 63] = s0 + s1;
    __syncthreads();
    if (zg == 0 && e < kGramBlocks * 256) {
        const double t = (red[0][threadIdx.x] + red[1][threadIdx.x]) +
                         (red[2][threadIdx.x] + red[3][threadIdx.x]);
        const int b = e >> 8, r = (e >> 6) & 3, ln = e & 63;
        const int row = 16 * kGramMap.bi[b] + (ln >> 4) + 4 * r;
        const int col = 16 * kGramMap.bj[b] + (ln & 15);
        if (col < n && row <= col) G[row + (size_t)col * ldg] = t;
    }
}

__global__ __launch_bounds__(256) void gram_reduce_kernel(int n, int nz,
                                                          const double* __restrict__ part,
                                                          double* __restrict__ G, int ldg) {
    gram_reduce_body(n, nz, part, G, ldg);
}

// blockIdx.y = problem; workgroups of blocks past its order leave at once.
__global__ __launch_bounds__(256) void gram_reduce_batched_kernel(
    const DhGramRegJob* __restrict__ jobs) {
    const DhGramRegJob& j = jobs[blockIdx.y];
    if (16 * kGramMap.bj[min((int)blockIdx.x * 64 / 256, kGramBlocks - 1)] >= j.n) return;
    gram_reduce_body(j.n, j.nslab, j.part, j.G, j.ldg);
}

}  // namespace
