// Mixed-precision K-passes: BF16 MFMA distance and projection with FP32
// accumulation and an FP32 exp epilogue; FP64 in and out.
//
// Precision scheme (per pass):
//   * X is centred by the mean of the column point set (distances are
//     translation invariant) and rounded to BF16 once per solve;
//   * the squared norms are FP32 sums of the ROUNDED coordinates, so
//     |xi|^2 + |xj|^2 - 2 xi.xj is the squared distance of the rounded
//     points up to FP32 accumulation error, and is clamped at 0;
//   * K = exp2(-gamma log2(e) d2) on v_exp_f32, rounded to BF16 in registers;
//   * Omega (or Q) is rounded to BF16 with the labels y_j folded in, stored
//     in the operand order of the projection MFMA ("packed");
//   * Y accumulates in FP32 and is written as FP64 for CholeskyQR2.
// Reference: DArray::LRA (dkernel.h:12-46) over rbf_kernel + tc_gemm
// (dkernel.cu:132-231), whose tensor-core path is an FP16 GemmEx with FP32
// accumulation (test_cuda.cu:111-132, dkernel.cu:161-163).
#include "dh_common.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr float kLog2e = 1.4426950408889634f;

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
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

// One wave per point row (rows >= n are written as zeros): Xb = bf16(X - mean),
// padded with zeros to dpad features, and a[i] = -gamma log2(e) |Xb_i|^2.
__global__ __launch_bounds__(256) void convert_rows_kernel(int n, int nrows, int d, int dpad,
                                                           const double* __restrict__ X, int ldx,
                                                           const double* __restrict__ mean,
                                                           float neg_gl2e,
                                                           __bf16* __restrict__ Xb,
                                                           float* __restrict__ a) {
    int row = (blockIdx.x * 256 + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (row >= nrows) return;
    float s = 0.f;
    for (int f = lane; f < dpad; f += 64) {
        __bf16 v = (__bf16)0.0f;
        if (row < n && f < d) v = (__bf16)(float)(X[f + (size_t)row * ldx] - mean[f]);
        Xb[(size_t)row * dpad + f] = v;
        float fv = (float)v;
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
__global__ __launch_bounds__(256) void pack_omega_kernel(int n, int l, int nchunks, int ncb_total,
                                                         const double* __restrict__ O, int ldo,
                                                         const double* __restrict__ y,
                                                         bf16x8* __restrict__ Op) {
    long idx = (long)blockIdx.x * 256 + threadIdx.x;
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
        for (int e = 0; e < 8; ++e) w[e] = (__bf16)v[16 * (e >> 2) + 4 * g + (e & 3)];
        Op[((size_t)t * ncb_total + cb) * 64 + 16 * g + li] = w;
    }
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
template <int DK, int NCB, int WAVES>
__global__ __launch_bounds__(WAVES * 64, 1) void rbf_project_bf16_kernel(
    int m, int nchunks, int l, int ncb_total, const bf16x8* __restrict__ Xi,
    const bf16x8* __restrict__ Xj, const float* __restrict__ ai, const float* __restrict__ bj,
    float c2, const bf16x8* __restrict__ Op, const double* __restrict__ yi,
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
    const int cb0 = blockIdx.y * NCB;
    const int i0 = blockIdx.x * (WAVES * 32) + wave * 32;

    bf16x8 xb[2][DK];
    float arow[2];
#pragma unroll
    for (int ib = 0; ib < 2; ++ib) {
        int row = i0 + 16 * ib + li;
#pragma unroll
        for (int ks = 0; ks < DK; ++ks) {
            bf16x8 z = {};
            xb[ib][ks] = (row < m) ? Xi[(size_t)row * XG + ks * 4 + g] : z;
        }
        arow[ib] = (row < m) ? ai[row] : 0.f;
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
            for (int ib = 0; ib < 2; ++ib) s[jb][ib] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < DK; ++ks)
#pragma unroll
            for (int jb = 0; jb < 2; ++jb) {
                bf16x8 a = sX[buf][(16 * jb + li) * XROW + ks * 4 + g];
#pragma unroll
                for (int ib = 0; ib < 2; ++ib) s[jb][ib] = mfma_bf16(a, xb[ib][ks], s[jb][ib]);
            }
    };

    // Prologue: chunks 0 and 1 resident, distance tile of chunk 0 computed.
    const int last = nchunks - 1;
    gload(0);
    sstore(0);
    gload(min(1, last));
    sstore(1);
    __syncthreads();
    f32x4 s_cur[2][2];
    distance(0, s_cur);
    int buf = 0;  // buffer of chunk t
    for (int t = 0; t < nchunks; ++t) {
        const int nbuf = (buf == NBUF - 1) ? 0 : buf + 1;     // chunk t+1
        const int sbuf = (nbuf == NBUF - 1) ? 0 : nbuf + 1;   // chunk t+2 (being staged)
        gload(min(t + 2, last));

        bf16x8 kb[2];
#pragma unroll
        for (int jb = 0; jb < 2; ++jb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float bv = sB[buf][16 * jb + 4 * g + r];
#pragma unroll
                for (int ib = 0; ib < 2; ++ib) {
                    float arg = fmaf(c2, s_cur[jb][ib][r], arow[ib] + bv);
                    kb[ib][4 * jb + r] = (__bf16)__builtin_amdgcn_exp2f(fminf(arg, 0.f));
                }
            }
        f32x4 s_next[2][2];
        distance(nbuf, s_next);

#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            bf16x8 bo = sO[buf][cb * 64 + lane];
#pragma unroll
            for (int ib = 0; ib < 2; ++ib) acc[ib][cb] = mfma_bf16(kb[ib], bo, acc[ib][cb]);
        }
        sstore(sbuf);
        __syncthreads();
#pragma unroll
        for (int jb = 0; jb < 2; ++jb)
#pragma unroll
            for (int ib = 0; ib < 2; ++ib) s_cur[jb][ib] = s_next[jb][ib];
        buf = nbuf;
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

// Waves per workgroup: 8 (two per SIMD, 256 registers each) while the Xi
// fragments are small; 4 (one per SIMD, 512 registers) for d > 128.
template <int DK>
constexpr int waves_for() { return DK >= 8 ? 4 : 8; }

template <int DK, int NCB>
void launch_bf16(dh_handle_t h, int m, int groups, int nchunks, int l, int ncb_total,
                 const bf16x8* Xi, const bf16x8* Xj, const float* ai, const float* bj, float c2,
                 const bf16x8* Op, const double* yi, double* Y, int ldy) {
    constexpr int W = waves_for<DK>();
    dim3 grid(dh_cdiv(m, W * 32), groups);
    rbf_project_bf16_kernel<DK, NCB, W><<<grid, W * 64, 0, h->stream>>>(
        m, nchunks, l, ncb_total, Xi, Xj, ai, bj, c2, Op, yi, Y, ldy);
}

int ncb_instance_bf16(int want) {
    static const int inst[] = {1, 2, 3, 5, 9, 17};
    for (int v : inst)
        if (v >= want) return v;
    return 17;
}

template <int DK>
int dispatch_bf16_ncb(dh_handle_t h, int ncb, int groups, int m, int nchunks, int l,
                      int ncb_total, const bf16x8* Xi, const bf16x8* Xj, const float* ai,
                      const float* bj, float c2, const bf16x8* Op, const double* yi, double* Y,
                      int ldy) {
#define DH_BF16_NCB(N)                                                                        \
    case N:                                                                                   \
        launch_bf16<DK, N>(h, m, groups, nchunks, l, ncb_total, Xi, Xj, ai, bj, c2, Op, yi, Y, ldy); \
        break;
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

int convert_points(dh_handle_t h, int n, int d, const double* X, int ldx, const double* mean,
                   float neg_gl2e, LowpPoints& p) {
    int dp = dpad_of(d);
    convert_rows_kernel<<<dh_cdiv(max(p.nrows_pad, 1), 4), 256, 0, h->stream>>>(
        n, p.nrows_pad, d, dp, X, ldx, mean, neg_gl2e, p.xb, p.a);
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

int pack_omega(dh_handle_t h, int n, int l, const double* O, int ldo, const double* y,
               bf16x8* Op) {
    int nchunks = dh_cdiv(n, 32);
    Bf16Cols cols = bf16_cols(l);
    long threads = (long)nchunks * cols.ncb_pack * 16;
    pack_omega_kernel<<<dh_cdiv(threads, 256), 256, 0, h->stream>>>(n, l, nchunks, cols.ncb_pack,
                                                                    O, ldo, y, Op);
    DH_CHECK_LAUNCH();
    return 0;
}

size_t packed_omega_bytes(int n, int l) {
    return (size_t)dh_cdiv(n, 32) * bf16_cols(l).ncb_pack * 64 * sizeof(bf16x8);
}

// Y (m x l, FP64) = K(Xi, Xj) * Op from prepared low-precision operands.
int project_bf16(dh_handle_t h, int m, int n, int d, int l, const LowpPoints& pi,
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
            return dispatch_bf16_ncb<1>(h, cols.ncb, grid, m, nchunks, l, cols.ncb_pack, xi, xj,
                                        pi.a, pj.a, c2, Op, yi, Y, ldy);
        case 64:
            return dispatch_bf16_ncb<2>(h, cols.ncb, grid, m, nchunks, l, cols.ncb_pack, xi, xj,
                                        pi.a, pj.a, c2, Op, yi, Y, ldy);
        case 128:
            return dispatch_bf16_ncb<4>(h, cols.ncb, grid, m, nchunks, l, cols.ncb_pack, xi, xj,
                                        pi.a, pj.a, c2, Op, yi, Y, ldy);
        default:
            return dispatch_bf16_ncb<8>(h, cols.ncb, grid, m, nchunks, l, cols.ncb_pack, xi, xj,
                                        pi.a, pj.a, c2, Op, yi, Y, ldy);
    }
}

// Carve the low-precision copies of one or two point sets out of DH_WS_LOWP_X
// and fill them (centred by the mean of the column set Xj).
int prepare_points(dh_handle_t h, int m, int n, int d, const double* Xi, int ldxi,
                   const double* Xj, int ldxj, double gamma, bool same, LowpPoints& pi,
                   LowpPoints& pj) {
    int dp = dpad_of(d);
    int mpad = dh_cdiv(m, 32) * 32, npad = dh_cdiv(n, 32) * 32;
    size_t bytes_j = (size_t)npad * dp * 2 + (size_t)npad * 4;
    size_t bytes_i = same ? 0 : (size_t)mpad * dp * 2 + (size_t)mpad * 4;
    char* base = (char*)dh_ws(h, DH_WS_LOWP_X, bytes_j + bytes_i + 512);
    if (!base) return DH_ERR_ALLOC;
    double* mean = (double*)dh_ws(h, DH_WS_NORMS, sizeof(double) * (size_t)d);
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
    float neg = (float)(-gamma) * kLog2e;
    DH_CHECK_STATUS(point_mean(h, n, d, Xj, ldxj, mean));
    DH_CHECK_STATUS(convert_points(h, n, d, Xj, ldxj, mean, neg, pj));
    if (!same) DH_CHECK_STATUS(convert_points(h, m, d, Xi, ldxi, mean, neg, pi));
    return 0;
}

}  // namespace

extern "C" int dh_rbf_project_bf16(dh_handle_t h, int m, int n, int d, int l, const double* Xi,
                                   int ldxi, const double* Xj, int ldxj, const double* yi,
                                   const double* yj, double gamma, const double* O, int ldo,
                                   double* Y, int ldy) {
    const char* name = "dh_rbf_project_bf16";
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
    DH_CHECK_STATUS(prepare_points(h, m, n, d, Xi, ldxi, Xj, ldxj, gamma, same, pi, pj));
    bf16x8* Op = (bf16x8*)dh_ws(h, DH_WS_LOWP_O, packed_omega_bytes(n, l));
    if (!Op) return DH_ERR_ALLOC;
    DH_CHECK_STATUS(pack_omega(h, n, l, O, ldo, yj, Op));
    return project_bf16(h, m, n, d, l, pi, pj, (float)(2.0 * gamma) * kLog2e, Op, yi, Y, ldy);
}

int dh_cholqr2_impl(dh_handle_t h, int m, int n, double* Y, int ldy, double* R, int ldr);

// Mixed-precision passes of dh_lra: BF16 K-passes, FP64 CholeskyQR2.
int dh_lra_bf16_passes(dh_handle_t h, int n, int d, int l, int q, const double* X, int ldx,
                       const double* y, double gamma, const double* Omega, int ldo, double* Q,
                       int ldq, double* W, int ldw) {
    if (d > 256) return dh_fail(DH_ERR_UNSUPPORTED, "dh_lra: BF16 passes need d <= 256");
    LowpPoints p, p2;
    DH_CHECK_STATUS(prepare_points(h, n, n, d, X, ldx, X, ldx, gamma, true, p, p2));
    bf16x8* Op = (bf16x8*)dh_ws(h, DH_WS_LOWP_O, packed_omega_bytes(n, l));
    if (!Op) return DH_ERR_ALLOC;
    const float c2 = (float)(2.0 * gamma) * kLog2e;
    DH_CHECK_STATUS(pack_omega(h, n, l, Omega, ldo, y, Op));
    DH_CHECK_STATUS(project_bf16(h, n, n, d, l, p, p, c2, Op, y, Q, ldq));
    DH_CHECK_STATUS(dh_cholqr2_impl(h, n, l, Q, ldq, nullptr, 0));
    for (int it = 0; it < q; ++it) {
        DH_CHECK_STATUS(pack_omega(h, n, l, Q, ldq, y, Op));
        DH_CHECK_STATUS(project_bf16(h, n, n, d, l, p, p, c2, Op, y, W, ldw));
        DH_CHECK_STATUS(dh_cholqr2_impl(h, n, l, W, ldw, nullptr, 0));
        DH_CHECK_HIP(hipMemcpy2DAsync(Q, sizeof(double) * (size_t)ldq, W,
                                      sizeof(double) * (size_t)ldw, sizeof(double) * (size_t)n, l,
                                      hipMemcpyDeviceToDevice, h->stream));
    }
    DH_CHECK_STATUS(pack_omega(h, n, l, Q, ldq, y, Op));
    return project_bf16(h, n, n, d, l, p, p, c2, Op, y, W, ldw);
}
