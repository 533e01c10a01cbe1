// RBF kernel on gfx950: squared norms, materialised tiles, the fused FP64
// K*Omega projection, and the FP64 error pass.
//
// Reference: rbf_kernel (dkernel.cu:132-190) = vecnorm (102-114) + distance
// GEMM (160-169) + elementwise rbf (116-129), with the K tile copied back to
// the host, replicated, uploaded again and multiplied by Omega in tc_gemm
// (dkernel.h:33-42, dkernel.cu:194-231).  Here the distance GEMM and the
// projection are two MFMA stages of one kernel and K stays in registers.
#include "dh_common.h"
#include "lds_dma.h"
#include "mfma_tile.h"

// ------------------------------------------------------------------ norms

template <typename T>
__global__ __launch_bounds__(256) void sqnorms_kernel(int n, int d, const T* __restrict__ X,
                                                       int ldx, T* __restrict__ nrm) {
    // One wave per point: lanes stride the (contiguous) features, then a
    // butterfly sum.  Summation order differs from the sequential vecnorm.
    int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wave >= n) return;
    const T* x = X + (size_t)wave * ldx;
    T s = 0;
    for (int f = lane; f < d; f += 64) s += x[f] * x[f];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) nrm[wave] = s;
}

template <typename T>
static int launch_sqnorms(dh_handle_t h, int n, int d, const T* X, int ldx, T* nrm) {
    if (n == 0) return 0;
    sqnorms_kernel<T><<<dh_cdiv(n, 4), 256, 0, h->stream>>>(n, d, X, ldx, nrm);
    DH_CHECK_LAUNCH();
    return 0;
}

extern "C" int dh_sqnorms_f64(dh_handle_t h, int n, int d, const double* X, int ldx,
                              double* nrm) {
    if (!h) return dh_fail(-1, "dh_sqnorms_f64: NULL handle");
    if (n < 0) return dh_fail(-2, "dh_sqnorms_f64: n < 0");
    if (d < 0) return dh_fail(-3, "dh_sqnorms_f64: d < 0");
    if (ldx < max(1, d)) return dh_fail(-5, "dh_sqnorms_f64: ldx < d");
    return launch_sqnorms<double>(h, n, d, X, ldx, nrm);
}

template <typename T>
__device__ __forceinline__ T rbf_exp(T x);
template <>
__device__ __forceinline__ double rbf_exp<double>(double x) {
    return exp(x);
}
template <>
__device__ __forceinline__ float rbf_exp<float>(float x) {
    return __expf(x);  // the reference's fast exp (dkernel.cu:125)
}

// ------------------------------------------------------- materialised tile
//
// Distance GEMM op(A) = Xi^T (point-major, contiguous features), op(B) = Xj,
// then the rbf epilogue on the accumulators: K = yi*yj*exp(-gamma*max(d2, 0)).
template <typename T, int BM, int BN>
__global__ __launch_bounds__(256) void rbf_tile_kernel(int m, int n, int d,
                                                        const T* __restrict__ Xi, int ldxi,
                                                        const T* __restrict__ Xj, int ldxj,
                                                        const T* __restrict__ ni,
                                                        const T* __restrict__ nj,
                                                        const T* __restrict__ yi,
                                                        const T* __restrict__ yj, T gamma,
                                                        T* __restrict__ K, int ldk) {
    constexpr int FM = BM / 32, FN = BN / 32;
    __shared__ T sA[kTileBK][BM + kTilePad];
    __shared__ T sB[kTileBK][BN + kTilePad];
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    typename Mfma<T>::acc_t acc[FM][FN];
    zero_acc<T, FM, FN>(acc);
    tile_mma<T, BM, BN, true, false>(acc, sA, sB, m, n, m0, n0, 0, d, Xi, ldxi, Xj, ldxj);
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int i = m0 + tile_row<T, BM, BN>(a, r);
                int j = n0 + tile_col<BM, BN>(b);
                if (i < m && j < n) {
                    T d2 = ni[i] + nj[j] - T(2) * acc[a][b][r];
                    d2 = d2 > T(0) ? d2 : T(0);
                    T v = rbf_exp<T>(-gamma * d2);
                    if (yi) v *= yi[i];
                    if (yj) v *= yj[j];
                    K[i + (size_t)j * ldk] = v;
                }
            }
}

template <typename T>
static int rbf_tile_api(const char* name, dh_handle_t h, int m, int n, int d, const T* Xi,
                        int ldxi, const T* Xj, int ldxj, const T* yi, const T* yj, T gamma, T* K,
                        int ldk) {
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (m < 0) return dh_fail(-2, "%s: m < 0", name);
    if (n < 0) return dh_fail(-3, "%s: n < 0", name);
    if (d < 0) return dh_fail(-4, "%s: d < 0", name);
    if (ldxi < max(1, d)) return dh_fail(-6, "%s: ldxi < d", name);
    if (ldxj < max(1, d)) return dh_fail(-8, "%s: ldxj < d", name);
    if (ldk < max(1, m)) return dh_fail(-13, "%s: ldk < m", name);
    if (m == 0 || n == 0) return 0;
    T* nrm = (T*)dh_ws(h, DH_WS_NORMS, sizeof(T) * ((size_t)m + n));
    if (!nrm) return DH_ERR_ALLOC;
    DH_CHECK_STATUS(launch_sqnorms<T>(h, m, d, Xi, ldxi, nrm));
    DH_CHECK_STATUS(launch_sqnorms<T>(h, n, d, Xj, ldxj, nrm + m));
    dim3 grid(dh_cdiv(m, 64), dh_cdiv(n, 64));
    rbf_tile_kernel<T, 64, 64><<<grid, 256, 0, h->stream>>>(m, n, d, Xi, ldxi, Xj, ldxj, nrm,
                                                             nrm + m, yi, yj, gamma, K, ldk);
    DH_CHECK_LAUNCH();
    return 0;
}

extern "C" int dh_rbf_tile_f64(dh_handle_t h, int m, int n, int d, const double* Xi, int ldxi,
                               const double* Xj, int ldxj, const double* yi, const double* yj,
                               double gamma, double* K, int ldk) {
    return rbf_tile_api<double>("dh_rbf_tile_f64", h, m, n, d, Xi, ldxi, Xj, ldxj, yi, yj, gamma,
                                K, ldk);
}
extern "C" int dh_rbf_tile_f32(dh_handle_t h, int m, int n, int d, const float* Xi, int ldxi,
                               const float* Xj, int ldxj, const float* yi, const float* yj,
                               float gamma, float* K, int ldk) {
    return rbf_tile_api<float>("dh_rbf_tile_f32", h, m, n, d, Xi, ldxi, Xj, ldxj, yi, yj, gamma, K,
                               ldk);
}

// ------------------------------------------------- fused FP64 K * Omega
//
// Workgroup: 4 waves, 64 rows i (16 per wave).  The j loop walks Xj and
// Omega in chunks of 16 points staged in LDS.  Per chunk each wave
//   1. forms S^T = Xj_chunk * Xi_wave^T (16 j x 16 i) with DK f64 MFMAs in two
//      independent accumulation chains.  Feature f = g*DK + kk is k-slot g of
//      MFMA kk, so lane (g, li) reads its A operands (point li of the chunk)
//      as consecutive doubles of a point-major LDS row, two per 16-byte read,
//      and keeps its B operands (row li of Xi, features g*DK ..) in registers;
//   2. turns its 4 accumulators into K values in place (exp epilogue);
//   3. uses them directly as the A operand of the projection MFMAs: lane L
//      holds S^T[L/16+4r][L%16] = K[i=L%16][j=4r+L/16] (f64 accumulator
//      layout), i.e. MFMA r contracts over j = 4r..4r+3; Omega is staged in
//      LDS so that lane L reads the B operands of MFMAs 2h and 2h+1 of block
//      cb as one 16-byte pair sO[cb][h][L ^ 4(L/16)] = (Omega[j0+4(2h)+L/16][c],
//      Omega[j0+4(2h+1)+L/16][c]), c = c0 + 16cb + L%16.  The XOR keeps each
//      aligned group of 8 lanes on 8 consecutive pairs (conflict-free reads)
//      and spreads the staging stores, whose 64 lanes cover 4 columns x 16
//      points, over all banks (2 cycles per store instead of 8).
// The chunks go through two LDS buffers: the global loads of chunk t+1 are
// issued into registers before chunk t is computed and stored to the other
// buffer after it, so one barrier per chunk orders both.  (With one buffer,
// single 8-byte operand reads consumed at once and the accumulators moved
// between register files every chunk, the pass ran at 38% of the FP64 MFMA
// peak; this layout runs at 51%.)
// Labels are folded: yj scales Omega as it is staged, yi scales the output.
// The projection accumulators (NCB blocks of 16 columns) stay in registers
// for the whole j loop; Y is written once.
// exp(x) without branches for the kernel's exponents x = -gamma d2 <= 0
// (valid up to x = 709): k = rint(x log2 e); r = (x - k LN2_HI) - k LN2_LO, a
// Cody-Waite reduction with LN2_HI = ln 2 rounded to 32 significant bits (k
// LN2_HI exact for |k| < 2^21) and LN2_LO = ln 2 - LN2_HI rounded to double;
// e^r by its Taylor series to degree 13 in Horner form (truncation below
// 2^-55 for |r| <= ln2 / 2); then 2^k by ldexp, which rounds subnormal
// results and gives 0 below 2^-1075.  Within 1 ulp of libm's exp over
// [-746, 0] (an exact-FMA emulation checked on 20000 points and the edges).
__device__ __forceinline__ double exp_kernel_f64(double x) {
    x = fmax(x, -746.0);
    const double k = rint(x * 1.4426950408889634);
    double r = fma(-k, 0x1.62e42ff000000p-1, x);
    r = fma(-k, -0x1.718432a1b0e26p-35, r);
    double p = 1.6059043836821613e-10;  // 1/13!
    p = fma(p, r, 2.08767569878681e-09);
    p = fma(p, r, 2.505210838544172e-08);
    p = fma(p, r, 2.755731922398589e-07);
    p = fma(p, r, 2.7557319223985893e-06);
    p = fma(p, r, 2.48015873015873e-05);
    p = fma(p, r, 1.984126984126984e-04);
    p = fma(p, r, 1.388888888888889e-03);
    p = fma(p, r, 8.333333333333333e-03);
    p = fma(p, r, 4.1666666666666664e-02);
    p = fma(p, r, 1.6666666666666666e-01);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}

typedef double d2v __attribute__((ext_vector_type(2)));

// Omega (or Q) with the labels folded in, in the kernel's LDS image: for
// 16-point chunk t, column block cb (of ncbp) and h, the 64 pairs
// Opk[((t ncbp + cb) 2 + h) 64 + (L ^ 4(L/16))] = (y_j O[j][c], y_{j+4} O[j+4][c]),
// j = 16t + 8h + L/16, c = 16cb + L%16, zero outside n x l; one thread per pair.
__global__ __launch_bounds__(256) void pack_omega_f64_kernel(int n, int l, int nch, int ncbp,
                                                             const double* __restrict__ O,
                                                             int ldo,
                                                             const double* __restrict__ yj,
                                                             d2v* __restrict__ Opk) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (long)nch * ncbp * 128) return;
    const int s = (int)(idx & 63), h = (int)((idx >> 6) & 1);
    const long tc = idx >> 7;
    const int cb = (int)(tc % ncbp), t = (int)(tc / ncbp);
    const int L = s ^ ((s >> 4) << 2);
    const int j = 16 * t + 8 * h + (L >> 4), c = 16 * cb + (L & 15);
    d2v v = {0.0, 0.0};
    if (c < l) {
        if (j < n) v[0] = O[j + (size_t)c * ldo] * (yj ? yj[j] : 1.0);
        if (j + 4 < n) v[1] = O[j + 4 + (size_t)c * ldo] * (yj ? yj[j + 4] : 1.0);
    }
    Opk[idx] = v;
}

template <int DK, int NCB, bool PK>
__global__ __launch_bounds__(256) void rbf_project_f64_kernel(
    int m, int n, int d, int l, const double* __restrict__ Xi, int ldxi,
    const double* __restrict__ Xj, int ldxj, const double* __restrict__ ni,
    const double* __restrict__ nj, const double* __restrict__ yi, const double* __restrict__ yj,
    double gamma, const double* __restrict__ O, int ldo, const d2v* __restrict__ Opk, int ncbp,
    double beta, double* __restrict__ Y, int ldy, int jps, long ystride) {
    constexpr int JB = 16;
    {  // split z of the j range: points [z jps, min((z + 1) jps, n)), partial into Y + z ystride
        const int j0 = blockIdx.z * jps;
        n = min(jps, n - j0);
        Xj += (size_t)j0 * ldxj;
        nj += j0;
        if (yj) yj += j0;
        O += j0;
        if (PK) Opk += (size_t)(j0 / JB) * ncbp * 128;
        Y += blockIdx.z * ystride;
    }
    constexpr int FD = DK * 4;  // padded feature count
    constexpr int XP = FD + 2;  // LDS pitch of a point (odd multiple of 16 bytes: no conflicts)
    __shared__ double sX[2][JB * XP];
    __shared__ d2v sO[2][NCB][2][64];
    __shared__ double sN[2][JB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, li = lane & 15;
    const int c0 = blockIdx.y * NCB * 16;
    const int rbase = blockIdx.x * 64 + wave * 16;
    const int row = rbase + li;

    double xr[DK];
#pragma unroll
    for (int kk = 0; kk < DK; ++kk) {
        const int f = g * DK + kk;
        xr[kk] = (row < m && f < d) ? Xi[f + (size_t)row * ldxi] : 0.0;
    }
    const double nrow = (row < m) ? ni[row] : 0.0;
    dh_d4 acc[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) acc[cb] = dh_d4{0, 0, 0, 0};

    // Staging of a chunk: every load is issued before any LDS store, from
    // addresses clamped into the operands (a load under a branch gets a wait
    // of its own).  Thread tid always stages point jj = tid % JB of the Omega
    // block, so its label is one load per chunk.
    constexpr int NXL = JB * FD / 256, NOL = JB * NCB * 16 / 256;
    static_assert(JB * FD % 256 == 0 && JB * NCB * 16 % 256 == 0, "staging split");
    double px[NXL], po[PK ? 1 : NOL], pn, yv;
    // PK: the chunk's packed Omega image (NCB x 2 KB, contiguous) goes to LDS
    // by direct copies, 1 KB per wave instruction.
    auto dma_omega = [&](int j0, int b) {
        const d2v* src = Opk + ((size_t)(j0 / JB) * ncbp + blockIdx.y * NCB) * 128;
#pragma unroll
        for (int k = 0; k < 2 * NCB; k += 4) {
            const int kk = k + wave;
            if ((2 * NCB) % 4 == 0 || kk < 2 * NCB)
                dma16(src + kk * 64, (unsigned)lane * 16u, &sO[b][kk >> 1][kk & 1][0]);
        }
    };
    auto load = [&](int j0) {
#pragma unroll
        for (int u = 0; u < NXL; ++u) {
            const int e = tid + 256 * u, jj = e / FD, f = e % FD;
            px[u] = Xj[min(f, d - 1) + (size_t)min(j0 + jj, n - 1) * ldxj];
        }
        const int jo = j0 + (tid % JB);
        if constexpr (!PK) {
            yv = yj ? yj[min(jo, n - 1)] : 1.0;
#pragma unroll
            for (int u = 0; u < NOL; ++u) {
                const int e = tid + 256 * u, cl = e / JB;
                po[u] = O[min(jo, n - 1) + (size_t)min(c0 + cl, l - 1) * ldo];
            }
        }
        pn = nj[min(j0 + (tid % JB), n - 1)];
    };
    auto store = [&](int j0, int b) {
        const int jo = j0 + (tid % JB);
#pragma unroll
        for (int u = 0; u < NXL; ++u) {
            const int e = tid + 256 * u, jj = e / FD, f = e % FD;
            sX[b][jj * XP + f] = (j0 + jj < n && f < d) ? px[u] : 0.0;
        }
        double* so = (double*)&sO[b][0][0][0];
#pragma unroll
        for (int u = 0; u < (PK ? 0 : NOL); ++u) {
            const int e = tid + 256 * u, jj = e % JB, cl = e / JB;
            const int r = jj >> 2, ln = (jj & 3) * 16 + ((cl & 15) ^ ((jj & 3) << 2));
            so[(((cl >> 4) * 2 + (r >> 1)) * 64 + ln) * 2 + (r & 1)] =
                (jo < n && c0 + cl < l) ? po[u] * yv : 0.0;
        }
        if (tid < JB) sN[b][tid] = (j0 + tid < n) ? pn : 0.0;
    };

    load(0);
    if constexpr (PK) dma_omega(0, 0);
    store(0, 0);
    if constexpr (PK) wait_vm();
    __syncthreads();
    int b = 0;
    for (int j0 = 0; j0 < n; j0 += JB) {
        const bool more = j0 + JB < n;  // uniform
        if (more) {
            load(j0 + JB);
            if constexpr (PK) dma_omega(j0 + JB, b ^ 1);
        }
        dh_d4 s0 = dh_d4{0, 0, 0, 0}, s1 = dh_d4{0, 0, 0, 0};
        const double* xrow = &sX[b][li * XP + g * DK];
#pragma unroll
        for (int kk = 0; kk < DK; kk += 2) {
            const d2v a2 = *(const d2v*)(xrow + kk);
            s0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[0], xr[kk], s0, 0, 0, 0);
            s1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[1], xr[kk + 1], s1, 0, 0, 0);
        }
        double kv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double d2 = nrow + sN[b][g + 4 * r] - 2.0 * (s0[r] + s1[r]);
            d2 = d2 > 0.0 ? d2 : 0.0;
            kv[r] = exp_kernel_f64(-gamma * d2);
        }
        const int sl = lane ^ ((lane >> 4) << 2);
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            const d2v o01 = sO[b][cb][0][sl], o23 = sO[b][cb][1][sl];
            acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(kv[0], o01[0], acc[cb], 0, 0, 0);
            acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(kv[1], o01[1], acc[cb], 0, 0, 0);
            acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(kv[2], o23[0], acc[cb], 0, 0, 0);
            acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(kv[3], o23[1], acc[cb], 0, 0, 0);
        }
        if (more) store(j0 + JB, b ^ 1);
        if constexpr (PK) wait_vm();
        __syncthreads();
        b ^= 1;
    }
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int i = rbase + g + 4 * r, c = c0 + 16 * cb + li;
            if (i < m && c < l) {
                double v = acc[cb][r];
                if (yi) v *= yi[i];
                double* yp = Y + i + (size_t)c * ldy;
                *yp = (beta == 0.0) ? v : v + beta * (*yp);
            }
        }
}

template <int DK, int NCB>
static void launch_project_f64(dh_handle_t h, dim3 grid, int m, int n, int d, int l,
                               const double* Xi, int ldxi, const double* Xj, int ldxj,
                               const double* ni, const double* nj, const double* yi,
                               const double* yj, double gamma, const double* O, int ldo,
                               const d2v* Opk, int ncbp, double beta, double* Y, int ldy, int jps,
                               long ystride) {
    if constexpr (DK < 64 || NCB <= 12) {  // d > 128 with 17 blocks is never selected (spills)
        dh_timing_begin(h);
        if (Opk)
            rbf_project_f64_kernel<DK, NCB, true><<<grid, 256, 0, h->stream>>>(
                m, n, d, l, Xi, ldxi, Xj, ldxj, ni, nj, yi, yj, gamma, O, ldo, Opk, ncbp, beta, Y,
                ldy, jps, ystride);
        else
            rbf_project_f64_kernel<DK, NCB, false><<<grid, 256, 0, h->stream>>>(
                m, n, d, l, Xi, ldxi, Xj, ldxj, ni, nj, yi, yj, gamma, O, ldo, nullptr, 0, beta, Y,
                ldy, jps, ystride);
        dh_timing_end(h);
    }
}

template <int DK>
static int dispatch_ncb(dh_handle_t h, int ncb, dim3 grid, int m, int n, int d, int l,
                        const double* Xi, int ldxi, const double* Xj, int ldxj, const double* ni,
                        const double* nj, const double* yi, const double* yj, double gamma,
                        const double* O, int ldo, const d2v* Opk, int ncbp, double beta,
                        double* Y, int ldy, int jps, long ystride) {
#define DH_NCB(N)                                                                              \
    case N:                                                                                    \
        launch_project_f64<DK, N>(h, grid, m, n, d, l, Xi, ldxi, Xj, ldxj, ni, nj, yi, yj, gamma, \
                                  O, ldo, Opk, ncbp, beta, Y, ldy, jps, ystride);              \
        break;
    switch (ncb) {
        DH_NCB(1)
        DH_NCB(2)
        DH_NCB(3)
        DH_NCB(4)
        DH_NCB(5)
        DH_NCB(6)
        DH_NCB(8)
        DH_NCB(9)
        DH_NCB(12)
        DH_NCB(17)
        default:
            return dh_fail(DH_ERR_UNSUPPORTED, "rbf_project_f64: no kernel for %d column blocks",
                           ncb);
    }
#undef DH_NCB
    DH_CHECK_LAUNCH();
    return 0;
}

// Above 256 features a wave's rows of Xi no longer fit in registers.  K is
// then materialised in row blocks of at most 2^27 entries (1 GiB), each by the
// distance-GEMM tile kernel (FP64 MFMA, exp epilogue), and multiplied into
// Omega by the FP64 GEMM: Y_blk = K_blk Omega + beta Y_blk.  The same split as
// rbf_kernel's explicit tiles (dkernel.cu:144-178), with the product fused in.
static int project_f64_blocked(dh_handle_t h, int m, int n, int d, int l, const double* Xi,
                               int ldxi, const double* Xj, int ldxj, const double* ni,
                               const double* nj, const double* yi, const double* yj, double gamma,
                               const double* O, int ldo, double beta, double* Y, int ldy) {
    if (n == 0)  // K has no columns: Y = beta Y
        return dh_gemm_impl<double>(h, false, false, m, l, 0, 1.0, O, 1, O, ldo, beta, Y, ldy, 0);
    const long cap = 1L << 27;
    int mb = (int)std::min<long>(m, std::max<long>(64, cap / n / 64 * 64));
    double* Kb = (double*)dh_ws(h, DH_WS_KBLOCK, sizeof(double) * (size_t)mb * n);
    if (!Kb) return DH_ERR_ALLOC;
    dh_timing_begin(h);
    for (int i0 = 0; i0 < m; i0 += mb) {
        const int ib = std::min(mb, m - i0);
        dim3 grid(dh_cdiv(ib, 64), dh_cdiv(n, 64));
        rbf_tile_kernel<double, 64, 64><<<grid, 256, 0, h->stream>>>(
            ib, n, d, Xi + (size_t)i0 * ldxi, ldxi, Xj, ldxj, ni + i0, nj, yi ? yi + i0 : nullptr,
            yj, gamma, Kb, ib);
        DH_CHECK_LAUNCH();
        DH_CHECK_STATUS(dh_gemm_impl<double>(h, false, false, ib, l, n, 1.0, Kb, ib, O, ldo, beta,
                                             Y + i0, ldy, 0));
    }
    dh_timing_end(h);
    return 0;
}

// Smallest instantiated column-block count >= want.
static int ncb_instance(int want) {
    static const int inst[] = {1, 2, 3, 4, 5, 6, 8, 9, 12, 17};
    for (int v : inst)
        if (v >= want) return v;
    return 17;
}

// Y (m x l) = K(Xi, Xj) O + beta Y with the fused kernel, the j range in
// `splits` parts (partials to P + z ystride, ld m, when splits > 1).
// With `pack`, Omega is first packed (labels folded in) into the kernel's LDS
// image in a workspace and staged by direct copies; without it (the
// allocation-free entry point) the kernel stages it through registers.  Both
// put the same values in LDS, so both give the same bits.
static int project_f64_fused(dh_handle_t h, int m, int n, int d, int l, const double* Xi,
                             int ldxi, const double* Xj, int ldxj, const double* ni,
                             const double* nj, const double* yi, const double* yj, double gamma,
                             const double* O, int ldo, double beta, double* Y, int ldy, int ncb,
                             int groups, int jps, int splits, long ystride, bool pack) {
    dim3 grid(dh_cdiv(m, 64), groups, splits);
    const d2v* Opk = nullptr;
    const int ncbp = groups * ncb, nch = dh_cdiv(n, 16);
    if (pack) {
        const long pairs = (long)nch * ncbp * 128;
        d2v* P = (d2v*)dh_ws(h, DH_WS_PACK64, sizeof(d2v) * (size_t)pairs);
        if (!P) return DH_ERR_ALLOC;
        pack_omega_f64_kernel<<<dh_cdiv(pairs, 256), 256, 0, h->stream>>>(n, l, nch, ncbp, O, ldo,
                                                                         yj, P);
        DH_CHECK_LAUNCH();
        Opk = P;
    }
#define DH_DK(K)                                                                              \
    dispatch_ncb<K>(h, ncb, grid, m, n, d, l, Xi, ldxi, Xj, ldxj, ni, nj, yi, yj, gamma, O, ldo, \
                    Opk, ncbp, beta, Y, ldy, jps, ystride)
    if (d <= 32) return DH_DK(8);
    if (d <= 64) return DH_DK(16);
    if (d <= 128) return DH_DK(32);
    return DH_DK(64);
#undef DH_DK
}

// Column blocking of the fused FP64 pass: ncb blocks per group.
static void f64_columns(int d, int l, int* ncb, int* groups) {
    const int ncb_total = dh_cdiv(l, 16);
    // Above 128 features the 17-block tile spills its accumulators: cap at 12.
    int g = dh_cdiv(ncb_total, d > 128 ? 12 : 17);
    *ncb = ncb_instance(dh_cdiv(ncb_total, g));
    *groups = dh_cdiv(ncb_total, *ncb);
}

// Whole j range in one launch; `pack` as in project_f64_fused (without it, no
// workspace for d <= 256).
int dh_project_f64_impl(dh_handle_t h, int m, int n, int d, int l, const double* Xi, int ldxi,
                        const double* Xj, int ldxj, const double* ni, const double* nj,
                        const double* yi, const double* yj, double gamma, const double* O,
                        int ldo, double beta, double* Y, int ldy, bool pack) {
    if (m == 0 || l == 0) return 0;
    if (n == 0)  // K has no columns: Y = beta Y
        return dh_gemm_impl<double>(h, false, false, m, l, 0, 1.0, O, 1, O, ldo, beta, Y, ldy, 0);
    if (d > 256)
        return project_f64_blocked(h, m, n, d, l, Xi, ldxi, Xj, ldxj, ni, nj, yi, yj, gamma, O,
                                   ldo, beta, Y, ldy);
    int ncb, groups;
    f64_columns(d, l, &ncb, &groups);
    return project_f64_fused(h, m, n, d, l, Xi, ldxi, Xj, ldxj, ni, nj, yi, yj, gamma, O, ldo,
                             beta, Y, ldy, ncb, groups, n, 1, 0, pack);
}

// Y = beta Y + sum over the splits z (in order) of P[z ystride + i + c m].
__global__ __launch_bounds__(256) void sum_splits_f64_kernel(int m, int l, int splits,
                                                             const double* __restrict__ P,
                                                             long ystride, double beta,
                                                             double* __restrict__ Y, int ldy) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)m * l) return;
    const int i = (int)(e % m), c = (int)(e / m);
    double v = 0.0;
    for (int z = 0; z < splits; ++z) v += P[z * ystride + e];
    double* yp = Y + i + (size_t)c * ldy;
    *yp = (beta == 0.0) ? v : v + beta * (*yp);
}

// As dh_project_f64_impl; a grid of at most half a workgroup per CU splits
// the j range over more workgroups (partials in a workspace, summed in order).
// The range finder's passes use it; dh_rbf_project_f64 keeps the one-launch
// form, so its allocation-free twin returns the same bits.
int dh_project_f64_split(dh_handle_t h, int m, int n, int d, int l, const double* Xi, int ldxi,
                         const double* Xj, int ldxj, const double* ni, const double* nj,
                         const double* yi, const double* yj, double gamma, const double* O,
                         int ldo, double beta, double* Y, int ldy) {
    if (m == 0 || l == 0 || n == 0 || d > 256)
        return dh_project_f64_impl(h, m, n, d, l, Xi, ldxi, Xj, ldxj, ni, nj, yi, yj, gamma, O,
                                   ldo, beta, Y, ldy, true);
    int ncb, groups;
    f64_columns(d, l, &ncb, &groups);
    const int nch = dh_cdiv(n, 32);
    const int splits = dh_j_splits(h->num_cu, (long)dh_cdiv(m, 64) * groups, nch);
    if (splits == 1)
        return project_f64_fused(h, m, n, d, l, Xi, ldxi, Xj, ldxj, ni, nj, yi, yj, gamma, O,
                                 ldo, beta, Y, ldy, ncb, groups, n, 1, 0, true);
    const int jps = dh_cdiv(nch, splits) * 32;
    const long ystride = (long)m * l;
    double* P = (double*)dh_ws(h, DH_WS_SPLITN, sizeof(double) * (size_t)ystride * splits);
    if (!P) return DH_ERR_ALLOC;
    DH_CHECK_STATUS(project_f64_fused(h, m, n, d, l, Xi, ldxi, Xj, ldxj, ni, nj, yi, yj, gamma, O,
                                      ldo, 0.0, P, m, ncb, groups, jps, splits, ystride, true));
    sum_splits_f64_kernel<<<dh_cdiv(ystride, 256), 256, 0, h->stream>>>(m, l, splits, P, ystride,
                                                                        beta, Y, ldy);
    DH_CHECK_LAUNCH();
    return 0;
}

extern "C" int dh_rbf_project_f64(dh_handle_t h, int m, int n, int d, int l, const double* Xi,
                                  int ldxi, const double* Xj, int ldxj, const double* yi,
                                  const double* yj, double gamma, const double* O, int ldo,
                                  double beta, double* Y, int ldy) {
    const char* name = "dh_rbf_project_f64";
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (m < 0) return dh_fail(-2, "%s: m < 0", name);
    if (n < 0) return dh_fail(-3, "%s: n < 0", name);
    if (d < 0) return dh_fail(-4, "%s: d < 0", name);
    if (l < 0) return dh_fail(-5, "%s: l < 0", name);
    if (ldxi < max(1, d)) return dh_fail(-7, "%s: ldxi < d", name);
    if (ldxj < max(1, d)) return dh_fail(-9, "%s: ldxj < d", name);
    if (ldo < max(1, n)) return dh_fail(-14, "%s: ldo < n", name);
    if (ldy < max(1, m)) return dh_fail(-17, "%s: ldy < m", name);
    if (m == 0 || l == 0) return 0;
    double* nrm = (double*)dh_ws(h, DH_WS_NORMS, sizeof(double) * ((size_t)m + n + 1));
    if (!nrm) return DH_ERR_ALLOC;
    DH_CHECK_STATUS(launch_sqnorms<double>(h, m, d, Xi, ldxi, nrm));
    DH_CHECK_STATUS(launch_sqnorms<double>(h, n, d, Xj, ldxj, nrm + m));
    return dh_project_f64_impl(h, m, n, d, l, Xi, ldxi, Xj, ldxj, nrm, nrm + m, yi, yj, gamma, O,
                               ldo, beta, Y, ldy, true);
}

// ------------------------------------------------------------ error pass
//
// Per 64 x 64 tile: S = Xi^T Xj (k over d) and P = Q_i W_j^T (k over l) in two
// MFMA accumulator sets, then sum (K - P)^2 and K^2 in FP64.  One pair of
// partial sums per workgroup, reduced in a fixed order by a second kernel.
__global__ __launch_bounds__(256) void lra_error_f64_kernel(
    int n, int d, int l, const double* __restrict__ X, int ldx, const double* __restrict__ nrm,
    const double* __restrict__ y, double gamma, const double* __restrict__ Q, int ldq,
    const double* __restrict__ W, int ldw, double* __restrict__ partial) {
    constexpr int BM = 64, BN = 64, FM = 2, FN = 2;
    __shared__ double sA[kTileBK][BM + kTilePad];
    __shared__ double sB[kTileBK][BN + kTilePad];
    __shared__ double red[2][4];
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    dh_d4 S[FM][FN], P[FM][FN];
    zero_acc<double, FM, FN>(S);
    zero_acc<double, FM, FN>(P);
    tile_mma<double, BM, BN, true, false>(S, sA, sB, n, n, m0, n0, 0, d, X, ldx, X, ldx);
    tile_mma<double, BM, BN, false, true>(P, sA, sB, n, n, m0, n0, 0, l, Q, ldq, W, ldw);
    double se = 0.0, sk = 0.0;
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int i = m0 + tile_row<double, BM, BN>(a, r);
                int j = n0 + tile_col<BM, BN>(b);
                if (i < n && j < n) {
                    double d2 = nrm[i] + nrm[j] - 2.0 * S[a][b][r];
                    d2 = d2 > 0.0 ? d2 : 0.0;
                    double kv = exp(-gamma * d2);
                    if (y) kv *= y[i] * y[j];
                    double e = kv - P[a][b][r];
                    se += e * e;
                    sk += kv * kv;
                }
            }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        se += __shfl_xor(se, o);
        sk += __shfl_xor(sk, o);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wave] = se;
        red[1][wave] = sk;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        partial[2 * b] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        partial[2 * b + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

// Fixed-order reduction of `count` (se, sk) pairs into out[0..1].
__global__ __launch_bounds__(256) void reduce_pairs_kernel(size_t count,
                                                            const double* __restrict__ partial,
                                                            double* __restrict__ out) {
    __shared__ double red[2][256];
    double se = 0.0, sk = 0.0;
    for (size_t b = threadIdx.x; b < count; b += 256) {
        se += partial[2 * b];
        sk += partial[2 * b + 1];
    }
    red[0][threadIdx.x] = se;
    red[1][threadIdx.x] = sk;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0][0];
        out[1] = red[1][0];
    }
}

extern "C" int dh_lra_error(dh_handle_t h, int n, int d, int l, const double* X, int ldx,
                            const double* y, double gamma, const double* Q, int ldq,
                            const double* W, int ldw, double* err2) {
    const char* name = "dh_lra_error";
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (n < 0) return dh_fail(-2, "%s: n < 0", name);
    if (d < 0) return dh_fail(-3, "%s: d < 0", name);
    if (l < 0) return dh_fail(-4, "%s: l < 0", name);
    if (ldx < max(1, d)) return dh_fail(-6, "%s: ldx < d", name);
    if (ldq < max(1, n)) return dh_fail(-10, "%s: ldq < n", name);
    if (ldw < max(1, n)) return dh_fail(-12, "%s: ldw < n", name);
    if (!err2) return dh_fail(-13, "%s: err2 is NULL", name);
    err2[0] = err2[1] = 0.0;
    if (n == 0) return 0;
    double* nrm = (double*)dh_ws(h, DH_WS_NORMS, sizeof(double) * (size_t)n);
    if (!nrm) return DH_ERR_ALLOC;
    DH_CHECK_STATUS(launch_sqnorms<double>(h, n, d, X, ldx, nrm));
    dim3 grid(dh_cdiv(n, 64), dh_cdiv(n, 64));
    size_t nblk = (size_t)grid.x * grid.y;
    double* part = (double*)dh_ws(h, DH_WS_PARTIAL, sizeof(double) * (2 * nblk + 2));
    if (!part) return DH_ERR_ALLOC;
    lra_error_f64_kernel<<<grid, 256, 0, h->stream>>>(n, d, l, X, ldx, nrm, y, gamma, Q, ldq, W,
                                                      ldw, part);
    DH_CHECK_LAUNCH();
    reduce_pairs_kernel<<<1, 256, 0, h->stream>>>(nblk, part, part + 2 * nblk);
    DH_CHECK_LAUNCH();
    DH_CHECK_HIP(hipMemcpyAsync(err2, part + 2 * nblk, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                h->stream));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    return 0;
}
