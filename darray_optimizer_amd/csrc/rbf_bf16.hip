// Mixed-precision K-passes, host side and preparation kernels: column means,
// 16-bit centred points with their squared norms, the packed projection
// operand, the one-call entry points dh_rbf_project_bf16 / _f16, the 16-bit
// passes of dh_lra, and the batched preparation launches of dh_lra_batched.
// The fused projection kernels themselves are in rbf16_kernels.h.
// Reference: DArray::LRA (dkernel.h:12-46) over rbf_kernel + tc_gemm
// (dkernel.cu:132-231).
#include <utility>

#include "rbf16_kernels.h"

namespace {

// ----------------------------------------------------------------- prep

// partial[b][f] = sum of X[f, i] over the points of slice b (FP64).
__global__ __launch_bounds__(256) void colsum_partial_kernel(int n, int d, const double* __restrict__ X,
                                                             int ldx, int slice,
                                                             double* __restrict__ partial) {
    int b = blockIdx.x;
    int i1 = min(n, (b + 1) * slice);
    if (d <= 128) {
        // 256 / d row groups (thread t: feature t % d, rows t / d + k * rpi),
        // four sums each, combined through LDS in a fixed order.
        __shared__ double red[256];
        const int rpi = 256 / d, r = threadIdx.x / d, f = threadIdx.x % d;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        if (r < rpi) {
            int i = b * slice + r;
            for (; i + 3 * rpi < i1; i += 4 * rpi) {
                s0 += X[f + (size_t)i * ldx];
                s1 += X[f + (size_t)(i + rpi) * ldx];
                s2 += X[f + (size_t)(i + 2 * rpi) * ldx];
                s3 += X[f + (size_t)(i + 3 * rpi) * ldx];
            }
            for (; i < i1; i += rpi) s0 += X[f + (size_t)i * ldx];
        }
        red[threadIdx.x] = (s0 + s1) + (s2 + s3);
        __syncthreads();
        if (threadIdx.x < d) {
            double t = 0.0;
            for (int gq = 0; gq < rpi; ++gq) t += red[gq * d + threadIdx.x];
            partial[(size_t)b * d + threadIdx.x] = t;
        }
        return;
    }
    for (int f = threadIdx.x; f < d; f += 256) {
        // Four independent sums so that the loads of consecutive points
        // overlap (fixed order: deterministic).
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int i = b * slice;
        for (; i + 4 <= i1; i += 4) {
            s0 += X[f + (size_t)i * ldx];
            s1 += X[f + (size_t)(i + 1) * ldx];
            s2 += X[f + (size_t)(i + 2) * ldx];
            s3 += X[f + (size_t)(i + 3) * ldx];
        }
        for (; i < i1; ++i) s0 += X[f + (size_t)i * ldx];
        partial[(size_t)b * d + f] = (s0 + s1) + (s2 + s3);
    }
}

__global__ __launch_bounds__(1024) void colmean_final_kernel(int n, int d, int nb,
                                                             const double* __restrict__ partial,
                                                             double* __restrict__ mean) {
    if (gridDim.x == 1 && d <= (int)blockDim.x) {
        // One workgroup: blockDim / d slab groups per feature, four sums each,
        // combined through LDS in a fixed order.
        __shared__ double red[1024];
        const int rpi = blockDim.x / d, r = threadIdx.x / d, f = threadIdx.x % d;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        if (r < rpi) {
            int b = r;
            for (; b + 3 * rpi < nb; b += 4 * rpi) {
                s0 += partial[(size_t)b * d + f];
                s1 += partial[(size_t)(b + rpi) * d + f];
                s2 += partial[(size_t)(b + 2 * rpi) * d + f];
                s3 += partial[(size_t)(b + 3 * rpi) * d + f];
            }
            for (; b < nb; b += rpi) s0 += partial[(size_t)b * d + f];
        }
        red[threadIdx.x] = (s0 + s1) + (s2 + s3);
        __syncthreads();
        if ((int)threadIdx.x < d) {
            double t = 0.0;
            for (int gq = 0; gq < rpi; ++gq) t += red[gq * d + threadIdx.x];
            mean[threadIdx.x] = n > 0 ? t / n : 0.0;
        }
        return;
    }
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < d; f += gridDim.x * blockDim.x) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int b = 0;
        for (; b + 4 <= nb; b += 4) {
            s0 += partial[(size_t)b * d + f];
            s1 += partial[(size_t)(b + 1) * d + f];
            s2 += partial[(size_t)(b + 2) * d + f];
            s3 += partial[(size_t)(b + 3) * d + f];
        }
        for (; b < nb; ++b) s0 += partial[(size_t)b * d + f];
        mean[f] = n > 0 ? ((s0 + s1) + (s2 + s3)) / n : 0.0;
    }
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
    if (d <= 256) {
        // Thread t keeps feature t % d of row offset t / d (256 / d rows per
        // block step; no per-element index division), four loads in flight.
        const int rpi = 256 / d, r = threadIdx.x / d, f = threadIdx.x % d;
        if (r < rpi) {
            const double mf = mean[f];
            const int step = rpi * gridDim.x;
            int i = blockIdx.x * rpi + r;
            for (; i + 3 * step < n; i += 4 * step) {
                const double v0 = X[f + (size_t)i * ldx], v1 = X[f + (size_t)(i + step) * ldx];
                const double v2 = X[f + (size_t)(i + 2 * step) * ldx];
                const double v3 = X[f + (size_t)(i + 3 * step) * ldx];
                mx = fmaxf(mx, fmaxf(fmaxf((float)fabs(v0 - mf), (float)fabs(v1 - mf)),
                                     fmaxf((float)fabs(v2 - mf), (float)fabs(v3 - mf))));
            }
            for (; i < n; i += step) mx = fmaxf(mx, (float)fabs(X[f + (size_t)i * ldx] - mf));
        }
    } else {
        for (int i = blockIdx.x; i < n; i += gridDim.x)
            for (int f = threadIdx.x; f < d; f += 256)
                mx = fmaxf(mx, (float)fabs(X[f + (size_t)i * ldx] - mean[f]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    // One atomic per workgroup: same-address atomics serialise in L2.
    __shared__ float wmax[4];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(maxabs, __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]))));
}

__global__ void zero_word_kernel(unsigned* w) { *w = 0u; }

// One wave per point row, features 4 lane .. 4 lane + 3 (dpad8 <= 256):
// q = e4m3((X - mean) * scale_i) with scale_i = 2^(8 - e) for the row's
// max |x - mean| = f 2^e (f in [1/2, 1)), so its largest magnitude lies in
// [128, 256) (e4m3 reaches 448); a[i] = neg_gl2e |q / scale_i|^2 from the
// dequantised values in FP32; sinv[i] = 1 / scale_i.  Rows >= n are zero.
__global__ __launch_bounds__(256) void convert_rows_fp8_kernel(int n, int nrows, int d, int dpad8,
                                                               const double* __restrict__ X,
                                                               int ldx,
                                                               const double* __restrict__ mean,
                                                               float neg_gl2e,
                                                               unsigned* __restrict__ Xb8,
                                                               float* __restrict__ a,
                                                               float* __restrict__ sinv) {
    const int row = (blockIdx.x * 256 + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (row >= nrows) return;
    float v[4], mx = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int f = 4 * lane + k;
        v[k] = (row < n && f < d) ? (float)(X[f + (size_t)row * ldx] - mean[f]) : 0.f;
        mx = fmaxf(mx, fabsf(v[k]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    float scale = 1.f, inv = 0.f;
    if (mx > 0.f) {
        int e;
        frexpf(mx, &e);
        scale = ldexpf(1.f, 8 - e);
        inv = ldexpf(1.f, e - 8);
    }
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(v[0] * scale, v[1] * scale, 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(v[2] * scale, v[3] * scale, w, true);
    const float q0 = __builtin_amdgcn_cvt_f32_fp8(w, 0) * inv, q1 = __builtin_amdgcn_cvt_f32_fp8(w, 1) * inv;
    const float q2 = __builtin_amdgcn_cvt_f32_fp8(w, 2) * inv, q3 = __builtin_amdgcn_cvt_f32_fp8(w, 3) * inv;
    float nrm = (q0 * q0 + q1 * q1) + (q2 * q2 + q3 * q3);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nrm += __shfl_xor(nrm, o);
    if (4 * lane < dpad8) Xb8[(size_t)row * (dpad8 / 4) + lane] = (unsigned)w;
    if (lane == 0) {
        a[row] = (row < n) ? neg_gl2e * nrm : 0.f;
        sinv[row] = (row < n) ? inv : 0.f;
    }
}


template <bool F16>
__global__ __launch_bounds__(256) void pack_omega_kernel(int n, int l, int nchunks, int ncb_total,
                                                         const double* __restrict__ O, int ldo,
                                                         const double* __restrict__ y,
                                                         bf16x8* __restrict__ Op) {
    pack_one<F16>((long)blockIdx.x * 256 + threadIdx.x, n, l, nchunks, ncb_total, O, ldo, y, Op);
}

int ncb_instance_bf16(int want) {
    static const int inst[] = {1, 2, 3, 5, 9, 17};
    for (int v : inst)
        if (v >= want) return v;
    return 17;
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
    if (d <= 256)
        colmean_final_kernel<<<1, 1024, 0, h->stream>>>(n, d, nb, part, mean);
    else
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
// Small grids split the j range over more workgroups: their partial products
// go to a workspace and are summed in order into Y.
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
    const int splits =
        dh_j_splits(h->num_cu, (long)dh_cdiv(m, dp > 128 ? 128 : 256) * cols.groups, nchunks);
    double* P = Y;
    int ldp = ldy;
    long ystride = 0;
    if (splits > 1) {
        ystride = (long)m * l;
        P = (double*)dh_ws(h, DH_WS_SPLITN, sizeof(double) * (size_t)ystride * splits);
        if (!P) return DH_ERR_ALLOC;
        ldp = m;
    }
    int st;
    switch (dp) {
        case 32:
            st = dh_project16_dk1(F16, h, cols.ncb, grid, m, nchunks, l, cols.ncb_pack, xi, xj,
                                  pi.a, pj.a, c2, Op, yi, P, ldp, splits, ystride);
            break;
        case 64:
            st = dh_project16_dk2(F16, h, cols.ncb, grid, m, nchunks, l, cols.ncb_pack, xi, xj,
                                  pi.a, pj.a, c2, Op, yi, P, ldp, splits, ystride);
            break;
        case 128:
            st = dh_project16_dk4(F16, h, cols.ncb, grid, m, nchunks, l, cols.ncb_pack, xi, xj,
                                  pi.a, pj.a, c2, Op, yi, P, ldp, splits, ystride);
            break;
        default:
            st = dh_project16_dk8(F16, h, cols.ncb, grid, m, nchunks, l, cols.ncb_pack, xi, xj,
                                  pi.a, pj.a, c2, Op, yi, P, ldp, splits, ystride);
            break;
    }
    if (st || splits == 1) return st;
    sum_splits_kernel<<<dh_cdiv(ystride, 256), 256, 0, h->stream>>>(m, l, splits, P, ystride, Y,
                                                                    ldy);
    DH_CHECK_LAUNCH();
    return 0;
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
        int blocks = max(1, min(dh_cdiv((long)n * d, 256), h->num_cu));
        maxabs_centred_kernel<<<blocks, 256, 0, h->stream>>>(n, d, Xj, ldxj, mean, word);
        DH_CHECK_LAUNCH();
        if (!same) {
            blocks = max(1, min(dh_cdiv((long)m * d, 256), h->num_cu));
            maxabs_centred_kernel<<<blocks, 256, 0, h->stream>>>(m, d, Xi, ldxi, mean, word);
            DH_CHECK_LAUNCH();
        }
        unsigned hbits = 0;
        unsigned* hb = (unsigned*)dh_host_status(h);  // pinned when available
        if (!hb) hb = &hbits;
        DH_CHECK_HIP(hipMemcpyAsync(hb, word, sizeof(unsigned), hipMemcpyDeviceToHost, h->stream));
        DH_CHECK_HIP(hipStreamSynchronize(h->stream));
        const unsigned bits = *hb;
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

namespace {

// FP8 copy of one point set inside a workspace (one byte per feature).
struct Fp8Points {
    unsigned* xb = nullptr;  // nrows_pad x dpad8 bytes
    float* a = nullptr;      // nrows_pad
    float* sinv = nullptr;   // nrows_pad
    int nrows_pad = 0;
};

int dpad8_of(int d) { return d <= 128 ? 128 : 256; }

// Carve and fill the FP8 copies of one or two point sets (centred by the
// mean of the column set Xj, rows scaled separately).
int prepare_points_fp8(dh_handle_t h, int m, int n, int d, const double* Xi, int ldxi,
                       const double* Xj, int ldxj, double gamma, bool same, Fp8Points& pi,
                       Fp8Points& pj) {
    const int dp = dpad8_of(d);
    const int mpad = dh_cdiv(m, 32) * 32, npad = dh_cdiv(n, 32) * 32;
    auto bytes_of = [&](int rows) { return ((size_t)rows * dp + 8 * (size_t)rows + 511) & ~size_t(255); };
    const size_t bj = bytes_of(npad), bi = same ? 0 : bytes_of(mpad);
    char* base = (char*)dh_ws(h, DH_WS_LOWP_X, bj + bi + 512);
    if (!base) return DH_ERR_ALLOC;
    double* mean = (double*)dh_ws(h, DH_WS_NORMS, sizeof(double) * (size_t)d + 64);
    if (!mean) return DH_ERR_ALLOC;
    auto carve = [&](char* p, int rows, Fp8Points& out) {
        out.xb = (unsigned*)p;
        out.a = (float*)(p + (size_t)rows * dp);
        out.sinv = out.a + rows;
        out.nrows_pad = rows;
    };
    carve(base, npad, pj);
    if (same)
        pi = pj;
    else
        carve(base + bj, mpad, pi);
    DH_CHECK_STATUS(point_mean(h, n, d, Xj, ldxj, mean));
    const float neg = (float)(-gamma) * kLog2e;
    convert_rows_fp8_kernel<<<dh_cdiv(npad, 4), 256, 0, h->stream>>>(n, npad, d, dp, Xj, ldxj,
                                                                     mean, neg, pj.xb, pj.a,
                                                                     pj.sinv);
    DH_CHECK_LAUNCH();
    if (!same) {
        convert_rows_fp8_kernel<<<dh_cdiv(mpad, 4), 256, 0, h->stream>>>(
            m, mpad, d, dp, Xi, ldxi, mean, neg, pi.xb, pi.a, pi.sinv);
        DH_CHECK_LAUNCH();
    }
    return 0;
}

int project_fp8(dh_handle_t h, int m, int n, int d, int l, const Fp8Points& pi,
                const Fp8Points& pj, float c2, const bf16x8* Op, const double* yi, double* Y,
                int ldy) {
    if (m == 0 || l == 0) return 0;
    Bf16Cols cols = bf16_cols(l);
    return dh_project_fp8(h, dpad8_of(d), cols.ncb, cols.groups, m, dh_cdiv(n, 32), l,
                          cols.ncb_pack, (const bf16x8*)pi.xb, (const bf16x8*)pj.xb, pi.a, pj.a,
                          pi.sinv, pj.sinv, c2, Op, yi, Y, ldy);
}

}  // namespace

extern "C" int dh_rbf_project_fp8(dh_handle_t h, int m, int n, int d, int l, const double* Xi,
                                  int ldxi, const double* Xj, int ldxj, const double* yi,
                                  const double* yj, double gamma, const double* O, int ldo,
                                  double* Y, int ldy) {
    const char* name = "dh_rbf_project_fp8";
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
    Fp8Points pi, pj;
    const bool same = (Xi == Xj && m == n && ldxi == ldxj);
    DH_CHECK_STATUS(prepare_points_fp8(h, m, n, d, Xi, ldxi, Xj, ldxj, gamma, same, pi, pj));
    bf16x8* Op = (bf16x8*)dh_ws(h, DH_WS_LOWP_O, packed_omega_bytes(n, l));
    if (!Op) return DH_ERR_ALLOC;
    DH_CHECK_STATUS(pack_omega<false>(h, n, l, O, ldo, yj, Op));
    return project_fp8(h, m, n, d, l, pi, pj, (float)(2.0 * gamma) * kLog2e, Op, yi, Y, ldy);
}

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
int dh_range_orth_begin(dh_handle_t h, int m, int n, double* Y, int ldy, int* pending);
int dh_range_orth_finish(dh_handle_t h, int m, int n, double* Y, int ldy, int* pending,
                         int* redone);

// Mixed-precision passes of dh_lra: 16-bit (BF16 or FP16) K-passes, FP64 CholeskyQR2.
// Each orthogonalisation's status is checked only after the next pass is queued
// (no idle GPU while the host waits); a breakdown, rare, redoes that pass.
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
    int pending = 0, redone = 0;
    DH_CHECK_STATUS(dh_range_orth_begin(h, n, l, Q, ldq, &pending));
    // Power iterations alternate the basis between Q and W (no copy for even q).
    double *cur = Q, *nxt = W;
    int ldc = ldq, ldn = ldw;
    for (int it = 0; it < q; ++it) {
        do {
            DH_CHECK_STATUS(pack_omega<F16>(h, n, l, cur, ldc, y, Op));
            DH_CHECK_STATUS(project_16<F16>(h, n, n, d, l, p, p, c2, Op, y, nxt, ldn));
            DH_CHECK_STATUS(dh_range_orth_finish(h, n, l, cur, ldc, &pending, &redone));
        } while (redone);
        DH_CHECK_STATUS(dh_range_orth_begin(h, n, l, nxt, ldn, &pending));
        std::swap(cur, nxt);
        std::swap(ldc, ldn);
    }
    // The copy is queued after the pending orthogonalisation, so Q holds its
    // result (or, after a breakdown, the state the fallback starts from) even
    // though the last pass then overwrites W.
    if (cur != Q)
        DH_CHECK_HIP(hipMemcpy2DAsync(Q, sizeof(double) * (size_t)ldq, cur,
                                      sizeof(double) * (size_t)ldc, sizeof(double) * (size_t)n, l,
                                      hipMemcpyDeviceToDevice, h->stream));
    do {
        DH_CHECK_STATUS(pack_omega<F16>(h, n, l, Q, ldq, y, Op));
        DH_CHECK_STATUS(project_16<F16>(h, n, n, d, l, p, p, c2, Op, y, W, ldw));
        DH_CHECK_STATUS(dh_range_orth_finish(h, n, l, Q, ldq, &pending, &redone));
    } while (redone);
    return 0;
}

int dh_lra_bf16_passes(dh_handle_t h, int n, int d, int l, int q, const double* X, int ldx,
                       const double* y, double gamma, const double* Omega, int ldo, double* Q,
                       int ldq, double* W, int ldw) {
    return lra_16_passes<false>(h, n, d, l, q, X, ldx, y, gamma, Omega, ldo, Q, ldq, W, ldw);
}

// FP8-distance passes: FP8 e4m3 distance, BF16 projection, FP64 CholeskyQR2.
int dh_lra_fp8_passes(dh_handle_t h, int n, int d, int l, int q, const double* X, int ldx,
                      const double* y, double gamma, const double* Omega, int ldo, double* Q,
                      int ldq, double* W, int ldw) {
    if (d > 256) return dh_fail(DH_ERR_UNSUPPORTED, "dh_lra: FP8 passes need d <= 256");
    Fp8Points p, p2;
    DH_CHECK_STATUS(prepare_points_fp8(h, n, n, d, X, ldx, X, ldx, gamma, true, p, p2));
    bf16x8* Op = (bf16x8*)dh_ws(h, DH_WS_LOWP_O, packed_omega_bytes(n, l));
    if (!Op) return DH_ERR_ALLOC;
    const float c2 = (float)(2.0 * gamma) * kLog2e;
    DH_CHECK_STATUS(pack_omega<false>(h, n, l, Omega, ldo, y, Op));
    DH_CHECK_STATUS(project_fp8(h, n, n, d, l, p, p, c2, Op, y, Q, ldq));
    DH_CHECK_STATUS(dh_range_orth(h, n, l, Q, ldq));
    // Power iterations alternate the basis between Q and W (no copy for even q).
    double *cur = Q, *nxt = W;
    int ldc = ldq, ldn = ldw;
    for (int it = 0; it < q; ++it) {
        DH_CHECK_STATUS(pack_omega<false>(h, n, l, cur, ldc, y, Op));
        DH_CHECK_STATUS(project_fp8(h, n, n, d, l, p, p, c2, Op, y, nxt, ldn));
        DH_CHECK_STATUS(dh_range_orth(h, n, l, nxt, ldn));
        std::swap(cur, nxt);
        std::swap(ldc, ldn);
    }
    if (cur != Q)
        DH_CHECK_HIP(hipMemcpy2DAsync(Q, sizeof(double) * (size_t)ldq, cur,
                                      sizeof(double) * (size_t)ldc, sizeof(double) * (size_t)n, l,
                                      hipMemcpyDeviceToDevice, h->stream));
    DH_CHECK_STATUS(pack_omega<false>(h, n, l, Q, ldq, y, Op));
    return project_fp8(h, n, n, d, l, p, p, c2, Op, y, W, ldw);
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

template <bool F16>
__global__ __launch_bounds__(256) void convert_rows_batched_kernel(const DhBf16Set* __restrict__ sets) {
    const DhBf16Set& s = sets[blockIdx.y];
    int row = (blockIdx.x * 256 + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (row >= s.npad) return;
    convert_row<F16>(row, lane, s.n, s.d, s.dpad, s.X, s.ldx, s.mean, F16 ? s.scale : 1.0,
                     (float)s.neg_gl2e, (__bf16*)s.xb, s.a);
}

template <bool F16>
__global__ __launch_bounds__(256) void pack_batched_kernel(const DhPackJob* __restrict__ jobs) {
    const DhPackJob& j = jobs[blockIdx.y];
    pack_one<F16>((long)blockIdx.x * 256 + threadIdx.x, j.n, j.src ? j.l : 0, j.nchunks,
                  j.ncb_pack, j.src, j.lds, j.y, (bf16x8*)j.Op);
}

// Per set: *maxabs = 0 (before maxabs_batched_kernel).
__global__ void clear_maxabs_batched_kernel(DhBf16Set* __restrict__ sets, int count) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < count) *sets[i].maxabs = 0u;
}

// max |X - mean| of set blockIdx.y (grid-strided over its elements; one
// atomic per workgroup).
__global__ __launch_bounds__(256) void maxabs_batched_kernel(const DhBf16Set* __restrict__ sets) {
    const DhBf16Set& s = sets[blockIdx.y];
    float mx = 0.f;
    const long total = (long)s.n * s.d;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const int i = (int)(e / s.d), f = (int)(e % s.d);
        mx = fmaxf(mx, (float)fabs(s.X[f + (size_t)i * s.ldx] - s.mean[f]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    __shared__ float wmax[4];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(s.maxabs, __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]))));
}

// Per set: scale = 2^-e with max|x - mean| = f 2^e, f in [1/2, 1); the gamma of
// the scaled points gamma / scale^2 goes into the set (-gamma log2 e) and into
// the c2 = 2 gamma log2 e of both projection jobs of the problem.
__global__ void f16_scale_batched_kernel(DhBf16Set* __restrict__ sets, DhProjJob* __restrict__ proj,
                                         int count) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= count) return;
    DhBf16Set& s = sets[i];
    const float mx = __uint_as_float(*s.maxabs);
    double scale = 1.0;
    if (mx > 0.f && isfinite(mx)) {
        int e = 0;
        (void)frexp((double)mx, &e);
        scale = ldexp(1.0, -e);
    }
    const double g = s.gamma / (scale * scale);
    s.scale = scale;
    s.neg_gl2e = -g * (double)kLog2e;
    const float c2 = (float)(2.0 * g) * kLog2e;
    proj[i].c2 = c2;
    proj[count + i].c2 = c2;
}

}  // namespace

int dh_bf16_dpad(int d) { return dpad_of(d); }
int dh_bf16_rows(int dpad) { return dpad > 128 ? 128 : 256; }

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
                            int max_npad, bool f16, DhProjJob* proj) {
    if (count == 0) return 0;
    colsum_partial_batched_kernel<<<dim3(dh_cdiv(max(max_n, 1), 256), count), 256, 0, h->stream>>>(sets);
    DH_CHECK_LAUNCH();
    colmean_final_batched_kernel<<<dim3(dh_cdiv(max_d, 256), count), 256, 0, h->stream>>>(sets);
    DH_CHECK_LAUNCH();
    if (f16) {
        DhBf16Set* msets = const_cast<DhBf16Set*>(sets);
        clear_maxabs_batched_kernel<<<dh_cdiv(count, 64), 64, 0, h->stream>>>(msets, count);
        DH_CHECK_LAUNCH();
        const int blocks = max(1, min(dh_cdiv((long)max_n * max_d, 256), 64));
        maxabs_batched_kernel<<<dim3(blocks, count), 256, 0, h->stream>>>(sets);
        DH_CHECK_LAUNCH();
        f16_scale_batched_kernel<<<dh_cdiv(count, 64), 64, 0, h->stream>>>(msets, proj, count);
        DH_CHECK_LAUNCH();
        convert_rows_batched_kernel<true>
            <<<dim3(dh_cdiv(max(max_npad, 1), 4), count), 256, 0, h->stream>>>(sets);
    } else {
        convert_rows_batched_kernel<false>
            <<<dim3(dh_cdiv(max(max_npad, 1), 4), count), 256, 0, h->stream>>>(sets);
    }
    DH_CHECK_LAUNCH();
    return 0;
}

int dh_bf16_pack_batched(dh_handle_t h, const DhPackJob* jobs, int count, long max_threads,
                         bool f16) {
    if (count == 0 || max_threads == 0) return 0;
    if (f16)
        pack_batched_kernel<true><<<dim3(dh_cdiv(max_threads, 256), count), 256, 0, h->stream>>>(jobs);
    else
        pack_batched_kernel<false><<<dim3(dh_cdiv(max_threads, 256), count), 256, 0, h->stream>>>(jobs);
    DH_CHECK_LAUNCH();
    return 0;
}

