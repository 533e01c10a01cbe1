// Fused projection kernels with the FP8 e4m3 distance (points padded to 128
// or 256 one-byte features) and the BF16 projection.
#include "rbf16_kernels.h"

namespace {

// An even number of 17-block column groups (l > 272) at 256 features runs as
// pairs: waves w and w + 4 share 32 rows and take two adjacent groups, each
// computing one half of every chunk's distances and K (no group recomputes
// another's).
template <int DK, int NCB>
void launch_fp8(dh_handle_t h, int m, int groups, int nchunks, int l, int ncb_total,
                const bf16x8* Xi, const bf16x8* Xj, const float* ai, const float* bj,
                const float* si, const float* sj, float c2, const bf16x8* Op, const double* yi, double* Y,
                int ldy) {
    if constexpr (NCB == 17 && DK == 4) {  // 256 features (pairs measured slower at 128)
        if (groups % 2 == 0) {
            rbf_project_bf16_kernel<DK, NCB, 8, false, 2, true, true>
                <<<dim3(dh_cdiv(m, 4 * 32), groups / 2), 512, 0, h->stream>>>(
                    m, nchunks, l, ncb_total, Xi, Xj, ai, bj, si, sj, c2, Op, yi, Y, ldy, nchunks,
                    0);
            return;
        }
    }
    rbf_project_bf16_kernel<DK, NCB, 8, false, 2, true>
        <<<dim3(dh_cdiv(m, 8 * 32), groups), 512, 0, h->stream>>>(
            m, nchunks, l, ncb_total, Xi, Xj, ai, bj, si, sj, c2, Op, yi, Y, ldy, nchunks, 0);
}

template <int DK>
int dispatch_fp8(dh_handle_t h, int ncb, int groups, int m, int nchunks, int l, int ncb_total,
                 const bf16x8* Xi, const bf16x8* Xj, const float* ai, const float* bj,
                 const float* si, const float* sj, float c2, const bf16x8* Op, const double* yi, double* Y,
                 int ldy) {
#define DH_FP8_NCB(N)                                                                       \
    case N:                                                                                 \
        launch_fp8<DK, N>(h, m, groups, nchunks, l, ncb_total, Xi, Xj, ai, bj, si, sj, c2, Op, \
                          yi, Y, ldy);                                                      \
        break;
    switch (ncb) {
        DH_FP8_NCB(1)
        DH_FP8_NCB(2)
        DH_FP8_NCB(3)
        DH_FP8_NCB(5)
        DH_FP8_NCB(9)
        DH_FP8_NCB(17)
        default:
            return dh_fail(DH_ERR_UNSUPPORTED, "rbf_project_fp8: no kernel for %d column blocks",
                           ncb);
    }
#undef DH_FP8_NCB
    return 0;
}

}  // namespace

int dh_project_fp8(dh_handle_t h, int dpad8, int ncb, int groups, int m, int nchunks, int l,
                   int ncb_total, const bf16x8* Xi, const bf16x8* Xj, const float* ai,
                   const float* bj, const float* si, const float* sj, float c2, const bf16x8* Op,
                   const double* yi, double* Y, int ldy) {
    dh_timing_begin(h);
    int st = dpad8 == 128 ? dispatch_fp8<2>(h, ncb, groups, m, nchunks, l, ncb_total, Xi, Xj, ai,
                                            bj, si, sj, c2, Op, yi, Y, ldy)
                          : dispatch_fp8<4>(h, ncb, groups, m, nchunks, l, ncb_total, Xi, Xj, ai,
                                            bj, si, sj, c2, Op, yi, Y, ldy);
    dh_timing_end(h);
    if (st) return st;
    DH_CHECK_LAUNCH();
    return 0;
}
