// Projection kernels for points padded to 32 features (one translation unit
// per width, so the build compiles the template instances in parallel).
#include "rbf16_kernels.h"

int dh_project16_dk1(bool f16, dh_handle_t h, int ncb, int groups, int m, int nchunks, int l, int ncb_total,
                      const bf16x8* Xi, const bf16x8* Xj, const float* ai, const float* bj, float c2,
                      const bf16x8* Op, const double* yi, double* Y, int ldy, int splits,
                      long ystride) {
    return f16 ? dispatch_bf16_ncb<1, true>(h, ncb, groups, m, nchunks, l, ncb_total, Xi, Xj, ai,
                                             bj, c2, Op, yi, Y, ldy, splits, ystride)
               : dispatch_bf16_ncb<1, false>(h, ncb, groups, m, nchunks, l, ncb_total, Xi, Xj, ai,
                                              bj, c2, Op, yi, Y, ldy, splits, ystride);
}
