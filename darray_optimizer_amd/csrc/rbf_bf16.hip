// Mixed-precision K-passes (BF16 MFMA); placeholder until the kernel lands.
#include "dh_common.h"

int dh_lra_bf16_passes(dh_handle_t h, int n, int d, int l, int q, const double* X, int ldx,
                       const double* y, double gamma, const double* Omega, int ldo, double* Q,
                       int ldq, double* W, int ldw) {
    return dh_fail(DH_ERR_UNSUPPORTED, "dh_lra: BF16 passes not built");
}
