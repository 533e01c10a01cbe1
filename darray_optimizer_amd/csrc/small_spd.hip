// Host launchers of the CholeskyQR step kernels (small_spd.h).
#include "small_spd.h"

int dh_chol_small_max_order() { return kMaxN; }

// G = upper(Y^T Y) for the CholeskyQR Gram when 128 < n <= 272 (returns 1,
// nothing launched, for other orders: the caller uses the tiled GEMM).
int dh_gram_small_launch(dh_handle_t h, int m, int n, const double* Y, int ldy, double* G,
                         int ldg) {
    if (n <= 128 || n > 16 * kRegBlocks || m < 16) return 1;
    // One slab per CU: eight waves at 256 registers fill a CU's register file.
    int nz = min(h->num_cu, dh_cdiv(m, 16));
    const int rows = 16 * dh_cdiv(dh_cdiv(m, nz), 16);
    nz = dh_cdiv(m, rows);
    double* part = (double*)dh_ws(h, DH_WS_SPLITK, sizeof(double) * 256 * kGramBlocks * (size_t)nz);
    if (!part) return DH_ERR_ALLOC;
    gram_reg_kernel<<<nz, NT, 0, h->stream>>>(m, n, Y, ldy, rows, part);
    DH_CHECK_LAUNCH();
    gram_reduce_kernel<<<dh_cdiv(kGramBlocks * 256, 64), 256, 0, h->stream>>>(n, nz, part, G, ldg);
    DH_CHECK_LAUNCH();
    return 0;
}

size_t dh_gram_reg_slab_doubles() { return (size_t)kGramBlocks * 256; }

int dh_gram_reg_batched(dh_handle_t h, const DhGramRegJob* jobs, int count,
                        const DhGramItem* items, int nitems) {
    if (count == 0 || nitems == 0) return 0;
    gram_reg_batched_kernel<<<nitems, NT, 0, h->stream>>>(jobs, items);
    DH_CHECK_LAUNCH();
    gram_reduce_batched_kernel<<<dim3(dh_cdiv(kGramBlocks * 256, 64), count), 256, 0, h->stream>>>(
        jobs);
    DH_CHECK_LAUNCH();
    return 0;
}

// One-workgroup Cholesky of G (+ 16 x 16 diagonal-block inverses into Dinv,
// which needs 256 * ceil(n/16) doubles); *info_dev receives the status.
int dh_chol_small_launch(dh_handle_t h, int n, double* G, int ldg, double* Dinv, int m_rows,
                         bool shift, bool rel_check, int* info_dev) {
    if (n > kMaxN) return dh_fail(DH_ERR_UNSUPPORTED, "chol_small: order %d > %d", n, kMaxN);
    if (n <= 16 * kRegBlocks) {
        chol_reg_kernel<<<1, NT, 0, h->stream>>>(n, G, ldg, Dinv, m_rows, shift ? 1 : 0,
                                                 rel_check ? 1 : 0, info_dev, nullptr);
        DH_CHECK_LAUNCH();
        return 0;
    }
    // 272 < n <= 544 as two register Cholesky factorisations: with G = [A11 A12;
    // . A22] and A11 of order 272, U11 from A11, U12 = U11^-T A12, then U22 from
    // the Schur complement A22 - U12^T U12.  Both use the breakdown tolerance
    // and shift of the whole G; a breakdown in the first leaves the rest to
    // run on garbage, and the combined status reports it (the caller redoes
    // the factorisation then).
    const int n1 = 16 * kRegBlocks, n2 = n - n1;
    double* sc = (double*)dh_ws(h, DH_WS_CHOL2, 64);
    if (!sc) return DH_ERR_ALLOC;
    int* parts = (int*)(sc + 2);
    chol_scales_kernel<<<1, NT, 0, h->stream>>>(n, G, ldg, m_rows, shift ? 1 : 0,
                                                rel_check ? 1 : 0, sc);
    DH_CHECK_LAUNCH();
    chol_reg_kernel<<<1, NT, 0, h->stream>>>(n1, G, ldg, Dinv, m_rows, shift ? 1 : 0,
                                             rel_check ? 1 : 0, parts, sc);
    DH_CHECK_LAUNCH();
    double* A12 = G + (size_t)n1 * ldg;
    double* A22 = A12 + n1;
    DH_CHECK_STATUS(dh_trsm_impl<double>(h, 'L', 'T', false, n1, n2, 1.0, G, ldg, A12, ldg));
    DH_CHECK_STATUS(dh_gemm_impl<double>(h, true, false, n2, n2, n1, -1.0, A12, ldg, A12, ldg,
                                         1.0, A22, ldg, 1));
    chol_reg_kernel<<<1, NT, 0, h->stream>>>(n2, A22, ldg, Dinv + 256 * kRegBlocks, m_rows,
                                             shift ? 1 : 0, rel_check ? 1 : 0, parts + 1, sc);
    DH_CHECK_LAUNCH();
    chol_info_combine_kernel<<<1, 64, 0, h->stream>>>(parts, n1, info_dev);
    DH_CHECK_LAUNCH();
    return 0;
}

// Batched CholeskyQR step: for each job, G := U (Cholesky with the relative
// breakdown test, optional shift), then Y := Y U^{-1} unless it broke down.
int dh_cholqr_step_batched(dh_handle_t h, const DhSpdJob* jobs, int count, int max_n, int max_m,
                           bool shift) {
    if (max_n > kMaxN) return dh_fail(DH_ERR_UNSUPPORTED, "chol_small: order %d > %d", max_n, kMaxN);
    constexpr int W = 2;
    static bool lds_attr_set = false;  // set outside any graph capture: first call has count 0
    if (!lds_attr_set) {
        DH_CHECK_HIP(hipFuncSetAttribute((const void*)trsm_rows_batched_kernel<W>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)(sizeof(double) * 16 * (16 * 34 + 1) * W)));
        lds_attr_set = true;
    }
    if (count == 0) return 0;
    if (max_n <= 16 * kRegBlocks)
        chol_reg_batched_kernel<<<count, NT, 0, h->stream>>>(jobs, shift ? 1 : 0, 1);
    else
        chol_small_batched_kernel<<<count, NT, 0, h->stream>>>(jobs, shift ? 1 : 0, 1);
    DH_CHECK_LAUNCH();
    if (max_n <= 16 * kRegBlocks) {
        constexpr int WR = 4;
        trsm_rows_reg_batched_kernel<WR>
            <<<dim3(dh_cdiv(max_m, 16 * WR), count), WR * 64, 0, h->stream>>>(jobs);
        DH_CHECK_LAUNCH();
        return 0;
    }
    int pitch = 16 * ((max_n + 15) / 16) + 1;
    size_t lds = sizeof(double) * 16 * (size_t)pitch * W;
    dim3 grid(dh_cdiv(max_m, 16 * W), count);
    trsm_rows_batched_kernel<W><<<grid, W * 64, lds, h->stream>>>(jobs);
    DH_CHECK_LAUNCH();
    return 0;
}

// *w = guard[0] | guard[1] (nguard words): one stop word for a GEMM.
__global__ void combine_guard_kernel(const int* __restrict__ guard, int nguard, int* __restrict__ w) {
    if (threadIdx.x == 0) *w = guard[0] | (nguard > 1 ? guard[1] : 0);
}

// Y := Y U^{-1} with U from chol_small (upper, order n <= 544) and its Dinv.
// With guard != nullptr the kernel does nothing unless guard[0 .. nguard-1]
// (device status words of the Cholesky steps before it, nguard <= 2) are
// all zero.
int dh_trsm_rows_launch(dh_handle_t h, int m, int n, double* Y, int ldy, const double* U,
                        int ldu, const double* Dinv, const int* guard, int nguard) {
    if (n > kMaxN) return dh_fail(DH_ERR_UNSUPPORTED, "trsm_rows: order %d > %d", n, kMaxN);
    if (m == 0 || n == 0) return 0;
    if (n <= 16 * kRegBlocks) {
        constexpr int WR = 8;
        trsm_rows_reg_kernel<WR><<<dh_cdiv(m, 16 * WR), WR * 64, 0, h->stream>>>(m, n, Y, ldy, U,
                                                                               ldu, Dinv, guard, nguard);
        DH_CHECK_LAUNCH();
        return 0;
    }
    {
        // 272 < n <= 544: U = [U11 U12; 0 U22] with U11 of order 272.  X1 = Y1 U11^-1
        // and X2 = (Y2 - X1 U12) U22^-1 by the register row solve around one FP64
        // GEMM (the single-kernel row solve below keeps the whole panel in LDS and
        // ran at ~13 TFLOP/s at n = 262144, l = 544).  The GEMM is skipped after
        // a breakdown, as the solves are, through one combined status word.
        constexpr int WR = 8;
        const int n1 = 16 * kRegBlocks, n2 = n - n1;
        trsm_rows_reg_kernel<WR><<<dh_cdiv(m, 16 * WR), WR * 64, 0, h->stream>>>(
            m, n1, Y, ldy, U, ldu, Dinv, guard, nguard);
        DH_CHECK_LAUNCH();
        const int* stop = nullptr;
        if (guard) {
            int* w = (int*)dh_ws(h, DH_WS_GUARD, sizeof(int));
            if (!w) return DH_ERR_ALLOC;
            combine_guard_kernel<<<1, 64, 0, h->stream>>>(guard, nguard, w);
            DH_CHECK_LAUNCH();
            stop = w;
        }
        DH_CHECK_STATUS(dh_gemm_stop_impl<double>(h, false, false, m, n2, n1, -1.0, Y, ldy,
                                                  U + (size_t)n1 * ldu, ldu, 1.0,
                                                  Y + (size_t)n1 * ldy, ldy, stop));
        trsm_rows_reg_kernel<WR><<<dh_cdiv(m, 16 * WR), WR * 64, 0, h->stream>>>(
            m, n2, Y + (size_t)n1 * ldy, ldy, U + n1 + (size_t)n1 * ldu, ldu,
            Dinv + 256 * kRegBlocks, guard, nguard);
        DH_CHECK_LAUNCH();
        return 0;
    }
    constexpr int W = 2;
    int pitch = 16 * ((n + 15) / 16) + 1;
    size_t lds = sizeof(double) * 16 * (size_t)pitch * W;
    static bool lds_attr_set = false;  // allow the full 160 KB of dynamic LDS once
    if (!lds_attr_set) {
        DH_CHECK_HIP(hipFuncSetAttribute((const void*)trsm_rows_kernel<W>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)(sizeof(double) * 16 * (16 * 34 + 1) * W)));
        lds_attr_set = true;
    }
    trsm_rows_kernel<W><<<dh_cdiv(m, 16 * W), W * 64, lds, h->stream>>>(m, n, Y, ldy, U, ldu,
                                                                          Dinv, pitch, guard, nguard);
    DH_CHECK_LAUNCH();
    return 0;
}
