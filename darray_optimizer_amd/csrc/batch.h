// Device-side tables of the batched range finder (batch.hip) and the batched
// BF16 kernels that read them (rbf_bf16.hip).
#pragma once

#include "dh_common.h"

// One point set prepared for 16-bit passes.
struct DhBf16Set {
    int n, d, dpad, npad;      // points, features, padded features, rows padded to 32
    const double* X;           // d x n, ld ldx
    int ldx;
    double neg_gl2e;           // -gamma log2(e) of the stored points (set on the device for FP16)
    double* part;              // ceil(n/256) x d column partial sums
    double* mean;              // d
    void* xb;                  // npad x dpad 16-bit (centred, for FP16 also scaled, points)
    float* a;                  // npad: -gamma log2(e) |xb_i|^2
    double gamma;              // kernel parameter of the unscaled points
    double scale;              // FP16: 2^-e with max |x - mean| 2^-e in (1/2, 1] (device-set)
    unsigned* maxabs;          // FP16: max |x - mean| as float bits (device word)
};

// Packing of one projection operand: src (n x l, ld lds; may be NULL for
// none), labels y (may be NULL) -> Op, ncb_pack 16-column blocks wide.
struct DhPackJob {
    int n, l, nchunks, ncb_pack;
    const double* src;
    int lds;
    const double* y;
    void* Op;
};

// One fused projection Y (m x l) = K(X, X) * Op of a prepared set.
struct DhProjJob {
    int m, nchunks, l, ncb_pack;
    const void* xb;
    const float* a;
    float c2;
    const void* Op;
    const double* y;
    double* Y;
    int ldy;
};

// A work item of a persistent projection kernel: row block, column group of a job.
struct DhProjItem {
    int job, rblk, cg, pad;
};

// Tile class of the 16-bit projection kernel.
struct DhBf16Class {
    int dk;     // 16-byte feature groups / 4 (dpad / 32)
    int ncb;    // 16-column blocks per workgroup
    int rows;   // rows per work item: 256, or 128 for dk = 8
    int wide;   // 4 waves x 64 rows instead of 8 waves x 32 rows (dk < 8)
    int f16;    // FP16 operands instead of BF16
};

int dh_bf16_dpad(int d);
// Rows of one BF16 projection workgroup for dpad (256, or 128 for dpad 256).
int dh_bf16_rows(int dpad);
// Column blocking of an l-wide pass: groups of ncb blocks, Op packed ncb_pack wide.
void dh_bf16_columns(int l, int* ncb, int* groups, int* ncb_pack);
size_t dh_bf16_packed_bytes(int n, int ncb_pack);

// Batched launches over device tables (count entries each).  With f16 the
// preparation also finds each set's power-of-two scale on the device and
// writes the scaled gamma into the sets and into both projection jobs of
// problem i (proj[i] and proj[count + i]), so no host synchronisation is needed.
int dh_bf16_prepare_batched(dh_handle_t h, const DhBf16Set* sets, int count, int max_n, int max_d,
                            int max_npad, bool f16, DhProjJob* proj);
int dh_bf16_pack_batched(dh_handle_t h, const DhPackJob* jobs, int count, long max_threads,
                         bool f16);
// Persistent projection of class c over nitems items, work handed out by
// atomically incrementing *counter (zero on entry); grid sized to fill the chip.
int dh_bf16_project_persistent(dh_handle_t h, DhBf16Class c, const DhProjJob* jobs,
                               const DhProjItem* items, int nitems, int* counter);
