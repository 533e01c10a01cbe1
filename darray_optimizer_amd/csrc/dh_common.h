// Internal declarations shared by the HIP sources of libdarray_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/darray_hip.h"

// Workspace slots owned by a handle.  Each slot is one device buffer that is
// grown on demand and reused by later calls; two calls that are live at the
// same time never share a slot.
enum dh_slot {
    DH_WS_SPLITK = 0,  // split-K partial products
    DH_WS_SPLITK_AUX,  // split-K partial products of GEMMs on the look-ahead stream
    DH_WS_NORMS,       // squared norms of point sets
    DH_WS_GRAM,        // CholeskyQR Gram matrix / R factors
    DH_WS_RBUF,        // second R factor
    DH_WS_YBUF,        // n x l panel buffer of the range finder
    DH_WS_LOWP_X,      // low-precision copy of X
    DH_WS_LOWP_O,      // low-precision packed Omega / Q
    DH_WS_LOWP_Y,      // FP32 output of a low-precision pass
    DH_WS_PARTIAL,     // per-workgroup partial sums
    DH_WS_INFO,        // device status words
    DH_WS_SOLVE_A,     // low-precision copy of a matrix being factored
    DH_WS_SOLVE_R,     // residual / correction panels of iterative refinement
    DH_WS_TRSM_INV,    // inverses of the 64 x 64 diagonal blocks of a triangular solve
    DH_WS_TRSM_X,      // solution blocks of the block-sweep triangular solve
    DH_WS_EIG,         // Jacobi SVD / Nystrom work matrices
    DH_WS_SVM,         // vectors and partial sums of the barrier SVM solver
    DH_WS_KBLOCK,      // materialised K row block of the d > 256 FP64 pass
    DH_WS_SPLITN,      // partial products of a K-pass whose j range is split
    DH_WS_PACK64,      // packed FP64 Omega / Q of the fused FP64 pass
    DH_WS_GUARD,       // combined status word that stops a GEMM after a breakdown
    DH_WS_CHOL2,       // scales and block statuses of the two-block small Cholesky
    DH_WS_COUNT
};

struct dh_handle_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    void* ws[DH_WS_COUNT] = {};
    size_t ws_size[DH_WS_COUNT] = {};
    int num_cu = 0;
    void* batch = nullptr;  // cached batched range-finder plan (batch.hip)
    void* tune = nullptr;   // autotuner decisions (batch.hip)
    void* timing = nullptr; // kernel timing events when enabled (handle.hip)
    // Second stream (default priority) and two events for the look-ahead of
    // the blocked Cholesky; created on first use, released by dh_destroy.
    hipStream_t aux = nullptr;
    hipEvent_t ev_panel = nullptr, ev_update = nullptr;
    hipEvent_t ev_orth = nullptr;  // end of a queued range-finder orthogonalisation
    // Pinned host words for device status reads (created on first use).
    int* host_status = nullptr;
    // Resident workgroups per CU of the persistent kernels, by kernel
    // (queried on the handle's device on first use; handle.hip).
    void* occupancy = nullptr;
};

// Creates h->aux and its events if they do not exist yet.
int dh_aux_stream(dh_handle_t h);

// Looks up, once per device, the device address of the static scratch array
// of dh_rbf_project_f64_noalloc (noalloc.hip); `device` must be current.
int dh_noalloc_setup(int device);

// *per_cu = resident workgroups per CU of `kernel` at `threads` threads per
// workgroup on the handle's device (at least 1), cached in the handle.
int dh_occupancy(dh_handle_t h, const void* kernel, int threads, int* per_cu);

// Pinned host buffer of 16 ints for status words copied back from the
// device (a copy into pageable memory takes a slower staged path); nullptr
// if it cannot be allocated.
int* dh_host_status(dh_handle_t h);

// Bracket one fused projection kernel launch with timing events (no-ops
// unless dh_timing_enable switched timing on for the handle).
void dh_timing_begin(dh_handle_t h);
void dh_timing_end(dh_handle_t h);

// Release the batched plan / tuning table of a handle (batch.hip).
void dh_batch_release(dh_handle_t h);
void dh_tune_release(dh_handle_t h);
// Autotuner table: choice for `key`, or -1 if not decided yet.
int dh_tune_lookup(dh_handle_t h, uint64_t key);
void dh_tune_store(dh_handle_t h, uint64_t key, int choice);
// False when DH_TUNE=0 in the environment (fixed defaults everywhere).
bool dh_tune_enabled();
// DH_SCHEDULE=<n> in the environment forces schedule n of the one-call 16-bit
// projection (for measurements); -1 when unset.
int dh_forced_schedule();
// Key of a BF16 projection schedule decision: tile class and row count
// (bucketed by powers of two: the choice depends on how full the chip is).
uint64_t dh_tune_key_bf16(int dk, int ncb, int waves, int rows);

// Records a formatted message for dh_last_error_string and returns `code`.
int dh_fail(int code, const char* fmt, ...);

// Device buffer of at least `bytes` in `slot`; nullptr on failure (message set).
void* dh_ws(dh_handle_t h, dh_slot slot, size_t bytes);

#define DH_CHECK_HIP(expr)                                                            \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return dh_fail(DH_ERR_HIP, "%s:%d: %s failed: %s", __FILE__, __LINE__, #expr, \
                           hipGetErrorString(e_));                                    \
    } while (0)

#define DH_CHECK_LAUNCH() DH_CHECK_HIP(hipGetLastError())

#define DH_CHECK_STATUS(expr)  \
    do {                       \
        int s_ = (expr);       \
        if (s_ != 0) return s_; \
    } while (0)

static inline bool dh_is_t(char c) { return c == 'T' || c == 't' || c == 'C' || c == 'c'; }
static inline bool dh_is_n(char c) { return c == 'N' || c == 'n'; }
static inline int dh_cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// Number of parts the j range (nchunks chunks) of a fused K-pass is split
// into: while the grid of `workgroups` has at most half a workgroup per CU,
// double the parts (up to 8, each at least 16 chunks).  Decided from the
// shape alone, so a shape always sums its partial products the same way.
// Returns the number of non-empty parts of dh_cdiv(nchunks, parts) chunks.
static inline int dh_j_splits(int num_cu, long workgroups, int nchunks) {
    int s = 1;
    while (s < 8 && workgroups * s * 2 <= num_cu && nchunks >= 32 * s) s *= 2;
    return dh_cdiv(nchunks, dh_cdiv(nchunks, s));
}

// Dense building blocks used across translation units (dense.hip).
template <typename T>
int dh_gemm_impl(dh_handle_t h, bool ta, bool tb, int m, int n, int k, T alpha, const T* A,
                 int lda, const T* B, int ldb, T beta, T* C, int ldc, int uplo_upper);
// As dh_gemm_impl (no triangle mask), doing nothing when *stop != 0 on the device.
template <typename T>
int dh_gemm_stop_impl(dh_handle_t h, bool ta, bool tb, int m, int n, int k, T alpha, const T* A,
                      int lda, const T* B, int ldb, T beta, T* C, int ldc, const int* stop);
template <typename T>
int dh_trsm_impl(dh_handle_t h, char side, char trans, bool unit, int m, int n, T alpha,
                 const T* A, int lda, T* B, int ldb);
template <typename T>
int dh_potrf_impl(dh_handle_t h, int n, T* A, int lda, int bs, int* info_host);
template <typename T>
int dh_symm_upper_impl(dh_handle_t h, int m, int n, T alpha, const T* A, int lda, const T* B,
                       int ldb, T beta, T* C, int ldc);

// One problem of a batched CholeskyQR step (small_spd.hip): Y (m x n, ld ldy)
// is orthogonalised with the Cholesky factor of G (n x n, ld ldg, holding
// Y^T Y on entry); Dinv receives 256 * ceil(n/16) doubles; *info the status
// of this step, and *sticky is set to 1 (never cleared) when it is non-zero.
struct DhSpdJob {
    int m, n, ldy, ldg;
    double* Y;
    double* G;
    double* Dinv;
    int* info;
    int* sticky;
};
int dh_cholqr_step_batched(dh_handle_t h, const DhSpdJob* jobs, int count, int max_n, int max_m,
                           bool shift);

// Register-blocked CholeskyQR Gram of one problem of a batch (order n <= 272):
// G = upper(Y^T Y) from nslab slabs of `rows` rows each, slab partials in part
// (dh_gram_reg_slab_doubles() doubles per slab), summed in slab order.
struct DhGramRegJob {
    int m, n, ldy, rows, nslab;
    const double* Y;
    double* part;
    double* G;
    int ldg;
};
struct DhGramItem {
    int job, slab;
};
size_t dh_gram_reg_slab_doubles();
// One launch over every (problem, slab) item, then one fixed-order reduction.
int dh_gram_reg_batched(dh_handle_t h, const DhGramRegJob* jobs, int count,
                        const DhGramItem* items, int nitems);
