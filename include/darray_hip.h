/*
 * darray_hip.h -- C ABI of the MI355X (gfx950) RBF low-rank-approximation library.
 *
 * Plain C types only: pointers, sizes, scalars.  Every matrix is column-major
 * with an explicit leading dimension, as DArray's LMatrix / DMatrix local
 * block (darray.h:257-421, 424-453).  Point sets are d x n column-major, one
 * point per column (DArray's training matrix X, testsvm.cpp:76).
 *
 * Memory: unless a function says otherwise, matrix arguments are DEVICE
 * pointers and the call is asynchronous on the handle's stream.  The handle
 * owns its device workspaces (allocated once, grown on demand).  Functions
 * that return a LAPACK-style info synchronise the stream before returning.
 *
 * Status codes: 0 = success; > 0 = LAPACK-style info (order of the failing
 * minor of a Cholesky, CholeskyQR breakdown); -1 .. -99 = -(index of the bad
 * argument, 1-based, the handle counting as argument 1); <= -1000: see
 * DH_ERR_*.  dh_last_error_string() describes the last failure on the
 * calling thread.
 *
 * Threading: one handle per host thread; independent handles may be used
 * concurrently.
 */
#ifndef DARRAY_HIP_H
#define DARRAY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DH_ERR_HIP (-1000)         /* a HIP runtime call failed */
#define DH_ERR_ALLOC (-1001)       /* device allocation failed */
#define DH_ERR_UNSUPPORTED (-1002) /* shape outside what the kernels handle */
#define DH_ERR_NOT_CONVERGED (-1003) /* an iterative solve missed its tolerance */

typedef struct dh_handle_s* dh_handle_t;

/* ---------------------------------------------------------------- handle */

/* Create a handle on `device`.  `stream` is a hipStream_t to run on, or NULL
 * for a stream owned by the handle.  Replaces cuda_init (dkernel.cu:28-34),
 * which created global cuBLAS/cuSOLVER handles on one stream. */
int dh_create(int device, void* stream, dh_handle_t* out);
/* Destroy the handle and its workspaces (cuda_finalize, dkernel.cu:38-39). */
int dh_destroy(dh_handle_t h);
/* Block until all work queued on the handle's stream has finished. */
int dh_synchronize(dh_handle_t h);
/* Message for the last non-zero status returned on this thread. */
const char* dh_last_error_string(void);
/* Bytes of device workspace currently held by the handle. */
size_t dh_workspace_bytes(dh_handle_t h);

/* Kernel timing for measurements (no reference counterpart; DArray times
 * phases with a host millisecond clock, utils.h:436-456).  With on != 0 the
 * handle records HIP events on its stream around every fused projection
 * kernel it launches (dh_rbf_project_*, and inside dh_lra); dh_timing_read
 * synchronises, returns the summed kernel milliseconds and the number of
 * launches since the previous read, and resets them. */
int dh_timing_enable(dh_handle_t h, int on);
int dh_timing_read(dh_handle_t h, double* ms, int* launches);

/* ------------------------------------------------- device memory helpers */
/* For hosts that include no HIP header (the DArray C++ mirror in darray_hip/).
 * The reference allocates with cudaMallocManaged per call (dkernel.cu:135-147)
 * and copies element by element through host bounce buffers (dkernel.cu:80-90). */

/* Device allocation of `bytes` on the handle's device (*ptr = NULL for 0 bytes). */
int dh_malloc(dh_handle_t h, size_t bytes, void** ptr);
/* Free a dh_malloc allocation after the handle's stream has drained. */
int dh_free(dh_handle_t h, void* ptr);

#define DH_COPY_H2D 0
#define DH_COPY_D2H 1
#define DH_COPY_D2D 2
/* Copy an m x n column-major block of `elem`-byte elements, src (ld lds) to
 * dst (ld ldd), direction DH_COPY_*.  Synchronous. */
int dh_copy_2d(dh_handle_t h, int dir, int m, int n, size_t elem, const void* src, int lds,
               void* dst, int ldd);
/* Device conversion B = A of an m x n block, FP32 <-> FP64. */
int dh_convert_f32_f64(dh_handle_t h, int m, int n, const float* A, int lda, double* B, int ldb);
int dh_convert_f64_f32(dh_handle_t h, int m, int n, const double* A, int lda, float* B, int ldb);

/* ------------------------------------------------------------- RBF kernel */

/* Squared norms nrm[i] = sum_f X[f + i*ldx]^2 (vecnorm, dkernel.cu:102-114). */
int dh_sqnorms_f64(dh_handle_t h, int n, int d, const double* X, int ldx, double* nrm);

/* Materialised RBF tile, K[i + j*ldk] = yi[i]*yj[j]*exp(-gamma*max(0, |xi|^2+|xj|^2-2 xi.xj)),
 * Xi d x m, Xj d x n (point per column); yi / yj may be NULL (all ones).
 * Device counterpart of rbf_kernel (dkernel.h:6, dkernel.cu:132-190): the
 * distance GEMM runs on MFMA and the exp in the GEMM epilogue, so the m x n
 * inner-product matrix never reaches memory. */
int dh_rbf_tile_f64(dh_handle_t h, int m, int n, int d, const double* Xi, int ldxi,
                    const double* Xj, int ldxj, const double* yi, const double* yj,
                    double gamma, double* K, int ldk);
int dh_rbf_tile_f32(dh_handle_t h, int m, int n, int d, const float* Xi, int ldxi,
                    const float* Xj, int ldxj, const float* yi, const float* yj, float gamma,
                    float* K, int ldk);

/* Fused Y[m x l] = K(Xi,Xj) * O[n x l] + beta*Y in FP64 (MFMA f64 for the distance
 * and the projection, FP64 exp).  K is never stored for d <= 256 (O, with yj folded
 * in, is first packed into a workspace of 8 * ceil(n/16) * 16 * 16 * ceil(l/16)
 * bytes in the kernel's operand order); above 256 features K is materialised in
 * row blocks of at most 2^27 entries (a 1 GiB workspace), as rbf_kernel does
 * (dkernel.cu:144-178), and multiplied into O by the FP64 GEMM.  Replaces the
 * rbf_kernel + replicate + tc_gemm sequence of DArray::LRA (dkernel.h:33-42). */
int dh_rbf_project_f64(dh_handle_t h, int m, int n, int d, int l, const double* Xi, int ldxi,
                       const double* Xj, int ldxj, const double* yi, const double* yj,
                       double gamma, const double* O, int ldo, double beta, double* Y, int ldy);

/* Allocation-free variant of dh_rbf_project_f64, same arguments and results.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  Its only scratch, the squared norms of Xi and Xj, lives
 * in a static device array of DH_NOALLOC_MAX_POINTS doubles (8 MiB, held by
 * every GPU the library runs on until the process ends; dh_create looks up
 * its address once per device), so m + n <= DH_NOALLOC_MAX_POINTS and
 * d <= 256 (no K block), else DH_ERR_UNSUPPORTED before any launch.  The call makes no hipMalloc,
 * hipFree or other allocation.  Calls that overlap on one GPU (any handle,
 * any stream) race on the array. */
#define DH_NOALLOC_MAX_POINTS (1 << 20)
int dh_rbf_project_f64_noalloc(dh_handle_t h, int m, int n, int d, int l, const double* Xi,
                               int ldxi, const double* Xj, int ldxj, const double* yi,
                               const double* yj, double gamma, const double* O, int ldo,
                               double beta, double* Y, int ldy);

/* As dh_rbf_project_bf16 with FP16 in place of BF16 (points also scaled by a
 * power of two so that max |x - mean| lies in (1/2, 1]; one synchronisation):
 * the same MFMA rate with 3 more mantissa bits for the points, K and O. */
int dh_rbf_project_f16(dh_handle_t h, int m, int n, int d, int l, const double* Xi, int ldxi,
                       const double* Xj, int ldxj, const double* yi, const double* yj,
                       double gamma, const double* O, int ldo, double* Y, int ldy);

/* As dh_rbf_project_bf16 with the distance in FP8 e4m3 (points centred by the
 * mean of Xj, each point scaled by a power of two so that its largest |x| lies
 * in [128, 256), one byte per feature; FP32 norms of the dequantised points;
 * 16x16x128 FP8 MFMA with FP32 accumulation) and the projection in BF16.
 * d <= 256. */
int dh_rbf_project_fp8(dh_handle_t h, int m, int n, int d, int l, const double* Xi, int ldxi,
                       const double* Xj, int ldxj, const double* yi, const double* yj,
                       double gamma, const double* O, int ldo, double* Y, int ldy);

/* Mixed-precision fused Y[m x l] = K(Xi,Xj) * O (Y overwritten): points centred by
 * the mean of Xj and rounded to BF16, FP32 squared norms of the rounded points,
 * BF16 MFMA distance with FP32 accumulation, K = exp2(..) on v_exp_f32 rounded
 * to BF16, O rounded to BF16 (labels folded in), BF16 MFMA projection with FP32
 * accumulation, Y written in FP64.  d <= 256.  FP64 inputs; conversions run
 * on the device.  Tensor-core counterpart of rbf_kernel + tc_gemm
 * (dkernel.cu:160-166, 209-213; FP16 GemmEx precedent test_cuda.cu:111-132). */
int dh_rbf_project_bf16(dh_handle_t h, int m, int n, int d, int l, const double* Xi, int ldxi,
                        const double* Xj, int ldxj, const double* yi, const double* yj,
                        double gamma, const double* O, int ldo, double* Y, int ldy);

/* ------------------------------------------------------------- dense BLAS */

/* C = alpha*op(A)*op(B) + beta*C, op in {'N','T'}: tc_gemm / tc_sgemm
 * (dkernel.h:7-8, dkernel.cu:194-275) and multiply_matrix (cblas_like.h:39-60).
 * Unlike tc_gemm, beta is honoured. */
int dh_gemm_f64(dh_handle_t h, char transa, char transb, int m, int n, int k, double alpha,
                const double* A, int lda, const double* B, int ldb, double beta, double* C,
                int ldc);
int dh_gemm_f32(dh_handle_t h, char transa, char transb, int m, int n, int k, float alpha,
                const float* A, int lda, const float* B, int ldb, float beta, float* C,
                int ldc);

/* Upper SYRK C = alpha*A^T A + beta*C on the upper triangle (incl. diagonal) of the
 * n x n C, A k x n; the strict lower triangle is not written
 * (symmetric_update -> ssyrk_("U","T"), cblas_like.h:61-69). */
int dh_syrk_ut_f64(dh_handle_t h, int n, int k, double alpha, const double* A, int lda,
                   double beta, double* C, int ldc);
int dh_syrk_ut_f32(dh_handle_t h, int n, int k, float alpha, const float* A, int lda,
                   float beta, float* C, int ldc);

/* Triangular solve with the upper triangle of A: side 'L' (B m x n, A m x m) or
 * 'R' (A n x n), trans 'N' or 'T', diag 'N' or 'U':
 *   L,U,T: B := alpha*U^{-T} B   (triangular_solve, blas_like.h:95-128)
 *   L,U,N: B := alpha*U^{-1} B   (no reference counterpart; blas_like.h:124-126)
 *   R,U,N: B := alpha*B U^{-1}   (no reference counterpart; CholeskyQR)
 * uplo must be 'U' (the reference implements only the upper case). */
int dh_trsm_f64(dh_handle_t h, char side, char uplo, char transa, char diag, int m, int n,
                double alpha, const double* A, int lda, double* B, int ldb);
int dh_trsm_f32(dh_handle_t h, char side, char uplo, char transa, char diag, int m, int n,
                float alpha, const float* A, int lda, float* B, int ldb);

/* Right-looking blocked Cholesky A = U^T U of the n x n upper triangle, block bs
 * (cholesky_factorize, lapack_like.h:9-56: potrf(A11) -> trsm(L,U,T) -> syrk).
 * The strict lower triangle is never written.  Returns 0 or info > 0, the
 * global 1-based order of the first non-positive pivot; A is then partially
 * factored (lapack_like.h:25-28).  Synchronises.  uplo must be 'U'.
 * For n > bs part of each trailing update runs on a second stream owned by
 * the handle (one-panel look-ahead); it is ordered after earlier work on the
 * handle's stream and joined back before the call returns. */
int dh_potrf_f64(dh_handle_t h, char uplo, int n, double* A, int lda, int bs);
int dh_potrf_f32(dh_handle_t h, char uplo, int n, float* A, int lda, int bs);

/* Solve A X = B (A n x n SPD, upper triangle used; B n x nrhs): A is overwritten
 * by U (dh_potrf_f64 with block bs), B by X through U^T Z = B (the L,U,T solve
 * of triangular_solve, blas_like.h:95-128) and U X = Z (L,U,N).  LAPACK posv
 * contract; returns 0 or the potrf info.  Synchronises. */
int dh_posv_f64(dh_handle_t h, char uplo, int n, int nrhs, double* A, int lda, double* B, int ldb,
                int bs);

/* Mixed-precision A X = B: FP32 Cholesky (block bs) of a rounded copy of A,
 * FP64 iterative refinement with residuals R = B - A X computed from the
 * upper triangle of A in FP64, until the normwise backward error
 * max_j ||r_j||_2 / (||A||_F ||x_j||_2 + ||b_j||_2) <= tol.  As LAPACK dsposv,
 * the call solves in FP64 instead (*iters = -1) when the FP32 factorisation
 * breaks down, or the refinement does not reach tol within max_iter
 * corrections, stalls (a correction that does not halve the error) or
 * produces a non-finite error.  A and B are not modified.  *iters =
 * corrections applied; *berr = backward error of the returned X.  Returns
 * DH_ERR_NOT_CONVERGED when even the FP64 solution misses tol and
 * 1e3 * n * eps64 (non-finite input).  Synchronises. */
int dh_posv_mixed(dh_handle_t h, char uplo, int n, int nrhs, const double* A, int lda,
                  const double* B, int ldb, double* X, int ldx, int bs, double tol, int max_iter,
                  int* iters, double* berr);

/* ------------------------------------------------------- range finder */

/* CholeskyQR2 in FP64: Y (m x n, m >= n) is overwritten by Q with Q^T Q = I, and
 * R (n x n upper, positive diagonal; may be NULL) receives Y = Q R.  A sweep
 * breaks down when a Cholesky pivot satisfies U_kk^2 <= n*eps*max_j G_jj
 * (Y rank deficient to working precision).  If the first sweep breaks down
 * the call switches to shifted CholeskyQR3 (Fukaya et al., SIAM J. Sci.
 * Comput. 2020).  Returns 0, or > 0 on breakdown: the 1-based pivot index,
 * plus 1000000 for the second sweep and 2000000 for the third.  Replaces
 * tall_skinny_qr_factorize (lapack_like.h:281-379) on the path.
 * Synchronises. */
int dh_cholqr2_f64(dh_handle_t h, int m, int n, double* Y, int ldy, double* R, int ldr);

/* Precision of the K-passes of dh_lra. */
#define DH_PREC_F64 0  /* FP64 MFMA distance + projection, FP64 exp */
#define DH_PREC_BF16 1 /* BF16 MFMA distance + projection, FP32 accumulate and exp */
#define DH_PREC_F16 2  /* as DH_PREC_BF16 with FP16 operands (dh_rbf_project_f16) */
#define DH_PREC_FP8 3  /* FP8 e4m3 distance, BF16 projection (dh_rbf_project_fp8) */

/* Randomised range finder on the RBF kernel of one point set (DArray::LRA,
 * dkernel.h:12-46, and the orthogonalisation of testsvm.cpp:102-108, fused):
 *   Y = K*Omega; Q = CholQR2(Y); q times { Y = K*Q; Q = CholQR2(Y) }; W = K*Q
 * so B = Q^T K = W^T.  X d x n, y labels (n, may be NULL), Omega n x l.
 * Outputs Q (n x l, FP64) and W (n x l, FP64).  `prec` selects the K-pass
 * arithmetic; CholeskyQR2 is FP64 in all.  If Y has rank < l (K of rank < l),
 * where CholeskyQR cannot orthonormalise it, the deficient columns are
 * replaced by pseudo-random ones, so Q always has l orthonormal columns
 * spanning range(Y) and more (as the reference's Householder QR would).
 * Above 256 features every `prec` runs the FP64 passes (the low-precision
 * kernels hold d <= 256).  Synchronises. */
int dh_lra(dh_handle_t h, int prec, int n, int d, int l, int q, const double* X, int ldx,
           const double* y, double gamma, const double* Omega, int ldo, double* Q, int ldq,
           double* W, int ldw);

/* One problem of dh_lra_batched: the arguments of dh_lra after `prec`. */
typedef struct dh_lra_problem {
    int n, d, l, q;
    const double* X;
    int ldx;
    const double* y;
    double gamma;
    const double* Omega;
    int ldo;
    double* Q;
    int ldq;
    double* W;
    int ldw;
} dh_lra_problem;

/* dh_lra for `count` independent problems of any n, d <= 256 and l <= 272
 * (one-vs-rest or per-gamma SVM problems; no reference counterpart: DArray
 * runs one LRA per process, testsvm.cpp:86-120).  With DH_PREC_BF16 or
 * DH_PREC_F16 (each problem's power-of-two scale found on the device) every
 * stage runs as one launch over all problems -- point preparation, operand
 * packing, CholeskyQR2 (batched Gram, one-workgroup Cholesky per problem,
 * row-parallel solve) -- and the K-passes as persistent kernels, one per
 * tile class, pulling (problem, row block) items from a queue longest
 * problem first; the whole sequence is captured in a hipGraph on the first
 * call and replayed while the problem table (sizes and pointers) stays the
 * same.  Tile schedules come from the autotuner.  A problem whose
 * CholeskyQR breaks down is redone alone by dh_lra (shifted CholeskyQR3,
 * completion of rank-deficient bases).
 * DH_PREC_F64 and DH_PREC_FP8 run dh_lra per problem.  status (host, may be NULL) receives
 * each problem's dh_lra status; returns 0 or the first non-zero status.
 * All problems must share q.  Synchronises. */
int dh_lra_batched(dh_handle_t h, int prec, int count, const dh_lra_problem* problems,
                   int* status);

/* FP64 error pass: err2[0] = ||K - Q W^T||_F^2, err2[1] = ||K||_F^2, where err2 is a
 * HOST pointer to 2 doubles.  K is regenerated on the fly from X.  This is the
 * quantity Knorm (dkernel.h:49-76) means to print.  Synchronises. */
int dh_lra_error(dh_handle_t h, int n, int d, int l, const double* X, int ldx, const double* y,
                 double gamma, const double* Q, int ldq, const double* W, int ldw,
                 double* err2);

/* ---------------------------------------------------- Nystrom factor */

/* One-sided (Hestenes) Jacobi SVD in FP64 of A (m x n, m >= n): column pairs
 * are rotated to orthogonality in round-robin order until a sweep rotates no
 * pair (|a_p.a_q| <= 1e-14 |a_p| |a_q| for all pairs).  On exit A = U diag(sva)
 * (columns not normalised), sva (n, DEVICE) the singular values in descending
 * order and V (n x n) the right singular vectors, columns in the same order;
 * *sweeps (host, may be NULL) the sweeps used.  Returns 0, or 1 if 60 sweeps
 * did not converge.  Replaces sgesvd_ in DArray::svd (lapack_like.h:388-413).
 * Synchronises. */
int dh_gesvj_f64(dh_handle_t h, int m, int n, double* A, int lda, double* sva, double* V,
                 int ldv, int* sweeps);

/* Nystrom factor of a range finder's output (Q n x l orthonormal, W = K Q as
 * dh_lra returns them): C = (Q^T W + W^T Q)/2 = Q^T K Q, C = V diag(lam) V^T
 * by dh_gesvj_f64 (lam_i = v_i.(C v_i)), and Z (n x k) = Q V_k
 * diag(sqrt(max(lam_k, 0))) for the k largest eigenvalues, so that
 * Z Z^T ~ K (its error through dh_lra_error with Q = W = Z).  lambda (k,
 * DEVICE, may be NULL) receives those eigenvalues in descending order.  The
 * reference forms C with tc_gemm and takes Z = Q diag(S) from sgesvd_'s
 * singular values alone (testsvm.cpp:126-172).  Returns 0, or 1 if the
 * Jacobi iteration did not converge.  Synchronises. */
int dh_nystrom_f64(dh_handle_t h, int n, int l, int k, const double* Q, int ldq, const double* W,
                   int ldw, double* Z, int ldz, double* lambda);

/* ------------------------------------------------- barrier SVM solver */

/* Preconditioned conjugate gradients for (Z Z^T + diag(D)) x = b, Z n x k
 * (ld ldz, k <= 2048), D, b, x (n; x holds the initial guess on entry and
 * the solution on exit), preconditioner diag(Z Z^T) + D, until
 * ||b - (Z Z^T + D) x||_2 <= tol or max_iter iterations: pre_conjgrad
 * (dbarrier.h:83-139), whose operator Z Z^T + D + n I is this one with D + n.
 * *iters and *rnorm (host, may be NULL) receive the iterations and the final
 * residual norm.  Synchronises. */
int dh_svm_pcg(dh_handle_t h, int n, int k, const double* Z, int ldz, const double* D,
               const double* b, double* x, double tol, int max_iter, int* iters, double* rnorm);

/* Barrier-method solver of the SVM dual on the Nystrom factor Z (n x k) of
 * diag(y) K diag(y), as testsvm.cpp:191-321 runs it: min a^T Z Z^T a / 2 -
 * sum(a) over 0 < a < C (no equality constraint), starting from the smaller
 * class at 0.9 C and the other scaled so that y^T a = 0; Newton steps on
 * g = f - sum(log a + log(C - a)) / t with the step from dh_svm_pcg's CG
 * ((Z Z^T + diag(D)) s = -grad, from s = 0, ||r|| <= cg_tol, max_cg
 * iterations), backtracking line search (0.1, 0.5), Newton stops at
 * decrement^2/2 < 1e-6, slow progress or max D / min D > 1e18, and
 * t *= 20 until 2n/t < 1e-5 |f|.  The reference solves its Newton system
 * with the right-hand side e_1 and an added n I (testsvm.cpp:234-235,
 * dbarrier.h:99); here it is the Newton system.  y (n, DEVICE, +-1) and a
 * (n, DEVICE, output).  stats (host, may be NULL) = barrier iterations,
 * Newton steps, CG iterations; *fval (host, may be NULL) = f(a).  Returns
 * -6 if a class is empty.  Synchronises. */
int dh_svm_barrier(dh_handle_t h, int n, int k, const double* Z, int ldz, const double* y,
                   double C, double cg_tol, int max_outer, int max_newton, int max_cg, double* a,
                   int* stats, double* fval);

/* ------------------------------------------------------ LIBSVM files */
/* Host functions (HOST pointers; no GPU use). */

/* Read n points of a LIBSVM data file (read_input_file, read_input.h:8-38):
 * line j is "<label> <index>:<value> ..." with 1-based feature indices; X
 * (d x n, one point per column, ld ldx) and y (n) are zeroed, then
 * X[index-1, j] = value and y[j] = +1 if the label is 1, else -1 (the
 * reference's rule).  Indices <= 0 are skipped as in the reference; an index
 * above d returns DH_ERR_UNSUPPORTED (the reference writes past the column)
 * and a token that is not index:value returns -1.  Returns the number of
 * points read (< n when the file has fewer lines). */
int dh_read_libsvm(const char* path, int n, int d, double* X, int ldx, double* y);

/* Write a LIBSVM "c_svc" / "rbf" model file for the dual solution a (n) of
 * the box-constrained SVM with bound C (writemodel, dbarrier.h:394-470):
 * support vectors are the points with a_i > 1e-6, the bias b is the mean of
 * y_j (1 - sum_i a_i (Z Z^T)_ij) over the first (up to) 50 free support
 * vectors j (1e-6 < a_j < C - 1e-6), Z (n x k, ld ldz) the Nystrom factor
 * with Z Z^T ~ diag(y) K diag(y); rho = -b, coefficients y_i a_i, points X
 * (d x n, ld ldx) written sparsely, labels "1 -1" with the class-+1 support
 * vectors first (the reference writes class sizes in "label" and uses the
 * diagonal (Z Z^T)_ii in the bias).  *bias (may be NULL) receives b.
 * Returns the number of support vectors. */
int dh_write_libsvm_model(const char* path, int n, int d, int k, const double* a, double C,
                          const double* Z, int ldz, const double* y, const double* X, int ldx,
                          double gamma, double* bias);

/* ---------------------------------------------------- host-side inputs */
/* Seeded host generators (HOST pointers; no GPU use). */

/* Raw 32-bit outputs of std::mt19937(seed). */
int dh_mt19937_raw(uint32_t seed, int count, uint32_t* out);
/* DMatrix::set_uniform with a fixed seed (darray.h:513-522): std::mt19937(seed) +
 * uniform_real_distribution<double>(a,b), filled row-outer / column-inner into the
 * column-major m x n A. */
int dh_fill_uniform(uint32_t seed, double a, double b, int m, int n, double* A, int lda);
/* DMatrix::set_normal_seed (darray.h:535-544): std::mt19937(seed) +
 * normal_distribution<double>(mean, stddev), same fill order. */
int dh_fill_normal(uint32_t seed, double mean, double stddev, int m, int n, double* A, int lda);
/* Gaussian mixture (d x n, point per column): std::mt19937_64(seed); centres
 * uniform in [-1,1]^d (centre-major), point -> centre assignment uniform, then
 * point = centre + sigma*N(0,I) (point-major).  labels (n, may be NULL) receive
 * +1 if the nearest centre has an even index, else -1; centres (d x ncentres,
 * may be NULL) receive the centres. */
int dh_fill_mixture(uint64_t seed, int d, int n, int ncentres, double sigma, double* X, int ldx,
                    double* labels, double* centres);

#ifdef __cplusplus
}
#endif
#endif
